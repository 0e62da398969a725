"""Kernel E with the merged AB2 history at 256^3 FP64.

The Fourier update needs V1 and N0 only through Y = (V1 + lap1) - K/2 N0
(exec/ns.cpp:171-186), so the merged form keeps Y in place of N0 and holds
the velocity as the x-inverse the next step reads.  Its parity with the C
oracle is in test_gpu_full_size.py (every schedule, start-up + 3 steps); here
the transitions between the two state forms: the getters recover V1 without
perturbing the run, switching kernels mid-run converts the state both ways,
and the layout switch, the snapshots and the scheme switch see (V1, N0).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
N, NU, DT, ROLE_FOURIER = 256, 1.0 / 800, 5e-4, 4


def _rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _variants(L, prec=64, n=N, plain_tag="grouped=true"):
    """(merged, plain): E schedules with and without the merged history (at
    256 FP64, plain: the same LDS-streamed kernel keeping V1 and N0)."""
    c = ctypes.c_int()
    L.sh_tune_variant_count(prec, n, ROLE_FOURIER, ctypes.byref(c))
    merged, plain = [], []
    buf = ctypes.create_string_buffer(8192)
    for v in range(c.value):
        L.sh_tune_reset()
        L.sh_tune_set_variant(prec, n, ROLE_FOURIER, v)
        L.sh_tune_report(buf, 8192)
        name = buf.value.decode()
        if "merged AB2 history" in name:
            merged.append(v)
        elif plain_tag in name and "no-emit" not in name:
            plain.append(v)
    L.sh_tune_reset()
    assert merged and plain
    return merged, plain


@pytest.fixture(scope="module")
def v0(hip):
    import torch
    V0 = torch.empty((3, N, N, N), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), N, N, N, 11, 4.0, 1.0, hip.stream())
    yield V0
    del V0
    torch.cuda.empty_cache()


def _solver(hip, V0, layout=1):
    ns = hip.NS((N,) * 3, NU, DT, 64, fused=True)
    hip.lib().sh_ns_set_layout(ns.h, layout)
    ns.set_initial(V0)
    ns.startup()
    return ns


def _use(L, v, prec=64, n=N):
    L.sh_tune_reset()
    L.sh_tune_set_variant(prec, n, ROLE_FOURIER, v)


def test_merged_getter_does_not_perturb(hip, v0):
    """velocity() mid-run recovers V1 from the x-inverse without touching the
    state: 3 + 3 steps with a getter between equal 6 steps bit for bit, and
    agree with the (V1, N0) kernel to rounding; both stay divergence-free."""
    L = hip.lib()
    merged, plain = _variants(L)
    try:
        for layout in (1, 0):
            _use(L, merged[0])
            a = _solver(hip, v0, layout)
            a.steps(6)
            va = a.velocity()
            da = ctypes.c_double()
            L.sh_ns_div_average(a.h, ctypes.byref(da))
            a.close()
            b = _solver(hip, v0, layout)
            b.steps(3)
            mid = b.velocity()
            b.steps(3)
            vb = b.velocity()
            b.close()
            assert bool((va == vb).all()), layout
            _use(L, plain[0])
            c = _solver(hip, v0, layout)
            c.steps(3)
            err_mid = _rel(mid, c.velocity())
            c.steps(3)
            vc = c.velocity()
            dc = ctypes.c_double()
            L.sh_ns_div_average(c.h, ctypes.byref(dc))
            c.close()
            err = _rel(va, vc)
            print("layout=%d merged vs (V1, N0): 3 steps %.2e, 6 steps %.2e, div %.3e %.3e" %
                  (layout, err_mid, err, da.value, dc.value), flush=True)
            assert err_mid < 1e-13 and err < 1e-13
            # |div V| is rounding noise in both forms (the projection is exact)
            assert da.value < 1e-12 and dc.value < 1e-12
    finally:
        L.sh_tune_reset()


def test_merged_switches_both_ways(hip, v0):
    """merged -> (V1, N0) -> merged kernels within one run (the state is
    converted at each switch) against the (V1, N0) kernel throughout."""
    L = hip.lib()
    merged, plain = _variants(L)
    try:
        _use(L, plain[0])
        ref = _solver(hip, v0)
        ref.steps(6)
        vr = ref.velocity()
        ref.close()
        ns = _solver(hip, v0)
        for v in (merged[0], plain[0], merged[-1]):
            _use(L, v)
            ns.steps(2)
        err = _rel(ns.velocity(), vr)
        ns.close()
        print("switched run vs (V1, N0) run: %.2e" % err, flush=True)
        assert err < 1e-13
    finally:
        L.sh_tune_reset()


def test_merged_layout_snapshot_and_scheme(hip, v0):
    """From the merged form: the box-filtered snapshot, leaving the half
    layout mid-run, and switching to exec/ns2.cpp's scheme (whose kernel E
    keeps V1 and N0; Y holds the old scheme's viscous factor) all match the
    same sequence run with the (V1, N0) kernel."""
    import torch
    L = hip.lib()
    merged, plain = _variants(L)
    nc = (N // 4) ** 3
    out = []
    try:
        for v in (merged[0], plain[0]):
            _use(L, v)
            ns = _solver(hip, v0, 1)
            ns.steps(3)
            snap = torch.empty((9, nc), dtype=torch.float32, device="cuda")
            assert L.sh_ns_snapshot(ns.h, 4, hip.ptr(snap)) == 0
            ns.steps(2)
            assert L.sh_ns_set_layout(ns.h, 0) == 0
            ns.steps(2)
            v_layout = ns.velocity()
            assert L.sh_ns_set_scheme(ns.h, 1) == 0
            ns.steps(2)
            out.append((snap, v_layout, ns.velocity()))
            ns.close()
        (sa, la, ba), (sb, lb, bb) = out
        e_snap, e_layout, e_scheme = _rel(sa, sb), _rel(la, lb), _rel(ba, bb)
        print("snapshot %.2e, after the layout switch %.2e, after the scheme switch %.2e" %
              (e_snap, e_layout, e_scheme), flush=True)
        assert e_snap < 1e-6 and e_layout < 1e-13 and e_scheme < 1e-13
        assert np.isfinite(e_snap)
    finally:
        L.sh_tune_reset()


def test_merged_fp32_register_kernel(hip, v0):
    """The register-loaded kernel E with the merged history in FP32 at 256^3:
    the getter leaves the run unperturbed, and the velocity agrees with the
    (V1, N0) kernel to FP32 rounding."""
    L = hip.lib()
    merged, plain = _variants(L, 32, N, "lines=")
    out = {}
    try:
        for tag, v in (("merged", merged[0]), ("plain", plain[0])):
            _use(L, v, 32)
            a = hip.NS((N,) * 3, NU, DT, 32, fused=True)
            a.set_initial(v0)
            a.startup()
            a.steps(3)
            mid = a.velocity()
            a.steps(3)
            out[tag] = (mid, a.velocity())
            f = ctypes.c_int()
            L.sh_ns_get_history_form(a.h, ctypes.byref(f))
            assert f.value == (1 if tag == "merged" else 0)  # the register kernel emits six fields
            a.close()
        b = hip.NS((N,) * 3, NU, DT, 32, fused=True)
        _use(L, merged[0], 32)
        b.set_initial(v0)
        b.startup()
        b.steps(6)
        assert bool((b.velocity() == out["merged"][1]).all())
        b.close()
        e3 = _rel(out["merged"][0], out["plain"][0])
        e6 = _rel(out["merged"][1], out["plain"][1])
        print("FP32 merged vs (V1, N0): 3 steps %.2e, 6 steps %.2e" % (e3, e6), flush=True)
        assert e3 < 1e-6 and e6 < 1e-6
    finally:
        L.sh_tune_reset()


@pytest.mark.parametrize("prec", [64, 32])
def test_merged_graph_capture(hip, prec):
    """Graphs of merged-history steps replay bit for bit like eager steps; a
    capture that starts from a (V1, N0) state runs the (V1, N0) kernel (no
    state conversion is recorded) and also replays bit for bit."""
    import torch
    from oracle import pyoracle as O
    L = hip.lib()
    n = 64
    merged, plain = _variants(L, prec, n, "lines=")
    V0 = torch.from_numpy(O.solenoidal((n,) * 3, 41, 3.0, 1.0)).cuda()
    torch.cuda.synchronize()
    st = ctypes.c_void_p()
    L.sh_stream_create(ctypes.byref(st))
    try:
        for warm_variant in (merged[0], plain[0]):
            hs = []
            for _ in range(2):
                h = ctypes.c_void_p()
                L.sh_ns_create(ctypes.byref(h), n, n, n, 0.01, 1e-3, prec, 1, st)
                L.sh_ns_set_initial(h, hip.ptr(V0))
                L.sh_ns_startup(h)
                hs.append(h)
            _use(L, warm_variant, prec, n)
            for h in hs:
                L.sh_ns_steps(h, 1)
            _use(L, merged[0], prec, n)
            L.sh_stream_synchronize(st)
            assert L.sh_graph_begin(st) == 0
            assert L.sh_ns_steps(hs[0], 4) == 0
            g = ctypes.c_void_p()
            assert L.sh_graph_end(st, ctypes.byref(g)) == 0
            for rep in range(3):
                L.sh_graph_launch(g, st)
                L.sh_ns_steps(hs[1], 4)
                L.sh_stream_synchronize(st)
                a = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
                b = torch.empty_like(a)
                L.sh_ns_get_velocity(hs[0], hip.ptr(a))
                L.sh_ns_get_velocity(hs[1], hip.ptr(b))
                L.sh_stream_synchronize(st)
                fa, fb = ctypes.c_int(), ctypes.c_int()
                L.sh_ns_get_history_form(hs[0], ctypes.byref(fa))
                L.sh_ns_get_history_form(hs[1], ctypes.byref(fb))
                print("fp%d warm=%d rep=%d forms %d %d" % (prec, warm_variant, rep, fa.value, fb.value), flush=True)
                if warm_variant == merged[0]:
                    assert bool((a == b).all()), rep
                else:
                    # the graph holds the (V1, N0) kernel, the eager twin converts
                    assert fa.value == 0 and fb.value == 1
                    assert _rel(a, b) < (1e-13 if prec == 64 else 1e-6), rep
            L.sh_graph_destroy(g)
            for h in hs:
                L.sh_ns_destroy(h)
    finally:
        L.sh_tune_reset()
        L.sh_stream_destroy(st)


@pytest.mark.parametrize("prec,n", [(64, 256), (32, 256), (64, 512), (32, 512)])
def test_merged_five_fields_every_y_curl(hip, prec, n):
    """Kernel E emitting five fields (V and d/dx Vy, d/dx Vz) with every
    schedule of the y pass that forms the vorticity from them, both layouts,
    against the (V1, N0) kernel E with the six-field y pass (the streamed
    kernel E at 256^3 FP64, the register-loaded one otherwise)."""
    import torch
    L = hip.lib()
    merged, plain = _variants(L, prec, n, "grouped=true" if (prec, n) == (64, 256) else "lines=")
    five = []
    buf = ctypes.create_string_buffer(8192)
    for v in merged:
        _use(L, v, prec, n)
        L.sh_tune_report(buf, 8192)
        if "five emitted fields" in buf.value.decode():
            five.append(v)
    assert five
    c = ctypes.c_int()
    L.sh_tune_variant_count(prec, n, 6, ctypes.byref(c))
    assert c.value > 0
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 12, 4.0, 1.0, hip.stream())
    tol = 1e-13 if prec == 64 else 1e-6
    steps = 5 if n == 256 else 3

    def run(layout):
        ns = hip.NS((n,) * 3, NU, DT, prec, fused=True)
        L.sh_ns_set_layout(ns.h, layout)
        ns.set_initial(V0)
        ns.startup()
        ns.steps(steps)
        form = ctypes.c_int()
        L.sh_ns_get_history_form(ns.h, ctypes.byref(form))
        v = ns.velocity()
        ns.close()
        return v, form.value

    try:
        for layout in (1, 0):
            _use(L, plain[0], prec, n)
            vr, f0 = run(layout)
            assert f0 == 0
            for e in five:
                for y in range(c.value):
                    _use(L, e, prec, n)
                    L.sh_tune_set_variant(prec, n, 6, y)
                    v, form = run(layout)
                    err = _rel(v, vr)
                    print("fp%d n=%d layout=%d E=%d y_curl=%d vs (V1, N0): %.2e" % (prec, n, layout, e, y, err),
                          flush=True)
                    assert form == 2
                    assert err < tol, (layout, e, y, err)
    finally:
        L.sh_tune_reset()
        del V0
        torch.cuda.empty_cache()
        L.sh_pool_trim()


@pytest.mark.parametrize("prec", [64, 32])
def test_streamed_and_wavefront_kernels_graph_capture(hip, prec):
    """At 256^3, steps captured in a graph with the LDS-streamed kernel E and
    with the wavefront-exchange one (five emitted fields; the vorticity y
    pass after them) replay bit for bit like the same steps run eagerly on a
    twin solver, replay after replay."""
    import torch
    L = hip.lib()
    n = N
    picks = []
    buf = ctypes.create_string_buffer(8192)
    c = ctypes.c_int()
    L.sh_tune_variant_count(prec, n, ROLE_FOURIER, ctypes.byref(c))
    for v in range(c.value):
        _use(L, v, prec, n)
        L.sh_tune_report(buf, 8192)
        name = buf.value.decode()
        if "LDS-streamed" in name and "five emitted fields" in name:
            picks.append((v, "wavefront" in name))
    L.sh_tune_reset()
    assert any(w for _, w in picks) and any(not w for _, w in picks)
    chosen = [next(v for v, w in picks if w), next(v for v, w in picks if not w)]
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 13, 4.0, 1.0, hip.stream())
    torch.cuda.synchronize()
    st = ctypes.c_void_p()
    L.sh_stream_create(ctypes.byref(st))
    try:
        for v in chosen:
            _use(L, v, prec, n)
            hs = []
            for _ in range(2):
                h = ctypes.c_void_p()
                assert L.sh_ns_create(ctypes.byref(h), n, n, n, NU, DT, prec, 1, st) == 0
                L.sh_ns_set_initial(h, hip.ptr(V0))
                L.sh_ns_startup(h)
                L.sh_ns_steps(h, 1)  # the merged form and its emitted fields are set up eagerly
                hs.append(h)
            L.sh_stream_synchronize(st)
            assert L.sh_graph_begin(st) == 0
            assert L.sh_ns_steps(hs[0], 2) == 0
            g = ctypes.c_void_p()
            assert L.sh_graph_end(st, ctypes.byref(g)) == 0
            for rep in range(3):
                L.sh_graph_launch(g, st)
                L.sh_ns_steps(hs[1], 2)
                L.sh_stream_synchronize(st)
                a = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
                b = torch.empty_like(a)
                L.sh_ns_get_velocity(hs[0], hip.ptr(a))
                L.sh_ns_get_velocity(hs[1], hip.ptr(b))
                L.sh_stream_synchronize(st)
                fa = ctypes.c_int()
                L.sh_ns_get_history_form(hs[0], ctypes.byref(fa))
                print("fp%d variant=%d rep=%d form=%d equal=%s" % (prec, v, rep, fa.value, bool((a == b).all())),
                      flush=True)
                assert fa.value == 2
                assert bool((a == b).all()), (v, rep)
                del a, b
            L.sh_graph_destroy(g)
            for h in hs:
                L.sh_ns_destroy(h)
    finally:
        L.sh_tune_reset()
        L.sh_stream_destroy(st)
        del V0
        torch.cuda.empty_cache()
        L.sh_pool_trim()
