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


def _variants(L):
    """(merged, plain): E schedules at 256 with and without the merged history
    (plain: the same LDS-streamed kernel keeping V1 and N0)."""
    c = ctypes.c_int()
    L.sh_tune_variant_count(64, N, ROLE_FOURIER, ctypes.byref(c))
    merged, plain = [], []
    buf = ctypes.create_string_buffer(8192)
    for v in range(c.value):
        L.sh_tune_reset()
        L.sh_tune_set_variant(64, N, ROLE_FOURIER, v)
        L.sh_tune_report(buf, 8192)
        name = buf.value.decode()
        if "merged AB2 history" in name:
            merged.append(v)
        elif "LDS-streamed" in name and "grouped=true" in name:
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


def _use(L, v):
    L.sh_tune_reset()
    L.sh_tune_set_variant(64, N, ROLE_FOURIER, v)


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
