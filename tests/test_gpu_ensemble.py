"""Ensembles of equal boxes in shared launches, box-filtered snapshots and
hipGraph capture: each box matches a single solver and the oracle; the filter
matches a numpy restatement; graph replays match eager execution."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _make_ensemble(hip, n, nus, K, prec, ics, stream=None):
    L = hip.lib()
    h = ctypes.c_void_p()
    arr = (ctypes.c_double * len(nus))(*nus)
    L.sh_ns_create_ensemble(ctypes.byref(h), n, len(nus), arr, K, prec, stream)
    for b, V0 in enumerate(ics):
        L.sh_ns_set_initial_box(h, b, hip.ptr(_dev(V0)))
    return h


def _box_velocity(hip, h, b, n):
    import torch
    V = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    hip.lib().sh_ns_get_velocity_box(h, b, hip.ptr(V))
    return _host(V)


def test_ensemble_boxes_match_single_and_oracle(hip):
    n, K = 32, 1e-3
    nus = [0.01, 0.003, 0.006]
    ics = [O.solenoidal((n, n, n), 40 + b, 2.0 + b, 1.0) for b in range(3)]
    h = _make_ensemble(hip, n, nus, K, 64, ics)
    hip.lib().sh_ns_steps(h, 10)
    for b in range(3):
        V = _box_velocity(hip, h, b, n)
        single = hip.NS((n, n, n), nus[b], K, 64, fused=True)
        single.set_initial(_dev(ics[b]))
        single.startup()
        single.steps(10)
        assert rel_l2(V, _host(single.velocity())) < 1e-14, b
        single.close()
        ref = O.NS((n, n, n), nus[b], K)
        ref.set_initial(ics[b])
        ref.startup()
        ref.steps(10)
        assert rel_l2(V, ref.velocity()) < 1e-10, b
        d = ctypes.c_double()
        hip.lib().sh_ns_div_average_box(h, b, ctypes.byref(d))
        assert d.value < 1e-12
    # the single-domain getters refuse an ensemble handle (its boxes are
    # field-major; a 3N copy would mix boxes)
    import torch
    buf = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    for call in (lambda: hip.lib().sh_ns_get_velocity(h, hip.ptr(buf)),
                 lambda: hip.lib().sh_ns_get_vorticity(h, hip.ptr(buf)),
                 lambda: hip.lib().sh_ns_div_average(h, ctypes.byref(ctypes.c_double()))):
        with pytest.raises(RuntimeError, match="box"):
            call()
    hip.lib().sh_ns_destroy(h)


def _box_filter_ref(u, r):
    """numpy restatement: ratio-r box averages of u_i and u_i u_j."""
    nz, ny, nx = u.shape[1:]
    cz, cy, cx = nz // r, ny // r, nx // r

    def avg(f):
        return f.reshape(cz, r, cy, r, cx, r).mean(axis=(1, 3, 5))

    ub = [avg(u[i]) for i in range(3)]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tau = [avg(u[i] * u[j]) - ub[i] * ub[j] for i, j in pairs]
    return np.stack(ub + tau).reshape(9, -1)


@pytest.mark.parametrize("prec", [64, 32])
def test_snapshot_box_filter(hip, prec):
    import torch
    n, r = 32, 4
    ics = [O.solenoidal((n, n, n), 50 + b, 3.0, 1.0) for b in range(2)]
    h = _make_ensemble(hip, n, [0.01, 0.005], 1e-3, prec, ics)
    hip.lib().sh_ns_steps(h, 3)
    out = torch.empty((2, 9, (n // r) ** 3), dtype=torch.float32, device="cuda")
    hip.lib().sh_ns_snapshot(h, r, hip.ptr(out))
    got = _host(out)
    tol = 1e-6 if prec == 64 else 2e-5
    for b in range(2):
        V = _box_velocity(hip, h, b, n)
        u = np.stack([O.inverse(c).real for c in V])
        ref = _box_filter_ref(u, r)
        assert rel_l2(got[b][:3], ref[:3]) < tol
        assert rel_l2(got[b][3:], ref[3:]) < 50 * tol
    hip.lib().sh_ns_destroy(h)


@pytest.mark.parametrize("overlap", [0, 1])
def test_graph_replay_matches_eager(hip, overlap):
    """Capture 5 steps + a snapshot; replay 3 times with eager work and
    synchronisation in between; each replay equals 5 more eager steps.  With
    overlap the capture spans the solver's second stream."""
    import torch
    L = hip.lib()
    n = 32
    nus = [0.01, 0.004]
    ics = [O.solenoidal((n, n, n), 60 + b, 3.0, 1.0) for b in range(2)]
    st = ctypes.c_void_p()
    L.sh_stream_create(ctypes.byref(st))
    hg = _make_ensemble(hip, n, nus, 1e-3, 32, ics, st)
    he = _make_ensemble(hip, n, nus, 1e-3, 32, ics, st)
    L.sh_ns_set_overlap(hg, overlap)
    L.sh_ns_set_overlap(he, overlap)
    snap = torch.empty((2, 9, 512), dtype=torch.float32, device="cuda")
    L.sh_ns_snapshot(hg, 4, hip.ptr(snap))
    L.sh_ns_steps(hg, 1)  # twiddles and W ready before capture
    L.sh_ns_steps(he, 1)
    L.sh_stream_synchronize(st)
    L.sh_graph_begin(st)
    L.sh_ns_steps(hg, 5)
    L.sh_ns_snapshot(hg, 4, hip.ptr(snap))
    g = ctypes.c_void_p()
    L.sh_graph_end(st, ctypes.byref(g))
    for rep in range(3):
        L.sh_graph_launch(g, st)
        L.sh_stream_synchronize(st)
        L.sh_ns_steps(he, 5)
        L.sh_stream_synchronize(st)
        for b in range(2):
            a = _box_velocity(hip, hg, b, n)
            e = _box_velocity(hip, he, b, n)
            assert rel_l2(a, e) == 0.0, (rep, b)
        assert bool(torch.isfinite(snap).all())
    L.sh_graph_destroy(g)
    L.sh_ns_destroy(hg)
    L.sh_ns_destroy(he)
    L.sh_stream_destroy(st)


def test_ensemble_run_and_snapshot_file(hip, tmp_path):
    from spectral_upcxx_amd import forms
    out = str(tmp_path / "snaps.bin")
    res = forms.ensemble_run(groups=((32, 3), (48, 2)), steps=20, every=10, ratio=4, dt=1e-3,
                             precision=32, seed=7, out=out)
    assert res["boxes"] == 5 and res["box_steps_per_s"] > 0
    header, arrays = forms.read_snapshots(out)
    assert [a.shape for a in arrays] == [(2, 3, 9, 8, 8, 8), (2, 2, 9, 12, 12, 12)]
    assert all(np.isfinite(a).all() for a in arrays)
    # the stresses' trace is positive (sum of variances)
    for a in arrays:
        assert (a[:, :, 3] + a[:, :, 4] + a[:, :, 5]).min() >= 0
    ks = [b["k0"] for g in header["arrays"] for b in g["boxes"]]
    nus = [b["nu"] for g in header["arrays"] for b in g["boxes"]]
    assert all(2 <= k <= 6 for k in ks) and all(1 / 1000 <= v <= 1 / 300 for v in nus)


@pytest.mark.parametrize("prec", [64, 32])
def test_ensemble_merged_history(hip, prec):
    """Ensembles keep the merged AB2 history in the register-loaded kernel E
    (box = blockIdx.y, each box with its own coefficients): each box matches
    the single solver on the same schedule and the oracle; installing one
    box's start converts the others back to (V1, N0) without changing them;
    sh_ns_prepare converts without advancing time, and a graph captured
    after it holds the merged kernel and replays like eager steps."""
    import torch
    L = hip.lib()
    n, K = 64, 1e-3
    nus = [0.01, 0.003, 0.006]
    ics = [O.solenoidal((n, n, n), 50 + b, 2.0 + b, 1.0) for b in range(3)]
    buf = ctypes.create_string_buffer(8192)
    c = ctypes.c_int()
    L.sh_tune_variant_count(prec, n, 4, ctypes.byref(c))
    merged = []
    for v in range(c.value):
        L.sh_tune_reset()
        L.sh_tune_set_variant(prec, n, 4, v)
        L.sh_tune_report(buf, 8192)
        if "merged AB2 history" in buf.value.decode() and "LDS-streamed" not in buf.value.decode():
            merged.append(v)
    assert merged
    tol = 1e-14 if prec == 64 else 1e-6

    def form(h):
        f = ctypes.c_int()
        L.sh_ns_get_history_form(h, ctypes.byref(f))
        return f.value

    try:
        L.sh_tune_reset()
        L.sh_tune_set_variant(prec, n, 4, merged[0])
        h = _make_ensemble(hip, n, nus, K, prec, ics)
        L.sh_ns_steps(h, 6)
        assert form(h) == 1
        vs = [_box_velocity(hip, h, b, n) for b in range(3)]
        for b in range(3):
            single = hip.NS((n, n, n), nus[b], K, prec, fused=True)
            single.set_initial(_dev(ics[b]))
            single.startup()
            single.steps(6)
            err = rel_l2(vs[b], _host(single.velocity()))
            print("fp%d box %d vs single: %.2e" % (prec, b, err), flush=True)
            assert err < tol, b
            single.close()
            if prec == 64:
                ref = O.NS((n, n, n), nus[b], K)
                ref.set_initial(ics[b])
                ref.startup()
                ref.steps(6)
                assert rel_l2(vs[b], ref.velocity()) < 1e-10, b
        # a new start for box 1: the others back to (V1, N0), unchanged
        L.sh_ns_set_initial_box(h, 1, hip.ptr(_dev(ics[0])))
        assert form(h) == 0
        for b in (0, 2):
            assert rel_l2(_box_velocity(hip, h, b, n), vs[b]) < (1e-13 if prec == 64 else 1e-6), b
        # prepare: merged without a step, the velocities unchanged to rounding
        before = [_box_velocity(hip, h, b, n) for b in range(3)]
        assert L.sh_ns_prepare(h) == 0
        assert form(h) == 1
        for b in range(3):
            assert rel_l2(_box_velocity(hip, h, b, n), before[b]) < (1e-13 if prec == 64 else 1e-6), b
        # a graph captured after prepare replays like an eager twin
        st = ctypes.c_void_p()
        L.sh_stream_create(ctypes.byref(st))
        twins = [_make_ensemble(hip, n, nus, K, prec, ics, st) for _ in range(2)]
        for t in twins:
            assert L.sh_ns_prepare(t) == 0
        L.sh_stream_synchronize(st)
        assert L.sh_graph_begin(st) == 0
        assert L.sh_ns_steps(twins[0], 3) == 0
        g = ctypes.c_void_p()
        assert L.sh_graph_end(st, ctypes.byref(g)) == 0
        for rep in range(2):
            L.sh_graph_launch(g, st)
            L.sh_ns_steps(twins[1], 3)
            L.sh_stream_synchronize(st)
            assert form(twins[0]) == 1 and form(twins[1]) == 1
            for b in range(3):
                a = _box_velocity(hip, twins[0], b, n)
                e = _box_velocity(hip, twins[1], b, n)
                assert np.array_equal(a, e), (rep, b)
        L.sh_graph_destroy(g)
        for t in twins:
            L.sh_ns_destroy(t)
        L.sh_stream_destroy(st)
        L.sh_ns_destroy(h)
    finally:
        L.sh_tune_reset()
