"""The kernels and shapes the measured numbers come from, against the oracle:

  * 256^3 FP64 fused step (the headline workload's grid and IC family) in both
    storage layouts: every compiled schedule of each fused role, the
    overlapped two-stream form and the autotuner's pick, each for the
    start-up + 3 steps at dt = 5e-4, against the C oracle at 1e-10 relative
    L2 of the velocity (exec/ns.cpp:132-217); every FP32 schedule of the same
    roles against it at 1e-5;
  * 512^3 FP32 fused step against the FP64 one from the same seed (20 steps
    at the stable dt 2.5e-4, both autotuned): velocity, shell spectrum and
    dissipation errors within stated bounds;
  * an FP32 ensemble of 2 boxes at 64^3 and 2 at 96^3 captured into one
    graph with snapshots: every box against its own single solver, and one
    box of each size against the oracle.
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
ROLE_CROSS, ROLE_FOURIER, ROLE_CROSS_H = 3, 4, 5
# (role, layout the solver must use for the role's kernel to run)
NS_ROLES = ((1, 1), (ROLE_CROSS_H, 1), (ROLE_FOURIER, 1), (1, 0), (ROLE_CROSS, 0), (ROLE_FOURIER, 0))
N256, NU, DT = 256, 1.0 / 800, 5e-4


def _count(L, prec, n, role):
    c = ctypes.c_int()
    L.sh_tune_variant_count(prec, n, role, ctypes.byref(c))
    return c.value


def _rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.fixture(scope="module")
def ref256(hip):
    """The oracle's velocity after the start-up and 3 steps from the config-2
    IC family at 256^3 (seed 9), on the device for the comparisons."""
    import torch
    L = hip.lib()
    n = N256
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 9, 4.0, 1.0, hip.stream())
    print("oracle 256^3: start-up + 3 steps on %d threads" % O.num_threads(), flush=True)
    ref = O.NS((n, n, n), NU, DT)
    ref.set_initial(V0.cpu().numpy())
    ref.startup()
    ref.steps(3)
    Vr = torch.from_numpy(ref.velocity()).cuda()
    ref.close()
    yield V0, Vr
    del V0, Vr
    torch.cuda.empty_cache()


def _run256(hip, V0, layout, overlap=0, tune=False, prec=64):
    L = hip.lib()
    ns = hip.NS((N256,) * 3, NU, DT, prec, fused=True)
    L.sh_ns_set_layout(ns.h, layout)
    L.sh_ns_set_overlap(ns.h, overlap)
    ns.set_initial(V0)
    ns.startup()
    half = ctypes.c_int()
    L.sh_ns_get_layout(ns.h, ctypes.byref(half))
    assert half.value == layout
    if tune:
        L.sh_ns_autotune(ns.h, 2)
    ns.steps(3)
    V = ns.velocity()
    ns.close()
    return V


def test_ns_256_every_schedule_vs_oracle(hip, ref256):
    V0, Vr = ref256
    L = hip.lib()
    try:
        for role, layout in NS_ROLES:
            for v in range(_count(L, 64, N256, role)):
                L.sh_tune_reset()
                L.sh_tune_set_variant(64, N256, role, v)
                err = _rel(_run256(hip, V0, layout), Vr)
                print("256^3 role=%d layout=%d variant=%d err=%.2e" % (role, layout, v, err), flush=True)
                assert err < 1e-10, (role, layout, v, err)
        L.sh_tune_reset()
        for layout in (1, 0):
            err = _rel(_run256(hip, V0, layout, overlap=1), Vr)
            print("256^3 overlapped layout=%d err=%.2e" % (layout, err), flush=True)
            assert err < 1e-10, (layout, err)
    finally:
        L.sh_tune_reset()


def test_ns_256_fp32_every_schedule_vs_oracle(hip, ref256):
    """The FP32 form's kernels at 256^3 (the bench's FP32 line): every
    compiled FP32 schedule of each fused role against the FP64 oracle, at the
    FP32 solver's 1e-5 relative L2."""
    V0, Vr = ref256
    L = hip.lib()
    try:
        for role, layout in NS_ROLES:
            for v in range(_count(L, 32, N256, role)):
                L.sh_tune_reset()
                L.sh_tune_set_variant(32, N256, role, v)
                err = _rel(_run256(hip, V0, layout, prec=32), Vr)
                print("256^3 fp32 role=%d layout=%d variant=%d err=%.2e" % (role, layout, v, err), flush=True)
                assert err < 1e-5, (role, layout, v, err)
    finally:
        L.sh_tune_reset()


def test_ns_256_autotuned_vs_oracle(hip, ref256):
    """The schedules bench.py times: sh_ns_autotune's pick per layout."""
    V0, Vr = ref256
    L = hip.lib()
    try:
        for layout in (1, 0):
            L.sh_tune_reset()
            err = _rel(_run256(hip, V0, layout, tune=True), Vr)
            buf = ctypes.create_string_buffer(8192)
            L.sh_tune_report(buf, 8192)
            print("256^3 autotuned layout=%d err=%.2e\n%s" % (layout, err, buf.value.decode()), flush=True)
            assert "n=256" in buf.value.decode()
            assert err < 1e-10, (layout, err)
    finally:
        L.sh_tune_reset()


def test_fp32_512_against_fp64(hip):
    """Config 3's form at 512^3 for 20 steps (stable dt): the FP32 solver's
    errors against FP64 stay within the bounds stated for it (100-step
    figures: velocity 2.7e-6, E(k) for k <= 16 1.1e-6, dissipation 4.6e-8)."""
    from spectral_upcxx_amd import forms
    r = forms.fp32_error(n=512, steps=20, dt=2.5e-4)
    print(r, flush=True)
    assert r["velocity_rel_l2"] < 1e-5
    assert r["spectrum_max_rel_err_k_le_16"] < 1e-5
    assert r["energy_rel_err"] < 1e-6
    assert r["dissipation_rel_err"] < 1e-6
    hip.lib().sh_pool_trim()


def test_ensemble_64_96_graph_vs_single_and_oracle(hip):
    import torch
    from spectral_upcxx_amd import forms
    L = hip.lib()
    L.sh_tune_reset()
    steps, every, ratio, dt = 20, 10, 4, 1e-3
    ens = forms.Ensemble(((64, 2), (96, 2)), dt, precision=32, seed=77, tune=True)
    try:
        snaps = [torch.empty((steps // every,) + s, dtype=torch.float32, device="cuda")
                 for s in ens.snapshot_shapes(ratio)]
        ens.snapshot([s[0] for s in snaps], ratio)  # the snapshot scratch exists before capture
        L.sh_stream_synchronize(ens.stream)
        g = ens.capture(steps, every, ratio, snaps)
        L.sh_graph_launch(g, ens.stream)
        L.sh_stream_synchronize(ens.stream)
        L.sh_graph_destroy(g)
        assert all(bool(torch.isfinite(s).all()) for s in snaps)
        for gi, (n, boxes) in enumerate(ens.params):
            for b, p in enumerate(boxes):
                V = ens.velocity(gi, b)
                V0 = forms.solenoidal(n, p["seed"], p["k0"])
                single = hip.NS((n, n, n), p["nu"], dt, 32, fused=True)
                single.set_initial(V0)
                single.startup()
                single.steps(steps)
                err = _rel(V, single.velocity())
                single.close()
                print("ensemble n=%d box=%d vs single err=%.2e" % (n, b, err), flush=True)
                assert err < 1e-6, (n, b, err)
                if b == 0:
                    ref = O.NS((n, n, n), p["nu"], dt)
                    ref.set_initial(V0.cpu().numpy())
                    ref.startup()
                    ref.steps(steps)
                    err = _rel(V, torch.from_numpy(ref.velocity()).cuda())
                    ref.close()
                    print("ensemble n=%d box=%d vs oracle err=%.2e" % (n, b, err), flush=True)
                    assert err < 1e-5, (n, b, err)
    finally:
        ens.close()
        L.sh_tune_reset()


@pytest.mark.parametrize("prec", [64, 32])
def test_ns_512_cross_schedules_agree(hip, prec):
    """Every half-spectrum kernel C schedule at 512^3 (the LDS-streamed ones
    with packed lines of 256 included) gives the default schedule's velocity
    after the start-up and 2 steps: 1e-12 in FP64, 1e-6 in FP32."""
    import torch
    L = hip.lib()
    n = 512
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 5, 4.0, 1.0, hip.stream())
    tol = 1e-12 if prec == 64 else 1e-6
    ref = None
    try:
        for v in range(_count(L, prec, n, ROLE_CROSS_H)):
            L.sh_tune_reset()
            L.sh_tune_set_variant(prec, n, ROLE_CROSS_H, v)
            ns = hip.NS((n,) * 3, NU, 2.5e-4, prec, fused=True)
            ns.set_initial(V0)
            ns.startup()
            ns.steps(2)
            V = ns.velocity()
            ns.close()
            if ref is None:
                ref = V
                continue
            err = _rel(V, ref)
            print("512^3 fp%d cross_zh variant=%d vs default: %.2e" % (prec, v, err), flush=True)
            assert err < tol, (v, err)
            del V
    finally:
        L.sh_tune_reset()
        del V0, ref
        torch.cuda.empty_cache()
        L.sh_pool_trim()


def test_ns_512_fp32_streamed_fourier_agree(hip):
    """The FP32 kernel E with LDS-streamed inputs at 512^3 (512-point lines,
    three radix-8 stages per wave), six and five emitted fields, gives the
    default schedule's velocity after the start-up and 3 steps (1e-6), and
    its history form is the merged one."""
    import torch
    L = hip.lib()
    n, prec = 512, 32
    names = []
    buf = ctypes.create_string_buffer(8192)
    for v in range(_count(L, prec, n, ROLE_FOURIER)):
        L.sh_tune_reset()
        L.sh_tune_set_variant(prec, n, ROLE_FOURIER, v)
        L.sh_tune_report(buf, 8192)
        if "LDS-streamed" in buf.value.decode():
            names.append((v, "five emitted fields" in buf.value.decode()))
    assert len(names) >= 2
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 6, 4.0, 1.0, hip.stream())

    def run(v):
        L.sh_tune_reset()
        if v is not None:
            L.sh_tune_set_variant(prec, n, ROLE_FOURIER, v)
        ns = hip.NS((n,) * 3, NU, 2.5e-4, prec, fused=True)
        ns.set_initial(V0)
        ns.startup()
        ns.steps(3)
        form = ctypes.c_int()
        L.sh_ns_get_history_form(ns.h, ctypes.byref(form))
        V = ns.velocity()
        ns.close()
        return V, form.value

    try:
        ref, f0 = run(None)
        assert f0 == 0
        for v, five in names:
            V, form = run(v)
            err = _rel(V, ref)
            print("512^3 fp32 streamed E variant=%d five=%s form=%d vs default: %.2e" % (v, five, form, err),
                  flush=True)
            assert form == (2 if five else 1), (v, form)
            assert err < 1e-6, (v, err)
            del V
    finally:
        L.sh_tune_reset()
        del V0, ref
        torch.cuda.empty_cache()
        L.sh_pool_trim()
