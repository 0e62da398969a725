"""Every tunable schedule computes the same transform: each variant of each pass
role is checked against the oracle (FFT bar 1e-12, step bar 1e-10), and the
autotuners leave results (and the solver state) unchanged."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
ROLES_FFT = (0, 1, 2)  # x, y, z
ROLE_CROSS, ROLE_FOURIER, ROLE_CROSS_H = 3, 4, 5
# (role, storage layout the solver must use for the role to run): the full-
# spectrum z kernel runs with layout 0, the half-spectrum one with layout 1
NS_ROLES = ((1, 1), (ROLE_CROSS_H, 1), (ROLE_FOURIER, 1), (1, 0), (ROLE_CROSS, 0), (ROLE_FOURIER, 0))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _count(L, prec, n, role):
    c = ctypes.c_int()
    L.sh_tune_variant_count(prec, n, role, ctypes.byref(c))
    return c.value


@pytest.mark.parametrize("n", [64, 96, 256])
@pytest.mark.parametrize("prec", [64, 32])
def test_fft_variants_vs_oracle(hip, n, prec):
    L = hip.lib()
    L.sh_tune_reset()
    dims = (n, n, n)
    x = O.prng_complex(500 + n, dims)
    ref = O.transform3d(x, -1)
    tol = 1e-12 if prec == 64 else 3e-6
    try:
        for role in ROLES_FFT:
            for v in range(_count(L, prec, n, role)):
                L.sh_tune_reset()
                L.sh_tune_set_variant(prec, n, role, v)
                t = _dev(x if prec == 64 else x.astype(np.complex64))
                hip.fft3d(t, -1, 0)
                err = rel_l2(_host(t).astype(np.complex128), ref)
                # the same schedule's inverse: back to x (normalised)
                hip.fft3d(t, 1, hip.SH_NORMALIZE)
                back = rel_l2(_host(t).astype(np.complex128), x)
                print("fp%d n=%d role=%d variant=%d err=%.2e round trip %.2e" % (prec, n, role, v, err, back),
                      flush=True)
                assert err < tol, (role, v, err)
                assert back < tol, (role, v, back)
    finally:
        L.sh_tune_reset()


@pytest.mark.parametrize("prec", [64, 32])
def test_fft_variants_512_round_trip(hip, prec):
    """512^3: each schedule against the default one, plus round trip (FP64
    1e-13, FP32 2e-6)."""
    import torch
    L = hip.lib()
    L.sh_tune_reset()
    n = 512
    tol = 1e-13 if prec == 64 else 2e-6
    x = torch.empty((n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_prng_complex_z(hip.ptr(x), x.numel(), 3, hip.stream())
    if prec == 32:
        x = x.to(torch.complex64)
    ref = x.clone()
    hip.fft3d(ref, -1, 0)
    try:
        for role in ROLES_FFT:
            for v in range(1, _count(L, prec, n, role)):
                L.sh_tune_reset()
                L.sh_tune_set_variant(prec, n, role, v)
                y = x.clone()
                hip.fft3d(y, -1, 0)
                err = float(torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref))
                print("fp%d n=512 role=%d variant=%d err=%.2e" % (prec, role, v, err), flush=True)
                assert err < tol, (role, v, err)
                hip.fft3d(y, 1, hip.SH_NORMALIZE)
                assert float(torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x)) < tol
                del y
    finally:
        L.sh_tune_reset()
        del x, ref
        torch.cuda.empty_cache()


@pytest.mark.parametrize("n,prec", [(64, 64), (64, 32), (96, 64), (96, 32), (128, 64)])
def test_ns_variants_vs_oracle(hip, n, prec):
    import torch
    L = hip.lib()
    dims = (n, n, n)
    V0 = O.solenoidal(dims, 31, 3.0, 1.0)
    ref = O.NS(dims, 0.01, 1e-3)
    ref.set_initial(V0)
    ref.startup()
    ref.steps(8)
    Vr = ref.velocity()
    tol = 1e-10 if prec == 64 else 1e-5
    try:
        for role, layout in NS_ROLES:
            for v in range(_count(L, prec, n, role)):
                L.sh_tune_reset()
                L.sh_tune_set_variant(prec, n, role, v)
                ns = hip.NS(dims, 0.01, 1e-3, prec, fused=True)
                L.sh_ns_set_layout(ns.h, layout)
                ns.set_initial(_dev(V0))
                ns.startup()
                ns.steps(8)
                err = rel_l2(_host(ns.velocity()), Vr)
                ns.close()
                print("fp%d n=%d role=%d layout=%d variant=%d err=%.2e" % (prec, n, role, layout, v, err),
                      flush=True)
                assert err < tol, (role, layout, v, err)
    finally:
        L.sh_tune_reset()


def test_ns_variants_256_agree(hip):
    """At 256^3 each schedule of the fused kernels reproduces the default's 3 steps."""
    import torch
    L = hip.lib()
    L.sh_tune_reset()
    n = 256
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 9, 4.0, 1.0, hip.stream())

    def run(layout):
        ns = hip.NS((n, n, n), 1 / 800, 5e-4, 64, fused=True)
        L.sh_ns_set_layout(ns.h, layout)
        ns.set_initial(V0)
        ns.startup()
        ns.steps(3)
        V = ns.velocity()
        ns.close()
        return V

    refs = {0: run(0), 1: run(1)}
    # the two layouts agree to rounding
    assert float(torch.linalg.vector_norm(refs[1] - refs[0]) / torch.linalg.vector_norm(refs[0])) < 1e-13
    try:
        for role, layout in NS_ROLES:
            ref = refs[layout]
            for v in range(1, _count(L, 64, n, role)):
                L.sh_tune_reset()
                L.sh_tune_set_variant(64, n, role, v)
                V = run(layout)
                err = float(torch.linalg.vector_norm(V - ref) / torch.linalg.vector_norm(ref))
                print("n=256 role=%d layout=%d variant=%d err=%.2e" % (role, layout, v, err), flush=True)
                assert err < 1e-13, (role, layout, v, err)
                del V
    finally:
        L.sh_tune_reset()
        del V0, refs
        torch.cuda.empty_cache()


def test_autotuners_keep_results(hip):
    import torch
    L = hip.lib()
    L.sh_tune_reset()
    n = 64
    dims = (n, n, n)
    L.sh_autotune_fft3d(n, n, n, 64, 2, hip.stream())
    x = O.prng_complex(77, dims)
    t = _dev(x)
    hip.fft3d(t, -1, 0)
    assert rel_l2(_host(t), O.transform3d(x, -1)) < 1e-12
    V0 = O.solenoidal(dims, 32, 3.0, 1.0)
    ns = hip.NS(dims, 0.01, 1e-3, 64, fused=True)
    ns.set_initial(_dev(V0))
    ns.startup()
    ns.steps(2)
    before = _host(ns.velocity())
    L.sh_ns_autotune(ns.h, 2)
    assert np.array_equal(_host(ns.velocity()), before)
    ns.steps(4)
    ref = O.NS(dims, 0.01, 1e-3)
    ref.set_initial(V0)
    ref.startup()
    ref.steps(6)
    assert rel_l2(_host(ns.velocity()), ref.velocity()) < 1e-10
    buf = ctypes.create_string_buffer(4096)
    L.sh_tune_report(buf, 4096)
    report = buf.value.decode()
    print(report)
    assert "ns_cross_z" in report and "n=64" in report
    ns.close()
    L.sh_tune_reset()
