"""HIP fractional-step solver (exec/ns.cpp) against the oracle and golden
vectors.  FP64 bar: velocity within 1e-10 relative L2 after 100 steps.
Also: the FP32 form's error against FP64, closed-form limits, and
full-size (256^3) invariants."""
import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
TOL_STEP = 1e-10


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _tg_device(hip, dims):
    import torch
    nx, ny, nz = dims
    V = torch.empty((3, nz, ny, nx), dtype=torch.complex128, device="cuda")
    hip.lib().sh_taylor_green_physical_z(hip.ptr(V), nx, ny, nz, hip.stream())
    hip.forward(V)
    return V


def test_taylor_green_ic(hip):
    dims = (16, 16, 16)
    assert rel_l2(_host(_tg_device(hip, dims)), O.taylor_green(dims)) < 1e-14


@pytest.mark.parametrize("fused", [False, True])
def test_tg16_golden(hip, fused):
    """TG 16^3, nu=0.01, dt=1e-3 against the numpy restatement's fixtures."""
    g = np.load(GOLDEN + "/tg_16.npz")
    dims = (16, 16, 16)
    ns = hip.NS(dims, 0.01, 1e-3, hip.SH_F64, fused=fused)
    ns.set_initial(_tg_device(hip, dims))
    ns.startup()
    assert rel_l2(_host(ns.velocity()), g["startup"]) < 1e-13
    done = 0
    for t in (1, 10, 100):
        ns.steps(t - done)
        done = t
        assert rel_l2(_host(ns.velocity()), g["step%d" % t]) < TOL_STEP, t
    assert ns.div_average() < 1e-12
    ns.close()


@pytest.mark.parametrize("fused", [False, True])
def test_tg32_vs_oracle_100_steps(hip, fused):
    """Config 1: TG 32^3, 100 steps, against the C oracle and the fixture norms."""
    dims = (32, 32, 32)
    g = np.load(GOLDEN + "/tg_32.npz")
    ref = O.NS(dims, 0.01, 1e-3)
    ref.set_initial(O.taylor_green(dims))
    ref.startup()
    ref.steps(100)
    Vr = ref.velocity()
    ns = hip.NS(dims, 0.01, 1e-3, hip.SH_F64, fused=fused)
    ns.set_initial(_tg_device(hip, dims))
    ns.startup()
    ns.steps(100)
    V = _host(ns.velocity())
    assert rel_l2(V, Vr) < TOL_STEP
    # the strided subsample is mostly rounding-level entries: normalise its
    # error by the whole field's norm, not by the subsample's
    sub_err = np.linalg.norm(V.reshape(3, -1)[:, ::97] - g["step100_sub"]) / np.linalg.norm(V)
    assert sub_err < TOL_STEP
    assert rel_l2(np.linalg.norm(V.reshape(3, -1), axis=1), g["step100_norm"]) < 1e-12


def test_solenoidal_ic(hip):
    import torch
    dims = (16, 16, 16)
    g = np.load(GOLDEN + "/solenoidal_16.npz")
    V = torch.empty((3, 16, 16, 16), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V), 16, 16, 16, 7, 4.0, 1.0, hip.stream())
    assert rel_l2(_host(V), g["V"]) < 1e-13
    assert rel_l2(_host(V), O.solenoidal(dims, 7, 4.0, 1.0)) < 1e-13


def test_random_ic_fused_vs_oracle(hip):
    """Config-2 IC family at 32^3 (nu=1/800, dt=2e-3), 20 steps, fused FP64 vs oracle."""
    import torch
    dims = (32, 32, 32)
    V0 = torch.empty((3, 32, 32, 32), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), 32, 32, 32, 1234, 4.0, 1.0, hip.stream())
    ref = O.NS(dims, 1 / 800, 2e-3)
    ref.set_initial(_host(V0))
    ref.startup()
    ref.steps(20)
    ns = hip.NS(dims, 1 / 800, 2e-3, hip.SH_F64, fused=True)
    ns.set_initial(V0)
    ns.startup()
    ns.steps(20)
    assert rel_l2(_host(ns.velocity()), ref.velocity()) < TOL_STEP
    # |div V| is rounding noise in both (the projection is exact per mode)
    assert ns.div_average() < 1e-12 and ref.div_average() < 1e-12


def test_fp32_form_error(hip):
    """FP32 solver vs FP64 from the same IC: relative L2 velocity error stated
    and bounded (FP32 rounding grows slowly over 100 steps)."""
    import torch
    dims = (32, 32, 32)
    V0 = torch.empty((3, 32, 32, 32), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), 32, 32, 32, 5, 4.0, 1.0, hip.stream())
    out = {}
    for prec in (hip.SH_F64, hip.SH_F32):
        ns = hip.NS(dims, 1 / 800, 2e-3, prec, fused=True)
        ns.set_initial(V0)
        ns.startup()
        ns.steps(100)
        out[prec] = _host(ns.velocity())
        ns.close()
    err = rel_l2(out[hip.SH_F32], out[hip.SH_F64])
    assert err < 1e-5, err


def test_stokes_limit_closed_form(hip):
    """Nonlinear term off: each step multiplies a solenoidal mode by
    (1 - nu K/2 k^2)/(1 + nu K/2 k^2); Forward Euler start-up by (1 - nu K k^2)."""
    dims = (16, 16, 16)
    nu, K, n = 0.05, 1e-2, 30
    V0 = O.solenoidal(dims, 3, 3.0, 1.0)
    ns = hip.NS(dims, nu, K, hip.SH_F64, fused=False)
    ns.set_nonlinear(False)
    ns.set_initial(_dev(V0))
    ns.startup()
    ns.steps(n)
    kx = np.array([O.freq(i, 16) for i in range(16)])
    k2 = kx[None, None, :] ** 2 + kx[None, :, None] ** 2 + kx[:, None, None] ** 2
    c = nu * K / 2
    expect = V0 * (1 - nu * K * k2) * ((1 - c * k2) / (1 + c * k2)) ** n
    assert rel_l2(_host(ns.velocity()), expect) < 1e-13


def test_beltrami_flow_closed_form(hip):
    """ABC flow (curl u = u, |k| = 1): u x omega = 0, pressure 0, so the full
    nonlinear step reduces to the per-step factor (1 - nu K/2)/(1 + nu K/2)."""
    import torch
    N = 16
    x = 2 * np.pi * np.arange(N) / N
    X = x[None, None, :]
    Y = x[None, :, None]
    Z = x[:, None, None]
    A, B, C = 1.0, 0.7, 0.4
    u = np.stack([np.broadcast_to(A * np.sin(Z) + C * np.cos(Y), (N, N, N)),
                  np.broadcast_to(B * np.sin(X) + A * np.cos(Z), (N, N, N)),
                  np.broadcast_to(C * np.sin(Y) + B * np.cos(X), (N, N, N))]).astype(np.complex128)
    V0 = np.stack([O.forward(c) for c in u])
    nu, K, n = 0.02, 1e-2, 50
    for fused in (False, True):
        ns = hip.NS((N, N, N), nu, K, hip.SH_F64, fused=fused)
        ns.set_initial(_dev(V0))
        ns.startup()
        ns.steps(n)
        expect = V0 * (1 - nu * K) * ((1 - nu * K / 2) / (1 + nu * K / 2)) ** n
        assert rel_l2(_host(ns.velocity()), expect) < 1e-12, fused
        ns.close()


def test_full_size_256_invariants(hip):
    """Config 2 at full size for 10 steps: divergence-free, energy decays, finite."""
    import ctypes
    import torch
    n = 256
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L = hip.lib()
    L.sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 2024, 4.0, 1.0, hip.stream())
    e0 = ctypes.c_double()
    Ek = (ctypes.c_double * 8)()
    L.sh_spectrum_z(hip.ptr(V0), n, n, n, 1 / 800, Ek, 8, ctypes.byref(e0), None, hip.stream())
    assert abs(e0.value - 1.5) < 1e-10  # u_rms = 1 -> E = 3/2
    ns = hip.NS((n, n, n), 1 / 800, 2e-3, hip.SH_F64, fused=True)
    ns.set_initial(V0)
    del V0
    ns.startup()
    ns.steps(10)
    V = ns.velocity()
    assert bool(torch.isfinite(torch.view_as_real(V)).all())
    e1 = ctypes.c_double()
    L.sh_spectrum_z(hip.ptr(V), n, n, n, 1 / 800, Ek, 8, ctypes.byref(e1), None, hip.stream())
    assert 0.9 * e0.value < e1.value < e0.value
    assert ns.div_average() < 1e-12
    ns.close()
    del V
    torch.cuda.empty_cache()


def _layout(hip, ns):
    import ctypes
    h = ctypes.c_int(-1)
    hip.lib().sh_ns_get_layout(ns.h, ctypes.byref(h))
    return h.value


@pytest.mark.parametrize("prec", [64, 32])
def test_half_spectrum_layout_matches_full(hip, prec):
    """A real initial velocity engages the half-spectrum layout; it gives the
    full-spectrum (C2C) fused step's velocity to rounding, and the oracle's
    within the FP64 bar."""
    import torch
    dims = (32, 32, 32)
    V0 = torch.empty((3, 32, 32, 32), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), 32, 32, 32, 4242, 4.0, 1.0, hip.stream())
    out = {}
    for layout in (0, 1):
        ns = hip.NS(dims, 1 / 800, 2e-3, hip.SH_F64 if prec == 64 else hip.SH_F32, fused=True)
        hip.lib().sh_ns_set_layout(ns.h, layout)
        ns.set_initial(V0)
        ns.startup()
        assert _layout(hip, ns) == layout
        ns.steps(40)
        out[layout] = (_host(ns.velocity()), ns.div_average())
        ns.close()
    assert rel_l2(out[1][0], out[0][0]) < (1e-12 if prec == 64 else 1e-5)
    # the expanded velocity is Hermitian and its Nyquist plane is zero (the
    # kz = 0 plane stores k and -k separately: Hermitian to the precision's rounding)
    v = out[1][0]
    assert np.abs(v[:, 16]).max() == 0
    mirror = np.conj(np.roll(v[:, ::-1, ::-1, ::-1], 1, axis=(1, 2, 3)))
    assert rel_l2(v, mirror) < (1e-14 if prec == 64 else 1e-6)
    if prec == 64:
        ref = O.NS(dims, 1 / 800, 2e-3)
        ref.set_initial(_host(V0))
        ref.startup()
        ref.steps(40)
        assert rel_l2(out[1][0], ref.velocity()) < TOL_STEP
        # |div V| averages of the two layouts agree (both rounding noise)
        assert out[1][1] < 1e-12 and out[0][1] < 1e-12


def test_non_hermitian_start_keeps_full_layout(hip):
    """A complex velocity that is not the transform of a real field stays on
    the full spectrum and matches the op-by-op schedule."""
    dims = (16, 16, 16)
    V = O.prng_complex(31, (16, 16, 48)).reshape(3, 16, 16, 16)
    # divergence-free projection, Nyquist planes zeroed (a legal solver state)
    kx = np.array([O.freq(i, 16) for i in range(16)])
    K3 = np.stack(np.broadcast_arrays(kx[None, None, :], kx[None, :, None], kx[:, None, None]))
    k2 = (K3 ** 2).sum(0)
    V = V - K3 * ((K3 * V).sum(0) / np.where(k2 == 0, 1, k2))
    V = np.stack([O.zero_high(c) for c in V])
    out = []
    for fused in (False, True):
        ns = hip.NS(dims, 0.01, 1e-3, hip.SH_F64, fused=fused)
        ns.set_initial(_dev(V))
        ns.startup()
        if fused:
            assert _layout(hip, ns) == 0
        ns.steps(20)
        out.append(_host(ns.velocity()))
        ns.close()
    assert rel_l2(out[1], out[0]) < 1e-12


def test_layout_switch_mid_run(hip):
    """Leaving the half layout mid-run (sh_ns_set_layout 0) expands the state
    in place and continues on the same trajectory."""
    import torch
    dims = (32, 32, 32)
    V0 = torch.empty((3, 32, 32, 32), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), 32, 32, 32, 77, 3.0, 1.0, hip.stream())
    a = hip.NS(dims, 1 / 400, 2e-3, hip.SH_F64, fused=True)
    a.set_initial(V0)
    a.startup()
    assert _layout(hip, a) == 1
    a.steps(10)
    hip.lib().sh_ns_set_layout(a.h, 0)
    assert _layout(hip, a) == 0
    a.steps(10)
    b = hip.NS(dims, 1 / 400, 2e-3, hip.SH_F64, fused=False)
    b.set_initial(V0)
    b.startup()
    b.steps(20)
    assert rel_l2(_host(a.velocity()), _host(b.velocity())) < 1e-12
    a.close()
    b.close()


@pytest.mark.parametrize("dims", [(32, 24, 48), (48, 16, 32), (24, 32, 16)])
def test_non_cubic_grids_vs_oracle(hip, dims):
    """Distinct extents per axis (each axis has its own plans; a transposed
    axis or a wrong stride fails): op-by-op, fused full spectrum and fused
    half spectrum against the oracle after 20 steps."""
    nx, ny, nz = dims
    V0 = O.solenoidal(dims, 606, 3.0, 1.0)
    ref = O.NS(dims, 0.005, 2e-3)
    ref.set_initial(V0)
    ref.startup()
    ref.steps(20)
    Vr = ref.velocity()
    for fused, layout in ((False, 0), (True, 0), (True, 1)):
        ns = hip.NS(dims, 0.005, 2e-3, hip.SH_F64, fused=fused)
        if fused:
            hip.lib().sh_ns_set_layout(ns.h, layout)
        ns.set_initial(_dev(V0))
        ns.startup()
        if fused:
            assert _layout(hip, ns) == layout
        ns.steps(20)
        assert rel_l2(_host(ns.velocity()), Vr) < TOL_STEP, (fused, layout)
        ns.close()
    ns = hip.NS(dims, 0.005, 2e-3, hip.SH_F32, fused=True)
    ns.set_initial(_dev(V0))
    ns.startup()
    ns.steps(20)
    assert rel_l2(_host(ns.velocity()), Vr) < 1e-5
    ns.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("dims", [(32, 24, 48), (24, 32, 32), (64, 64, 64)])
def test_overlapped_schedule_is_identical(hip, prec, dims):
    """Two column halves of kernels B-D on two streams compute exactly what
    the single launches compute (same arithmetic per column), in both layouts."""
    import torch
    nx, ny, nz = dims
    V0 = _dev(O.solenoidal(dims, 808, 3.0, 1.0))
    for layout in (0, 1):
        out = []
        for ov in (0, 1):
            ns = hip.NS(dims, 0.005, 2e-3, hip.SH_F64 if prec == 64 else hip.SH_F32, fused=True)
            hip.lib().sh_ns_set_layout(ns.h, layout)
            hip.lib().sh_ns_set_overlap(ns.h, ov)
            ns.set_initial(V0)
            ns.startup()
            ns.steps(12)
            out.append(_host(ns.velocity()))
            ns.close()
        assert np.array_equal(out[0], out[1]), layout


@pytest.mark.parametrize("dims", [(24, 12, 8), (8, 8, 12), (12, 24, 24)])
def test_small_grids_layout_choice(hip, dims):
    """Lengths 8, 12 and 24: nz = 8 and 12 have no half-length fused plan, so
    the solver keeps the full spectrum; nz = 24 steps the half layout (a
    length-12 transform of packed pairs).  Both against the oracle."""
    nx, ny, nz = dims
    V0 = O.solenoidal(dims, 515, 2.0, 1.0)
    ref = O.NS(dims, 0.01, 2e-3)
    ref.set_initial(V0)
    ref.startup()
    ref.steps(15)
    ns = hip.NS(dims, 0.01, 2e-3, hip.SH_F64, fused=True)
    ns.set_initial(_dev(V0))
    ns.startup()
    assert _layout(hip, ns) == (1 if nz == 24 else 0)
    ns.steps(15)
    assert rel_l2(_host(ns.velocity()), ref.velocity()) < TOL_STEP
    ns.close()
