"""exec/ns2.cpp on the device: the backward-Euler viscous step, the
Gaussian-vortex initial conditions with the drivers' explicit omega0, the
vorticity diagnostics, and the marker particles of src/Particle.cpp.

Bars: velocity within 1e-10 relative L2 of the oracle after the stated
steps (FP64), initial conditions within 1e-13, particle positions,
velocities and energies bit-identical to the oracle (both do the same IEEE
double operations in the reference's order with no contraction)."""
import numpy as np
import pytest

from conftest import rel_l2
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


def _k2(n):
    kx = np.array([O.freq(i, n) for i in range(n)])
    return kx[None, None, :] ** 2 + kx[None, :, None] ** 2 + kx[:, None, None] ** 2


@pytest.mark.parametrize("fused", [False, True])
def test_backward_euler_tg32_vs_oracle(hip, fused):
    dims = (32, 32, 32)
    ref = O.NS(dims, 0.01, 1e-3)
    ref.set_scheme(1)
    ref.set_initial(O.taylor_green(dims))
    ref.startup()
    ref.steps(50)
    ns = hip.NS(dims, 0.01, 1e-3, hip.SH_F64, fused=fused)
    ns.set_scheme(1)
    ns.set_initial(_dev(O.taylor_green(dims)))
    ns.startup()
    ns.steps(50)
    assert rel_l2(_host(ns.velocity()), ref.velocity()) < TOL_STEP
    assert ns.div_average() < 1e-12
    ns.close()


def test_backward_euler_random_ic_fused_vs_oracle(hip):
    import torch
    dims = (32, 32, 32)
    V0 = torch.empty((3, 32, 32, 32), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), 32, 32, 32, 99, 4.0, 1.0, hip.stream())
    ref = O.NS(dims, 1 / 800, 2e-3)
    ref.set_scheme(1)
    ref.set_initial(_host(V0))
    ref.startup()
    ref.steps(20)
    ns = hip.NS(dims, 1 / 800, 2e-3, hip.SH_F64, fused=True)
    ns.set_scheme(1)
    ns.set_initial(V0)
    ns.startup()
    ns.steps(20)
    v = _host(ns.velocity())
    assert rel_l2(v, ref.velocity()) < TOL_STEP
    # switching back to Crank-Nicolson changes the trajectory
    cn = hip.NS(dims, 1 / 800, 2e-3, hip.SH_F64, fused=True)
    cn.set_initial(V0)
    cn.startup()
    cn.steps(20)
    assert rel_l2(_host(cn.velocity()), v) > 1e-9
    ns.close()
    cn.close()


def test_backward_euler_stokes_closed_form(hip):
    dims = (16, 16, 16)
    nu, K, n = 0.05, 1e-2, 30
    V0 = O.solenoidal(dims, 3, 3.0, 1.0)
    ns = hip.NS(dims, nu, K, hip.SH_F64, fused=False)
    ns.set_scheme(1)
    ns.set_nonlinear(False)
    ns.set_initial(_dev(V0))
    ns.startup()
    ns.steps(n)
    k2 = _k2(16)
    expect = V0 * (1 - nu * K * k2) / (1 + nu * K * k2) ** n
    assert rel_l2(_host(ns.velocity()), expect) < 1e-13


@pytest.mark.parametrize("prec", [64, 32])
def test_backward_euler_beltrami_closed_form(hip, prec):
    N = 16
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    A, B, C = 1.0, 0.7, 0.4
    u = [A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z), C * np.sin(Y) + B * np.cos(X)]
    V0 = np.stack([O.forward(np.broadcast_to(c, (N, N, N)).astype(np.complex128)) for c in u])
    nu, K, n = 0.02, 1e-2, 50
    expect = V0 * (1 - nu * K) / (1 + nu * K) ** n
    for fused in ((False, True) if prec == 64 else (True,)):
        ns = hip.NS((N, N, N), nu, K, hip.SH_F64 if prec == 64 else hip.SH_F32, fused=fused)
        ns.set_scheme(1)
        ns.set_initial(_dev(V0))
        ns.startup()
        ns.steps(n)
        assert rel_l2(_host(ns.velocity()), expect) < (1e-12 if prec == 64 else 1e-5), fused
        ns.close()


def test_set_scheme_rejects_unknown(hip):
    ns = hip.NS((16, 16, 16), 0.01, 1e-3, hip.SH_F64, fused=True)
    try:
        with pytest.raises(RuntimeError):
            hip.lib().sh_ns_set_scheme(ns.h, 2)
    finally:
        ns.close()


def _vortex_device(hip, n, which):
    import torch
    V = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    W = torch.empty_like(V)
    hip.lib().sh_vortex_ic_z(hip.ptr(V), hip.ptr(W), n, n, n, which, hip.stream())
    return V, W


def test_vortex_ic_vs_oracle(hip):
    n = 32
    for which in (0, 1):
        V, W = _vortex_device(hip, n, which)
        Vr, Wr = O.vortices((n, n, n), which)
        assert rel_l2(_host(W), Wr) < 1e-13, which
        assert rel_l2(_host(V), Vr) < 1e-13, which


@pytest.mark.parametrize("fused", [False, True])
def test_ns2_vortex_run_vs_oracle(hip, fused):
    """exec/ns2.cpp's own run at 32^3 (two vortices, nu = K = 1e-3, backward
    Euler, omega0 kept from the IC), 30 steps."""
    n = 32
    V, W = _vortex_device(hip, n, 1)
    Vr, Wr = O.vortices((n, n, n), 1)
    ref = O.NS((n, n, n), 1e-3, 1e-3)
    ref.set_scheme(1)
    ref.set_initial(Vr)
    ref.set_vorticity(Wr)
    ref.startup()
    ref.steps(30)
    ns = hip.NS((n, n, n), 1e-3, 1e-3, hip.SH_F64, fused=fused)
    ns.set_scheme(1)
    ns.set_initial(V)
    hip.lib().sh_ns_set_vorticity(ns.h, hip.ptr(W))
    ns.startup()
    ns.steps(30)
    assert rel_l2(_host(ns.velocity()), ref.velocity()) < TOL_STEP
    ns.close()


def test_vorticity_stats_vs_oracle_and_closed_form(hip):
    import ctypes
    import torch
    n = 32
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V0), n, n, n, 5, 4.0, 1.0, hip.stream())
    e, dn = ctypes.c_double(), ctypes.c_double()
    hip.lib().sh_vorticity_stats_z(hip.ptr(V0), n, n, n, ctypes.byref(e), ctypes.byref(dn), hip.stream())
    er, dr = O.vorticity_stats(_host(V0))
    assert abs(e.value - er) < 1e-12 * er
    assert dn.value < 1e-12 and dr < 1e-12
    # Taylor-Green: enstrophy 3/8
    tg = _dev(O.taylor_green((n, n, n)))
    hip.lib().sh_vorticity_stats_z(hip.ptr(tg), n, n, n, ctypes.byref(e), None, hip.stream())
    assert abs(e.value - 3 / 8) < 1e-14
    # u = (sin x, 0, 0): divergence norm sqrt(1/2), no vorticity
    x = 2 * np.pi * np.arange(n) / n
    u = np.zeros((3, n, n, n), np.complex128)
    u[0] = O.forward(np.broadcast_to(np.sin(x)[None, None, :], (n, n, n)).astype(np.complex128))
    hip.lib().sh_vorticity_stats_z(hip.ptr(_dev(u)), n, n, n, ctypes.byref(e), ctypes.byref(dn), hip.stream())
    assert e.value < 1e-28 and abs(dn.value - np.sqrt(0.5)) < 1e-14


def _particles_device(hip, npart, seed):
    import torch
    loc = torch.empty((npart, 3), dtype=torch.float64, device="cuda")
    hip.lib().sh_particles_seed(hip.ptr(loc), npart, seed, 0, hip.stream())
    return loc


def test_particles_seed_bit_exact(hip):
    import torch
    loc = _particles_device(hip, 10000, 42)
    ref = O.particles_seed(10000, 42)
    assert np.array_equal(_host(loc), ref)
    # an offset batch continues the same sequence
    part = torch.empty((10, 3), dtype=torch.float64, device="cuda")
    hip.lib().sh_particles_seed(hip.ptr(part), 10, 42, 777, hip.stream())
    assert np.array_equal(_host(part), ref[777:787])


def test_particles_move_bit_exact_vs_oracle(hip):
    """10000 particles (exec/ns2.cpp:159), 5 moves through a random physical
    field on a non-cubic grid: positions, velocities and energies equal."""
    import torch
    dims = (24, 20, 16)
    nx, ny, nz = dims
    U = O.prng_complex(17, (nx, ny, 3 * nz)).reshape(3, nz, ny, nx) * 3.0
    Ud = _dev(U)
    npart = 10000
    loc_d = _particles_device(hip, npart, 7)
    vel_d = torch.empty((npart, 3), dtype=torch.float64, device="cuda")
    en_d = torch.empty(npart, dtype=torch.float64, device="cuda")
    loc = O.particles_seed(npart, 7)
    # some particles in the top cells and on the boundaries
    loc[:50, 0] = 2 * np.pi * (nx - 1) / nx + 1e-3
    loc[50:60, 1] = 0.0
    loc_d.copy_(torch.from_numpy(loc))
    K = 0.05
    for step in range(5):
        hip.lib().sh_particles_move(hip.ptr(loc_d), hip.ptr(vel_d), hip.ptr(en_d), npart, hip.ptr(Ud[0]),
                                    hip.ptr(Ud[1]), hip.ptr(Ud[2]), nx, ny, nz, K, hip.stream())
        vel, en = O.particles_move(loc, U, K)
        assert np.array_equal(_host(loc_d), loc), step
        assert np.array_equal(_host(vel_d), vel), step
        assert np.array_equal(_host(en_d), en), step


def test_particles_multilinear_exact(hip):
    """Without the oracle: a multilinear field is interpolated exactly."""
    import torch
    rng = np.random.default_rng(5)
    nx, ny, nz = 8, 6, 5
    x = 2 * np.pi * np.arange(nx) / nx
    y = 2 * np.pi * np.arange(ny) / ny
    z = 2 * np.pi * np.arange(nz) / nz
    X, Y, Z = x[None, None, :], y[None, :, None], z[:, None, None]
    c = rng.uniform(-1, 1, 8)
    P = lambda X, Y, Z: c[0] + c[1] * X + c[2] * Y + c[3] * Z + c[4] * X * Y + c[5] * X * Z + c[6] * Y * Z + c[7] * X * Y * Z
    U = np.zeros((3, nz, ny, nx), np.complex128)
    U[2].real = np.broadcast_to(P(X, Y, Z), (nz, ny, nx))
    U[2].imag = 9.0
    npart = 2048
    loc_d = _particles_device(hip, npart, 11)
    start = _host(loc_d).copy()
    vel_d = torch.empty((npart, 3), dtype=torch.float64, device="cuda")
    Ud = _dev(U)
    hip.lib().sh_particles_move(hip.ptr(loc_d), hip.ptr(vel_d), None, npart, hip.ptr(Ud[0]), hip.ptr(Ud[1]),
                                hip.ptr(Ud[2]), nx, ny, nz, 0.0, hip.stream())
    v = _host(vel_d)
    exact = P(start[:, 0], start[:, 1], start[:, 2])
    assert np.abs(v[:, 2] - exact).max() < 1e-11 * np.abs(exact).max()
    assert np.abs(v[:, :2]).max() == 0


def test_particles_reject_bad_arguments(hip):
    import torch
    loc = _particles_device(hip, 4, 1)
    vel = torch.empty((4, 3), dtype=torch.float64, device="cuda")
    U = torch.zeros((3, 4, 4, 1), dtype=torch.complex128, device="cuda")
    with pytest.raises(RuntimeError):
        hip.lib().sh_particles_move(hip.ptr(loc), hip.ptr(vel), None, 4, hip.ptr(U[0]), hip.ptr(U[1]),
                                    hip.ptr(U[2]), 1, 4, 4, 0.1, hip.stream())
    with pytest.raises(RuntimeError):
        hip.lib().sh_particles_move(hip.ptr(loc), hip.ptr(vel), None, -1, hip.ptr(U[0]), hip.ptr(U[1]),
                                    hip.ptr(U[2]), 4, 4, 4, 0.1, hip.stream())
