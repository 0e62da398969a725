"""Pins the oracle's exec/ns2.cpp pieces before they are trusted as the checker:
the backward-Euler viscous step (exec/ns2.cpp:171-189), the Gaussian-vortex
initial conditions (exec/ns.cpp:84-131, exec/ns2.cpp:84-130), the vorticity
diagnostics, and the marker particles (src/Particle.cpp:5-55).

Each check is chosen so that a plausible mistake fails one of them:
  * closed forms of the step (Stokes limit, Beltrami flow) whose per-step
    factor differs between Crank-Nicolson and backward Euler;
  * first-order temporal convergence of the backward-Euler step;
  * an independent numpy (pocketfft) construction of the vortex IC;
  * the invariant that omega0's mean mode only adds a gradient (absorbed by
    the projection), so the velocity does not depend on it;
  * enstrophy and divergence closed forms (Taylor-Green, ABC, sin x);
  * trilinear interpolation reproduces multilinear fields exactly, including
    the reference's extrapolating clamp in the top cell (a periodic wrap of
    the corner index would fail), and ignores imaginary parts;
  * the periodic wrap cases of src/Particle.cpp:38-45.
"""
import numpy as np

from conftest import rel_l2
from oracle import pyoracle as O


def _k2(n):
    kx = np.array([O.freq(i, n) for i in range(n)])
    return kx[None, None, :] ** 2 + kx[None, :, None] ** 2 + kx[:, None, None] ** 2


def test_backward_euler_stokes_closed_form():
    """Nonlinear term off: Forward-Euler start-up (1 - nu K k^2), then each
    backward-Euler step divides by (1 + nu K k^2)."""
    dims = (16, 16, 16)
    nu, K, n = 0.05, 1e-2, 30
    V0 = O.solenoidal(dims, 3, 3.0, 1.0)
    ns = O.NS(dims, nu, K)
    ns.set_scheme(1)
    ns.set_nonlinear(False)
    ns.set_initial(V0)
    ns.startup()
    ns.steps(n)
    k2 = _k2(16)
    expect = V0 * (1 - nu * K * k2) / (1 + nu * K * k2) ** n
    assert rel_l2(ns.velocity(), expect) < 1e-13
    # and the Crank-Nicolson factor is far from it
    cn = V0 * (1 - nu * K * k2) * ((1 - nu * K / 2 * k2) / (1 + nu * K / 2 * k2)) ** n
    assert rel_l2(ns.velocity(), cn) > 1e-4


def _abc(N, A=1.0, B=0.7, C=0.4):
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    u = [A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z), C * np.sin(Y) + B * np.cos(X)]
    return np.stack([O.forward(np.broadcast_to(c, (N, N, N)).astype(np.complex128)) for c in u])


def test_backward_euler_beltrami_closed_form():
    """ABC flow (curl u = u, |k| = 1): u x omega = 0 and the full nonlinear
    backward-Euler step is the factor 1/(1 + nu K)."""
    N = 16
    V0 = _abc(N)
    nu, K, n = 0.02, 1e-2, 50
    ns = O.NS((N, N, N), nu, K)
    ns.set_scheme(1)
    ns.set_initial(V0)
    ns.startup()
    ns.steps(n)
    expect = V0 * (1 - nu * K) / (1 + nu * K) ** n
    assert rel_l2(ns.velocity(), expect) < 1e-12


def test_backward_euler_first_order_in_time():
    """Backward Euler makes the step first order: halving dt halves the error
    at a fixed time (the Crank-Nicolson step quarters it)."""
    dims = (16, 16, 16)
    V0 = O.solenoidal(dims, 11, 2.0, 1.0)
    T = 0.2

    def run(K):
        ns = O.NS(dims, 0.05, K)
        ns.set_scheme(1)
        ns.set_initial(V0)
        ns.startup()
        ns.steps(int(round(T / K)) - 1)
        return ns.velocity()

    v1, v2, v4 = run(0.01), run(0.005), run(0.0025)
    r = np.linalg.norm(v1 - v2) / np.linalg.norm(v2 - v4)
    assert 1.7 < r < 2.3, r


def _vortex_numpy(n, which):
    x = 2 * np.pi * np.arange(n) / n
    X, Y = x[None, None, :], x[None, :, None]
    w = np.pi / 8
    g = lambda xc, yc: np.exp(-(X - xc) ** 2 / w) * np.exp(-(Y - yc) ** 2 / w)
    d = np.pi / 2
    if which == 0:
        wz = 50 * (g(np.pi - d, np.pi - d) + g(np.pi + d, np.pi - d) + g(np.pi, np.pi + d))
    else:
        wz = 50 * (g(np.pi - d, np.pi) + g(np.pi + d, np.pi))
    wz = np.fft.fftn(np.broadcast_to(wz, (n, n, n)).astype(np.complex128))
    h = n // 2
    wz[h, :, :] = 0
    wz[:, h, :] = 0
    wz[:, :, h] = 0
    kx = np.array([O.freq(i, n) for i in range(n)])
    KX, KY = kx[None, None, :], kx[None, :, None]
    q = KX ** 2 + KY ** 2 + 0 * kx[:, None, None]
    psi = np.where(q != 0, -1.0 / np.where(q != 0, q, 1) * wz, 0)
    V = np.stack([-(1j * KY * psi), 1j * KX * psi, np.zeros_like(psi)])
    W = np.stack([np.zeros_like(wz), np.zeros_like(wz), wz])
    return V, W


def test_vortex_ic_matches_numpy_construction():
    n = 16
    for which in (0, 1):
        V, W = O.vortices((n, n, n), which)
        Vn, Wn = _vortex_numpy(n, which)
        assert rel_l2(W, Wn) < 1e-13, which
        assert rel_l2(V, Vn) < 1e-13, which
        # divergence-free; curl(V0) equals omega0 except the mean mode
        assert np.abs(O.divergence(V)).max() < 1e-9
        c = O.curl(V)
        diff = c - W
        assert abs(W[2, 0, 0, 0]) > 1.0
        diff[2, 0, 0, 0] = 0
        assert np.linalg.norm(diff) < 1e-12 * np.linalg.norm(W)
    # the two sets differ
    assert rel_l2(O.vortices((n, n, n), 0)[0], O.vortices((n, n, n), 1)[0]) > 0.1


def test_mean_vorticity_only_adds_a_gradient():
    """u x (mean w_z z) = mean w_z grad psi is a gradient, so the vortex IC's
    velocity does not depend on omega0's mean mode (to rounding), while the
    explicit omega0 is what the reference drivers use."""
    n = 16
    V, W = O.vortices((n, n, n), 1)
    out = []
    for explicit in (False, True):
        ns = O.NS((n, n, n), 1e-3, 1e-3)
        ns.set_scheme(1)
        ns.set_initial(V)
        if explicit:
            ns.set_vorticity(W)
        ns.startup()
        ns.steps(10)
        out.append(ns.velocity())
    assert rel_l2(out[1], out[0]) < 1e-12


def test_vorticity_stats_closed_forms():
    N = 16
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    tg = O.taylor_green((N, N, N))
    e, dn = O.vorticity_stats(tg)
    assert abs(e - 3 / 8) < 1e-14 and dn < 1e-14  # |k|^2 = 3 on every mode, energy 1/8
    A, B, C = 1.0, 0.7, 0.4
    e, dn = O.vorticity_stats(_abc(N, A, B, C))
    assert abs(e - (A * A + B * B + C * C) / 2) < 1e-13 and dn < 1e-14
    u = np.stack([O.forward(np.broadcast_to(np.sin(X), (N, N, N)).astype(np.complex128)),
                  np.zeros((N, N, N), np.complex128), np.zeros((N, N, N), np.complex128)])
    e, dn = O.vorticity_stats(u)
    assert e < 1e-28 and abs(dn - np.sqrt(0.5)) < 1e-14
    # a shear u_x = sin y: |curl u|^2 = cos^2 y -> enstrophy 1/4
    u[0] = O.forward(np.broadcast_to(np.sin(Y), (N, N, N)).astype(np.complex128))
    e, dn = O.vorticity_stats(u)
    assert abs(e - 0.25) < 1e-14 and dn < 1e-14


def test_particles_seed():
    loc = O.particles_seed(1000, 42)
    assert loc.shape == (1000, 3)
    assert loc.min() >= 0 and loc.max() < 2 * np.pi
    assert np.array_equal(loc.ravel(), O.prng_uniform(42, 0, 3000) * 2 * np.pi)


def _multilinear_field(dims, rng):
    """Real parts: P_d(x, y, z) = multilinear polynomial with random
    coefficients on the grid x_i = 2 pi i / n; imaginary parts: noise that
    Particle::move must ignore."""
    nx, ny, nz = dims
    x = 2 * np.pi * np.arange(nx) / nx
    y = 2 * np.pi * np.arange(ny) / ny
    z = 2 * np.pi * np.arange(nz) / nz
    X, Y, Z = x[None, None, :], y[None, :, None], z[:, None, None]
    coefs = rng.uniform(-1, 1, (3, 8))
    polys = []
    U = np.empty((3, nz, ny, nx), np.complex128)
    for d in range(3):
        c = coefs[d]
        P = lambda X, Y, Z, c=c: (c[0] + c[1] * X + c[2] * Y + c[3] * Z + c[4] * X * Y + c[5] * X * Z
                                  + c[6] * Y * Z + c[7] * X * Y * Z)
        polys.append(P)
        U[d].real = np.broadcast_to(P(X, Y, Z), (nz, ny, nx))
        U[d].imag = rng.uniform(-5, 5, (nz, ny, nx))
    return U, polys


def test_particles_interpolate_multilinear_exactly():
    """Trilinear interpolation is exact for multilinear fields, also in the
    top cell where the reference clamps the index and extrapolates (the
    field is not periodic, so wrapping the corner index would be wrong)."""
    rng = np.random.default_rng(3)
    dims = (8, 6, 5)  # distinct extents catch transposed axes
    U, polys = _multilinear_field(dims, rng)
    loc = O.particles_seed(4000, 9)
    # force some particles into the top cell of each axis
    for d, n in enumerate(dims):
        h = 2 * np.pi / n
        loc[d * 100:(d + 1) * 100, d] = 2 * np.pi - h * rng.uniform(0.01, 0.99, 100)
    start = loc.copy()
    K = 1e-3
    vel, en = O.particles_move(loc, U, K)
    for d in range(3):
        exact = polys[d](start[:, 0], start[:, 1], start[:, 2])
        assert np.abs(vel[:, d] - exact).max() < 1e-11 * max(1.0, np.abs(exact).max()), d
    assert np.array_equal(en, 0.5 * (vel[:, 0] * vel[:, 0] + vel[:, 1] * vel[:, 1] + vel[:, 2] * vel[:, 2]))
    moved = start + vel * K
    moved = np.where(moved >= 2 * np.pi, moved - 2 * np.pi, np.where(moved <= 0, 2 * np.pi + moved, moved))
    assert np.array_equal(loc, moved)


def test_particles_periodic_wrap():
    """src/Particle.cpp:38-45: >= 2 pi subtracts 2 pi, <= 0 adds 2 pi."""
    dims = (4, 4, 4)
    U = np.zeros((3, 4, 4, 4), np.complex128)
    U[0].real = 2.0   # uniform +x
    U[1].real = -2.0  # uniform -y
    U[2].real = 0.0
    loc = np.array([[2 * np.pi - 1e-3, 1e-3, 0.0],
                    [1.0, 2e-3, 3.0],
                    [2 * np.pi - 2e-3, 1.0, 2 * np.pi - 1e-9]])
    K = 1e-3
    vel, _ = O.particles_move(loc, U, K)
    assert np.allclose(vel, [[2, -2, 0]] * 3, atol=1e-14)
    assert abs(loc[0, 0] - (2 * np.pi - 1e-3 + 2e-3 - 2 * np.pi)) < 1e-15  # wrapped to ~1e-3
    assert abs(loc[0, 1] - (2 * np.pi + (1e-3 - 2e-3))) < 1e-15           # wrapped to ~2 pi - 1e-3
    assert loc[0, 2] == 2 * np.pi                                            # exactly 0 maps to 2 pi
    assert loc[1, 1] == 2 * np.pi                                            # 2e-3 - 2e-3 = 0 -> 2 pi
    assert abs(loc[2, 0] - 0.0) < 1e-15 or loc[2, 0] == 2 * np.pi
