"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake (wrong sign or index,
transposed axis, dropped term, wrong AB2/CN coefficient) fails one of them:
  * golden vectors written by an independent numpy restatement
    (tests/golden/make_golden.py, numpy's pocketfft);
  * brute-force DFT sums in extended precision for tiny sizes;
  * the reference's own known answers (test/operators.cpp, test/freq.cpp,
    test/cn_heat.cpp) at 1e-12 instead of abs 1e-3;
  * closed forms of the step (Stokes limit, Beltrami flow), temporal
    convergence order of AB2/CN, exact projection (div V = 0).
"""
import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from oracle import pyoracle as O
from tests.golden import make_golden as G


def test_prng_matches_numpy_restatement():
    for seed in (0, 1, 20240611, 2 ** 63 + 5):
        assert np.array_equal(O.prng_complex(seed, (8, 4, 6)), G.prng_complex(seed, (8, 4, 6)))
    u = O.prng_uniform(9, 100, 1000)
    assert np.array_equal(u, G.uniform01(9, 1100)[100:])
    assert u.min() >= 0 and u.max() < 1


def _dft_matrix(n, sign):
    k = np.arange(n, dtype=np.longdouble)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    ang = sign * 2 * pi * np.outer(k, k) / n
    return np.cos(ang) + 1j * np.sin(ang)


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_brute_force(n, sign):
    x = O.prng_complex(n, (n, 1, 1)).ravel()
    M = _dft_matrix(n, sign)
    ref = (M.real.astype(np.float64) + 1j * M.imag.astype(np.float64)) @ x
    assert rel_l2(O.fft1d(x, sign), ref) < 1e-15


def test_fft3d_brute_force_small():
    """3D transform = DFT along x, then y, then z; non-cubic so an axis mix-up fails."""
    dims = (4, 6, 8)
    x = O.prng_complex(42, dims)
    nx, ny, nz = dims
    y = np.einsum("ai,kji->kja", _dft_matrix(nx, -1).astype(np.clongdouble), x.astype(np.clongdouble))
    y = np.einsum("bj,kja->kba", _dft_matrix(ny, -1).astype(np.clongdouble), y)
    y = np.einsum("ck,kba->cba", _dft_matrix(nz, -1).astype(np.clongdouble), y)
    assert rel_l2(O.transform3d(x, -1), y.astype(np.complex128)) < 1e-15


@pytest.mark.parametrize("n", [8, 12, 16, 24, 32, 48, 64, 96, 128, 256, 512])
def test_fft1d_vs_numpy(n):
    x = O.prng_complex(n + 1, (n, 1, 1)).ravel()
    assert rel_l2(O.fft1d(x, -1), np.fft.fft(x)) < 1e-14
    assert rel_l2(O.fft1d(x, 1), np.fft.ifft(x) * n) < 1e-14


@pytest.mark.parametrize("name,dims,seed", [("fft_16", (16, 16, 16), 20240611), ("fft_24x8x12", (24, 8, 12), 777)])
def test_fft3d_golden(name, dims, seed):
    g = np.load("%s/%s.npz" % (GOLDEN, name))
    x = O.prng_complex(seed, dims)
    assert rel_l2(O.transform3d(x, -1), g["fwd"]) < 1e-14
    assert rel_l2(O.transform3d(x, +1) / x.size, g["inv"]) < 1e-14


def test_fft3d_invariants():
    dims = (32, 16, 24)
    x = O.prng_complex(3, dims)
    y = O.transform3d(x, -1)
    # Parseval and round trip
    assert abs(np.sum(np.abs(y) ** 2) / x.size / np.sum(np.abs(x) ** 2) - 1) < 1e-14
    assert rel_l2(O.transform3d(y, 1) / x.size, x) < 1e-15
    # linearity
    z = O.prng_complex(4, dims)
    assert rel_l2(O.transform3d(2 * x - 3j * z, -1), 2 * y - 3j * O.transform3d(z, -1)) < 1e-14
    # an impulse at the origin transforms to all ones
    d = np.zeros((24, 16, 32), np.complex128)
    d[0, 0, 0] = 1
    assert np.abs(O.transform3d(d, -1) - 1).max() < 1e-15
    # a single mode lands on its own index: exp(+2 pi i (3x/32 + 2y/16 + 5z/24)) -> N^3 at (5,2,3)
    i = np.arange(32)[None, None, :]
    j = np.arange(16)[None, :, None]
    k = np.arange(24)[:, None, None]
    m = np.exp(2j * np.pi * (3 * i / 32 + 2 * j / 16 + 5 * k / 24))
    f = O.transform3d(m, -1)
    assert abs(f[5, 2, 3] - m.size) < 1e-9 and np.sum(np.abs(f) > 1e-8) == 1


def test_freq_map():
    """test/freq.cpp: N = 16 prints 0..8 then -7..-1 (Nyquist maps to +N/2)."""
    assert [O.freq(i, 16) for i in range(16)] == list(range(9)) + list(range(-7, 0))
    assert [O.freq(i, 12) for i in range(12)] == list(range(7)) + list(range(-5, 0))


def test_zero_high_planes():
    dims = (8, 6, 4)
    x = O.prng_complex(1, dims)
    y = O.zero_high(x)
    expect = x.copy()
    expect[:, :, 4] = 0
    expect[:, 3, :] = 0
    expect[2] = 0
    assert np.array_equal(y, expect)


def test_ops_golden():
    g = np.load(GOLDEN + "/ops_16.npz")
    dims = (16, 16, 16)
    s = O.forward(O.prng_complex(11, dims))
    v = np.stack([O.forward(O.prng_complex(12 + d, dims)) for d in range(3)])
    w = np.stack([O.forward(O.prng_complex(20 + d, dims)) for d in range(3)])
    got = {"grad": O.gradient(s), "lap": O.laplacian(s), "lapinv": O.laplacian_inverse(s),
           "lapinv_cn": O.laplacian_inverse(s, 1.0, -0.3), "div": O.divergence(v), "curl": O.curl(v),
           "vlap": O.vlaplacian(v), "cross": O.cross(v, w)}
    for k, val in got.items():
        assert rel_l2(val, g[k]) < 1e-14, k


def test_operators_known_answers():
    """test/operators.cpp:20-198 at 1e-12 (the reference checks abs 1e-3)."""
    N = 32
    dims = (N, N, N)
    x = 2 * np.pi * np.arange(N) / N
    s = np.broadcast_to(np.sin(x)[None, None, :], (N, N, N)).astype(np.complex128)
    c = np.broadcast_to(np.cos(x)[None, None, :], (N, N, N))
    sf = O.forward(s)
    g = np.stack([O.inverse(a) for a in O.gradient(sf)])
    assert rel_l2(g[0].real, c) < 1e-13 and np.abs(g[1]).max() < 1e-15 and np.abs(g[2]).max() < 1e-15
    lap = O.laplacian(sf)
    assert rel_l2(O.inverse(lap).real, -s.real) < 1e-13
    assert rel_l2(O.inverse(O.laplacian_inverse(lap)).real, s.real) < 1e-13
    v = np.stack([sf, np.zeros_like(sf), np.zeros_like(sf)])
    assert rel_l2(O.inverse(O.divergence(v)).real, c) < 1e-13
    assert rel_l2(O.inverse(O.vlaplacian(v)[0]).real, -s.real) < 1e-13
    one, zero = np.ones((N, N, N), np.complex128), np.zeros((N, N, N), np.complex128)
    a = np.stack([O.forward(zero), O.forward(zero), O.forward(one)])
    b = np.stack([O.forward(zero), O.forward(one), O.forward(zero)])
    cr = np.stack([O.inverse(q) for q in O.cross(a, b)])
    assert np.abs(cr[0] + 1).max() < 1e-15 and np.abs(cr[1]).max() < 1e-15 and np.abs(cr[2]).max() < 1e-15
    grad = O.gradient(sf)
    assert np.abs(O.curl(grad)).max() < 1e-12


def test_vector_identities():
    dims = (16, 12, 8)
    v = np.stack([O.forward(O.prng_complex(50 + d, dims)) for d in range(3)])
    s = O.forward(O.prng_complex(60, dims))
    # div curl = 0, curl grad = 0, div grad = lap, lap^-1 lap = id off k=0
    assert np.abs(O.divergence(O.curl(v))).max() < 1e-12
    assert np.abs(O.curl(O.gradient(s))).max() < 1e-12
    assert rel_l2(O.divergence(O.gradient(s)), O.laplacian(s)) < 1e-15
    li = O.laplacian_inverse(O.laplacian(s))
    mask = np.ones(s.shape, bool)
    mask[0, 0, 0] = False
    assert rel_l2(li[mask], s[mask]) < 1e-15
    # laplacian_inverse leaves the k = 0 mode untouched (src/Scalar.cpp:327)
    assert O.laplacian_inverse(s)[0, 0, 0] == s[0, 0, 0]


def test_cn_heat_factor():
    """test/cn_heat.cpp:43-53: u = u0 + K/2 lap u0; u.laplacian_inverse(1, -K/2)
    is the per-mode CN factor (1 - K/2 k^2)/(1 + K/2 k^2)."""
    dims = (16, 16, 16)
    u0 = O.forward(O.prng_complex(70, dims))
    K = 5.0 / 2000
    u = u0 + K / 2 * O.laplacian(u0)
    u = O.laplacian_inverse(u, 1.0, -K / 2)
    kx = np.array([O.freq(i, 16) for i in range(16)])
    k2 = kx[None, None, :] ** 2 + kx[None, :, None] ** 2 + kx[:, None, None] ** 2
    assert rel_l2(u, u0 * (1 - K / 2 * k2) / (1 + K / 2 * k2)) < 1e-15


def test_tg16_golden():
    g = np.load(GOLDEN + "/tg_16.npz")
    dims = (16, 16, 16)
    ns = O.NS(dims, 0.01, 1e-3)
    ns.set_initial(O.taylor_green(dims))
    ns.startup()
    assert rel_l2(ns.velocity(), g["startup"]) < 1e-14
    done = 0
    for t in (1, 10, 100):
        ns.steps(t - done)
        done = t
        assert rel_l2(ns.velocity(), g["step%d" % t]) < 1e-12, t
    assert ns.div_average() < 1e-13


def test_tg32_golden_norms():
    g = np.load(GOLDEN + "/tg_32.npz")
    dims = (32, 32, 32)
    ns = O.NS(dims, 0.01, 1e-3)
    ns.set_initial(O.taylor_green(dims))
    ns.startup()
    ns.steps(100)
    V = ns.velocity()
    assert rel_l2(np.linalg.norm(V.reshape(3, -1), axis=1), g["step100_norm"]) < 1e-13
    assert np.linalg.norm(V.reshape(3, -1)[:, ::97] - g["step100_sub"]) / np.linalg.norm(V) < 1e-12


def test_solenoidal_ic():
    g = np.load(GOLDEN + "/solenoidal_16.npz")
    V = O.solenoidal((16, 16, 16), 7, 4.0, 1.0)
    assert rel_l2(V, g["V"]) < 1e-14
    # divergence-free, Hermitian (real in physical space), u_rms = 1
    assert np.abs(O.divergence(V)).max() < 1e-15 * 16 * np.abs(V).max()
    u = np.stack([O.inverse(c) for c in V])
    assert np.abs(u.imag).max() < 1e-15
    assert abs(np.sqrt(np.mean(np.sum(u.real ** 2, axis=0)) / 3) - 1) < 1e-13
    assert abs(O.energy(V) - 1.5) < 1e-13


def test_stokes_limit_closed_form():
    dims = (16, 16, 16)
    nu, K, n = 0.05, 1e-2, 30
    V0 = O.solenoidal(dims, 3, 3.0, 1.0)
    ns = O.NS(dims, nu, K)
    ns.set_nonlinear(False)
    ns.set_initial(V0)
    ns.startup()
    ns.steps(n)
    kx = np.array([O.freq(i, 16) for i in range(16)])
    k2 = kx[None, None, :] ** 2 + kx[None, :, None] ** 2 + kx[:, None, None] ** 2
    c = nu * K / 2
    expect = V0 * (1 - nu * K * k2) * ((1 - c * k2) / (1 + c * k2)) ** n
    assert rel_l2(ns.velocity(), expect) < 1e-13


def test_beltrami_closed_form():
    N = 16
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    A, B, C = 1.0, 0.7, 0.4
    u = [A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z), C * np.sin(Y) + B * np.cos(X)]
    V0 = np.stack([O.forward(np.broadcast_to(c, (N, N, N)).astype(np.complex128)) for c in u])
    nu, K, n = 0.02, 1e-2, 50
    ns = O.NS((N, N, N), nu, K)
    ns.set_initial(V0)
    ns.startup()
    ns.steps(n)
    expect = V0 * (1 - nu * K) * ((1 - nu * K / 2) / (1 + nu * K / 2)) ** n
    assert rel_l2(ns.velocity(), expect) < 1e-12


def test_second_order_in_time():
    """AB2 + CN with a Forward-Euler start is second order: halving dt cuts the
    error at a fixed time by ~4 (a first-order slip gives ~2)."""
    dims = (16, 16, 16)
    V0 = O.solenoidal(dims, 11, 2.0, 1.0)
    T = 0.2

    def run(K):
        ns = O.NS(dims, 0.02, K)
        ns.set_initial(V0)
        ns.startup()
        ns.steps(int(round(T / K)) - 1)
        return ns.velocity()

    v1, v2, v4, v8 = run(0.02), run(0.01), run(0.005), run(0.0025)
    r1 = np.linalg.norm(v1 - v2) / np.linalg.norm(v2 - v4)
    r2 = np.linalg.norm(v2 - v4) / np.linalg.norm(v4 - v8)
    assert 3.3 < r2 < 4.7, (r1, r2)
