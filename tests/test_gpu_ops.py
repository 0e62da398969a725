"""HIP spectral operators against the oracle, the golden vectors and the
single-mode closed forms of the reference's test/operators.cpp (there at abs
1e-3; here at 1e-12 relative)."""
import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _rand_fourier(seed, dims):
    return O.forward(O.prng_complex(seed, dims))


@pytest.fixture(scope="module")
def fields():
    dims = (16, 16, 16)
    s = _rand_fourier(11, dims)
    v = np.stack([_rand_fourier(12 + d, dims) for d in range(3)])
    w = np.stack([_rand_fourier(20 + d, dims) for d in range(3)])
    return dims, s, v, w


def _empty_like(a):
    import torch
    return torch.empty(a.shape, dtype=torch.complex128, device="cuda")


def run_ops(hip, dims, s, v, w):
    L, st = hip.lib(), hip.stream()
    nx, ny, nz = dims
    P = hip.ptr
    S, V, W = _dev(s), _dev(v), _dev(w)
    out = {}
    g = _empty_like(v)
    L.sh_gradient_z(P(S), P(g[0]), P(g[1]), P(g[2]), nx, ny, nz, st)
    out["grad"] = g
    lap = _empty_like(s)
    L.sh_laplacian_z(P(S), P(lap), nx, ny, nz, st)
    out["lap"] = lap
    li = S.clone()
    L.sh_laplacian_inverse_z(P(li), 0.0, 1.0, nx, ny, nz, st)
    out["lapinv"] = li
    lc = S.clone()
    L.sh_laplacian_inverse_z(P(lc), 1.0, -0.3, nx, ny, nz, st)
    out["lapinv_cn"] = lc
    dv = _empty_like(s)
    L.sh_divergence_z(P(V[0]), P(V[1]), P(V[2]), P(dv), nx, ny, nz, st)
    out["div"] = dv
    c = _empty_like(v)
    L.sh_curl_z(P(V[0]), P(V[1]), P(V[2]), P(c[0]), P(c[1]), P(c[2]), nx, ny, nz, st)
    out["curl"] = c
    vl = _empty_like(v)
    L.sh_vlaplacian_z(P(V[0]), P(V[1]), P(V[2]), P(vl[0]), P(vl[1]), P(vl[2]), nx, ny, nz, st)
    out["vlap"] = vl
    cr = _empty_like(v)
    work = _empty_like(np.zeros((6,) + s.shape))
    L.sh_cross_z(P(V), P(W), P(cr), P(work), nx, ny, nz, st)
    out["cross"] = cr
    return {k: _host(t) for k, t in out.items()}


def test_ops_golden(hip, fields):
    dims, s, v, w = fields
    g = np.load(GOLDEN + "/ops_16.npz")
    got = run_ops(hip, dims, s, v, w)
    for k in ["grad", "lap", "lapinv", "lapinv_cn", "div", "curl", "vlap", "cross"]:
        assert rel_l2(got[k], g[k]) < TOL, k


@pytest.mark.parametrize("dims", [(32, 32, 32), (24, 8, 12), (64, 32, 16)])
def test_ops_vs_oracle(hip, dims):
    s = _rand_fourier(31, dims)
    v = np.stack([_rand_fourier(32 + d, dims) for d in range(3)])
    w = np.stack([_rand_fourier(40 + d, dims) for d in range(3)])
    got = run_ops(hip, dims, s, v, w)
    ref = {"grad": O.gradient(s), "lap": O.laplacian(s), "lapinv": O.laplacian_inverse(s),
           "lapinv_cn": O.laplacian_inverse(s, 1.0, -0.3), "div": O.divergence(v), "curl": O.curl(v),
           "vlap": O.vlaplacian(v), "cross": O.cross(v, w)}
    for k, r in ref.items():
        assert rel_l2(got[k], r) < TOL, k


def test_operators_known_answers(hip):
    """test/operators.cpp:20-198 restated: s = sin(2 pi i / N) and friends."""
    import torch
    N = 32
    dims = (N, N, N)
    L, st, P = hip.lib(), hip.stream(), hip.ptr
    i = np.arange(N)
    s = np.broadcast_to(np.sin(2 * np.pi * i / N)[None, None, :], (N, N, N)).astype(np.complex128)
    cosx = np.broadcast_to(np.cos(2 * np.pi * i / N)[None, None, :], (N, N, N))
    sf = _dev(s)
    hip.forward(sf)
    g = torch.empty((3, N, N, N), dtype=torch.complex128, device="cuda")
    L.sh_gradient_z(P(sf), P(g[0]), P(g[1]), P(g[2]), N, N, N, st)
    hip.inverse(g)
    gh = _host(g)
    assert rel_l2(gh[0].real, cosx) < TOL
    assert np.abs(gh[1]).max() < 1e-14 and np.abs(gh[2]).max() < 1e-14
    lap = torch.empty_like(sf)
    L.sh_laplacian_z(P(sf), P(lap), N, N, N, st)
    lap_f = lap.clone()
    hip.inverse(lap)
    assert rel_l2(_host(lap).real, -s.real) < TOL
    L.sh_laplacian_inverse_z(P(lap_f), 0.0, 1.0, N, N, N, st)
    hip.inverse(lap_f)
    assert rel_l2(_host(lap_f).real, s.real) < TOL
    # cross of constants: (0,0,1) x (0,1,0) = (-1,0,0)
    one = np.ones((N, N, N), np.complex128)
    zero = np.zeros((N, N, N), np.complex128)
    a = _dev(np.stack([zero, zero, one]))
    b = _dev(np.stack([zero, one, zero]))
    hip.forward(a)
    hip.forward(b)
    c = torch.empty_like(a)
    work = torch.empty((6, N, N, N), dtype=torch.complex128, device="cuda")
    L.sh_cross_z(P(a), P(b), P(c), P(work), N, N, N, st)
    hip.inverse(c)
    ch = _host(c)
    assert np.abs(ch[0] + 1).max() < 1e-14 and np.abs(ch[1]).max() < 1e-14 and np.abs(ch[2]).max() < 1e-14
    # curl(grad s) = 0 and 2*curl (test/operators.cpp:172-198)
    hip.forward(g)
    cu = torch.empty_like(g)
    L.sh_curl_z(P(g[0]), P(g[1]), P(g[2]), P(cu[0]), P(cu[1]), P(cu[2]), N, N, N, st)
    assert float(cu.abs().max()) < 1e-12
    two = torch.empty_like(g)
    L.sh_scale_z(P(two), 2.0, P(g), g.numel(), st)
    assert rel_l2(_host(two), 2.0 * _host(g)) == 0.0


def test_elementwise_and_average(hip):
    import torch
    L, st, P = hip.lib(), hip.stream(), hip.ptr
    rng = np.random.default_rng(4)
    a = rng.standard_normal(1000) + 1j * rng.standard_normal(1000)
    b = rng.standard_normal(1000) + 1j * rng.standard_normal(1000)
    A, B = _dev(a), _dev(b)
    o = torch.empty_like(A)
    L.sh_mul_z(P(o), P(A), P(B), 1000, st)
    assert rel_l2(_host(o), a * b) < 1e-15
    L.sh_add_z(P(o), P(A), P(B), 1000, st)
    assert np.array_equal(_host(o), a + b)
    L.sh_sub_z(P(o), P(A), P(B), 1000, st)
    assert np.array_equal(_host(o), a - b)
    L.sh_scale_z(P(o), 0.37, P(A), 1000, st)
    assert np.array_equal(_host(o), 0.37 * a)
    import ctypes
    r = ctypes.c_double()
    L.sh_average_z(P(A), 1000, ctypes.byref(r), st)
    assert abs(r.value - O.average(a)) < 1e-14 * abs(r.value)
    L.sh_sumsq_z(P(A), 1000, ctypes.byref(r), st)
    assert abs(r.value - np.sum(np.abs(a) ** 2)) < 1e-12 * r.value


def test_zero_high(hip):
    dims = (16, 8, 12)
    x = O.prng_complex(8, dims)
    t = _dev(x)
    hip.lib().sh_zero_high_z(hip.ptr(t), *dims, hip.stream())
    assert np.array_equal(_host(t), O.zero_high(x))


@pytest.mark.parametrize("dims", [(256, 256, 256), (512, 256, 128)])
def test_operator_identities_full_size(hip, dims):
    """At the benchmark sizes, on random fields generated on the device:
    curl(grad f) = 0, div(curl V) = 0, lap f = div(grad f), vlap V is lap per
    component, lap^-1 (lap f) = f off k = 0, and the Crank-Nicolson solve
    (a, b) = (1, -c) inverts 1 - c lap (src/Scalar.cpp:245-334,
    src/Vector.cpp:92-181).  Both sides round their products in different
    orders, so the bars are a few ulps of the magnitudes involved."""
    import torch
    L, st, P = hip.lib(), hip.stream(), hip.ptr
    nx, ny, nz = dims
    kmax2 = (nx / 2) ** 2 + (ny / 2) ** 2 + (nz / 2) ** 2
    f = torch.empty((nz, ny, nx), dtype=torch.complex128, device="cuda")
    V = torch.empty((3, nz, ny, nx), dtype=torch.complex128, device="cuda")
    L.sh_prng_complex_z(P(f), f.numel(), 101, st)
    L.sh_prng_complex_z(P(V), V.numel(), 102, st)
    nrm = lambda t: float(torch.linalg.vector_norm(t))
    g = torch.empty_like(V)
    L.sh_gradient_z(P(f), P(g[0]), P(g[1]), P(g[2]), nx, ny, nz, st)
    c = torch.empty_like(V)
    L.sh_curl_z(P(g[0]), P(g[1]), P(g[2]), P(c[0]), P(c[1]), P(c[2]), nx, ny, nz, st)
    torch.cuda.synchronize()
    assert nrm(c) < 1e-15 * kmax2 * nrm(f)
    L.sh_curl_z(P(V[0]), P(V[1]), P(V[2]), P(c[0]), P(c[1]), P(c[2]), nx, ny, nz, st)
    d = torch.empty_like(f)
    L.sh_divergence_z(P(c[0]), P(c[1]), P(c[2]), P(d), nx, ny, nz, st)
    torch.cuda.synchronize()
    assert nrm(d) < 1e-15 * kmax2 * nrm(V)
    lap = torch.empty_like(f)
    L.sh_laplacian_z(P(f), P(lap), nx, ny, nz, st)
    L.sh_divergence_z(P(g[0]), P(g[1]), P(g[2]), P(d), nx, ny, nz, st)
    torch.cuda.synchronize()
    assert nrm(d - lap) < 1e-15 * nrm(lap)
    vl = torch.empty_like(V)
    L.sh_vlaplacian_z(P(V[0]), P(V[1]), P(V[2]), P(vl[0]), P(vl[1]), P(vl[2]), nx, ny, nz, st)
    for a in range(3):
        L.sh_laplacian_z(P(V[a]), P(lap), nx, ny, nz, st)
        torch.cuda.synchronize()
        assert torch.equal(vl[a], lap), a
    L.sh_laplacian_z(P(f), P(lap), nx, ny, nz, st)
    back = lap.clone()
    L.sh_laplacian_inverse_z(P(back), 0.0, 1.0, nx, ny, nz, st)
    torch.cuda.synchronize()
    assert back[0, 0, 0] == 0  # lap f has no mean and laplacian_inverse skips |k|^2 = 0
    want = f.clone()
    want[0, 0, 0] = 0
    assert nrm(back - want) < 1e-15 * nrm(want)
    cfac = 0.37 / kmax2
    u = f.clone()
    L.sh_laplacian_inverse_z(P(u), 1.0, -cfac, nx, ny, nz, st)
    L.sh_laplacian_z(P(u), P(lap), nx, ny, nz, st)
    torch.cuda.synchronize()
    assert nrm(u - cfac * lap - f) < 1e-15 * nrm(f)
    del f, V, g, c, d, lap, vl, back, want, u
    torch.cuda.empty_cache()
