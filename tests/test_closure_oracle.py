"""CPU pins of the closure oracle (oracle/closure.py): the backward pass
against finite differences of the loss in exact mode, bf16 rounding known
answers, spectral gradients and the stress divergence against closed forms,
the packed parameter layout, and the box filter against direct sums."""
import numpy as np
import pytest

from oracle import closure as C


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, C.NIN))
    T = rng.normal(size=(n, C.NOUT)) * 0.3
    P = C.init_params(seed).astype(np.float64)
    P[C.OFF_B1:C.OFF_W2] = rng.normal(size=64) * 0.1
    P[C.OFF_B2:C.OFF_W3] = rng.normal(size=64) * 0.1
    P[C.OFF_B3:C.OFF_B3 + 6] = rng.normal(size=6) * 0.1
    return P, X, T


def _valid_indices():
    mask = C.pack(*[np.ones_like(a) for a in C.unpack(np.zeros(C.NPARAM))])
    return np.nonzero(mask)[0]


def test_param_layout():
    assert C.NPARAM == 16 * 64 + 64 + 64 * 64 + 64 + 64 * 16 + 16
    valid = _valid_indices()
    assert len(valid) == 9 * 64 + 64 + 64 * 64 + 64 + 64 * 6 + 6
    P = np.arange(C.NPARAM, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = C.unpack(P)
    assert W1[2, 5] == 2 * 64 + 5 and W3[10, 4] == C.OFF_W3 + 10 * 16 + 4 and b3[5] == C.OFF_B3 + 5
    Q = C.pack(W1, b1, W2, b2, W3, b3)
    assert np.array_equal(Q[valid], P[valid])


def test_backward_matches_finite_differences():
    """Exact mode: every bias and a sample of weights in each layer against
    central differences of the loss (a transposed operand, a dropped
    (1 - h^2) factor or a wrong scale each fail this)."""
    P, X, T = _data(24, 3)
    sc_in = np.linspace(0.5, 1.5, 9)
    sc_out = np.linspace(0.8, 1.3, 6)
    _, g = C.loss_grad(P, X, T, sc_in, sc_out, bf=False)
    rng = np.random.default_rng(0)
    idx = list(range(C.OFF_B1, C.OFF_W2)) + list(range(C.OFF_B2, C.OFF_W3)) + list(range(C.OFF_B3, C.OFF_B3 + 6))
    for lo, hi, rows, cols, vr, vc in ((C.OFF_W1, C.OFF_B1, 16, 64, 9, 64), (C.OFF_W2, C.OFF_B2, 64, 64, 64, 64),
                                       (C.OFF_W3, C.OFF_B3, 64, 16, 64, 6)):
        for _ in range(25):
            idx.append(lo + rng.integers(vr) * cols + rng.integers(vc))
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (C.loss_grad(Pp, X, T, sc_in, sc_out, bf=False)[0] - C.loss_grad(Pm, X, T, sc_in, sc_out, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    # padding entries carry no gradient
    pad = np.setdiff1d(np.arange(C.NPARAM), _valid_indices())
    assert np.all(g[pad] == 0)


def test_loss_is_mean_square_error():
    P, X, T = _data(10, 5)
    y = C.forward(P, X, bf=False)
    loss, _ = C.loss_grad(P, X, T, bf=False)
    assert loss == pytest.approx(np.mean((y - T) ** 2), rel=1e-14)
    # zero weights: the output is the bias
    Z = np.zeros(C.NPARAM)
    Z[C.OFF_B3:C.OFF_B3 + 6] = np.arange(6)
    assert np.array_equal(C.forward(Z, X, bf=False), np.tile(np.arange(6.0), (10, 1)))


def test_bf16_rounding_known_answers():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 3.140625, 1e-3], dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5, 3.140625, 0.00099945068359375], dtype=np.float32)
    assert np.array_equal(C.bf16(vals), want)
    # rounding is idempotent and keeps 8 significant bits
    r = C.bf16(np.random.default_rng(1).normal(size=1000).astype(np.float32))
    assert np.array_equal(C.bf16(r), r)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_bf16_mode_close_to_exact():
    P, X, T = _data(64, 7)
    l0, g0 = C.loss_grad(P, X, T, bf=False)
    l1, g1 = C.loss_grad(P, X, T, bf=True)
    assert abs(l1 - l0) / l0 < 2e-2
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 5e-2


def test_adam_first_step_is_signed_lr():
    P, X, T = _data(16, 9)
    _, g = C.loss_grad(P, X, T, bf=False)
    m = np.zeros(C.NPARAM)
    v = np.zeros(C.NPARAM)
    P1, m1, v1 = C.adam(P, m, v, g, 1, 1e-3)
    valid = _valid_indices()
    big = valid[np.abs(g[valid]) > 1e-4]
    # bias-corrected first step moves each weight by lr against its gradient's sign
    assert np.allclose(P1[big] - P[big], -1e-3 * np.sign(g[big]), rtol=1e-3)


def test_init_params_glorot_and_reproducible():
    P = C.init_params(11)
    assert np.array_equal(P, C.init_params(11)) and not np.array_equal(P, C.init_params(12))
    W1, b1, W2, b2, W3, b3 = C.unpack(P)
    assert np.abs(W1).max() <= np.sqrt(6 / 73) and np.abs(W2).max() <= np.sqrt(6 / 128)
    assert np.abs(W3).max() <= np.sqrt(6 / 70)
    assert not b1.any() and not b2.any() and not b3.any()
    pad = np.setdiff1d(np.arange(C.NPARAM), _valid_indices())
    assert not P[pad].any()


def _grid(n):
    x = 2 * np.pi * np.arange(n) / n
    return np.meshgrid(x, x, x, indexing="ij")  # z, y, x


def test_gradient_features_closed_form():
    n = 16
    z, y, x = _grid(n)
    u = np.stack([np.sin(2 * y) * np.cos(z), np.cos(3 * x), np.sin(x + y)])
    g = C.gradient_features(u)
    want = np.zeros((n ** 3, 9))
    want[:, 1] = (2 * np.cos(2 * y) * np.cos(z)).reshape(-1)   # d ux / dy
    want[:, 2] = (-np.sin(2 * y) * np.sin(z)).reshape(-1)      # d ux / dz
    want[:, 3] = (-3 * np.sin(3 * x)).reshape(-1)              # d uy / dx
    want[:, 6] = np.cos(x + y).reshape(-1)                     # d uz / dx
    want[:, 7] = np.cos(x + y).reshape(-1)                     # d uz / dy
    assert np.abs(g - want).max() < 1e-12
    # the Nyquist mode carries no derivative
    un = np.stack([np.cos(8 * x), 0 * x, 0 * x])
    assert np.abs(C.gradient_features(un)).max() < 1e-12


def test_sgs_force_closed_form():
    """tau_xx = cos(2x), tau_xy = sin(y): f_x = -d tau_xx/dx - d tau_xy/dy
    = 2 sin(2x) - cos(y); f_y = -d tau_yx/dx = 0."""
    n = 16
    z, y, x = _grid(n)
    tau = np.zeros((6, n, n, n))
    tau[3] = np.sin(y)
    tau[0] = np.cos(2 * x)
    f = C.sgs_force(tau)
    fx = np.fft.ifftn(f[0]).real
    assert np.abs(fx - (2 * np.sin(2 * x) - np.cos(y))).max() < 1e-12
    assert np.abs(np.fft.ifftn(f[1])).max() < 1e-12 and np.abs(np.fft.ifftn(f[2])).max() < 1e-12


def test_clip_zeroes_backscatter_only():
    g = np.zeros((2, 9))
    g[:, 1] = g[:, 3] = 1.0   # S_xy = 1
    tau = np.zeros((2, 6))
    tau[0, 3] = -0.5          # tau:S < 0: forward transfer, kept
    tau[1, 3] = 0.5           # tau:S > 0: backscatter, zeroed
    out = C.clip(tau, g)
    assert out[0, 3] == -0.5 and not out[1].any()
    # an isotropic stress does no work on a traceless strain and is kept
    g2 = np.zeros((1, 9))
    g2[0, 0], g2[0, 4] = 1.0, -1.0
    iso = np.array([[2.0, 2.0, 2.0, 0, 0, 0]])
    assert np.array_equal(C.clip(iso, g2), iso)


def _force_case(n, seed):
    rng = np.random.default_rng(seed)
    P = C.init_params(seed).astype(np.float64)
    P[C.OFF_B3:C.OFF_B3 + 6] = rng.normal(size=6) * 0.1
    X = rng.normal(size=(n ** 3, 9))
    ft = np.stack([np.fft.fftn(rng.normal(size=(n, n, n))) for _ in range(3)]) * 0.05
    return P, X, ft


@pytest.mark.parametrize("dealias,weight,kw", [(0, 1.0, 0.0), (1, 1.0, 0.0), (1, 3.0, 2.5), (2, 1.0, 0.0)])
def test_force_loss_backward_matches_finite_differences(dealias, weight, kw):
    """The adjoint of f = -div tau (symmetric stress, off-diagonals counted in
    both force components), the projection/mask and the spectral weight
    against central differences of the force-matching loss, exact mode."""
    n = 8
    P, X, ft = _force_case(n, 13)
    sc_in = np.linspace(0.6, 1.2, 9)
    sc_out = np.linspace(0.05, 0.1, 6)

    def lossg(Pv):
        return C.force_loss_grad(Pv, X, ft, sc_in, sc_out, dealias=dealias, bf=False, weight=weight, kw=kw)

    _, g = lossg(P)
    rng = np.random.default_rng(1)
    idx = list(range(C.OFF_B3, C.OFF_B3 + 6)) + [C.OFF_B1 + i for i in rng.integers(0, 64, 6)]
    for lo, cols, vr, vc in ((C.OFF_W1, 64, 9, 64), (C.OFF_W2, 64, 64, 64), (C.OFF_W3, 16, 64, 6)):
        idx += [lo + rng.integers(vr) * cols + rng.integers(vc) for _ in range(8)]
    h = 1e-5
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (lossg(Pp)[0] - lossg(Pm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-9 + 1e-5 * abs(fd), (i, fd, g[i])


def test_expfilter_closed_form():
    """exp(-36 (|k|/(n/2))^36): 1 at k = 0, e^-36 at |k| = n/2 on one axis,
    and a product over axes."""
    n = 16
    s = C.expfilter(n)
    assert s[0, 0, 0] == 1.0
    assert np.isclose(s[0, 0, 7], np.exp(-36 * (7 / 8) ** 36))
    assert np.isclose(s[3, 5, 0], np.exp(-36 * (3 / 8) ** 36) * np.exp(-36 * (5 / 8) ** 36))
    assert s.max() <= 1.0 and s.min() > 0


def test_force_loss_zero_at_own_force():
    """With the target equal to the network's own force the loss and the
    gradient vanish; a pure gradient target is invisible to the loss."""
    n = 8
    P, X, _ = _force_case(n, 17)
    tau = C.forward(P, X, bf=False)
    f = C.sgs_force(tau.T.reshape(6, n, n, n))
    loss, g = C.force_loss_grad(P, X, f, dealias=False, bf=False)
    assert loss < 1e-28 and np.abs(g).max() < 1e-12
    k = C._k(n)
    phi = np.fft.fftn(np.random.default_rng(2).normal(size=(n, n, n)))
    grad_phi = np.stack([1j * k[None, None, :] * phi, 1j * k[None, :, None] * phi, 1j * k[:, None, None] * phi])
    loss2, _ = C.force_loss_grad(P, X, f + grad_phi, dealias=False, bf=False)
    assert loss2 < 1e-20


def test_box_filter_direct_sums():
    rng = np.random.default_rng(4)
    u = rng.normal(size=(3, 8, 8, 8))
    out = C.box_filter(u, 4)
    # cell (cz, cy, cx) = (1, 0, 1)
    blk = u[:, 4:8, 0:4, 4:8].reshape(3, -1)
    c = 1 * 4 + 0 * 2 + 1
    assert np.allclose(out[:3, c], blk.mean(1), rtol=1e-14)
    assert np.isclose(out[3 + 3, c], (blk[0] * blk[1]).mean() - blk[0].mean() * blk[1].mean(), rtol=1e-12)
    # the stress trace is a sum of variances
    assert np.all(out[3] + out[4] + out[5] >= 0)


def test_smagorinsky_closed_forms():
    """The mixed model's eddy-viscosity term: a pure shear g_xy = gamma gives
    S_xy = gamma/2, |S| = |gamma|, so tau_xy = -c |gamma| gamma and nothing
    else; for trace-free gradients the added stress is trace-free and its
    work on the strain is tau:S = -c |S|^3 (S:S = |S|^2 / 2)."""
    rng = np.random.default_rng(3)
    c = 0.013
    gam = rng.normal(size=5)
    g = np.zeros((5, 9))
    g[:, 1] = gam
    t = C.smagorinsky(np.zeros((5, 6)), g, c)
    assert np.allclose(t[:, 3], -c * np.abs(gam) * gam, rtol=1e-14, atol=0)
    assert not t[:, [0, 1, 2, 4, 5]].any()
    g = rng.normal(size=(200, 9))
    g[:, 8] = -(g[:, 0] + g[:, 4])
    base = rng.normal(size=(200, 6))
    t = C.smagorinsky(base, g, c) - base
    assert np.abs(t[:, :3].sum(1)).max() < 1e-13
    s = np.stack([g[:, 0], g[:, 4], g[:, 8], 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]),
                  0.5 * (g[:, 5] + g[:, 7])], 1)
    work = (t[:, :3] * s[:, :3]).sum(1) + 2 * (t[:, 3:] * s[:, 3:]).sum(1)
    mag = np.sqrt(2 * ((s[:, :3] ** 2).sum(1) + 2 * (s[:, 3:] ** 2).sum(1)))
    assert np.allclose(work, -c * mag ** 3, rtol=1e-12)
