"""CPU oracle of the learned subgrid-stress closure (TEST INFRASTRUCTURE).

Only tests/ import this module, as the checker of the GPU network and LES
kernels.  The closure has no counterpart in the reference; this restatement
is pinned by what the mathematics fixes (tests/test_closure_oracle.py):
finite differences of the loss in exact (FP64) mode, bf16 rounding known
answers, closed-form spectral gradients, and the Kolmogorov-flow steady state
for the body-force hook.

Network: 9 -> 64 -> 64 -> 6, tanh hidden layers.  Packed parameters (float):
W1[16][64] (rows 9..15 zero), b1[64], W2[64][64], b2[64], W3[64][16]
(columns 6..15 zero), b3[16]; W[i][o] maps input i to output o.

bf16 mode mirrors the GPU kernels' roundings: inputs (after scaling),
weights, hidden activations and the back-propagated deltas are rounded to
bf16 before each matrix product; bias gradients and the (1 - h^2) factors use
the unrounded values.  Products accumulate in float64 here (FP32 on the GPU).
"""
import numpy as np

NIN, NOUT, NH = 9, 6, 64
OFF_W1 = 0
OFF_B1 = 16 * 64
OFF_W2 = OFF_B1 + 64
OFF_B2 = OFF_W2 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 64 * 16
NPARAM = OFF_B3 + 16

# stress components in output order and their index pairs
STRESS = ["xx", "yy", "zz", "xy", "xz", "yz"]
PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def bf16(x):
    """Round float32 values to the nearest bf16 (ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[OFF_W1:OFF_B1].reshape(16, 64)[:NIN]
    b1 = P[OFF_B1:OFF_W2]
    W2 = P[OFF_W2:OFF_B2].reshape(64, 64)
    b2 = P[OFF_B2:OFF_W3]
    W3 = P[OFF_W3:OFF_B3].reshape(64, 16)[:, :NOUT]
    b3 = P[OFF_B3:OFF_B3 + NOUT]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    P = np.zeros(NPARAM, dtype=np.float64)
    P[OFF_W1:OFF_B1].reshape(16, 64)[:NIN] = W1
    P[OFF_B1:OFF_W2] = b1
    P[OFF_W2:OFF_B2] = np.asarray(W2).reshape(-1)
    P[OFF_B2:OFF_W3] = b2
    P[OFF_W3:OFF_B3].reshape(64, 16)[:, :NOUT] = W3
    P[OFF_B3:OFF_B3 + NOUT] = b3
    return P


def _u01_stream(seed):
    m = (1 << 64) - 1
    c = 0
    while True:
        c += 1
        z = (seed + c * 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        z ^= z >> 31
        yield (z >> 11) * 2.0 ** -53


def init_params(seed):
    """Glorot-uniform weights (splitmix64 draws in row-major order of W1, W2,
    W3), zero biases and padding."""
    g = _u01_stream(seed)

    def draw(r, c):
        a = np.sqrt(6.0 / (r + c))
        return np.array([[(2 * next(g) - 1) * a for _ in range(c)] for _ in range(r)])

    W1 = draw(NIN, NH)
    W2 = draw(NH, NH)
    W3 = draw(NH, NOUT)
    return pack(W1, np.zeros(NH), W2, np.zeros(NH), W3, np.zeros(NOUT)).astype(np.float32)


def _r(x, bf):
    return bf16(x).astype(np.float64) if bf else np.asarray(x, dtype=np.float64)


def forward(P, X, sc_in=None, sc_out=None, bf=True):
    """X: [n, 9] -> network output [n, 6] times out_scale."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    sc_in = np.ones(NIN) if sc_in is None else np.asarray(sc_in, dtype=np.float64)
    sc_out = np.ones(NOUT) if sc_out is None else np.asarray(sc_out, dtype=np.float64)
    x = _r(np.asarray(X, dtype=np.float64) * sc_in, bf)
    h1 = np.tanh(x @ _r(W1, bf) + b1)
    h2 = np.tanh(_r(h1, bf) @ _r(W2, bf) + b2)
    y = _r(h2, bf) @ _r(W3, bf) + b3
    return y * sc_out


def _net(P, X, sc_in, bf):
    """Forward pass keeping what the backward pass needs; y is the raw
    network output (before the output scale)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    sc_in = np.ones(NIN) if sc_in is None else np.asarray(sc_in, dtype=np.float64)
    x = _r(np.asarray(X, dtype=np.float64) * sc_in, bf)
    w1, w2, w3 = _r(W1, bf), _r(W2, bf), _r(W3, bf)
    h1 = np.tanh(x @ w1 + b1)
    h1r = _r(h1, bf)
    h2 = np.tanh(h1r @ w2 + b2)
    h2r = _r(h2, bf)
    y = h2r @ w3 + b3
    return {"x": x, "w2": w2, "w3": w3, "h1": h1, "h1r": h1r, "h2": h2, "h2r": h2r, "y": y}


def _backward(c, dy, bf):
    """Packed parameter gradient for dL/dy (raw outputs) dy [n, 6]."""
    dyr = _r(dy, bf)
    dz2 = (dyr @ c["w3"].T) * (1 - c["h2"] * c["h2"])
    dz2r = _r(dz2, bf)
    dz1 = (dz2r @ c["w2"].T) * (1 - c["h1"] * c["h1"])
    dz1r = _r(dz1, bf)
    return pack(c["x"].T @ dz1r, dz1.sum(0), c["h1r"].T @ dz2r, dz2.sum(0), c["h2r"].T @ dyr, dy.sum(0))


def loss_grad(P, X, T, sc_in=None, sc_out=None, bf=True):
    """Loss mean((net(X) - T / out_scale)^2) and its gradient (packed)."""
    sc_out = np.ones(NOUT) if sc_out is None else np.asarray(sc_out, dtype=np.float64)
    n = X.shape[0]
    c = _net(P, X, sc_in, bf)
    e = c["y"] - np.asarray(T, dtype=np.float64) / sc_out
    loss = float(np.mean(e * e))
    return loss, _backward(c, 2.0 * e / (n * NOUT), bf)


def _mask(n, dealias):
    """Modes the force residual keeps: not on a Nyquist plane, not k = 0 and,
    with dealias mode 1, every |k_i| <= n/3."""
    k = np.fft.fftfreq(n, 1.0 / n)
    keep = np.ones((n, n, n), dtype=bool)
    h = n // 2
    keep[h, :, :] = keep[:, h, :] = keep[:, :, h] = False
    keep[0, 0, 0] = False
    if dealias == 1:
        big = 3 * np.abs(k) > n
        keep &= ~big[:, None, None] & ~big[None, :, None] & ~big[None, None, :]
    return keep


def expfilter(n):
    """The coarse solver's exponential dealiasing factor (mode 2),
    prod_i exp(-36 (|k_i| / (n/2))^36), on the (z, y, x) mode grid."""
    s = np.exp(-36.0 * (np.abs(_k(n)) / (0.5 * n)) ** 36)
    return s[:, None, None] * s[None, :, None] * s[None, None, :]


def project(v):
    """Divergence-free part of Fourier fields v [3, n, n, n] (Nyquist
    wavenumbers as zero; k = 0 kept)."""
    n = v.shape[-1]
    k = _k(n)
    ks = (k[None, None, :], k[None, :, None], k[:, None, None])
    k2 = ks[0] ** 2 + ks[1] ** 2 + ks[2] ** 2
    k2 = np.where(k2 == 0, 1.0, k2)
    kv = sum(ks[a] * v[a] for a in range(3)) / k2
    return np.stack([v[a] - ks[a] * kv for a in range(3)])


def force_loss_grad(P, X, ft, sc_in=None, sc_out=None, dealias=True, bf=True, weight=1.0, kw=0.0):
    """Force-matching loss weight * sum_k w(k) |r(k)|^2 / N^2 (N = n^3; with
    w = 1 the mean of |r(x)|^2 over the grid), r = M P (-div tau - ft),
    tau = out_scale * net(X) on an n^3 grid (X: [n^3, 9], ft: Fourier
    [3, n, n, n]), w = 1/(1 + |k|^2/kw^2) (kw = 0: w = 1), and its packed
    parameter gradient through the adjoint of the divergence."""
    sc_out = np.ones(NOUT) if sc_out is None else np.asarray(sc_out, dtype=np.float64)
    n = ft.shape[-1]
    N = n ** 3
    c = _net(P, X, sc_in, bf)
    tau = c["y"] * sc_out
    f = sgs_force(tau.T.reshape(6, n, n, n))
    r_hat = project(f - ft) * _mask(n, dealias)
    w = np.ones((n, n, n))
    if kw > 0:
        k = _k(n)
        w = 1.0 / (1.0 + (k[None, None, :] ** 2 + k[None, :, None] ** 2 + k[:, None, None] ** 2) / kw ** 2)
    if dealias == 2:
        w = w * expfilter(n) ** 2
    loss = float(weight * np.sum(w * np.abs(r_hat) ** 2) / N ** 2)
    s = w * r_hat * weight
    g = gradient_features(np.stack([np.fft.ifftn(x).real for x in s]))  # g[3a+b] = d s_a / d x_b
    dtau = (2.0 / N) * np.stack([g[:, 0], g[:, 4], g[:, 8], g[:, 1] + g[:, 3], g[:, 2] + g[:, 6],
                                 g[:, 5] + g[:, 7]], 1)
    return loss, _backward(c, dtau * sc_out, bf)


def adam(P, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); padding entries stay zero."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    P = P - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    mask = pack(*[np.ones_like(a) for a in unpack(np.zeros(NPARAM))])
    return P * mask, m, v


def clip(tau, g):
    """Zero the stress where its deviatoric part has tau_ij S_ij > 0.
    tau: [n, 6], g: [n, 9] with g[3i+j] = d u_i / d x_j."""
    tr = tau[:, :3].sum(1) / 3
    sxy = 0.5 * (g[:, 1] + g[:, 3])
    sxz = 0.5 * (g[:, 2] + g[:, 6])
    syz = 0.5 * (g[:, 5] + g[:, 7])
    ts = ((tau[:, 0] - tr) * g[:, 0] + (tau[:, 1] - tr) * g[:, 4] + (tau[:, 2] - tr) * g[:, 8]
          + 2 * (tau[:, 3] * sxy + tau[:, 4] * sxz + tau[:, 5] * syz))
    out = tau.copy()
    out[ts > 0] = 0
    return out


def smagorinsky(tau, g, c):
    """tau - 2 c |S| S: the Smagorinsky stress of the strain rate
    S = (g + g^T)/2 added to tau ([n, 6], xx yy zz xy xz yz), |S| =
    sqrt(2 S:S), c = (Cs Delta)^2.  Textbook eddy-viscosity closure, no
    counterpart in the reference."""
    s = np.stack([g[:, 0], g[:, 4], g[:, 8], 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]),
                  0.5 * (g[:, 5] + g[:, 7])], 1)
    ss = (s[:, :3] ** 2).sum(1) + 2 * (s[:, 3:] ** 2).sum(1)
    return tau - 2 * c * np.sqrt(2 * ss)[:, None] * s


def _k(n):
    k = np.fft.fftfreq(n, 1.0 / n)
    k[n // 2] = 0.0  # the Nyquist wavenumber counts as zero in derivatives
    return k


def gradient_features(u):
    """u: physical velocity [3, n, n, n] (z, y, x) -> [n^3, 9] with
    g[3i+j] = d u_i / d x_j by spectral derivatives (Nyquist dropped)."""
    n = u.shape[-1]
    k = _k(n)
    kx, ky, kz = k[None, None, :], k[None, :, None], k[:, None, None]
    out = np.empty((n ** 3, 9))
    for i in range(3):
        U = np.fft.fftn(u[i])
        for j, kj in enumerate((kx, ky, kz)):
            out[:, 3 * i + j] = np.fft.ifftn(1j * kj * U).real.reshape(-1)
    return out


def sgs_force(tau):
    """tau: physical stresses [6, n, n, n] -> Fourier force
    f_i = -sum_j i k_j tau_ij (Nyquist planes zero), [3, n, n, n]."""
    n = tau.shape[-1]
    k = _k(n)
    ks = (k[None, None, :], k[None, :, None], k[:, None, None])
    T = [np.fft.fftn(t) for t in tau]
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4, (1, 2): 5, (2, 1): 5}
    f = np.zeros((3, n, n, n), dtype=complex)
    for i in range(3):
        for j in range(3):
            f[i] -= 1j * ks[j] * T[idx[(i, j)]]
    nyq = np.zeros((n, n, n), dtype=bool)
    h = n // 2
    nyq[h, :, :] = nyq[:, h, :] = nyq[:, :, h] = True
    f[:, nyq] = 0
    return f


def box_filter(u, r):
    """Ratio-r box averages of u_i and tau_ij = <u_i u_j> - ubar_i ubar_j:
    [9, (n/r)^3] in the order ubar_x ubar_y ubar_z, tau xx yy zz xy xz yz."""
    nz, ny, nx = u.shape[1:]
    cz, cy, cx = nz // r, ny // r, nx // r

    def avg(f):
        return f.reshape(cz, r, cy, r, cx, r).mean(axis=(1, 3, 5))

    ub = [avg(u[i]) for i in range(3)]
    tau = [avg(u[i] * u[j]) - ub[i] * ub[j] for i, j in PAIRS]
    return np.stack(ub + tau).reshape(9, -1)
