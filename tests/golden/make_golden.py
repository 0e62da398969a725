"""Golden-vector generator: a numpy restatement of the spectral_upcxx hot path.

Test infrastructure only.  It writes the small fixtures under tests/golden/
that pin both the C oracle (oracle/) and the HIP library.  It is independent of
both: the FFT is numpy's pocketfft, every operator is written again here from
the reference sources, and the PRNG is the counter-based splitmix64 that all
three implementations share.

Reference anchors (paths relative to the reference root):
  * 3D FFT sign/normalisation: src/FFT1D.cpp:72-101 (DIR=-1 -> sign -1),
    src/Scalar.cpp:208-229 (forward = transform(-1) + zeroHighModes,
    inverse = transform(+1) then /= Nx*Ny*Nz).  Same as numpy fftn/ifftn.
  * freq: src/Scalar.cpp:198-206 (index <= N/2 -> +index, else index-N).
  * zeroHighModes: src/Scalar.cpp:336-370 (planes i=Nx/2, j=Ny/2, k=Nz/2).
  * gradient/laplacian/laplacian_inverse: src/Scalar.cpp:245-334.
  * cross/curl/divergence/vector laplacian: src/Vector.cpp:66-181.
  * Operators: src/Operators.cpp:9-99.
  * NS start-up and fractional step: exec/ns.cpp:133-217.

Run:  python tests/golden/make_golden.py   (rewrites the .npz fixtures)
"""
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = 0x9E3779B97F4A7C15


# ---------------------------------------------------------------- PRNG
def splitmix64(seed, n):
    """z_n = mix(seed + (n+1)*golden), n = 0..count-1 (counter form of splitmix64)."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def prng_complex(seed, dims):
    """Complex field with Re, Im i.i.d. U[-1,1]; element e = i + nx*(j + ny*k)
    takes draws 2e (real) and 2e+1 (imaginary).  Returned as (nz, ny, nx)."""
    nx, ny, nz = dims
    n = nx * ny * nz
    u = 2.0 * uniform01(seed, 2 * n) - 1.0
    return (u[0::2] + 1j * u[1::2]).reshape(nz, ny, nx)


# ---------------------------------------------------------------- spectral helpers
def freq_axis(n):
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n).astype(np.float64)


def wavenumbers(dims):
    nx, ny, nz = dims
    kx = freq_axis(nx)[None, None, :]
    ky = freq_axis(ny)[None, :, None]
    kz = freq_axis(nz)[:, None, None]
    return kx, ky, kz


def zero_high(f):
    nz, ny, nx = f.shape
    f = f.copy()
    f[:, :, nx // 2] = 0
    f[:, ny // 2, :] = 0
    f[nz // 2, :, :] = 0
    return f


def forward(f):
    return zero_high(np.fft.fftn(f))


def inverse(f):
    return np.fft.ifftn(f)


def ksq(dims):
    kx, ky, kz = wavenumbers(dims)
    return kx * kx + ky * ky + kz * kz


def gradient(s):
    dims = s.shape[::-1]
    kx, ky, kz = wavenumbers(dims)
    return [(1j * kx) * s, (1j * ky) * s, (1j * kz) * s]


def laplacian(s):
    return -ksq(s.shape[::-1]) * s


def laplacian_inverse(s, a=0.0, b=1.0):
    q = ksq(s.shape[::-1])
    q = np.broadcast_to(q, s.shape)
    out = s.copy()
    m = q != 0
    out[m] = s[m] / (a - b * q[m])
    return out


def divergence(v):
    dims = v[0].shape[::-1]
    k = wavenumbers(dims)
    d = np.zeros_like(v[0])
    for c in range(3):
        d = d + (1j * k[c]) * v[c]
    return d


def curl(v):
    kx, ky, kz = wavenumbers(v[0].shape[::-1])
    return [(1j * ky) * v[2] - (1j * kz) * v[1],
            (1j * kz) * v[0] - (1j * kx) * v[2],
            (1j * kx) * v[1] - (1j * ky) * v[0]]


def vlaplacian(v):
    return [laplacian(c) for c in v]


def cross(a, b):
    u = [inverse(c) for c in a]
    w = [inverse(c) for c in b]
    r = [u[1] * w[2] - u[2] * w[1],
         u[2] * w[0] - u[0] * w[2],
         u[0] * w[1] - u[1] * w[0]]
    return [forward(c) for c in r]


def vadd(a, b):
    return [x + y for x, y in zip(a, b)]


def vsub(a, b):
    return [x - y for x, y in zip(a, b)]


def vscale(al, a):
    return [al * x for x in a]


# ---------------------------------------------------------------- NS (exec/ns.cpp:133-217)
def ns_startup(V0, omega0, nu, K):
    C0 = cross(V0, omega0)
    Pi0 = divergence(vadd(C0, vscale(nu, vlaplacian(V0))))
    Pi0 = laplacian_inverse(Pi0)
    V1 = vadd(V0, vscale(K, vsub(vadd(cross(V0, omega0), vscale(nu, vlaplacian(V0))),
                                 gradient(Pi0))))
    omega1 = curl(V1)
    cross0 = cross(V0, omega0)
    return V1, omega1, cross0


def ns_step(V1, omega1, cross0, nu, K):
    cross1 = cross(V1, omega1)
    F = vscale(K / 2, vsub(vscale(3.0, cross1), cross0))
    lap1 = vscale(nu * K / 2, vlaplacian(V1))
    G = divergence(vadd(vadd(V1, lap1), F))
    G = laplacian_inverse(G)
    Vhalf = vsub(vadd(V1, F), gradient(G))
    V = vadd(Vhalf, lap1)
    V = [laplacian_inverse(c, 1.0, -nu * K / 2) for c in V]
    omega = curl(V)
    return V, omega, cross1


# ---------------------------------------------------------------- initial conditions
def grid(dims):
    nx, ny, nz = dims
    x = (2.0 * np.pi / nx) * np.arange(nx)
    y = (2.0 * np.pi / ny) * np.arange(ny)
    z = (2.0 * np.pi / nz) * np.arange(nz)
    return x[None, None, :], y[None, :, None], z[:, None, None]


def taylor_green(dims):
    """u = (sin x cos y cos z, -cos x sin y cos z, 0), forward-transformed."""
    x, y, z = grid(dims)
    shape = dims[::-1]
    u = np.broadcast_to(np.sin(x) * np.cos(y) * np.cos(z), shape).astype(np.complex128)
    v = np.broadcast_to(-np.cos(x) * np.sin(y) * np.cos(z), shape).astype(np.complex128)
    w = np.zeros(shape, np.complex128)
    return [forward(u), forward(v), forward(w)]


def solenoidal_ic(dims, seed, k0=4.0, urms=1.0):
    """Random solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2).

    Mode m (linear index, x fastest), component d: amplitude
    A = sqrt(E(|k|) / (4 pi |k|^2)), phase 2 pi U(seed, 3m+d).  Then the
    divergence-free projection P = I - k k^T/|k|^2, Hermitian symmetrisation
    (u(k) + conj(u(-k)))/2, k=0 and Nyquist planes zeroed, and a final scale so
    that sqrt(<|u|^2>/3) = urms in physical space."""
    nx, ny, nz = dims
    n = nx * ny * nz
    kx, ky, kz = wavenumbers(dims)
    shape = (nz, ny, nx)
    kx = np.broadcast_to(kx, shape)
    ky = np.broadcast_to(ky, shape)
    kz = np.broadcast_to(kz, shape)
    q = kx * kx + ky * ky + kz * kz
    kmag = np.sqrt(q)
    with np.errstate(divide="ignore", invalid="ignore"):
        E = kmag ** 4 * np.exp(-2.0 * (kmag / k0) ** 2)
        A = np.where(q > 0, np.sqrt(E / (4.0 * np.pi * q)), 0.0)
    U = uniform01(seed, 3 * n)
    ph = [2.0 * np.pi * U[d::3].reshape(shape) for d in range(3)]
    u = [A * (np.cos(p) + 1j * np.sin(p)) for p in ph]
    with np.errstate(divide="ignore", invalid="ignore"):
        kdotu = kx * u[0] + ky * u[1] + kz * u[2]
        inv = np.where(q > 0, 1.0 / q, 0.0)
    k = [kx, ky, kz]
    u = [u[d] - k[d] * (kdotu * inv) for d in range(3)]
    # Hermitian symmetrisation: index of -k along each axis is (N - i) mod N
    def neg(f):
        return f[(-np.arange(nz)) % nz][:, (-np.arange(ny)) % ny][:, :, (-np.arange(nx)) % nx]
    u = [0.5 * (c + np.conj(neg(c))) for c in u]
    u = [zero_high(c) for c in u]
    for c in u:
        c[0, 0, 0] = 0
    e2 = sum(float(np.sum(np.abs(c) ** 2)) for c in u) / float(n) ** 2
    s = urms / np.sqrt(e2 / 3.0)
    return [s * c for c in u]


def main():
    out = {}
    # 1. raw 3D FFT (no Nyquist zeroing), cubic and non-cubic with a radix-3 factor
    for name, dims, seed in [("fft_16", (16, 16, 16), 20240611),
                             ("fft_24x8x12", (24, 8, 12), 777)]:
        f = prng_complex(seed, dims)
        np.savez(os.path.join(HERE, name + ".npz"), dims=np.array(dims), seed=np.uint64(seed),
                 fwd=np.fft.fftn(f), inv=np.fft.ifftn(f))
        out[name] = dims
    # 2. operators on a random Fourier field (forward with Nyquist zeroing)
    dims = (16, 16, 16)
    s = forward(prng_complex(11, dims))
    v = [forward(prng_complex(12 + d, dims)) for d in range(3)]
    w = [forward(prng_complex(20 + d, dims)) for d in range(3)]
    np.savez(os.path.join(HERE, "ops_16.npz"), dims=np.array(dims),
             seeds=np.array([11, 12, 13, 14, 20, 21, 22], dtype=np.uint64),
             grad=np.stack(gradient(s)), lap=laplacian(s),
             lapinv=laplacian_inverse(s), lapinv_cn=laplacian_inverse(s, 1.0, -0.3),
             div=divergence(v), curl=np.stack(curl(v)), vlap=np.stack(vlaplacian(v)),
             cross=np.stack(cross(v, w)))
    # 3. Taylor-Green, nu=0.01, dt=1e-3 (the parity configuration at 16^3 and 32^3)
    nu, K = 0.01, 1e-3
    for n, keep_fields in [(16, True), (32, False)]:
        dims = (n, n, n)
        V0 = taylor_green(dims)
        om0 = curl(V0)
        V1, om1, c0 = ns_startup(V0, om0, nu, K)
        snaps = {"startup": V1}
        for t in range(1, 101):
            V, om, c1 = ns_step(V1, om1, c0, nu, K)
            V1, om1, c0 = V, om, c1
            if t in (1, 10, 100):
                snaps["step%d" % t] = V
        rec = {"dims": np.array(dims), "nu": nu, "K": K}
        for key, val in snaps.items():
            st = np.stack(val)
            rec[key + "_norm"] = np.array([np.linalg.norm(c) for c in st])
            if keep_fields:
                rec[key] = st
            else:
                rec[key + "_sub"] = st.reshape(3, -1)[:, ::97]
        np.savez(os.path.join(HERE, "tg_%d.npz" % n), **rec)
    # 4. random solenoidal initial condition
    dims = (16, 16, 16)
    u = solenoidal_ic(dims, 7, 4.0, 1.0)
    np.savez(os.path.join(HERE, "solenoidal_16.npz"), dims=np.array(dims), seed=np.uint64(7),
             k0=4.0, urms=1.0, V=np.stack(u))
    print("wrote fixtures:", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
