"""Pins the oracle's reductions and diagnostics, its transforms of any length
and its virtual slab ranks (oracle/oracle.c), before they serve as checkers.

  * Scalar::average (src/Scalar.cpp:15-22): closed forms (constant modulus,
    one nonzero element) and numpy's mean of abs;
  * the shell spectrum, energy and dissipation: single modes land in shell
    round(|k|) with their energy, Taylor-Green's closed form (all energy 1/8
    in shell 2, dissipation 3 nu / 4), the sum rule, and the identity
    dissipation = 2 nu enstrophy for solenoidal fields;
  * FFT1D of any length (src/FFT1D.cpp:14-54 plans FFTW for any N): brute-force
    extended-precision DFTs for primes and smooth lengths with factors 5 and 7,
    numpy's pocketfft for long lengths (1000, 1009, 1024);
  * slab ranks (src/Scalar.cpp:446-626): the R-rank transform moves data only,
    so it equals the one-rank transform bit for bit.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as O


# ---------------------------------------------------------------- average
def test_average_constant_modulus():
    """|f| = c everywhere (random phases) gives exactly c."""
    rng = np.random.default_rng(1)
    for c in (0.5, 3.0, 1e-7):
        f = c * np.exp(2j * np.pi * rng.random((6, 10, 12)))
        assert abs(O.average(f) - c) <= 4e-16 * c * 10


def test_average_single_element():
    """one nonzero v among n elements gives |v| / n (a sum of squares, a real
    part only or a divide by anything but n fails)."""
    for dims, idx, v in (((8, 8, 8), (3, 1, 7), 3 - 4j), ((12, 6, 10), (9, 5, 0), -2.5j)):
        f = np.zeros(dims[::-1], np.complex128)
        f[idx[2], idx[1], idx[0]] = v
        assert O.average(f) == abs(v) / f.size


def test_average_matches_numpy():
    f = O.prng_complex(5, (24, 8, 12))
    assert abs(O.average(f) - np.mean(np.abs(f))) < 1e-15 * np.mean(np.abs(f)) * 10


# ---------------------------------------------------------------- spectrum, energy, dissipation
def _mode_field(dims, k, v):
    """a field with the single Fourier coefficient v at wavenumber k (freq index map)"""
    nx, ny, nz = dims
    V = np.zeros((3, nz, ny, nx), np.complex128)
    ix, iy, iz = (k[0] % nx, k[1] % ny, k[2] % nz)
    V[:, iz, iy, ix] = v
    return V


@pytest.mark.parametrize("k", [(1, 0, 0), (1, 1, 0), (1, 1, 1), (2, 1, 0), (-2, 2, -1), (0, -3, 4),
                               (8, 0, 0), (5, -5, 5), (3, 3, 3)])
def test_spectrum_single_mode(k):
    dims = (16, 16, 16)
    v = np.array([1 + 2j, -0.5j, 0.25])
    V = _mode_field(dims, k, v)
    n = 16 ** 3
    e = 0.5 * np.sum(v.real ** 2 + v.imag ** 2) / n ** 2
    q = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    b = int(np.floor(np.sqrt(q) + 0.5))
    Ek = O.spectrum(V, 20)
    assert abs(Ek[b] - e) <= 4e-16 * e, (k, b)
    assert not np.delete(Ek, b).any(), (k, b)
    assert abs(O.energy(V) - e) <= 4e-16 * e
    assert abs(O.dissipation(V, 0.01) - 2 * 0.01 * q * e) <= 1e-16 * 2 * 0.01 * q * e * 4


def test_spectrum_drops_shells_beyond_nbins():
    V = _mode_field((16, 16, 16), (6, 0, 0), np.array([1.0, 0, 0]))
    assert not O.spectrum(V, 6).any()


def test_taylor_green_closed_forms():
    """u = (sin x cos y cos z, -cos x sin y cos z, 0): 8 modes at |k| = sqrt 3,
    energy 1/8 (all in shell round(1.73) = 2), enstrophy 3/8, dissipation
    2 nu 3 (1/8) = 3 nu / 4."""
    V = O.taylor_green((16, 16, 16))
    Ek = O.spectrum(V, 16)
    assert abs(Ek[2] - 0.125) < 1e-15
    assert np.abs(np.delete(Ek, 2)).max() < 1e-30
    assert abs(O.energy(V) - 0.125) < 1e-15
    assert abs(O.dissipation(V, 0.01) - 0.0075) < 1e-17


def test_spectrum_sum_rule_and_numpy():
    """every shell covered: sum of bins = energy; bins against a numpy
    restatement (bincount over round(|k|)); dissipation = 2 nu enstrophy."""
    dims = (24, 16, 20)
    V = O.solenoidal(dims, 3, 4.0, 1.0)
    nbins = 25
    Ek = O.spectrum(V, nbins)
    assert abs(Ek.sum() - O.energy(V)) < 1e-14 * O.energy(V)
    nx, ny, nz = dims
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kz = np.fft.fftfreq(nz, 1.0 / nz)
    # Scalar::freq maps N/2 to +N/2, numpy's fftfreq to -N/2: same |k|
    K2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    e = 0.5 * np.sum(np.abs(V) ** 2, axis=0) / (nx * ny * nz) ** 2
    ref = np.bincount(np.floor(np.sqrt(K2) + 0.5).astype(int).ravel(), weights=e.ravel(), minlength=nbins)
    assert rel_l2(Ek, ref[:nbins]) < 1e-14
    nu = 1 / 800
    ens, _ = O.vorticity_stats(V)
    assert abs(O.dissipation(V, nu) - 2 * nu * ens) < 1e-13 * 2 * nu * ens
    assert abs(O.dissipation(V, nu) - 2 * nu * np.sum(K2 * e)) < 1e-14 * O.dissipation(V, nu)


# ---------------------------------------------------------------- FFT1D of any length
def _dft(x, sign):
    n = x.size
    k = np.arange(n, dtype=np.longdouble)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    # reduce j*k mod n exactly before the angle, as an exact DFT needs
    jk = (np.outer(np.arange(n), np.arange(n)) % n).astype(np.longdouble)
    ang = sign * 2 * pi * jk / n
    M = np.cos(ang) + 1j * np.sin(ang)
    del k
    return (M @ x.astype(np.clongdouble)).astype(np.complex128)


@pytest.mark.parametrize("n", [1, 5, 7, 10, 11, 13, 14, 20, 25, 35, 40, 49, 97, 100, 121, 210])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_any_length_brute_force(n, sign):
    x = O.prng_complex(n + 17, (n, 1, 1)).ravel()
    assert rel_l2(O.fft1d(x, sign), _dft(x, sign)) < 5e-16 * max(1, np.log2(n))


@pytest.mark.parametrize("n", [80, 1000, 1009, 1024, 2 * 3 * 5 * 7 * 11])
def test_fft1d_long_lengths_vs_numpy(n):
    x = O.prng_complex(n, (n, 1, 1)).ravel()
    assert rel_l2(O.fft1d(x, -1), np.fft.fft(x)) < 1e-14
    assert rel_l2(O.fft1d(x, 1), np.fft.ifft(x) * n) < 1e-14


def test_fft3d_any_dims_vs_numpy():
    for dims in ((10, 14, 11), (20, 5, 7), (13, 12, 25)):
        x = O.prng_complex(sum(dims), dims)
        assert rel_l2(O.transform3d(x, -1), np.fft.fftn(x)) < 1e-14, dims
        assert rel_l2(O.transform3d(x, 1), np.fft.ifftn(x) * x.size) < 1e-14, dims


# ---------------------------------------------------------------- virtual slab ranks
@pytest.mark.parametrize("ranks", [2, 4, 8])
def test_slab_ranks_equal_one_rank(ranks):
    dims = (12, 16, 24)
    x = O.prng_complex(8, dims)
    O.set_ranks(1)
    ref = [O.transform3d(x, -1), O.transform3d(x, 1)]
    try:
        O.set_ranks(ranks)
        assert O.get_ranks() == ranks
        got = [O.transform3d(x, -1), O.transform3d(x, 1)]
    finally:
        O.set_ranks(1)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert rel_l2(ref[0], np.fft.fftn(x)) < 1e-14


def test_slab_ranks_step_equals_one_rank():
    dims = (16, 16, 16)
    V0 = O.solenoidal(dims, 9, 4.0, 1.0)
    out = []
    try:
        for r in (1, 4):
            O.set_ranks(r)
            ns = O.NS(dims, 1 / 800, 2e-3)
            ns.set_initial(V0)
            ns.startup()
            ns.steps(3)
            out.append(ns.velocity())
            ns.close()
    finally:
        O.set_ranks(1)
    assert np.array_equal(out[0], out[1])
