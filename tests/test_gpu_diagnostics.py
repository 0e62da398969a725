"""Reductions and spectral diagnostics on the device against the oracle and
closed forms: Scalar::average (src/Scalar.cpp:15-22), the shell spectrum
E(k), the energy and the dissipation rate; the spectrum is deterministic
(repeat calls give the same bits) and obeys the sum rule at 256^3."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _spectrum(hip, V, nu, nbins):
    nz, ny, nx = V.shape[-3:]
    Ek = (ctypes.c_double * max(nbins, 1))()
    e, d = ctypes.c_double(), ctypes.c_double()
    hip.lib().sh_spectrum_z(hip.ptr(V), nx, ny, nz, float(nu), Ek, nbins, ctypes.byref(e),
                            ctypes.byref(d), hip.stream())
    return np.array(Ek[:nbins]), e.value, d.value


def _average(hip, f):
    r = ctypes.c_double()
    hip.lib().sh_average_z(hip.ptr(f), f.numel(), ctypes.byref(r), hip.stream())
    return r.value


def test_average_closed_forms(hip):
    rng = np.random.default_rng(3)
    c = 2.5
    f = c * np.exp(2j * np.pi * rng.random((12, 10, 6)))
    assert abs(_average(hip, _dev(f)) - c) < 1e-14 * c
    g = np.zeros((8, 8, 8), np.complex128)
    g[3, 1, 7] = 3 - 4j
    assert abs(_average(hip, _dev(g)) - 5.0 / 512) < 1e-16
    h = O.prng_complex(11, (32, 16, 24))
    assert abs(_average(hip, _dev(h)) - np.mean(np.abs(h))) < 1e-14 * np.mean(np.abs(h))
    assert abs(_average(hip, _dev(h)) - O.average(h)) < 1e-14 * O.average(h)


@pytest.mark.parametrize("k", [(1, 0, 0), (1, 1, 1), (2, 1, 0), (-2, 2, -1), (8, 0, 0), (5, -5, 5)])
def test_spectrum_single_mode(hip, k):
    n = 16
    V = np.zeros((3, n, n, n), np.complex128)
    v = np.array([1 + 2j, -0.5j, 0.25])
    V[:, k[2] % n, k[1] % n, k[0] % n] = v
    e = 0.5 * np.sum(v.real ** 2 + v.imag ** 2) / n ** 6
    q = sum(c * c for c in k)
    b = int(np.floor(np.sqrt(q) + 0.5))
    Ek, en, diss = _spectrum(hip, _dev(V), 0.01, 20)
    assert abs(Ek[b] - e) <= 4e-16 * e and not np.delete(Ek, b).any()
    assert abs(en - e) <= 4e-16 * e
    assert abs(diss - 2 * 0.01 * q * e) <= 1e-15 * 2 * 0.01 * q * e


def test_spectrum_taylor_green(hip):
    V = O.taylor_green((32, 32, 32))
    Ek, en, diss = _spectrum(hip, _dev(V), 0.01, 30)
    assert abs(Ek[2] - 0.125) < 1e-15 and np.abs(np.delete(Ek, 2)).max() < 1e-30
    assert abs(en - 0.125) < 1e-15 and abs(diss - 0.0075) < 1e-17


@pytest.mark.parametrize("dims", [(32, 32, 32), (24, 16, 20)])
def test_spectrum_vs_oracle(hip, dims):
    import torch
    nx, ny, nz = dims
    V = torch.empty((3, nz, ny, nx), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V), nx, ny, nz, 17, 4.0, 1.0, hip.stream())
    Vh = V.cpu().numpy()
    nu = 1 / 800
    nbins = 30
    Ek, en, diss = _spectrum(hip, V, nu, nbins)
    ref = O.spectrum(Vh, nbins)
    assert rel_l2(Ek, ref) < 1e-14
    # every shell with energy above rounding agrees bin by bin
    big = ref > 1e-12 * ref.max()
    assert np.all(np.abs(Ek[big] - ref[big]) <= 1e-13 * ref[big])
    assert abs(en - O.energy(Vh)) < 1e-14 * en
    assert abs(diss - O.dissipation(Vh, nu)) < 1e-14 * diss


def test_spectrum_256_deterministic_and_sum_rule(hip):
    """256^3: repeat calls give the same bits; with every shell covered
    (round(|k|) <= round(sqrt(3) 128) = 222) the bins add up to the energy;
    dissipation = 2 nu enstrophy (solenoidal field)."""
    import torch
    n = 256
    V = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V), n, n, n, 2024, 4.0, 1.0, hip.stream())
    nu = 1 / 800
    a = _spectrum(hip, V, nu, 223)
    b = _spectrum(hip, V, nu, 223)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]
    Ek, en, diss = a
    assert abs(Ek.sum() - en) < 1e-13 * en
    assert abs(en - 1.5) < 1e-12  # u_rms = 1: 0.5 <|u|^2> = 3/2
    ens, dn = ctypes.c_double(), ctypes.c_double()
    hip.lib().sh_vorticity_stats_z(hip.ptr(V), n, n, n, ctypes.byref(ens), ctypes.byref(dn), hip.stream())
    assert abs(diss - 2 * nu * ens.value) < 1e-12 * diss
    del V
    torch.cuda.empty_cache()
