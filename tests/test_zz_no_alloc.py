"""sh_spectrum_z_noalloc (static device scratch, no allocation) against
sh_spectrum_z (pooled scratch): the same bits on two calls in a row, the
second on a smaller field with fewer bins, through the C ABI; nbins beyond
the static scratch is refused with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(hip, fn, V, nu, nbins):
    nz, ny, nx = V.shape[-3:]
    Ek = (ctypes.c_double * max(nbins, 1))()
    e, d = ctypes.c_double(), ctypes.c_double()
    getattr(hip.lib(), fn)(hip.ptr(V), nx, ny, nz, float(nu), Ek, nbins, ctypes.byref(e), ctypes.byref(d),
                           hip.stream())
    return np.array(Ek[:nbins]), e.value, d.value


def _field(hip, dims, seed):
    import torch
    nx, ny, nz = dims
    V = torch.empty((3, nz, ny, nx), dtype=torch.complex128, device="cuda")
    hip.lib().sh_solenoidal_ic_z(hip.ptr(V), nx, ny, nz, seed, 4.0, 1.0, hip.stream())
    return V


def test_noalloc_spectrum_matches_pooled(hip):
    cases = [((128, 96, 64), 111, 500, 1 / 800), ((24, 16, 20), 5, 17, 0.01)]
    for dims, seed, nbins, nu in cases:
        V = _field(hip, dims, seed)
        got = _call(hip, "sh_spectrum_z_noalloc", V, nu, nbins)
        ref = _call(hip, "sh_spectrum_z", V, nu, nbins)
        print(dims, nbins, "energy", got[1], ref[1], flush=True)
        assert np.array_equal(got[0], ref[0]), dims
        assert got[1] == ref[1] and got[2] == ref[2], dims
        assert got[1] > 0 and got[0].sum() > 0
    # energy only (no bins)
    V = _field(hip, (16, 16, 16), 3)
    assert _call(hip, "sh_spectrum_z_noalloc", V, 0.01, 0)[1:] == _call(hip, "sh_spectrum_z", V, 0.01, 0)[1:]


def test_noalloc_spectrum_rejects_too_many_bins(hip):
    V = _field(hip, (16, 16, 16), 4)
    with pytest.raises(RuntimeError, match="510"):
        _call(hip, "sh_spectrum_z_noalloc", V, 0.01, 511)
