"""HIP FFT parity: batched 1D and 3D C2C transforms against the oracle and the
committed golden vectors (FP64 bar: 1e-12 relative L2), plus size-independent
properties at the headline sizes (round trip, Parseval, linearity)."""
import ctypes

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


def test_prng_bit_exact(hip):
    import torch
    for dims, seed in [((16, 16, 16), 20240611), ((24, 8, 12), 777)]:
        nx, ny, nz = dims
        t = torch.empty((nz, ny, nx), dtype=torch.complex128, device="cuda")
        hip.lib().sh_prng_complex_z(hip.ptr(t), t.numel(), seed, hip.stream())
        assert np.array_equal(_host(t), O.prng_complex(seed, dims))


@pytest.mark.parametrize("name,dims,seed", [("fft_16", (16, 16, 16), 20240611),
                                            ("fft_24x8x12", (24, 8, 12), 777)])
def test_fft3d_golden(hip, name, dims, seed):
    g = np.load("%s/%s.npz" % (GOLDEN, name))
    x = O.prng_complex(seed, dims)
    t = _dev(x)
    hip.fft3d(t, -1, 0)
    assert rel_l2(_host(t), g["fwd"]) < TOL
    t = _dev(x)
    hip.fft3d(t, +1, hip.SH_NORMALIZE)
    assert rel_l2(_host(t), g["inv"]) < TOL


@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 12, 12), (32, 32, 32), (64, 64, 64), (96, 96, 96),
                                  (128, 128, 128), (48, 16, 24), (16, 64, 32), (256, 256, 256)])
@pytest.mark.parametrize("direction", [-1, 1])
def test_fft3d_vs_oracle(hip, dims, direction):
    x = O.prng_complex(1000 + sum(dims), dims)
    ref = O.transform3d(x, direction)
    t = _dev(x)
    hip.fft3d(t, direction, 0)
    assert rel_l2(_host(t), ref) < TOL


def test_forward_inverse_flags(hip):
    """Scalar::forward zeroes the Nyquist planes; Scalar::inverse divides by N^3."""
    dims = (16, 8, 12)
    x = O.prng_complex(5, dims)
    t = _dev(x)
    hip.forward(t)
    assert rel_l2(_host(t), O.forward(x)) < TOL
    y = _host(t)
    nx, ny, nz = dims
    assert np.all(y[:, :, nx // 2] == 0) and np.all(y[:, ny // 2, :] == 0) and np.all(y[nz // 2] == 0)
    hip.inverse(t)
    assert rel_l2(_host(t), O.inverse(O.forward(x))) < TOL


@pytest.mark.parametrize("n,stride,dist,batch", [(8, 1, 8, 37), (256, 1, 300, 9), (512, 1, 512, 5),
                                                 (64, 3, 1, 3), (96, 7, 1, 7), (16, 40, 2, 19),
                                                 (128, 1, 128, 1000)])
def test_fft1d_batched_strided(hip, n, stride, dist, batch):
    """FFT1D::transform semantics on strided, batched lines (src/FFT1D.cpp:95-101)."""
    size = (batch - 1) * dist + (n - 1) * stride + 1
    rng = np.random.default_rng(n * 131 + stride)
    buf = rng.uniform(-1, 1, size) + 1j * rng.uniform(-1, 1, size)
    idx = np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride
    for direction in (-1, 1):
        ref = buf.copy()
        lines = ref[idx]
        ref[idx] = np.fft.fft(lines, axis=1) if direction < 0 else np.fft.ifft(lines, axis=1) * n
        t = _dev(buf)
        hip.lib().sh_fft1d_z2z(hip.ptr(t), n, stride, dist, batch, direction, hip.stream())
        assert rel_l2(_host(t), ref) < TOL


@pytest.mark.parametrize("n", [256, 512])
def test_fft3d_properties_full_size(hip, n):
    """Round trip and Parseval at the benchmark sizes (the oracle-free bar)."""
    import torch
    x = torch.empty((n, n, n), dtype=torch.complex128, device="cuda")
    hip.lib().sh_prng_complex_z(hip.ptr(x), x.numel(), 99, hip.stream())
    x0 = x.clone()
    hip.fft3d(x, -1, 0)
    e_phys = float((x0.abs() ** 2).sum())
    e_four = float((x.abs() ** 2).sum()) / n ** 3
    assert abs(e_four / e_phys - 1) < 1e-12
    # single mode check: DC term equals the sum
    assert abs(complex(x[0, 0, 0].item()) - complex(x0.sum().item())) < 1e-9 * n ** 1.5
    hip.fft3d(x, +1, hip.SH_NORMALIZE)
    err = float(torch.linalg.vector_norm(x - x0) / torch.linalg.vector_norm(x0))
    assert err < 1e-13
    del x, x0
    torch.cuda.empty_cache()


def test_fft3d_c2c_fp32(hip):
    """complex64 transform against the FP64 oracle: FP32 rounding only."""
    import torch
    dims = (64, 64, 64)
    x = O.prng_complex(3, dims)
    ref = O.transform3d(x, -1)
    t = _dev(x.astype(np.complex64))
    hip.fft3d(t, -1, 0)
    assert rel_l2(_host(t).astype(np.complex128), ref) < 2e-6


def test_unsupported_size_is_an_error(hip):
    import torch
    t = torch.zeros((7, 7, 7), dtype=torch.complex128, device="cuda")
    with pytest.raises(RuntimeError, match="unsupported"):
        hip.fft3d(t, -1, 0)
