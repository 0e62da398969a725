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


# lengths of every path: compiled plans with radix 5 / 7 and 1024; the
# mixed-radix kernel (factors 8 4 2 3 5 7 11 13, up to 2048); Bluestein (a
# larger prime factor, up to 1024)
ANY_COMPILED = [10, 14, 20, 28, 40, 56, 80, 100, 1024]
ANY_MIXED = [6, 11, 13, 22, 30, 44, 60, 77, 120, 143, 1000, 2048]
ANY_BLUESTEIN = [17, 31, 101, 127, 509, 1009]


@pytest.mark.parametrize("n", ANY_COMPILED + ANY_MIXED + ANY_BLUESTEIN)
@pytest.mark.parametrize("prec", [64, 32])
def test_fft1d_any_length_vs_oracle(hip, n, prec):
    """FFT1D of any length (src/FFT1D.cpp:14-54 plans FFTW for any N): contiguous
    and strided batches against the oracle (itself pinned by brute-force DFTs
    and numpy), both signs."""
    import torch
    batch = 5
    x = O.prng_complex(n + 3, (n * batch, 1, 1)).reshape(batch, n)
    tol = TOL if prec == 64 else 3e-6
    dt = np.complex128 if prec == 64 else np.complex64
    fn = hip.lib().sh_fft1d_z2z if prec == 64 else hip.lib().sh_fft1d_c2c
    for direction in (-1, 1):
        ref = np.stack([O.fft1d(row, direction) for row in x])
        t = _dev(x.astype(dt))
        fn(hip.ptr(t), n, 1, n, batch, direction, hip.stream())
        err = rel_l2(_host(t).astype(np.complex128), ref)
        assert err < tol, ("contiguous", n, direction, err)
        # the same lines stored column-wise (stride batch, distance 1)
        t = _dev(np.ascontiguousarray(x.T).astype(dt))
        fn(hip.ptr(t), n, batch, 1, batch, direction, hip.stream())
        err = rel_l2(_host(t).T.astype(np.complex128), ref)
        assert err < tol, ("strided", n, direction, err)


@pytest.mark.parametrize("dims", [(10, 20, 40), (100, 80, 14), (11, 13, 17), (30, 22, 31), (6, 1024, 2)])
def test_fft3d_any_dims_vs_oracle(hip, dims):
    x = O.prng_complex(sum(dims), dims)
    for direction in (-1, 1):
        t = _dev(x)
        hip.fft3d(t, direction, 0)
        assert rel_l2(_host(t), O.transform3d(x, direction)) < TOL, (dims, direction)
    # Scalar::forward / inverse: Nyquist planes at index n/2 (floor for odd n)
    t = _dev(x)
    hip.forward(t)
    assert rel_l2(_host(t), O.forward(x)) < TOL
    hip.inverse(t)
    assert rel_l2(_host(t), O.inverse(O.forward(x))) < TOL


def test_ns_op_by_op_any_grid_vs_oracle(hip):
    """exec/ns.cpp's step on a 20 x 20 x 20 Taylor-Green grid (radix-5
    plans), the op-by-op schedule, 10 steps against the oracle; the fused
    schedule reports that it has no kernels for the grid."""
    dims = (20, 20, 20)
    V0 = O.taylor_green(dims)
    ref = O.NS(dims, 0.01, 1e-3)
    ref.set_initial(V0)
    ref.startup()
    ref.steps(10)
    ns = hip.NS(dims, 0.01, 1e-3, hip.SH_F64, fused=False)
    ns.set_initial(_dev(V0))
    ns.startup()
    ns.steps(10)
    assert rel_l2(_host(ns.velocity()), ref.velocity()) < 1e-10
    ns.close()
    with pytest.raises(RuntimeError, match="fused"):
        hip.NS(dims, 0.01, 1e-3, hip.SH_F64, fused=True)


def test_unsupported_size_is_an_error(hip):
    """a length beyond both the mixed-radix kernel (2048) and Bluestein (1024):
    the prime 2053"""
    import torch
    t = torch.zeros((2053,), dtype=torch.complex128, device="cuda")
    with pytest.raises(RuntimeError, match="unsupported"):
        hip.lib().sh_fft1d_z2z(hip.ptr(t), 2053, 1, 2053, 1, -1, hip.stream())
