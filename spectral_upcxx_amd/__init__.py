"""MI355X-native hot path of spectral_upcxx: 3D FFTs, spectral operators and
the ns.cpp fractional step as gfx950 HIP kernels behind a C ABI
(include/spectral_hip.h), with the reference's C++ Scalar/Vector/Operators API
on top (spectral_upcxx_amd/cpp).

This Python module only binds the C ABI (ctypes) and moves torch-allocated
device buffers into it; torch is plumbing for device memory and streams.
"""
import ctypes

from ._abi import lib, parse_header, LIB_PATH  # noqa: F401

SH_F64 = 64
SH_F32 = 32
SH_ZERO_NYQUIST = 1
SH_NORMALIZE = 2


def _torch():
    import torch
    return torch


def ptr(t):
    """Device pointer of a contiguous CUDA/HIP tensor."""
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def dims_of(t):
    """(nx, ny, nz) of a (..., nz, ny, nx) field tensor (x fastest, as the reference)."""
    nz, ny, nx = t.shape[-3:]
    return nx, ny, nz


def fft3d(t, direction, flags=0):
    """In-place 3D C2C transform of a complex128/complex64 (nz, ny, nx) tensor
    or a stack of them (count, nz, ny, nx)."""
    torch = _torch()
    nx, ny, nz = dims_of(t)
    count = t.numel() // (nx * ny * nz)
    L = lib()
    if t.dtype == torch.complex128:
        L.sh_fft3d_many_z2z(ptr(t), count, nx * ny * nz, nx, ny, nz, int(direction), int(flags), stream())
    elif t.dtype == torch.complex64:
        L.sh_fft3d_many_c2c(ptr(t), count, nx * ny * nz, nx, ny, nz, int(direction), int(flags), stream())
    else:
        raise TypeError("fft3d expects complex128 or complex64")
    return t


def forward(t):
    """Scalar::forward: transform(-1) then zeroHighModes."""
    return fft3d(t, -1, SH_ZERO_NYQUIST)


def inverse(t):
    """Scalar::inverse: transform(+1) then divide by Nx*Ny*Nz."""
    return fft3d(t, +1, SH_NORMALIZE)


class NS:
    """exec/ns.cpp's solver on the device: set_initial -> startup -> steps."""

    def __init__(self, dims, nu, K, precision=SH_F64, fused=True):
        self.dims = tuple(int(d) for d in dims)
        self.h = ctypes.c_void_p()
        nx, ny, nz = self.dims
        lib().sh_ns_create(ctypes.byref(self.h), nx, ny, nz, float(nu), float(K), int(precision),
                           int(bool(fused)), stream())

    def close(self):
        if self.h and self.h.value:
            lib().sh_ns_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_initial(self, V0):
        lib().sh_ns_set_initial(self.h, ptr(V0))

    def set_nonlinear(self, on):
        lib().sh_ns_set_nonlinear(self.h, int(bool(on)))

    def set_scheme(self, scheme):
        """0: exec/ns.cpp (Crank-Nicolson viscous term), 1: exec/ns2.cpp (backward Euler)."""
        lib().sh_ns_set_scheme(self.h, int(scheme))

    def startup(self):
        lib().sh_ns_startup(self.h)

    def steps(self, n):
        lib().sh_ns_steps(self.h, int(n))

    def velocity(self):
        torch = _torch()
        nx, ny, nz = self.dims
        out = torch.empty((3, nz, ny, nx), dtype=torch.complex128, device="cuda")
        lib().sh_ns_get_velocity(self.h, ptr(out))
        return out

    def vorticity(self):
        torch = _torch()
        nx, ny, nz = self.dims
        out = torch.empty((3, nz, ny, nx), dtype=torch.complex128, device="cuda")
        lib().sh_ns_get_vorticity(self.h, ptr(out))
        return out

    def div_average(self):
        r = ctypes.c_double()
        lib().sh_ns_div_average(self.h, ctypes.byref(r))
        return r.value
