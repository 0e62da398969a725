"""ctypes access to the C oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / the timed CPU baseline.  Arrays are
numpy complex128 of shape (nz, ny, nx), C order -- the reference's layout
n = i + Nx*j + Nx*Ny*k (src/Scalar.cpp:682-690).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
# the same source built with -march=native on the host that runs it (make native)
NATIVE_LIB_PATH = os.path.join(HERE, "build", "native", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def use_library(path):
    """Load the oracle from `path` (e.g. NATIVE_LIB_PATH) instead of the
    portable build; only before the first call into the oracle."""
    global LIB_PATH
    if _lib is not None and path != LIB_PATH:
        raise RuntimeError("the oracle library is already loaded from %s" % LIB_PATH)
    LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        i, d, u64, sz, vp = ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_void_p
        sig = {
            "or_splitmix64": (u64, [u64, u64]),
            "or_prng_uniform": (None, [u64, u64, sz, _dp]),
            "or_prng_complex": (None, [u64, i, i, i, _dp]),
            "or_fft1d": (i, [_dp, i, ctypes.c_long, i]),
            "or_transform3d": (i, [_dp, _dp, i, i, i, i]),
            "or_forward": (i, [_dp, _dp, i, i, i]),
            "or_inverse": (i, [_dp, _dp, i, i, i]),
            "or_zero_high": (None, [_dp, i, i, i]),
            "or_freq": (d, [i, i]),
            "or_gradient": (None, [_dp, _dp, _dp, _dp, i, i, i]),
            "or_laplacian": (None, [_dp, _dp, i, i, i]),
            "or_laplacian_inverse": (None, [_dp, d, d, i, i, i]),
            "or_divergence": (None, [_dp, _dp, _dp, _dp, i, i, i]),
            "or_curl": (None, [_dp] * 6 + [i, i, i]),
            "or_vlaplacian": (None, [_dp] * 6 + [i, i, i]),
            "or_cross": (i, [_dp, _dp, _dp, _dp, i, i, i]),
            "or_average": (d, [_dp, sz]),
            "or_ns_create": (vp, [i, i, i, d, d]),
            "or_ns_destroy": (None, [vp]),
            "or_ns_set_initial": (i, [vp, _dp]),
            "or_ns_startup": (i, [vp]),
            "or_ns_steps": (i, [vp, i]),
            "or_ns_prime_for_timing": (i, [vp]),
            "or_ns_get_velocity": (None, [vp, _dp]),
            "or_ns_get_vorticity": (None, [vp, _dp]),
            "or_ns_div_average": (d, [vp]),
            "or_ns_set_nonlinear": (None, [vp, i]),
            "or_ns_set_scheme": (None, [vp, i]),
            "or_ns_set_vorticity": (None, [vp, _dp]),
            "or_vorticity_stats": (None, [_dp, i, i, i, _dp, _dp]),
            "or_ic_vortices": (i, [_dp, _dp, _dp, i, i, i, i]),
            "or_particles_seed": (None, [_dp, ctypes.c_long, u64]),
            "or_particles_move": (None, [_dp, _dp, _dp, ctypes.c_long, _dp, i, i, i, d]),
            "or_ic_taylor_green": (i, [_dp, _dp, i, i, i]),
            "or_ic_solenoidal": (None, [_dp, i, i, i, u64, d, d]),
            "or_energy": (d, [_dp, i, i, i]),
            "or_spectrum": (None, [_dp, i, i, i, _dp, i]),
            "or_dissipation": (d, [_dp, i, i, i, d]),
            "or_set_ranks": (None, [i]),
            "or_get_ranks": (i, []),
            "or_num_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _dims(a):
    nz, ny, nx = a.shape[-3:]
    return nx, ny, nz


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): unsupported length?" % (what, rc))


def prng_complex(seed, dims):
    nx, ny, nz = dims
    out = np.empty((nz, ny, nx), np.complex128)
    lib().or_prng_complex(seed, nx, ny, nz, _p(out))
    return out


def prng_uniform(seed, first, n):
    out = np.empty(n, np.float64)
    lib().or_prng_uniform(seed, first, n, _p(out))
    return out


def transform3d(f, direction):
    f = np.ascontiguousarray(f, np.complex128).copy()
    _check(lib().or_transform3d(_p(f), _p(np.empty_like(f)), *_dims(f), int(direction)), "transform3d")
    return f


def fft1d(x, direction):
    x = np.ascontiguousarray(x, np.complex128).copy()
    _check(lib().or_fft1d(_p(x), x.size, 1, int(direction)), "fft1d")
    return x


def forward(f):
    f = np.ascontiguousarray(f, np.complex128).copy()
    _check(lib().or_forward(_p(f), _p(np.empty_like(f)), *_dims(f)), "forward")
    return f


def inverse(f):
    f = np.ascontiguousarray(f, np.complex128).copy()
    _check(lib().or_inverse(_p(f), _p(np.empty_like(f)), *_dims(f)), "inverse")
    return f


def zero_high(f):
    f = np.ascontiguousarray(f, np.complex128).copy()
    lib().or_zero_high(_p(f), *_dims(f))
    return f


def freq(ind, n):
    return lib().or_freq(ind, n)


def gradient(s):
    s = np.ascontiguousarray(s, np.complex128)
    out = np.empty((3,) + s.shape, np.complex128)
    lib().or_gradient(_p(s), _p(out[0]), _p(out[1]), _p(out[2]), *_dims(s))
    return out


def laplacian(s):
    s = np.ascontiguousarray(s, np.complex128)
    out = np.empty_like(s)
    lib().or_laplacian(_p(s), _p(out), *_dims(s))
    return out


def laplacian_inverse(s, a=0.0, b=1.0):
    s = np.ascontiguousarray(s, np.complex128).copy()
    lib().or_laplacian_inverse(_p(s), a, b, *_dims(s))
    return s


def divergence(v):
    v = np.ascontiguousarray(v, np.complex128)
    out = np.empty(v.shape[1:], np.complex128)
    lib().or_divergence(_p(v[0]), _p(v[1]), _p(v[2]), _p(out), *_dims(out))
    return out


def curl(v):
    v = np.ascontiguousarray(v, np.complex128)
    out = np.empty_like(v)
    lib().or_curl(_p(v[0]), _p(v[1]), _p(v[2]), _p(out[0]), _p(out[1]), _p(out[2]), *_dims(out))
    return out


def vlaplacian(v):
    v = np.ascontiguousarray(v, np.complex128)
    out = np.empty_like(v)
    lib().or_vlaplacian(_p(v[0]), _p(v[1]), _p(v[2]), _p(out[0]), _p(out[1]), _p(out[2]), *_dims(out))
    return out


def cross(a, b):
    a = np.ascontiguousarray(a, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    out = np.empty_like(a)
    work = np.empty((7,) + a.shape[1:], np.complex128)
    _check(lib().or_cross(_p(a), _p(b), _p(out), _p(work), *_dims(out)), "cross")
    return out


def average(f):
    f = np.ascontiguousarray(f, np.complex128)
    return lib().or_average(_p(f), f.size)


def taylor_green(dims):
    nx, ny, nz = dims
    V = np.empty((3, nz, ny, nx), np.complex128)
    _check(lib().or_ic_taylor_green(_p(V), _p(np.empty((nz, ny, nx), np.complex128)), nx, ny, nz), "tg")
    return V


def solenoidal(dims, seed, k0=4.0, urms=1.0):
    nx, ny, nz = dims
    V = np.empty((3, nz, ny, nx), np.complex128)
    lib().or_ic_solenoidal(_p(V), nx, ny, nz, seed, k0, urms)
    return V


def energy(V):
    V = np.ascontiguousarray(V, np.complex128)
    return lib().or_energy(_p(V), *_dims(V))


def spectrum(V, nbins):
    V = np.ascontiguousarray(V, np.complex128)
    Ek = np.zeros(nbins)
    lib().or_spectrum(_p(V), *_dims(V), _p(Ek), nbins)
    return Ek


def dissipation(V, nu):
    """2 nu sum_modes |k|^2 e(k) (e(k) = 0.5 |u(k)|^2 / N^2)."""
    V = np.ascontiguousarray(V, np.complex128)
    return lib().or_dissipation(_p(V), *_dims(V), float(nu))


def set_ranks(r):
    """Virtual slab ranks of the 3D transform (1 = OpenMP threads over each phase)."""
    lib().or_set_ranks(int(r))


def get_ranks():
    return lib().or_get_ranks()


def vorticity_stats(V):
    """(enstrophy 0.5<|curl u|^2>, divergence norm sqrt(<|div u|^2>))."""
    V = np.ascontiguousarray(V, np.complex128)
    e, dn = ctypes.c_double(), ctypes.c_double()
    lib().or_vorticity_stats(_p(V), *_dims(V), ctypes.byref(e), ctypes.byref(dn))
    return e.value, dn.value


def vortices(dims, which):
    """Gaussian-vortex IC: which 0 = exec/ns.cpp (three), 1 = exec/ns2.cpp (two).
    Returns (V0, omega0), Fourier, 3 fields each."""
    nx, ny, nz = dims
    V = np.empty((3, nz, ny, nx), np.complex128)
    W = np.empty_like(V)
    _check(lib().or_ic_vortices(_p(V), _p(W), _p(np.empty((nz, ny, nx), np.complex128)), nx, ny, nz,
                                int(which)), "vortices")
    return V, W


def particles_seed(npart, seed):
    loc = np.empty((npart, 3), np.float64)
    lib().or_particles_seed(_p(loc), npart, seed)
    return loc


def particles_move(loc, U, K):
    """Particle::move for every row of loc (modified in place); U is the
    physical velocity (3, nz, ny, nx) complex128.  Returns (vel, energy)."""
    assert loc.dtype == np.float64 and loc.flags["C_CONTIGUOUS"]
    U = np.ascontiguousarray(U, np.complex128)
    n = loc.shape[0]
    vel = np.empty((n, 3), np.float64)
    en = np.empty(n, np.float64)
    lib().or_particles_move(_p(loc), _p(vel), _p(en), n, _p(U), *_dims(U), float(K))
    return vel, en


def num_threads():
    return lib().or_num_threads()


class NS:
    """exec/ns.cpp's solver state: set_initial -> startup -> steps."""

    def __init__(self, dims, nu, K):
        self.dims = tuple(dims)
        nx, ny, nz = dims
        self.h = lib().or_ns_create(nx, ny, nz, nu, K)
        if not self.h:
            raise MemoryError("or_ns_create")

    def close(self):
        if self.h:
            lib().or_ns_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_nonlinear(self, on):
        lib().or_ns_set_nonlinear(self.h, int(bool(on)))

    def set_scheme(self, scheme):
        """0: exec/ns.cpp (Crank-Nicolson viscous term), 1: exec/ns2.cpp (backward Euler)."""
        lib().or_ns_set_scheme(self.h, int(scheme))

    def set_initial(self, V0):
        V0 = np.ascontiguousarray(V0, np.complex128)
        _check(lib().or_ns_set_initial(self.h, _p(V0)), "ns_set_initial")

    def set_vorticity(self, W0):
        W0 = np.ascontiguousarray(W0, np.complex128)
        lib().or_ns_set_vorticity(self.h, _p(W0))

    def startup(self):
        _check(lib().or_ns_startup(self.h), "ns_startup")

    def prime_for_timing(self):
        _check(lib().or_ns_prime_for_timing(self.h), "ns_prime_for_timing")

    def steps(self, n):
        _check(lib().or_ns_steps(self.h, n), "ns_steps")

    def velocity(self):
        nx, ny, nz = self.dims
        out = np.empty((3, nz, ny, nx), np.complex128)
        lib().or_ns_get_velocity(self.h, _p(out))
        return out

    def vorticity(self):
        nx, ny, nz = self.dims
        out = np.empty((3, nz, ny, nx), np.complex128)
        lib().or_ns_get_vorticity(self.h, _p(out))
        return out

    def div_average(self):
        return lib().or_ns_div_average(self.h)
