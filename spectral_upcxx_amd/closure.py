"""The learned subgrid-stress closure and the coarse-grid LES through the C ABI
(sh_mlp_*, sh_les_*).  torch allocates the device buffers; every computation
runs in the library's HIP kernels."""
import ctypes

import numpy as np

from . import lib, ptr, stream

NPARAM = None


def param_count():
    global NPARAM
    if NPARAM is None:
        c = ctypes.c_int()
        lib().sh_mlp_param_count(ctypes.byref(c))
        NPARAM = c.value
    return NPARAM


def _fp(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _offset(t, elems):
    """Device pointer `elems` float32 elements into tensor t."""
    return ctypes.c_void_p(t.data_ptr() + 4 * int(elems))


class MLP:
    """9 -> 64 -> 64 -> 6 network on BF16 MFMA (FP32 master weights)."""

    def __init__(self, seed=1, strm=None):
        self.h = ctypes.c_void_p()
        self.stream = stream() if strm is None else strm
        lib().sh_mlp_create(ctypes.byref(self.h), int(seed), self.stream)

    def close(self):
        if self.h:
            lib().sh_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def params(self):
        out = np.empty(param_count(), dtype=np.float32)
        lib().sh_mlp_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out

    def set_params(self, P):
        a, p = _fp(P)
        assert a.size == param_count()
        lib().sh_mlp_set_params(self.h, p)

    def set_scaling(self, in_scale, out_scale):
        a, pa = _fp(in_scale)
        b, pb = _fp(out_scale)
        assert a.size == 9 and b.size == 6
        lib().sh_mlp_set_scaling(self.h, pa, pb)

    def forward(self, X, n=None):
        """X: float32 device tensor [9, ld] -> [6, n]."""
        import torch
        ld = X.shape[1]
        n = ld if n is None else n
        Y = torch.empty((6, n), dtype=torch.float32, device=X.device)
        lib().sh_mlp_forward(self.h, ptr(X), ld, n, ptr(Y), n)
        return Y

    def gradient(self, X, T, n=None, start=0):
        import torch
        ld = X.shape[1]
        n = ld - start if n is None else n
        g = torch.empty(param_count(), dtype=torch.float32, device=X.device)
        loss = ctypes.c_double()
        lib().sh_mlp_gradient(self.h, _offset(X, start), _offset(T, start), ld, n, ptr(g), ctypes.byref(loss))
        return loss.value, g

    def train_step(self, X, T, lr, n=None, start=0, want_loss=False):
        """One Adam step on samples [start, start + n) of X [9, ld], T [6, ld]."""
        ld = X.shape[1]
        n = ld - start if n is None else n
        loss = ctypes.c_double()
        lib().sh_mlp_train_step(self.h, _offset(X, start), _offset(T, start), ld, n, float(lr),
                                ctypes.byref(loss) if want_loss else None)
        return loss.value if want_loss else None


def features(ubar, strm=None):
    """Network inputs d ubar_i / d x_j ([9, n^3] float32) of a physical
    filtered velocity ubar ([3, n, n, n] float32 device tensor)."""
    import torch
    n = ubar.shape[-1]
    X = torch.empty((9, n ** 3), dtype=torch.float32, device=ubar.device)
    lib().sh_les_features(n, ptr(ubar), ptr(X), n ** 3, stream() if strm is None else strm)
    return X


class LES:
    """FP32 coarse-grid solver with the network's stress as body force
    (model=None: the no-model solve)."""

    def __init__(self, n, nu, K, model=None, clip=True, strm=None, dealias=0, lag=0):
        """dealias: 0 none (the reference scheme), 1 the 2/3 rule, 2 the exponential filter;
        lag: 0 synchronous closure, 1 lagged closure (sh_les_set_lag)"""
        self.n = n
        self.h = ctypes.c_void_p()
        self.stream = stream() if strm is None else strm
        lib().sh_les_create(ctypes.byref(self.h), int(n), float(nu), float(K), model.h if model else None,
                            int(bool(clip)), self.stream)
        if dealias:
            lib().sh_les_set_dealias(self.h, int(dealias))
        if lag:
            lib().sh_les_set_lag(self.h, int(lag))

    def set_lag(self, lag):
        lib().sh_les_set_lag(self.h, int(lag))

    def set_smagorinsky(self, cs):
        """mixed model: the network's stress plus the Smagorinsky stress with constant cs"""
        lib().sh_les_set_smagorinsky(self.h, float(cs))

    def close(self):
        if self.h:
            lib().sh_les_destroy(self.h)
            self.h = ctypes.c_void_p()

    def set_initial(self, ubar):
        lib().sh_les_set_initial_physical(self.h, ptr(ubar))

    def steps(self, k):
        lib().sh_les_steps(self.h, int(k))

    def autotune(self, reps=2):
        lib().sh_les_autotune(self.h, int(reps))

    def velocity(self):
        import torch
        V = torch.empty((3, self.n, self.n, self.n), dtype=torch.complex128, device="cuda")
        lib().sh_les_get_velocity(self.h, ptr(V))
        return V

    def stress(self):
        import torch
        t = torch.empty((6, self.n ** 3), dtype=torch.float32, device="cuda")
        lib().sh_les_stress(self.h, ptr(t))
        return t

    def force(self):
        import torch
        f = torch.empty((3, self.n, self.n, self.n), dtype=torch.complex64, device="cuda")
        lib().sh_les_force(self.h, ptr(f))
        return f

    def pending_force(self):
        import torch
        f = torch.empty((3, self.n, self.n, self.n), dtype=torch.complex64, device="cuda")
        lib().sh_les_pending_force(self.h, ptr(f))
        return f
