"""ctypes binding of include/spectral_hip.h.

The argument types are read from the header itself, so every entry point the
header declares is bound (and the CPU test suite checks each one is exported).
Loading fails loudly when the library has not been built: there is no
fallback implementation.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HEADER = os.path.join(ROOT, "include", "spectral_hip.h")
LIB_PATH = os.path.join(PKG, "lib", "libspectral_hip.so")

_CTYPES = {
    "void*": ctypes.c_void_p,
    "const void*": ctypes.c_void_p,
    "int": ctypes.c_int,
    "long": ctypes.c_long,
    "double": ctypes.c_double,
    "size_t": ctypes.c_size_t,
    "uint64_t": ctypes.c_uint64,
    "double*": ctypes.POINTER(ctypes.c_double),
    "const double*": ctypes.POINTER(ctypes.c_double),
    "long*": ctypes.POINTER(ctypes.c_long),
    "int*": ctypes.POINTER(ctypes.c_int),
    "size_t*": ctypes.POINTER(ctypes.c_size_t),
    "void**": ctypes.POINTER(ctypes.c_void_p),
    "char*": ctypes.c_char_p,
    "const char*": ctypes.c_char_p,
    "sh_ns_t": ctypes.c_void_p,
    "sh_ns_t*": ctypes.POINTER(ctypes.c_void_p),
    "sh_mlp_t": ctypes.c_void_p,
    "sh_mlp_t*": ctypes.POINTER(ctypes.c_void_p),
    "sh_les_t": ctypes.c_void_p,
    "sh_les_t*": ctypes.POINTER(ctypes.c_void_p),
    "float": ctypes.c_float,
    "float*": ctypes.POINTER(ctypes.c_float),
    "const float*": ctypes.POINTER(ctypes.c_float),
}


def _norm(t):
    t = re.sub(r"\s+", " ", t.strip())
    t = re.sub(r"\s*\*", "*", t)
    return t


def parse_header(path=HEADER):
    """Return {name: (restype_str, [argtype_str, ...])} for every prototype."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    protos = {}
    for m in re.finditer(r"\b(int|const char\*|const char \*)\s+(sh_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        argl = []
        if args.strip() and args.strip() != "void":
            for a in args.split(","):
                a = _norm(a)
                # drop the parameter name
                mm = re.match(r"^(.*?[\*\s])(\w+)$", a)
                ty = _norm(mm.group(1)) if mm else a
                argl.append(ty)
        protos[name] = (_norm(ret), argl)
    return protos


class SpectralLib:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError(
                "libspectral_hip.so is not built (%s). Run `python -m spectral_upcxx_amd.build` "
                "(or __graft_entry__.build()); there is no CPU fallback." % path)
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.protos = parse_header()
        for name, (ret, args) in self.protos.items():
            fn = getattr(self.lib, name)
            fn.restype = ctypes.c_char_p if "char" in ret else ctypes.c_int
            fn.argtypes = [_CTYPES[a] for a in args]

    def error(self):
        return self.lib.sh_last_error().decode()

    def __getattr__(self, name):
        if not name.startswith("sh_"):
            raise AttributeError(name)
        fn = getattr(self.lib, name)

        def call(*args):
            rc = fn(*args)
            if isinstance(rc, int) and rc != 0:
                raise RuntimeError("%s failed (%d): %s" % (name, rc, self.error()))
            return rc

        call.__name__ = name
        return call


_singleton = None


def lib():
    global _singleton
    if _singleton is None:
        _singleton = SpectralLib()
    return _singleton
