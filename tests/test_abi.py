"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every
symbol include/*.h declares; the C++ API layer and drivers are built and link
against it; host-side argument checking; bench.py refuses multi-GPU runs.
No kernel is launched here."""
import ctypes
import glob
import os
import re
import subprocess
import sys

import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "spectral_upcxx_amd", "lib")


def _declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(sh_\w+)\s*\(", text))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_headers_exist():
    hs = glob.glob(os.path.join(INCLUDE, "*.h"))
    assert os.path.join(INCLUDE, "spectral_hip.h") in hs


def test_library_exports_every_declared_symbol():
    lib = os.path.join(LIBDIR, "libspectral_hip.so")
    assert os.path.exists(lib), "build the library first (__graft_entry__.build())"
    declared = set()
    for h in glob.glob(os.path.join(INCLUDE, "*.h")):
        declared |= _declared(h)
    assert len(declared) >= 40
    missing = declared - _exported(lib)
    assert not missing, sorted(missing)


def test_ctypes_binding_covers_header():
    import spectral_upcxx_amd as S
    L = S.lib()
    assert set(L.protos) == _declared(os.path.join(INCLUDE, "spectral_hip.h"))
    for name in L.protos:
        assert hasattr(L.lib, name)


def test_library_is_gfx950_code():
    """The shared object carries a gfx950 code object (no other targets)."""
    lib = os.path.join(LIBDIR, "libspectral_hip.so")
    data = open(lib, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
    for other in (b"--gfx90a", b"--gfx942", b"--gfx1100"):
        assert other not in data


def test_host_argument_errors_without_gpu():
    """Argument validation happens before any device call."""
    import spectral_upcxx_amd as S
    L = S.lib()
    with pytest.raises(RuntimeError, match="invalid"):
        L.sh_fft3d_z2z(None, 16, 16, 16, -1, 0, None)
    with pytest.raises(RuntimeError, match="invalid"):
        L.sh_fft1d_z2z(ctypes.c_void_p(16), 16, 1, 16, 1, 3, None)
    with pytest.raises(RuntimeError):
        L.sh_gradient_z(None, None, None, None, 8, 8, 8, None)
    assert "null" in L.error() or "invalid" in L.error()
    # the solver, closure and tracer options reject what they do not define
    with pytest.raises(RuntimeError, match="scheme must be"):
        L.sh_ns_set_scheme(None, 2)
    with pytest.raises(RuntimeError, match="mode must be"):
        L.sh_ns_set_layout(None, 1)
    with pytest.raises(RuntimeError, match="lag must be"):
        L.sh_les_set_lag(None, 1)
    with pytest.raises(RuntimeError, match="lag must be"):
        L.sh_les_set_lag(None, 2)
    with pytest.raises(RuntimeError, match="invalid argument"):
        L.sh_les_pending_force(None, None)
    with pytest.raises(RuntimeError, match="invalid argument"):
        L.sh_les_force(None, None)
    with pytest.raises(RuntimeError, match="Cs >= 0"):
        L.sh_les_set_smagorinsky(None, 0.1)
    with pytest.raises(RuntimeError, match="invalid arguments"):
        L.sh_particles_seed(None, 10, 1, 0, None)
    with pytest.raises(RuntimeError, match="invalid arguments"):
        L.sh_particles_move(None, None, None, -1, None, None, None, 8, 8, 8, 0.1, None)
    with pytest.raises(RuntimeError, match="null argument"):
        L.sh_vortex_ic_z(None, None, 8, 8, 8, 1, None)


def test_cpp_api_built():
    """The reference's C++ API (DataObjects.H) is built on top of the C ABI."""
    api = os.path.join(LIBDIR, "libspectral_api.so")
    assert os.path.exists(api)
    needed = subprocess.run(["readelf", "-d", api], stdout=subprocess.PIPE, text=True).stdout
    assert "libspectral_hip.so" in needed
    for exe in ("ns", "fft", "ilap", "test_operators", "test_3d_fft", "test_freq", "test_cn_heat"):
        assert os.access(os.path.join(LIBDIR, "bin", exe), os.X_OK), exe


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_oracle_library_exports():
    from oracle import pyoracle as O
    L = O.lib()
    for name in ("or_transform3d", "or_cross", "or_ns_steps", "or_ic_solenoidal"):
        assert hasattr(L, name)


def test_ensemble_parameters_are_seeded_and_in_range():
    """Config-4 box parameters: k0 ~ U[2,6], nu ~ U[1/1000, 1/300], reproducible."""
    from spectral_upcxx_amd import forms
    p1 = forms.ensemble_params(((64, 32), (96, 16)), 4242)
    p2 = forms.ensemble_params(((64, 32), (96, 16)), 4242)
    assert p1 == p2
    boxes = [b for _, bx in p1 for b in bx]
    assert len(boxes) == 48 and len({b["seed"] for b in boxes}) == 48
    assert all(2 <= b["k0"] <= 6 and 1 / 1000 <= b["nu"] <= 1 / 300 for b in boxes)
    # the draws use the IC generator's splitmix64 stream
    from oracle import pyoracle as O
    assert forms._u01(4242, 5) == O.prng_uniform(4242, 5, 1)[0]
