"""The drop-in C++ API (Scalar/Vector/Operators, DataObjects.H) on the GPU: the
reference's own test programs restated with assertions, and the ns driver's
op-by-op and fused modes against the golden vectors and the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel_l2
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "spectral_upcxx_amd", "lib", "bin")


def run(args, timeout=300):
    print("running", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=timeout)
    return r.returncode, r.stdout


def test_operators_program():
    rc, out = run(["test_operators", 32])
    assert rc == 0, out
    assert "failed" not in out and out.count("passed") >= 14


def test_3d_fft_program():
    rc, out = run(["test_3d_fft", 64])
    assert rc == 0 and "my man!" in out, out


def test_freq_program():
    rc, out = run(["test_freq"])
    assert rc == 0, out
    assert [float(v) for v in out.split()] == list(range(9)) + list(range(-7, 0))


def test_cn_heat_program():
    rc, out = run(["test_cn_heat", 32])
    assert rc == 0 and "cn heat passed" in out, out


def test_heat_program():
    rc, out = run(["test_heat", 32])
    assert rc == 0 and "heat passed" in out, out


def test_constructor_program():
    rc, out = run(["test_constructor"])
    assert rc == 0 and "constructor passed" in out, out


def test_particle_program():
    rc, out = run(["test_particle"])
    assert rc == 0 and "particle passed" in out, out


def test_fft_bench_program():
    rc, out = run(["fft", 64, 3])
    assert rc == 0 and "my man!" in out, out


@pytest.mark.parametrize("n", [20, 40, 17])
def test_programs_any_length(n):
    """the drop-in API at lengths outside the power-of-two family: radix-5
    plans (20, 40) and Bluestein (17): the fft driver's round trip, the
    restated test/3d_fft.cpp and test/operators.cpp"""
    rc, out = run(["fft", n, 2])
    assert rc == 0 and "my man!" in out, out
    rc, out = run(["test_3d_fft", n])
    assert rc == 0 and "my man!" in out, out
    rc, out = run(["test_operators", n])
    assert rc == 0, out
    assert "failed" not in out and out.count("passed") >= 14


def test_ns_driver_any_grid_vs_oracle(tmp_path):
    """exec/ns.cpp through the API on a 20^3 grid (op-by-op), Taylor-Green,
    20 steps, against the oracle"""
    n = 20
    out = tmp_path / "v.bin"
    rc, txt = run(["ns", n, 20, "tg", 0.01, 0.001, "api", out])
    assert rc == 0, txt
    ref = O.NS((n, n, n), 0.01, 0.001)
    ref.set_initial(O.taylor_green((n, n, n)))
    ref.startup()
    ref.steps(20)
    assert rel_l2(_read_v(out, n), ref.velocity()) < 1e-10


def test_ilap_program():
    rc, out = run(["ilap", 32, 3])
    assert rc == 0 and "min time" in out, out


def _read_v(path, n):
    return np.fromfile(path, dtype=np.complex128).reshape(3, n, n, n)


@pytest.mark.parametrize("mode", ["api", "fused"])
def test_ns_driver_tg16_golden(tmp_path, mode):
    out = tmp_path / ("v_%s.bin" % mode)
    rc, txt = run(["ns", 16, 100, "tg", 0.01, 0.001, mode, out])
    assert rc == 0, txt
    g = np.load(GOLDEN + "/tg_16.npz")
    assert rel_l2(_read_v(out, 16), g["step100"]) < 1e-10


@pytest.mark.parametrize("mode", ["api", "fused"])
def test_ns_driver_vortex_ic_vs_oracle(tmp_path, mode):
    """The reference's own IC (three Gaussian vortices, exec/ns.cpp:97-131,
    omega0 kept from the IC), nu = 1e-3, K = 1e-3, 20 steps, against the oracle."""
    n = 32
    out = tmp_path / "v.bin"
    rc, txt = run(["ns", n, 20, "vortex", 0.001, 0.001, mode, out])
    assert rc == 0, txt
    V0, W0 = O.vortices((n, n, n), 0)
    ref = O.NS((n, n, n), 0.001, 0.001)
    ref.set_initial(V0)
    ref.set_vorticity(W0)
    ref.startup()
    ref.steps(20)
    assert rel_l2(_read_v(out, n), ref.velocity()) < 1e-10


def _read_particles(path):
    with open(path) as f:
        assert f.readline().split() == ["x", "y", "z", "E"]
        return np.loadtxt(f)


def test_ns2_driver_vs_oracle_with_particles(tmp_path):
    """exec/ns2.cpp at 32^3 (two vortices, backward Euler, 1000 particles) for
    40 steps: both modes against the oracle's velocity; the particle files of
    the two modes agree and match the oracle's tracers moved through the
    oracle's fields (printed with 6 significant digits, as the reference)."""
    n, steps, npart = 32, 40, 1000
    outs = {}
    for mode in ("api", "fused"):
        d = tmp_path / mode
        d.mkdir()
        out = tmp_path / ("v_%s.bin" % mode)
        rc, txt = run(["ns2", n, steps, 0.001, 0.001, mode, npart, d, out])
        assert rc == 0, txt
        outs[mode] = (_read_v(out, n), d)
    V0, W0 = O.vortices((n, n, n), 1)
    ref = O.NS((n, n, n), 0.001, 0.001)
    ref.set_scheme(1)
    ref.set_initial(V0)
    ref.set_vorticity(W0)
    ref.startup()
    loc = O.particles_seed(npart, 1)
    files = []
    for t in range(1, steps + 1):
        ref.steps(1)
        U = np.stack([O.inverse(c) for c in ref.velocity()])
        _, en = O.particles_move(loc, U, 0.001)
        if (t - 1) % max(1, steps // 300) == 0:
            files.append(np.column_stack([loc, en]))
    for mode in ("api", "fused"):
        assert rel_l2(outs[mode][0], ref.velocity()) < 1e-10, mode
    assert len(files) == steps
    for i in (0, steps // 2, steps - 1):
        a = _read_particles(outs["api"][1] / ("p%d.3D" % i))
        b = _read_particles(outs["fused"][1] / ("p%d.3D" % i))
        assert a.shape == (npart, 4)
        # the two modes' fields differ at rounding level; 6 printed digits
        # can then differ in the last place only
        assert np.allclose(a, b, rtol=1.1e-5, atol=1e-12), i
        assert np.allclose(a, files[i], rtol=1.1e-5, atol=1e-12), i


def test_ns_driver_api_matches_fused(tmp_path):
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    assert run(["ns", 32, 30, "sol:77:3", 0.002, 0.002, "api", a])[0] == 0
    assert run(["ns", 32, 30, "sol:77:3", 0.002, 0.002, "fused", b])[0] == 0
    assert rel_l2(_read_v(a, 32), _read_v(b, 32)) < 1e-12
