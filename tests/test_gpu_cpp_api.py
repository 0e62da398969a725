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


def test_fft_bench_program():
    rc, out = run(["fft", 64, 3])
    assert rc == 0 and "my man!" in out, out


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


def test_ns_driver_vortex_ic_vs_oracle(tmp_path):
    """The reference's own IC (three Gaussian vortices, exec/ns.cpp:97-131),
    nu = 1e-3, K = 1e-3, 20 steps: API mode against the oracle."""
    n = 32
    out = tmp_path / "v.bin"
    rc, txt = run(["ns", n, 20, "vortex", 0.001, 0.001, "api", out])
    assert rc == 0, txt
    # the same IC on the host
    x = 2 * np.pi * np.arange(n) / n
    X, Y = x[None, None, :], x[None, :, None]
    g = lambda xc, yc: np.exp(-(X - xc) ** 2 / (np.pi / 8)) * np.exp(-(Y - yc) ** 2 / (np.pi / 8))
    d = np.pi / 2
    wz = 50 * (g(np.pi - d, np.pi - d) + g(np.pi + d, np.pi - d) + g(np.pi, np.pi + d))
    wz = O.forward(np.broadcast_to(wz, (n, n, n)).astype(np.complex128))
    kx = np.array([O.freq(i, n) for i in range(n)])
    k2 = kx[None, None, :] ** 2 + kx[None, :, None] ** 2
    psi = np.where(k2 != 0, -1.0 / np.where(k2 != 0, k2, 1) * wz, 0)
    grad = O.gradient(np.ascontiguousarray(psi))
    V0 = np.stack([-1.0 * grad[1], grad[0], np.zeros_like(psi)])
    ref = O.NS((n, n, n), 0.001, 0.001)
    ref.set_initial(V0)
    ref.startup()
    ref.steps(20)
    assert rel_l2(_read_v(out, n), ref.velocity()) < 1e-10


def test_ns_driver_api_matches_fused(tmp_path):
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    assert run(["ns", 32, 30, "sol:77:3", 0.002, 0.002, "api", a])[0] == 0
    assert run(["ns", 32, 30, "sol:77:3", 0.002, 0.002, "fused", b])[0] == 0
    assert rel_l2(_read_v(a, 32), _read_v(b, 32)) < 1e-12
