"""Build the gfx950 HIP library and the C++ Scalar/Vector/Operators layer in-tree.

    python -m spectral_upcxx_amd.build        (or __graft_entry__.build())

Outputs (git-ignored, shipped with the tree):
    spectral_upcxx_amd/lib/libspectral_hip.so   C ABI of include/spectral_hip.h
    spectral_upcxx_amd/lib/libspectral_api.so   C++ API (DataObjects.H) on top of it
    spectral_upcxx_amd/lib/bin/*                C++ drivers and API tests
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CPP = os.path.join(PKG, "cpp")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(LIB, "obj")
BIN = os.path.join(LIB, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["runtime.hip", "fft3d.hip", "fft_any.hip", "ops.hip", "ns_step.hip", "tune.hip", "mlp.hip", "les.hip", "particles.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # packed FP32 instructions are kept off (v_pk_* forms are not used)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
API_SOURCES = ["Scalar.cpp", "Vector.cpp", "Operators.cpp", "FFT1D.cpp", "Particle.cpp"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-I" + CPP, "-I" + os.path.join(ROOT, "include")]
EXECS = {
    "ns": "exec/ns.cpp",
    "ns2": "exec/ns2.cpp",
    "fft": "exec/fft.cpp",
    "ilap": "exec/ilap.cpp",
    "test_operators": "test/operators.cpp",
    "test_3d_fft": "test/3d_fft.cpp",
    "test_freq": "test/freq.cpp",
    "test_cn_heat": "test/cn_heat.cpp",
    "test_heat": "test/heat.cpp",
    "test_constructor": "test/constructor.cpp",
    "test_particle": "test/particle.cpp",
}

LIB_HIP = os.path.join(LIB, "libspectral_hip.so")
LIB_API = os.path.join(LIB, "libspectral_api.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), r.stdout[-8000:]))
    return r.stdout


def build_hip(jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB_HIP] + objs)
    return LIB_HIP


def build_api():
    os.makedirs(BIN, exist_ok=True)
    srcs = [os.path.join(CPP, s) for s in API_SOURCES]
    _run([CXX] + CXX_FLAGS + ["-shared", "-o", LIB_API] + srcs +
         ["-L" + LIB, "-lspectral_hip", "-Wl,-rpath,$ORIGIN"])
    cmds = []
    for name, src in EXECS.items():
        cmds.append([CXX] + CXX_FLAGS + ["-o", os.path.join(BIN, name), os.path.join(CPP, src),
                                        "-L" + LIB, "-lspectral_api", "-lspectral_hip",
                                        "-Wl,-rpath,$ORIGIN/.."])
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_run, cmds))
    return LIB_API


def build_all():
    build_hip()
    if all(os.path.exists(os.path.join(CPP, s)) for s in API_SOURCES):
        build_api()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "hip":
        build_hip()
    elif what == "api":
        build_api()
    else:
        build_all()
    print("built", LIB_HIP)
