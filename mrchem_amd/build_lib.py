"""Build recipe for the native libraries (run by __graft_entry__.build()).

libmwgreen.so  host C++ (g++) + HIP kernels for gfx950 (hipcc), C ABI in
               include/mwgreen.h.
The oracle (test infrastructure) is built by oracle/build_oracle.py.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["mw_math.cpp", "mw_kernel.cpp", "mw_operator.cpp", "mw_tree.cpp", "mw_plan.cpp",
                "mwgreen_capi.cpp"]
HIP_SOURCES = ["mw_device.hip", "mw_opbuild.hip", "mw_multiply.hip"]
# gfx950 only; no packed FP32 instructions.  MFMA results go to VGPRs directly
# (the AGPR form costs four v_accvgpr_read per tile: -9 % on the FP64 apply).
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-mllvm", "-amdgpu-mfma-vgpr-form"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-march=x86-64-v3", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        _run([os.path.join(ROCM, "bin", "hipcc")] + HIPCC_FLAGS +
             ["-I" + CSRC, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    out = os.path.join(HERE, "libmwgreen.so")
    tmp = out + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
