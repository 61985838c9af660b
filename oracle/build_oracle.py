"""Build recipe for the oracle (test infrastructure only).

Builds oracle/libmworacle_v3.so (AVX2) and oracle/libmworacle_v4.so (AVX-512)
from oracle/mw_oracle.cpp; oracle/__init__.py picks the one the host CPU runs.
The upstream implementation of the apply (MRCPP, external/upstream/
fetch_mrcpp.cmake:36-43) is not in the reference tree, so there is no
oracle/_ref build: the restatement is pinned by closed forms instead.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    src = os.path.join(HERE, "mw_oracle.cpp")
    outs = []
    for tag, march in (("v3", "x86-64-v3"), ("v4", "x86-64-v4")):
        out = os.path.join(HERE, "libmworacle_%s.so" % tag)
        tmp = out + ".tmp"
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=" + march,
               src, "-o", tmp]
        print("+", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    build()
