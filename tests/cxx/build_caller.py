"""Build recipe of tests/cxx/mwg_caller (run by __graft_entry__.build()).

A C++ program that includes include/mwgreen.hpp and links
mrchem_amd/libmwgreen.so, the way a C++ code such as MRChem would; the rpath
finds the library relative to the program, so the built tree can move.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def build():
    src = os.path.join(HERE, "mwg_caller.cpp")
    out = os.path.join(HERE, "mwg_caller")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), src, "-o",
           out + ".tmp", "-L" + os.path.join(ROOT, "mrchem_amd"), "-lmwgreen", "-Wl,-rpath,$ORIGIN/../../mrchem_amd"]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
