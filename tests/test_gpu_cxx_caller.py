"""The compiled C++ caller of the C ABI on the GPU.

tests/cxx/mwg_caller.cpp drives libmwgreen.so through include/mwgreen.hpp on
a stand-in tree type, as a C++ code such as MRChem would: the Coulomb form
(CoulombPotential.cpp:126-131) against a direct mwg_apply and the closed-form
energy, the exchange form with precTrees on both parts of a complex pair
density (ExchangePotential.cpp:196-205), HelmholtzVector::apply over three
orbitals with distinct mu, one complex and conjugated, as one batch with the
-1/(2 pi) and sign/(2 pi) scales (HelmholtzVector.cpp:141-159), and an
argument error through mwgreen::Error.  It runs as its own process.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cxx_caller_on_the_gpu():
    exe = os.path.join(ROOT, "tests", "cxx", "mwg_caller")
    assert os.path.exists(exe), "tests/cxx/mwg_caller is not built; run __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mwg_caller ok" in r.stdout and "FAIL" not in r.stdout
