"""The reference's golden fixture for the Coulomb apply
(tests/qmoperators/coulomb_operator.cpp:42-84): five hydrogenic orbitals
(1s, 2s, 2p; Z = 1, doubly occupied) projected at prec 1e-3 on the
reference test's MRA (tests/unit_tests.cpp:66-84: order 5, interpolating,
root scale -5, boxes {2,2,2}, corner {-1,-1,-1}; maximum depth 20 here
against the reference's 25, which the trees never reach), the density
rho = 2 sum_i |phi_i|^2, the potential V = P[rho] applied as
CoulombPotential::setupGlobalPotential does (CoulombPotential.cpp:115-133:
relative precision prec / ||rho||), and <phi_i|V|phi_i>.

The fixture's values are MRCPP's at prec 1e-3 (1.3 %, 0.3 % and 0.07 % from
the exact values: the discretisation at that precision).  The GPU path
reproduces them to about 1e-6 relative, so the same grids, thresholds and
screening decisions are taken; the test holds it to 1e-5 of the fixture (the
reference test's own tolerance, 1e-8, would need MRCPP's bits) and to 2 % of
the exact values, the Slater-Condon F0/F2 integrals of hydrogenic orbitals, <1s|J|1s> = 2(5/8 + 17/81 + 3 59/243),
<2s|J|2s> = 2(17/81 + 77/512 + 3 83/512),
<2p|J|2p> = 2(59/243 + 83/512 + (93 + 4 45/25)/512 + 2 (93 - 2 45/25)/512).
The orbitals come from an adaptive point-value projection restated in numpy
(tests/mw_project_ref.py), which reproduces the library's own Gaussian
projection exactly.
"""
import numpy as np
import pytest

import mrchem_amd as mw
import mw_project_ref as pr

pytestmark = pytest.mark.gpu

# coulomb_operator.cpp:78-84, E_P(i, i) at prec 1e-3
FIXTURE = {"1s": 3.1676468518, "2s": 1.6980679074, "2p": 1.8983578764}
EXACT = {
    "1s": 2 * (5 / 8 + 17 / 81 + 3 * 59 / 243),
    "2s": 2 * (17 / 81 + 77 / 512 + 3 * 83 / 512),
    "2p": 2 * (59 / 243 + 83 / 512 + (93 + 4 * 45 / 25) / 512 + 2 * (93 - 2 * 45 / 25) / 512),
}


def _orbitals():
    """Hydrogenic orbitals (Z = 1) at the origin, as HydrogenFunction builds
    them (src/analyticfunctions/HydrogenFunction.cpp:32-80): R_nl Y_lm with
    real spherical harmonics."""
    r = lambda X, Y, Z: np.sqrt(X * X + Y * Y + Z * Z)
    y00 = 1.0 / np.sqrt(4 * np.pi)
    y1 = np.sqrt(3.0 / (4 * np.pi))
    with np.errstate(invalid="ignore", divide="ignore"):
        return {
            "1s": lambda X, Y, Z: 2.0 * np.exp(-r(X, Y, Z)) * y00,
            "2s": lambda X, Y, Z: (2.0 - r(X, Y, Z)) * np.exp(-r(X, Y, Z) / 2) / (2 * np.sqrt(2.0)) * y00,
            "2px": lambda X, Y, Z: X * np.exp(-r(X, Y, Z) / 2) / (2 * np.sqrt(6.0)) * y1,
            "2py": lambda X, Y, Z: Y * np.exp(-r(X, Y, Z) / 2) / (2 * np.sqrt(6.0)) * y1,
            "2pz": lambda X, Y, Z: Z * np.exp(-r(X, Y, Z) / 2) / (2 * np.sqrt(6.0)) * y1,
        }


def test_coulomb_hydrogen_fixture(ctx):
    prec = 1e-3
    mra = mw.MultiResolutionAnalysis(order=5, basis="interpolating", min_scale=-5, corner=(-1, -1, -1),
                                     boxes=(2, 2, 2), scaling_factor=1.0, max_depth=20)
    P = mw.PoissonOperator(mra, prec)
    phi = {k: mw.FunctionTree(mra).set_nodes(*pr.project(mra, f, prec, norm=1.0)) for k, f in _orbitals().items()}
    for k, t in phi.items():
        assert abs(t.norm() - 1.0) < 10 * prec, (k, t.norm())
    sq = {k: mw.multiply(prec / 10, mw.FunctionTree(mra), t, t, abs_prec=True, ctx=ctx) for k, t in phi.items()}
    rho = mw.add(prec / 10, mw.FunctionTree(mra), [2.0] * len(sq), list(sq.values()), abs_prec=True, ctx=ctx)
    V = mw.FunctionTree(mra)
    mw.apply(prec / rho.norm(), V, P, rho, ctx=ctx)
    got = {"1s": V.dot(sq["1s"]), "2s": V.dot(sq["2s"]),
           "2p": np.mean([V.dot(sq[k]) for k in ("2px", "2py", "2pz")])}
    for k in ("1s", "2s", "2p"):
        print("<%s|J|%s> = %.10f: fixture %.10f (%+.2e), exact %.10f (%+.2f %%; fixture %+.2f %%)" %
              (k, k, got[k], FIXTURE[k], got[k] / FIXTURE[k] - 1, EXACT[k], 100 * (got[k] / EXACT[k] - 1),
               100 * (FIXTURE[k] / EXACT[k] - 1)))
    print("orbitals %s nodes, density %d, potential %d nodes, ||rho|| = %.4f (10 electrons)" %
          ({k: t.size() for k, t in phi.items()}, rho.size(), V.size(), rho.norm()))
    for k in ("1s", "2s", "2p"):
        assert abs(got[k] / EXACT[k] - 1) < 0.02
        assert abs(got[k] / FIXTURE[k] - 1) < 1e-5
