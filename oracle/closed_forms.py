"""ORACLE - test infrastructure only.

Closed forms that pin the apply without the reference implementation.  All
Gaussians are normalised, g_a(r) = (a/pi)^{3/2} exp(-a r^2); for a pair
(a, A), (b, B): p = ab/(a+b), R = |A - B|.

  Poisson (kernel 1/r):
    V_a(r) = erf(sqrt(a) r) / r
    <g_a(.-A)| 1/r |g_b(.-B)> = erf(sqrt(p) R) / R   (R -> 0: 2 sqrt(p/pi))
  Helmholtz / Yukawa (kernel exp(-mu r)/r):
    <g_a|G|g_b> = U_mu(R; p)
      = [exp(mu^2/(4p) - mu R) erfc(mu/(2 sqrt p) - sqrt(p) R)
         - exp(-p R^2) erfcx(mu/(2 sqrt p) + sqrt(p) R)] / (2R)
    (R -> 0 limit by a symmetric difference); mu -> 0 reproduces erf(sqrt(p)R)/R.
  Gaussian product theorem:
    g_a(r-A) g_b(r-B) = exp(-p R^2) (ab/pi^2)^{3/2} exp(-(a+b)|r-P|^2),
    P = (aA + bB)/(a+b)  ->  c g_{a+b}(r - P) with c = exp(-pR^2)(ab/(pi(a+b)))^{3/2}.

The Yukawa form follows from the 3D convolution of two normalised Gaussians
(a normalised Gaussian of exponent p) with exp(-mu r)/r, done in spherical
coordinates; it is checked against direct radial quadrature in
tests/test_oracle.py.
"""
import math

import numpy as np
from scipy import integrate, special


def coulomb_pair(a, b, R):
    p = a * b / (a + b)
    if R < 1e-12:
        return 2.0 * math.sqrt(p / math.pi)
    return math.erf(math.sqrt(p) * R) / R


def yukawa_pair(a, b, R, mu):
    p = a * b / (a + b)
    if mu == 0.0:
        return coulomb_pair(a, b, R)
    sp = math.sqrt(p)
    x = mu / (2.0 * sp)
    if R < 1e-10:
        # potential at the centre of a normalised Gaussian of exponent p
        return 2.0 * math.sqrt(p / math.pi) - mu * special.erfcx(x)
    y = x - sp * R
    if y > 0.0:
        # exp(x^2 - mu R) erfc(y) = exp(-p R^2) erfcx(y) exactly
        t1 = math.exp(-p * R * R) * special.erfcx(y)
    else:
        t1 = math.exp(x * x - mu * R) * special.erfc(y)
    t2 = math.exp(-p * R * R) * special.erfcx(x + sp * R)
    return (t1 - t2) / (2.0 * R)


def yukawa_pair_quadrature(a, b, R, mu):
    """Direct radial quadrature of <g_a(.-A)|exp(-mu r)/r|g_b(.-B)> (independent check)."""
    p = a * b / (a + b)
    # rho_p(s) at distance s from the centre-of-interaction; potential of the
    # Yukawa kernel from a normalised Gaussian of exponent p at distance R
    def integrand(s):
        if s == 0:
            return 0.0
        # angular average of exp(-mu|R - s|)/|R - s| over directions = shell theorem
        if R == 0:
            shell = math.exp(-mu * s) / s
        else:
            shell = (math.exp(-mu * abs(R - s)) - math.exp(-mu * (R + s))) / (2 * mu * R * s) if mu > 0 else \
                (1.0 / max(R, s))
        return 4 * math.pi * s * s * (p / math.pi) ** 1.5 * math.exp(-p * s * s) * shell
    val, _ = integrate.quad(integrand, 0, 12.0 / math.sqrt(p) + R, limit=400, epsabs=1e-14, epsrel=1e-13)
    return val


def pair_energy(coef, exp, centers, mu=0.0):
    """Sum_ij c_i c_j <g_i|K|g_j> for a Gaussian sum (K = 1/r or exp(-mu r)/r)."""
    coef = np.asarray(coef, dtype=float)
    exp = np.asarray(exp, dtype=float)
    centers = np.asarray(centers, dtype=float).reshape(-1, 3)
    E = 0.0
    for i in range(len(coef)):
        for j in range(len(coef)):
            R = float(np.linalg.norm(centers[i] - centers[j]))
            E += coef[i] * coef[j] * (yukawa_pair(exp[i], exp[j], R, mu) if mu > 0 else
                                      coulomb_pair(exp[i], exp[j], R))
    return E


def coulomb_potential(coef, exp, centers, points):
    """V(r) = sum_i c_i erf(sqrt(a_i)|r - R_i|)/|r - R_i|."""
    pts = np.asarray(points, dtype=float).reshape(-1, 3)
    centers = np.asarray(centers, dtype=float).reshape(-1, 3)
    V = np.zeros(len(pts))
    for c, a, R in zip(coef, exp, centers):
        d = np.linalg.norm(pts - R, axis=1)
        with np.errstate(invalid="ignore", divide="ignore"):
            v = np.where(d > 1e-12, special.erf(np.sqrt(a) * d) / np.where(d > 1e-12, d, 1.0),
                         2.0 * np.sqrt(a / np.pi))
        V += c * v
    return V


def gaussian_product(a, A, b, B):
    """Normalised-Gaussian product g_a(r-A) g_b(r-B) = c g_{a+b}(r-P): returns (c, a+b, P)."""
    A = np.asarray(A, dtype=float)
    B = np.asarray(B, dtype=float)
    p = a * b / (a + b)
    R2 = float(np.sum((A - B) ** 2))
    c = math.exp(-p * R2) * (a * b / (math.pi * (a + b))) ** 1.5
    return c, a + b, (a * A + b * B) / (a + b)
