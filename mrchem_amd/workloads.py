"""Synthetic Gaussian-sum workloads of the five benchmark configurations.

One project PRNG, SplitMix64 (Steele, Lea, Flood 2014), gives doubles from
the top 53 bits and normals by Box-Muller, so the inputs are reproducible from
the committed seeds.  All Gaussians are normalised,
g_a(r) = (a/pi)^{3/2} exp(-a |r - R|^2); coefficients multiply them.  The world
box is [-20, 20]^3: root scale -4, 2x2x2 root boxes from corner (-1,-1,-1),
scaling factor 1.25 (the reference's world_size 5 box of src/mrenv.cpp:99-105
stretched by 1.25).
"""
import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def normal(self):
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


WORLD = dict(min_scale=-4, corner=(-1, -1, -1), boxes=(2, 2, 2), scaling_factor=1.25)
SEEDS = {"c1": 1001, "c2": 1002, "c3": 1003, "c4": 1004, "c5": 1005}


def gaussian_sum(rng, n, exp_lo, exp_hi, box, coef="normal"):
    c = np.array([rng.normal() if coef == "normal" else 1.0 for _ in range(n)])
    a = np.array([rng.log_uniform(exp_lo, exp_hi) for _ in range(n)])
    R = np.array([[rng.uniform(-box, box) for _ in range(3)] for _ in range(n)])
    return c, a, R


def config(name):
    """Returns the workload description of one benchmark configuration."""
    rng = SplitMix64(SEEDS[name])
    if name == "c1":
        r0 = np.array([[rng.uniform(-1, 1) for _ in range(3)]])
        return dict(name="c1_poisson_single_gaussian_k7", kind="poisson", order=7, prec=1e-5, abs_prec=False,
                    densities=[(np.array([1.0]), np.array([10.0]), r0)])
    if name == "c2":
        c, a, R = gaussian_sum(rng, 10, 0.5, 50.0, 3.0)
        return dict(name="c2_helmholtz_mu1.2_10gauss_k9", kind="helmholtz", mu=[1.2], order=9, prec=1e-6,
                    abs_prec=True, densities=[(c, a, R)])
    if name == "c3":
        orbs = [gaussian_sum(rng, 6, 0.3, 30.0, 4.0) for _ in range(16)]
        dens = [pair_density(orbs[i], orbs[j]) for i in range(16) for j in range(i, 16)]
        return dict(name="c3_exchange_136_pairs_k7", kind="poisson", order=7, prec=1e-5, abs_prec=True,
                    densities=dens, orbitals=orbs)
    if name == "c4":
        c, a, R = gaussian_sum(rng, 8, 1.0, 100.0, 3.0)
        return dict(name="c4_poisson_8gauss_k9", kind="poisson", order=9, prec=1e-6, abs_prec=False,
                    densities=[(c, a, R)])
    if name == "c5":
        orbs = [gaussian_sum(rng, 8, 0.3, 30.0, 6.0) for _ in range(64)]
        mus = [rng.uniform(0.5, 2.5) for _ in range(64)]
        return dict(name="c5_helmholtz_64_orbitals_k11", kind="helmholtz", mu=mus, order=11, prec=1e-7,
                    abs_prec=True, densities=orbs)
    raise KeyError(name)


def pair_density(f, g):
    """rho = f g for two Gaussian sums, exactly, by the Gaussian product theorem."""
    (cf, af, Rf), (cg, ag, Rg) = f, g
    c, a, R = [], [], []
    for i in range(len(cf)):
        for j in range(len(cg)):
            p = af[i] * ag[j] / (af[i] + ag[j])
            d2 = float(np.sum((Rf[i] - Rg[j]) ** 2))
            c.append(cf[i] * cg[j] * math.exp(-p * d2) * (af[i] * ag[j] / (math.pi * (af[i] + ag[j]))) ** 1.5)
            a.append(af[i] + ag[j])
            R.append((af[i] * Rf[i] + ag[j] * Rg[j]) / (af[i] + ag[j]))
    return np.array(c), np.array(a), np.array(R)
