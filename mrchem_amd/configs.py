"""Seeded synthetic workloads for the Poisson / Helmholtz apply.

Random numbers come from a splitmix64 stream (one 64-bit state per workload,
seeded with the workload seed); ``uniform()`` takes the top 53 bits.  Draw
order is part of the workload definition and is documented per builder.

Workloads (box, order, precision as used by the callers of ``mrcpp::apply``):
  * ``density3``: 3 normalised isotropic Gaussian charges, centres U[-2,2]^3,
    exponents log-U[1,50], charges U[0.5,2]; box [-16,16]^3, k=5, prec 1e-4,
    Poisson.  Closed form: sum_i q_i erf(sqrt(a_i) r_i)/r_i.
  * ``orbitals``: N orbitals, each a sum of 6 s/p primitives on seeded nuclei,
    exponents log-U[0.3,30], coefficients N(0,1), Gram-Schmidt orthonormalised
    with exact overlaps; Helmholtz mu ~ U[0.5,2] (5 orbitals, 3 nuclei within
    2 bohr, k=7, prec 1e-6, box [-20,20]^3).
"""
from __future__ import annotations

import math

import numpy as np

from .gaussians import GaussSum, compact, gram_schmidt, normalized_s, primitive

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def normal(self) -> float:
        # Box-Muller, first output only
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


# MRA descriptions: (root scale, corner, boxes, scaling factor)
BOX16 = dict(n0=-4, corner=(-1, -1, -1), boxes=(2, 2, 2), sfac=1.0)      # [-16,16]^3
BOX20 = dict(n0=-4, corner=(-1, -1, -1), boxes=(2, 2, 2), sfac=1.25)     # [-20,20]^3
BOX32 = dict(n0=-5, corner=(-1, -1, -1), boxes=(2, 2, 2), sfac=1.0)      # [-32,32]^3


def density3(seed: int):
    """3 normalised Gaussian charges.  Draws per charge: centre x,y,z; exponent; charge.

    Returns (GaussSum, list of (q, a, R))."""
    rng = SplitMix64(seed)
    parts = []
    g = GaussSum()
    for _ in range(3):
        R = np.array([rng.uniform(-2.0, 2.0) for _ in range(3)])
        a = rng.log_uniform(1.0, 50.0)
        q = rng.uniform(0.5, 2.0)
        parts.append((q, a, R))
        g = g + normalized_s(q, a, R)
    return g, parts


def orbitals(seed: int, n_orb: int = 5, n_nuc: int = 3, nuc_box: float = 2.0, n_prim: int = 6):
    """Synthetic orthonormal orbitals.

    Draws: nuclei (x,y,z each U[-nuc_box/2, nuc_box/2]... see below), then per
    orbital and primitive: nucleus index, type (s,px,py,pz with prob 1/4 each),
    exponent log-U[0.3,30], coefficient N(0,1); finally one mu ~ U[0.5,2] per
    orbital.  ``nuc_box`` is the edge of the cube holding the nuclei (2 bohr:
    nuclei within 2 bohr of each other along each axis).

    Returns (list of GaussSum orbitals, nuclei array, list of mu)."""
    rng = SplitMix64(seed)
    half = 0.5 * nuc_box
    nuc = np.array([[rng.uniform(-half, half) for _ in range(3)] for _ in range(n_nuc)])
    kinds = ("s", "px", "py", "pz")
    raw = []
    for _ in range(n_orb):
        f = GaussSum()
        for _ in range(n_prim):
            c = min(int(rng.uniform() * n_nuc), n_nuc - 1)
            kind = kinds[min(int(rng.uniform() * 4), 3)]
            a = rng.log_uniform(0.3, 30.0)
            coef = rng.normal()
            f = f + primitive(kind, a, nuc[c], coef)
        raw.append(compact(f))
    orbs = [compact(o) for o in gram_schmidt(raw)]
    mus = [rng.uniform(0.5, 2.0) for _ in range(n_orb)]
    return orbs, nuc, mus
