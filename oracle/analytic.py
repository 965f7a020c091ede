"""Closed-form potentials that pin the apply (test infrastructure only).

Kernels follow the reference's normalisation: the Poisson kernel is 1/r
(MRChem uses P[rho] directly as the Coulomb potential,
src/qmoperators/two_electron/CoulombPotential.cpp:85) and the Helmholtz kernel
is exp(-mu r)/r (the -1/(2 pi) rescale of src/scf_solver/HelmholtzVector.cpp:81,135
is applied by the caller).

For an unnormalised s Gaussian g = exp(-a |r-R|^2) with rho = |r-R|:
  (1/r) * g            = (pi/a)^{3/2} erf(sqrt(a) rho) / rho
  (exp(-mu r)/r) * g   = (pi/a)^{3/2} E(rho)/(2 rho),
  E(rho) = exp(c^2) [exp(-mu rho) erfc(c - sqrt(a) rho) - exp(mu rho) erfc(c + sqrt(a) rho)],
  c = mu/(2 sqrt(a)),
which reduces to the erf form as mu -> 0.  A p-type primitive
(x-X) g = (1/(2a)) dg/dX, and convolution commutes with d/dX, so its potential
is -(1/(2a)) Phi'(rho) (x-X)/rho.

Hydrogenic Slater-Condon integrals (Z=1) pin the Coulomb expectation values of
tests/qmoperators/coulomb_operator.cpp:76-82: with F0(1s,1s)=5/8,
F0(1s,2s)=17/81, F0(1s,2p)=59/243, F0(2s,2s)=77/512, F0(2s,2p)=83/512,
F0(2p,2p)=93/512, F2(2p,2p)=45/512.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf, erfcx


def _phi_poisson(a, rho):
    rho = np.asarray(rho, float)
    sa = math.sqrt(a)
    small = rho < 1e-8
    r = np.where(small, 1.0, rho)
    val = erf(sa * r) / r
    return np.where(small, 2.0 * sa / math.sqrt(math.pi), val)


def _dphi_poisson(a, rho):
    rho = np.asarray(rho, float)
    sa = math.sqrt(a)
    return 2.0 * sa / math.sqrt(math.pi) * np.exp(-a * rho * rho) / rho - erf(sa * rho) / (rho * rho)


def _E_helm(a, mu, rho):
    sa = math.sqrt(a)
    c = mu / (2.0 * sa)
    x1 = c - sa * rho
    x2 = c + sa * rho
    g = np.exp(-a * rho * rho)
    # term1 = exp(c^2 - mu rho) erfc(x1), stable for either sign of x1
    t1 = np.where(x1 >= 0.0, g * erfcx(np.maximum(x1, 0.0)),
                  np.exp(c * c - mu * rho) * (2.0 - np.exp(-np.minimum(x1, 0.0) ** 2)
                                              * erfcx(-np.minimum(x1, 0.0))))
    t2 = g * erfcx(x2)
    return t1, t2


def _phi_helm(a, mu, rho):
    rho = np.asarray(rho, float)
    t1, t2 = _E_helm(a, mu, rho)
    return (t1 - t2) / (2.0 * rho)


def _dphi_helm(a, mu, rho):
    rho = np.asarray(rho, float)
    t1, t2 = _E_helm(a, mu, rho)
    E = t1 - t2
    dE = -mu * (t1 + t2) + 4.0 * math.sqrt(a) / math.sqrt(math.pi) * np.exp(-a * rho * rho)
    return dE / (2.0 * rho) - E / (2.0 * rho * rho)


def potential(g, pts, mu: float = 0.0):
    """Kernel (1/r or exp(-mu r)/r) convolved with a GaussSum whose terms are s
    or single-axis p primitives (degree <= 1 in one dimension)."""
    pts = np.atleast_2d(np.asarray(pts, float))
    out = np.zeros(len(pts))
    for i in range(g.n):
        a = g.alpha[i]
        R = g.center[i]
        u = pts - R
        rho = np.sqrt((u * u).sum(1))
        pref = g.coef[i] * (math.pi / a) ** 1.5
        P = g.poly[i]
        degs = [int(np.max(np.nonzero(P[d])[0])) if np.any(P[d]) else -1 for d in range(3)]
        if any(dg < 0 for dg in degs):
            continue
        if all(dg == 0 for dg in degs):
            c = P[0, 0] * P[1, 0] * P[2, 0]
            phi = _phi_poisson(a, rho) if mu == 0.0 else _phi_helm(a, mu, rho)
            out += pref * c * phi
        else:
            lin = [d for d in range(3) if degs[d] == 1]
            if len(lin) != 1 or any(degs[d] > 1 for d in range(3)):
                raise ValueError("closed form covers s and single-axis p primitives only")
            d = lin[0]
            others = [q for q in range(3) if q != d]
            c0 = P[others[0], 0] * P[others[1], 0]
            # poly_d = p0 + p1 (x-X): s part + p part
            s_part = c0 * P[d, 0]
            p_part = c0 * P[d, 1]
            phi = _phi_poisson(a, rho) if mu == 0.0 else _phi_helm(a, mu, rho)
            dphi = _dphi_poisson(a, rho) if mu == 0.0 else _dphi_helm(a, mu, rho)
            out += pref * (s_part * phi - p_part / (2.0 * a) * dphi * u[:, d] / rho)
    return out


def erf_potential(parts, pts):
    """sum_i q_i erf(sqrt(a_i) |r-R_i|)/|r-R_i| for normalised charges (q, a, R)."""
    pts = np.atleast_2d(pts)
    out = np.zeros(len(pts))
    for q, a, R in parts:
        rho = np.sqrt(((pts - R) ** 2).sum(1))
        out += q * _phi_poisson(a, rho)
    return out


# Slater-Condon F0 integrals for Z=1 hydrogenic orbitals
F0 = {("1s", "1s"): 5 / 8, ("1s", "2s"): 17 / 81, ("1s", "2p"): 59 / 243, ("2s", "2s"): 77 / 512,
      ("2s", "2p"): 83 / 512, ("2p", "2p"): 93 / 512}
F2_2p2p = 45 / 512


def _f0(a, b):
    return F0.get((a, b), F0.get((b, a)))


def coulomb_expectations():
    """<phi|J|phi> for rho = 2(1s^2 + 2s^2 + 2px^2 + 2py^2 + 2pz^2), J = (1/r) * rho.

    Returns dict with keys '1s', '2s', '2p'.  The 2p-2p' contributions use the
    angular coupling: <2p_m 2p_m | 2p_n 2p_n> = F0 + 4/25 F2 (m=n), F0 - 2/25 F2 (m!=n)."""
    e1s = 2 * (_f0("1s", "1s") + _f0("1s", "2s") + 3 * _f0("1s", "2p"))
    e2s = 2 * (_f0("2s", "1s") + _f0("2s", "2s") + 3 * _f0("2s", "2p"))
    same = _f0("2p", "2p") + 4.0 / 25.0 * F2_2p2p
    diff = _f0("2p", "2p") - 2.0 / 25.0 * F2_2p2p
    e2p = 2 * (_f0("2p", "1s") + _f0("2p", "2s") + same + 2 * diff)
    return {"1s": e1s, "2s": e2s, "2p": e2p}


# reference regression values (tests/qmoperators/coulomb_operator.cpp:76-82), prec 1e-3
REFERENCE_E_P = {"1s": 3.1676468518, "2s": 1.6980679074, "2p": 1.8983578764, "1s2s": 0.262570199}
