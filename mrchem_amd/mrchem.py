"""MRChem's callers of the Green's-function apply, over the batched device apply.

Mirrors (same names, argument meaning and error behaviour):
  * ``HelmholtzVector``  — src/scf_solver/HelmholtzVector.cpp:50-184
      lambda > 0 is clamped to -0.5 (:50-56); mu_i = sqrt(-2 lambda_i) (:169);
      operator()(Phi): out_i = -1/(2 pi) H_i[phi_i] (:67-104, rescale :81);
      operator()(V, Phi, Psi): out_i = H_i[-1/(2 pi) (V phi_i + psi_i)] (:118-159).
    The reference builds one operator and runs one apply per orbital; here the
    operators are built once per mu and all orbitals go through one batched apply.
  * ``CoulombPotential`` — src/qmoperators/two_electron/CoulombPotential.cpp:44-90
      rho = sum_i occ_i phi_i^2 (density_utils.cpp:101-133); abs_prec = prec/||rho||
      (:80); V = P[rho] (:85).
  * ``ExchangePotential`` — src/qmoperators/two_electron/ExchangePotential.cpp:172-424
      pair products phi_i phi_j (i <= j), pair precision min(scaled(i,j), scaled(j,i)),
      skipped above 1 and capped at 0.1 (:306-308, :379-387), spin factors (:415-424),
      K[i] += f_ij phi_j V_ij / ||phi_j||^2 and K[j] += f_ji phi_i V_ij / ||phi_i||^2
      (:287-347); all pair Poissons of one setup in one batched apply.
  * orbital vector algebra — src/qmfunctions/orbital_utils.cpp: ``add`` (:128-144),
      ``rotate`` (:151-183; out_i = sum_j U_ij phi_j), ``calc_overlap_matrix``,
      ``calc_lowdin_matrix`` (:436-449, S^-1/2 by math_utils::hermitian_matrix_pow,
      eigenvalues below MachineZero dropped), ``orthonormalize`` (:578-582), ``get_norms``.
      The linear combinations are exact sums on the union grid of the terms (the
      reference's numerically exact mode, no crop).
  * ``helmholtz_step`` — one SCF orbital update of OrbitalOptimizer::optimize
      (src/scf_solver/OrbitalOptimizer.cpp:148-170): Lambda = diag(F) (HelmholtzVector
      clamps positive entries to -0.5), Psi = rotate(Lambda - F, Phi), Phi' = H(V, Phi, Psi),
      U = orthonormalize(Phi'), F <- U F U^T, dPhi = Phi' - Phi, errors = ||dPhi_i||,
      Phi <- Phi + dPhi (no KAIN acceleration).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import mrcpp as M


@dataclass
class Orbital:
    """A real orbital: a function tree plus occupation and spin ('p' paired, 'a', 'b')."""
    tree: M.FunctionTree
    occ: float = 2.0
    spin: str = "p"

    def squared_norm(self) -> float:
        return self.tree.getSquareNorm()


class HelmholtzVector:
    def __init__(self, prec: float, lambdas, mra: M.MultiResolutionAnalysis, flags: int = M.MWG_APPLY_FP64):
        self.prec = prec
        self.mra = mra
        self.flags = flags
        self.lambdas = np.array([(-0.5 if l > 0.0 else l) for l in lambdas], float)
        self._ops = None
        self.stats = None

    def mu(self, i: int) -> float:
        lam = self.lambdas[i]
        if -2.0 * lam < 0.0:
            raise M.MwgError("Mu cannot be complex")
        return math.sqrt(-2.0 * lam)

    def operators(self):
        if self._ops is None:
            self._ops = [M.HelmholtzOperator(self.mra, self.mu(i), self.prec) for i in range(len(self.lambdas))]
        return self._ops

    def apply(self, inputs):
        outs = [M.FunctionTree(self.mra) for _ in inputs]
        self.stats = M.apply_batch(self.prec, outs, self.operators(), inputs, flags=self.flags)
        return outs

    def __call__(self, Phi, V: M.FunctionTree | None = None, Psi=None):
        if len(Phi) != len(self.lambdas):
            raise M.MwgError("OrbitalVector size mismatch")
        if V is None:
            outs = self.apply([p.tree for p in Phi])
            for o in outs:
                o.rescale(-1.0 / (2.0 * math.pi))
        else:
            if Psi is not None and len(Psi) != len(Phi):
                raise M.MwgError("OrbitalVector size mismatch")
            rhs = []
            for i, p in enumerate(Phi):
                vphi = M.multiply(self.prec, M.FunctionTree(self.mra), 1.0, V, p.tree)
                if Psi is not None:
                    vphi = M.add(M.FunctionTree(self.mra), 1.0, vphi, 1.0, Psi[i].tree)
                vphi.rescale(-1.0 / (2.0 * math.pi))
                rhs.append(vphi)
            outs = self.apply(rhs)
        return [Orbital(o, p.occ, p.spin) for o, p in zip(outs, Phi)]


def density(prec: float, Phi, mra) -> M.FunctionTree:
    """rho = sum_i occ_i phi_i^2 (src/qmfunctions/density_utils.cpp:65-133)."""
    rho = None
    for p in Phi:
        sq = M.multiply(prec, M.FunctionTree(mra), p.occ, p.tree, p.tree)
        rho = sq if rho is None else M.add(M.FunctionTree(mra), 1.0, rho, 1.0, sq)
    return rho


class CoulombPotential:
    def __init__(self, P: M.PoissonOperator, Phi, mra):
        self.poisson = P
        self.orbitals = Phi
        self.mra = mra
        self.potential = None
        self.density = None
        self.stats = None

    def setup(self, prec: float, flags: int = M.MWG_APPLY_FP64):
        if self.poisson is None:
            raise M.MwgError("Poisson operator not initialized")
        if self.potential is not None:
            raise M.MwgError("Potential not properly cleared")
        self.density = density(prec, self.orbitals, self.mra)
        abs_prec = prec / math.sqrt(self.density.getSquareNorm())
        self.potential = M.FunctionTree(self.mra)
        self.stats = M.apply(abs_prec, self.potential, self.poisson, self.density, flags=flags)
        return self.potential

    def clear(self):
        self.potential = None
        self.density = None

    def __call__(self, phi: Orbital, prec: float) -> Orbital:
        if self.potential is None:
            raise M.MwgError("Coulomb potential not set up")
        return Orbital(M.multiply(prec, M.FunctionTree(self.mra), 1.0, self.potential, phi.tree), phi.occ, phi.spin)


def spin_factor(phi_i: Orbital, phi_j: Orbital) -> float:
    """ExchangePotential::getSpinFactor (ExchangePotential.cpp:415-424)."""
    if phi_j.spin == "p":
        return 1.0
    if phi_i.spin == "p":
        return 0.5
    return 1.0 if phi_i.spin == phi_j.spin else 0.0


class ExchangePotential:
    def __init__(self, P: M.PoissonOperator, Phi, mra, screen: bool = True):
        self.poisson = P
        self.orbitals = Phi
        self.mra = mra
        self.screen = screen
        n = len(Phi)
        self.tot_norms = np.zeros(n)
        self.part_norms = np.zeros((n, n))
        self.exchange = None
        self.apply_prec = -1.0
        self.stats = None
        self.pair_potentials = {}

    def scaled_precision(self, i: int, j: int) -> float:
        p = self.apply_prec
        if self.screen:
            t = self.tot_norms[i]
            pn = max(self.part_norms[i, j], self.part_norms[j, i])
            if t > 0.0:
                p *= t / pn if pn > 0.0 else math.inf
        return p

    def pair_list(self):
        """(i, j, prec) for i <= j in the order of the reference's setupInternal."""
        n = len(self.orbitals)
        pairs = []
        for i in range(n):
            pairs.append((i, i, min(self.scaled_precision(i, i), 1e-1)))
        for i in range(n):
            for j in range(i + 1, n):
                fi = spin_factor(self.orbitals[i], self.orbitals[j])
                fj = spin_factor(self.orbitals[j], self.orbitals[i])
                if abs(fi) < 1e-14 or abs(fj) < 1e-14:
                    continue
                p = min(self.scaled_precision(i, j), self.scaled_precision(j, i))
                if p > 1.0:
                    continue
                pairs.append((i, j, min(p, 1e-1)))
        return pairs

    def pair_densities(self, pairs):
        """phi_i phi_j for every pair, as one batched multiply (the reference multiplies the
        pairs one at a time, ExchangePotential.cpp:312-335; each product is the same)."""
        outs = [M.FunctionTree(self.mra) for _ in pairs]
        return M.multiply_batch([p for _, _, p in pairs], outs, 1.0, [self.orbitals[i].tree for i, _, _ in pairs],
                                [self.orbitals[j].tree for _, j, _ in pairs])

    def setup(self, prec: float, flags: int = M.MWG_APPLY_FP64, pair_densities=None):
        if self.exchange is not None:
            raise M.MwgError("Exchange not properly cleared")
        self.apply_prec = prec
        Phi = self.orbitals
        pairs = self.pair_list()
        rho = pair_densities if pair_densities is not None else self.pair_densities(pairs)
        outs = [M.FunctionTree(self.mra) for _ in pairs]
        self.stats = M.apply_batch([p for _, _, p in pairs], outs, [self.poisson] * len(pairs), rho, flags=flags)
        ex = [None] * len(Phi)

        def acc(k, f, t):
            if ex[k] is None:
                if f != 1.0:
                    t.rescale(f)
                ex[k] = t
            else:
                ex[k] = M.add(M.FunctionTree(self.mra), 1.0, ex[k], f, t)

        # the products V_ij phi_j (and V_ij phi_i) of all pairs in one batched multiply, then
        # summed into the exchange functions in pair order
        spec = []
        for (i, j, p), V in zip(pairs, outs):
            self.pair_potentials[(i, j)] = V
            spec.append((p, 1.0 / Phi[j].squared_norm(), Phi[j].tree, V))
            if i != j:
                spec.append((p, 1.0 / Phi[i].squared_norm(), Phi[i].tree, V))
        prods = M.multiply_batch([q[0] for q in spec], [M.FunctionTree(self.mra) for _ in spec],
                                 [q[1] for q in spec], [q[2] for q in spec], [q[3] for q in spec])
        n = 0
        for (i, j, p), V in zip(pairs, outs):
            if i == j:
                t = prods[n]
                n += 1
                self.part_norms[i, i] = math.sqrt(t.getSquareNorm())
                acc(i, 1.0, t)
            else:
                fi = spin_factor(Phi[i], Phi[j])
                fj = spin_factor(Phi[j], Phi[i])
                tj, ti = prods[n], prods[n + 1]
                n += 2
                self.part_norms[j, i] = math.sqrt(tj.getSquareNorm())
                self.part_norms[i, j] = math.sqrt(ti.getSquareNorm())
                acc(i, fi, tj)
                acc(j, fj, ti)
        self.exchange = ex
        for i in range(len(Phi)):
            self.tot_norms[i] = math.sqrt(ex[i].getSquareNorm()) if ex[i] is not None else 0.0
        return ex

    def clear(self):
        self.exchange = None


# ------------------------------------------------------------------ orbital vector algebra

MACHINE_ZERO = 1e-14  # mrcpp::MachineZero


def _lincomb(mra, terms, grid=None):
    """sum_k c_k f_k for (c_k, tree_k) terms, exact on the union grid (pairwise adds in
    term order); coefficients below MachineZero are skipped (qmfunction_utils.cpp:152-165)
    and no remaining term gives the zero function on the grid of `grid`."""
    kept = [(c, t) for c, t in terms if abs(c) > MACHINE_ZERO]
    if not kept:
        g = grid if grid is not None else terms[0][1]
        return M.add(M.FunctionTree(mra), 0.0, g, 0.0, g)
    terms = kept
    c0, t0 = terms[0]
    acc = M.add(M.FunctionTree(mra), c0, t0, 0.0, t0)
    for c, t in terms[1:]:
        acc = M.add(M.FunctionTree(mra), 1.0, acc, c, t)
    return acc


def add(a: float, Phi_a, b: float, Phi_b):
    """orbital::add: out_i = a (Phi_a)_i + b (Phi_b)_i."""
    if len(Phi_a) != len(Phi_b):
        raise M.MwgError("Size mismatch")
    out = []
    for pa, pb in zip(Phi_a, Phi_b):
        mra = pa.tree.mra
        out.append(Orbital(_lincomb(mra, [(a, pa.tree), (b, pb.tree)]), pa.occ, pa.spin))
    return out


def rotate(U, Phi):
    """orbital::rotate: out_i = sum_j U_ij phi_j (occupation and spin of phi_i)."""
    U = np.asarray(U, float)
    if U.shape != (len(Phi), len(Phi)):
        raise M.MwgError("rotation matrix size mismatch")
    out = []
    for i, p in enumerate(Phi):
        out.append(Orbital(_lincomb(p.tree.mra, [(U[i, j], Phi[j].tree) for j in range(len(Phi))], p.tree),
                           p.occ, p.spin))
    return out


def get_norms(Phi) -> np.ndarray:
    return np.array([math.sqrt(p.squared_norm()) for p in Phi])


def calc_overlap_matrix(Phi) -> np.ndarray:
    n = len(Phi)
    S = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            S[i, j] = S[j, i] = M.dot(Phi[i].tree, Phi[j].tree)
    return S


def hermitian_matrix_pow(A, b: float) -> np.ndarray:
    """math_utils::hermitian_matrix_pow: U diag(lambda^b) U^T, eigenvalues below
    MachineZero set to 0."""
    w, V = np.linalg.eigh(np.asarray(A, float))
    d = np.array([0.0 if abs(x) < MACHINE_ZERO else x ** b for x in w])
    return (V * d) @ V.T


def calc_lowdin_matrix(Phi) -> np.ndarray:
    return hermitian_matrix_pow(calc_overlap_matrix(Phi), -0.5)


def orthonormalize(prec: float, Phi):
    """orbital::orthonormalize: Phi <- S^-1/2 Phi; returns (Phi, U)."""
    U = calc_lowdin_matrix(Phi)
    return rotate(U, Phi), U


def helmholtz_step(prec: float, F, Phi, V: M.FunctionTree, mra, flags: int = M.MWG_APPLY_FP64):
    """One orbital update of OrbitalOptimizer::optimize (OrbitalOptimizer.cpp:148-170) with a
    local potential V: returns (Phi_new, F_new, errors)."""
    F = np.asarray(F, float)
    H = HelmholtzVector(prec, np.diag(F), mra, flags)
    L = np.diag(H.lambdas)
    Psi = rotate(L - F, Phi)
    Phi_np1 = H(Phi, V=V, Psi=Psi)
    Phi_np1, U = orthonormalize(prec, Phi_np1)
    F_new = U @ F @ U.T
    dPhi = add(1.0, Phi_np1, -1.0, Phi)
    errors = get_norms(dPhi)
    Phi_new = add(1.0, Phi, 1.0, dPhi)
    return Phi_new, F_new, errors
