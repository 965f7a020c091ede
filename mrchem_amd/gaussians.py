"""Separable polynomial-Gaussian sums: the analytic inputs of the apply.

A term is ``coef * prod_d p_d(x_d - c_d) * exp(-alpha |x - c|^2)`` with a
per-dimension polynomial ``p_d`` given by its coefficients in powers of
``(x_d - c_d)``.  s- and p-type primitives, their products (Gaussian product
theorem) and the exact overlap integrals are all closed under this form, which
is what the synthetic orbitals, pair densities and charge densities of the
benchmark workloads need.

The layout matches the C ABI of ``mwg_project_gauss``: ``coef[n]``,
``alpha[n]``, ``center[n,3]``, ``poly[n,3,5]`` (degree <= 4 per dimension).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MAXDEG = 8  # internal polynomial length (products of two degree-4 factors)


@dataclass
class GaussSum:
    coef: np.ndarray = field(default_factory=lambda: np.zeros(0))
    alpha: np.ndarray = field(default_factory=lambda: np.zeros(0))
    center: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    poly: np.ndarray = field(default_factory=lambda: np.zeros((0, 3, MAXDEG + 1)))

    @property
    def n(self) -> int:
        return len(self.coef)

    def copy(self) -> "GaussSum":
        return GaussSum(self.coef.copy(), self.alpha.copy(), self.center.copy(), self.poly.copy())

    def scaled(self, c: float) -> "GaussSum":
        g = self.copy()
        g.coef = g.coef * c
        return g

    def __add__(self, other: "GaussSum") -> "GaussSum":
        return GaussSum(np.concatenate([self.coef, other.coef]), np.concatenate([self.alpha, other.alpha]),
                        np.concatenate([self.center, other.center]), np.concatenate([self.poly, other.poly]))

    def packed(self):
        """(coef, alpha, center[n*3], poly[n*15]) as contiguous float64 arrays for the C ABI."""
        if self.n and np.any(np.abs(self.poly[:, :, 5:]) > 0):
            raise ValueError("polynomial degree > 4 cannot cross the C ABI")
        poly = np.ascontiguousarray(self.poly[:, :, :5], dtype=np.float64).reshape(-1)
        return (np.ascontiguousarray(self.coef, dtype=np.float64), np.ascontiguousarray(self.alpha, dtype=np.float64),
                np.ascontiguousarray(self.center, dtype=np.float64).reshape(-1), poly)

    def __call__(self, pts: np.ndarray) -> np.ndarray:
        pts = np.atleast_2d(pts)
        out = np.zeros(len(pts))
        for i in range(self.n):
            u = pts - self.center[i]
            v = self.coef[i] * np.exp(-self.alpha[i] * (u * u).sum(1))
            for d in range(3):
                v = v * np.polynomial.polynomial.polyval(u[:, d], self.poly[i, d])
            out += v
        return out


def primitive(kind: str, alpha: float, center, coef: float = 1.0) -> GaussSum:
    """Unnormalised s, px, py or pz Gaussian primitive."""
    poly = np.zeros((1, 3, MAXDEG + 1))
    poly[0, :, 0] = 1.0
    if kind in ("px", "py", "pz"):
        d = "xyz".index(kind[1])
        poly[0, d, 0] = 0.0
        poly[0, d, 1] = 1.0
    elif kind != "s":
        raise ValueError(kind)
    return GaussSum(np.array([coef], float), np.array([alpha], float), np.array([center], float), poly)


def normalized_s(charge: float, alpha: float, center) -> GaussSum:
    """charge * (alpha/pi)^{3/2} exp(-alpha |r-R|^2): integrates to ``charge``."""
    return primitive("s", alpha, center, charge * (alpha / math.pi) ** 1.5)


def _shift_poly(p: np.ndarray, s: float) -> np.ndarray:
    """Coefficients of p(u + s) in powers of u."""
    out = np.zeros_like(p)
    for e, c in enumerate(p):
        if c == 0.0:
            continue
        for j in range(e + 1):
            out[j] += c * math.comb(e, j) * s ** (e - j)
    return out


def product(a: GaussSum, b: GaussSum) -> GaussSum:
    """Pointwise product of two Gaussian sums (Gaussian product theorem)."""
    n = a.n * b.n
    coef = np.zeros(n)
    alpha = np.zeros(n)
    center = np.zeros((n, 3))
    poly = np.zeros((n, 3, MAXDEG + 1))
    k = 0
    for i in range(a.n):
        for j in range(b.n):
            al, be = a.alpha[i], b.alpha[j]
            p = al + be
            A, B = a.center[i], b.center[j]
            P = (al * A + be * B) / p
            coef[k] = a.coef[i] * b.coef[j] * math.exp(-al * be / p * float(((A - B) ** 2).sum()))
            alpha[k] = p
            center[k] = P
            for d in range(3):
                pa = _shift_poly(a.poly[i, d], P[d] - A[d])
                pb = _shift_poly(b.poly[j, d], P[d] - B[d])
                pr = np.polynomial.polynomial.polymul(pa, pb)
                if len(pr) > MAXDEG + 1 and np.any(np.abs(pr[MAXDEG + 1:]) > 0):
                    raise ValueError("polynomial degree too high")
                poly[k, d, :min(len(pr), MAXDEG + 1)] = pr[:MAXDEG + 1]
            k += 1
    return GaussSum(coef, alpha, center, poly)


def _moment(e: int, a: float) -> float:
    """int u^e exp(-a u^2) du over the real line."""
    if e % 2:
        return 0.0
    return math.sqrt(math.pi / a) * math.prod(range(e - 1, 0, -2)) / (2.0 * a) ** (e // 2)


def integral(g: GaussSum) -> float:
    s = 0.0
    for i in range(g.n):
        v = g.coef[i]
        for d in range(3):
            v *= sum(c * _moment(e, g.alpha[i]) for e, c in enumerate(g.poly[i, d]) if c != 0.0)
        s += v
    return s


_GH_T, _GH_W = np.polynomial.hermite.hermgauss(8)  # exact for polynomial degree <= 15


def overlap_matrix(a: GaussSum, b: GaussSum) -> np.ndarray:
    """S[i, j] = integral of term i of a times term j of b (exact: Gauss-Hermite in the
    combined Gaussian, polynomial degree <= 8 + 8 per dimension)."""
    al = a.alpha[:, None]
    be = b.alpha[None, :]
    p = al + be
    S = np.outer(a.coef, b.coef)
    for d in range(3):
        A = a.center[:, d][:, None]
        B = b.center[:, d][None, :]
        P = (al * A + be * B) / p
        pref = np.exp(-al * be / p * (A - B) ** 2) / np.sqrt(p)
        x = P[..., None] + _GH_T[None, None, :] / np.sqrt(p)[..., None]      # (na, nb, 8)
        ua = x - A[..., None]
        ub = x - B[..., None]
        pa = np.zeros_like(x)
        pb = np.zeros_like(x)
        for e in range(MAXDEG, -1, -1):  # Horner
            pa = pa * ua + a.poly[:, d, e][:, None, None]
            pb = pb * ub + b.poly[:, d, e][None, :, None]
        S = S * pref * (pa * pb * _GH_W).sum(-1)
    return S


def overlap(a: GaussSum, b: GaussSum) -> float:
    return float(overlap_matrix(a, b).sum())


def gram_schmidt(funcs: list[GaussSum]) -> list[GaussSum]:
    """Orthonormalise with exact overlaps in FP64 (modified Gram-Schmidt, two passes),
    in coefficient space over the union of all primitives."""
    prim = GaussSum()
    rows = []
    for f in funcs:
        rows.append((prim.n, f.n))
        unit = f.copy()
        unit.coef = np.ones(f.n)
        prim = prim + unit
    n = prim.n
    S = overlap_matrix(prim, prim)
    C = np.zeros((len(funcs), n))
    for i, (o, m) in enumerate(rows):
        C[i, o:o + m] = funcs[i].coef
    out = np.zeros_like(C)
    for i in range(len(funcs)):
        v = C[i].copy()
        for _ in range(2):
            for j in range(i):
                v -= (out[j] @ S @ v) * out[j]
        out[i] = v / np.sqrt(v @ S @ v)
    res = []
    for i in range(len(funcs)):
        g = prim.copy()
        g.coef = out[i].copy()
        res.append(compact(g, tol=0.0))
    return res


def compact(g: GaussSum, tol: float = 0.0) -> GaussSum:
    """Merge terms with identical (alpha, center, poly) and drop exact zeros."""
    keys: dict = {}
    order = []
    for i in range(g.n):
        key = (round(g.alpha[i], 15), tuple(np.round(g.center[i], 15)), tuple(np.round(g.poly[i].reshape(-1), 15)))
        if key in keys:
            keys[key][0] += g.coef[i]
        else:
            keys[key] = [g.coef[i], i]
            order.append(key)
    idx = [keys[k][1] for k in order if abs(keys[k][0]) > tol]
    coef = np.array([keys[k][0] for k in order if abs(keys[k][0]) > tol])
    return GaussSum(coef, g.alpha[idx].copy(), g.center[idx].copy(), g.poly[idx].copy())
