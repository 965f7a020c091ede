"""ctypes front end of the oracle library ``oracle/libmwref.so``.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``; never by the product package.

The oracle is the project's FP64 CPU restatement of ``mrcpp::apply`` with the
Poisson and Helmholtz operators (reference call sites:
src/scf_solver/HelmholtzVector.cpp:168-184,
src/qmoperators/two_electron/CoulombPotential.cpp:69-90,
src/qmoperators/two_electron/ExchangePotential.cpp:244-347).  MRCPP itself is
not vendored (external/mrcpp is an empty submodule, .gitmodules:4-6), so the
oracle is pinned by these checks (tests/test_oracle.py):

  * basis: orthonormality of the scaling functions, orthogonality of the
    two-scale matrix, exact reproduction of polynomials of degree <= k;
  * kernel fit: max relative error vs 1/r and exp(-mu r)/r on [r_min, r_max];
  * operator: direct quadrature of T^n vs the two-scale NS construction, the
    constant-kernel closed form, symmetry T_{-l} = T_l^T;
  * projection: point values and norms vs the analytic Gaussian sums;
  * apply: erf closed form (Poisson on Gaussian charges), erfc closed form
    (Helmholtz), the p-type derivative closed forms, linearity, symmetry
    <a|Pb> = <b|Pa>, mu -> 0 limit, and the hydrogenic Slater-Condon values
    behind tests/qmoperators/coulomb_operator.cpp:76-82.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Stats(C.Structure):
    _fields_ = [("tuples", C.c_int64), ("gen_nodes", C.c_int64), ("nodes_out", C.c_int64), ("iters", C.c_int64),
                ("t_total", C.c_double), ("tuples_iter", C.c_int64 * 64), ("nodes_iter", C.c_int64 * 64)]

    def as_dict(self):
        it = int(self.iters)
        return dict(tuples=int(self.tuples), gen_nodes=int(self.gen_nodes), nodes_out=int(self.nodes_out),
                    iters=it, t_total=float(self.t_total), tuples_iter=list(self.tuples_iter)[:it],
                    nodes_iter=list(self.nodes_iter)[:it])


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmwref.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, d, i, i64 = C.c_void_p, C.c_double, C.c_int, C.c_int64
        P = np.ctypeslib.ndpointer
        fd = P(dtype=np.float64, flags="C")
        fi = P(dtype=np.int32, flags="C")
        fl = P(dtype=np.int64, flags="C")
        sig = {
            "mwref_mra_new": (vp, [i, fi, fi, d, i, i, i]),
            "mwref_mra_free": (None, [vp]),
            "mwref_mra_filters": (None, [vp, fd]),
            "mwref_basis_eval": (None, [vp, d, fd]),
            "mwref_tree_new": (vp, [vp]),
            "mwref_tree_free": (None, [vp]),
            "mwref_tree_nnodes": (i, [vp]),
            "mwref_tree_sqnorm": (d, [vp]),
            "mwref_tree_maxdepth": (i, [vp]),
            "mwref_project_gauss": (None, [vp, d, i, fd, fd, fd, fd]),
            "mwref_project_hydrogen": (None, [vp, d, i, fd, fi, d, fd]),
            "mwref_eval": (None, [vp, i, fd, fd]),
            "mwref_dot": (d, [vp, vp]),
            "mwref_tree_rescale": (None, [vp, d]),
            "mwref_tree_export": (None, [vp, fi, fl, fd]),
            "mwref_tree_import": (i, [vp, i, fi, fl, fd]),
            "mwref_oper_new": (vp, [vp, d, d]),
            "mwref_oper_free": (None, [vp]),
            "mwref_oper_nterms": (i, [vp]),
            "mwref_oper_terms": (None, [vp, fd, fd, fd]),
            "mwref_oper_blocks": (None, [vp, i, i, i64, i64, fd]),
            "mwref_oper_tblock": (None, [vp, i, i, i64, fd]),
            "mwref_oper_level": (None, [vp, i, fi, fi, fd]),
            "mwref_apply": (i, [d, vp, vp, vp, i, C.POINTER(Stats)]),
            "mwref_set_threads": (None, [i]),
            "mwref_get_threads": (i, []),
            "mwref_set_screen_factor": (None, [d]),
            "mwref_multiply": (None, [vp, d, d, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class MRA:
    def __init__(self, n0, corner, boxes, sfac=1.0, k=5, basis="L", max_depth=16):
        self.k = k
        self.K = k + 1
        self.n0, self.corner, self.boxes, self.sfac, self.max_depth = n0, tuple(corner), tuple(boxes), sfac, max_depth
        self.basis = basis
        self.h = lib().mwref_mra_new(n0, np.array(corner, np.int32), np.array(boxes, np.int32), float(sfac), k,
                                     0 if basis == "L" else 1, max_depth)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.mwref_mra_free(self.h)
            self.h = None

    def filters(self):
        N = 2 * self.K
        W = np.zeros(N * N)
        lib().mwref_mra_filters(self.h, W)
        return W.reshape(N, N)

    def phi(self, x):
        out = np.zeros(self.K)
        lib().mwref_basis_eval(self.h, float(x), out)
        return out

    @property
    def lower(self):
        return np.array([self.sfac * c * 2.0 ** (-self.n0) for c in self.corner])

    @property
    def upper(self):
        return np.array([self.sfac * (c + b) * 2.0 ** (-self.n0) for c, b in zip(self.corner, self.boxes)])


class Tree:
    def __init__(self, mra: MRA):
        self.mra = mra
        self.h = lib().mwref_tree_new(mra.h)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.mwref_tree_free(self.h)
            self.h = None

    @property
    def n_nodes(self):
        return lib().mwref_tree_nnodes(self.h)

    @property
    def sqnorm(self):
        return lib().mwref_tree_sqnorm(self.h)

    @property
    def max_depth(self):
        return lib().mwref_tree_maxdepth(self.h)

    def project_gauss(self, prec, g):
        coef, alpha, center, poly = g.packed()
        lib().mwref_project_gauss(self.h, float(prec), g.n, coef, alpha, center, poly)
        return self

    def project_hydrogen(self, prec, terms, Z=1.0, origin=(0.0, 0.0, 0.0)):
        """terms: list of (coef, (n,l,m)) or (coef, (n,l,m), (n',l',m')) products."""
        coef = np.array([t[0] for t in terms], float)
        q = np.zeros((len(terms), 6), np.int32)
        for i, t in enumerate(terms):
            q[i, :3] = t[1]
            if len(t) > 2:
                q[i, 3:] = t[2]
        lib().mwref_project_hydrogen(self.h, float(prec), len(terms), coef, q.reshape(-1), float(Z),
                                     np.array(origin, float))
        return self

    def eval(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pts))
        lib().mwref_eval(self.h, len(pts), pts.reshape(-1), out)
        return out

    def dot(self, other):
        return lib().mwref_dot(self.h, other.h)

    def rescale(self, c):
        lib().mwref_tree_rescale(self.h, float(c))

    def export(self):
        n = self.n_nodes
        K3 = self.mra.K ** 3
        depth = np.zeros(n, np.int32)
        lr = np.zeros(3 * n, np.int64)
        coefs = np.zeros(n * 8 * K3)
        lib().mwref_tree_export(self.h, depth, lr, coefs)
        return depth, lr.reshape(n, 3), coefs.reshape(n, 8, K3)

    def load(self, depth, lr, coefs):
        rc = lib().mwref_tree_import(self.h, len(depth), np.ascontiguousarray(depth, np.int32),
                                     np.ascontiguousarray(lr, np.int64).reshape(-1),
                                     np.ascontiguousarray(coefs, np.float64).reshape(-1))
        if rc != 0:
            raise ValueError("tree import failed: orphan node")
        return self


class Oper:
    def __init__(self, mra: MRA, mu: float, prec: float):
        self.mra = mra
        self.mu = mu
        self.prec = prec
        self.h = lib().mwref_oper_new(mra.h, float(mu), float(prec))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.mwref_oper_free(self.h)
            self.h = None

    @property
    def n_terms(self):
        return lib().mwref_oper_nterms(self.h)

    def terms(self):
        n = self.n_terms
        a = np.zeros(n)
        w = np.zeros(n)
        info = np.zeros(5)
        lib().mwref_oper_terms(self.h, a, w, info)
        return a, w, dict(rmin=info[0], rmax=info[1], eps=info[2], h=info[3], err=info[4])

    def blocks(self, term, depth, dlo, dhi):
        K2 = self.mra.K ** 2
        out = np.zeros((dhi - dlo + 1) * 4 * K2)
        lib().mwref_oper_blocks(self.h, term, depth, dlo, dhi, out)
        return out.reshape(dhi - dlo + 1, 4, self.mra.K, self.mra.K)

    def tblock(self, term, depth, delta):
        K = self.mra.K
        out = np.zeros(K * K)
        lib().mwref_oper_tblock(self.h, term, depth, delta, out)
        return out.reshape(K, K)

    def level(self, depth):
        n = self.n_terms
        act = np.zeros(n, np.int32)
        band = np.zeros(n, np.int32)
        maxn = np.zeros(4 * n)
        lib().mwref_oper_level(self.h, depth, act, band, maxn)
        return act, band, maxn.reshape(n, 4)


def poisson(mra, prec):
    return Oper(mra, 0.0, prec)


def helmholtz(mra, mu, prec):
    return Oper(mra, mu, prec)


def apply(prec, out: Tree, oper: Oper, inp: Tree, max_iter=-1):
    st = Stats()
    lib().mwref_apply(float(prec), out.h, oper.h, inp.h, int(max_iter), C.byref(st))
    return st.as_dict()


def set_threads(n):
    lib().mwref_set_threads(int(n))


def multiply(prec, out: Tree, coef, a: Tree, b: Tree):
    """out = coef * a * b (pointwise), union grid of a and b refined by the split rule."""
    lib().mwref_multiply(out.h, float(prec), float(coef), a.h, b.h)
    return out


def set_screen_factor(f):
    lib().mwref_set_screen_factor(float(f))


def get_threads():
    return lib().mwref_get_threads()
