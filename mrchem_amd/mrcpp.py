"""Python mirror of the mrcpp interface that MRChem calls on its hot path.

Same names, argument meaning and error behaviour as the calls in the reference:
  * ``MultiResolutionAnalysis(BoundingBox(min_scale, corner, boxes), basis, max_depth)``
    (src/mrenv.cpp:46-72, tests/unit_tests.cpp:28-44)
  * ``FunctionTree(MRA)`` with ``getNNodes``, ``getSquareNorm``, ``rescale``
    (src/qmfunctions/QMFunction.cpp:68-77, src/scf_solver/HelmholtzVector.cpp:81)
  * ``PoissonOperator(MRA, prec)`` (src/SCFDriver.cpp:279) and
    ``HelmholtzOperator(MRA, mu, prec)`` (src/scf_solver/HelmholtzVector.cpp:171)
  * ``apply(prec, out, oper, inp, max_iter=-1)`` (HelmholtzVector.cpp:176,180,
    CoulombPotential.cpp:85, ExchangePotential.cpp:137,259,318)
  * ``project(prec, tree, func)`` and ``dot(a, b)``.
Every call goes through libmwgreen.so (C ABI, include/mwg.h); the compute runs on
the GPU.  Errors raise ``MwgError`` (the reference aborts via MSG_ERROR/MSG_FATAL).
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import MWG_APPLY_FP64, MWG_APPLY_MIXED, MWG_APPLY_PREDICT, ApplyStats, MwgError, check, lib
from .gaussians import GaussSum

__all__ = ["BoundingBox", "LegendreBasis", "InterpolatingBasis", "MultiResolutionAnalysis", "FunctionTree",
           "PoissonOperator", "HelmholtzOperator", "apply", "apply_batch", "project", "project_hydrogen", "dot",
           "multiply", "multiply_batch", "add", "Predictor", "trace_enable", "trace_get",
           "MwgError", "MWG_APPLY_FP64", "MWG_APPLY_MIXED", "MWG_APPLY_PREDICT", "Context", "default_context"]


class Context:
    """One HIP stream + scratch pools on one device (mwg_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().mwg_ctx_create(int(device), C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        check(lib().mwg_sync(self.h), self.h)

    def set_tuning(self, fp64_chunk: int, slice_chunk: int):
        """Contraction task granularity: tuples per FP64 wave task, MFMA slices per wave task."""
        check(lib().mwg_ctx_set_tuning(self.h, int(fp64_chunk), int(slice_chunk)), self.h)

    def get_tuning(self) -> tuple[int, int]:
        a, b = C.c_int(), C.c_int()
        check(lib().mwg_ctx_get_tuning(self.h, C.byref(a), C.byref(b)), self.h)
        return int(a.value), int(b.value)

    def trim(self) -> int:
        """Return the device memory kept for reuse to the device; the bytes released."""
        r = C.c_int64()
        check(lib().mwg_ctx_trim(self.h, C.byref(r)), self.h)
        return int(r.value)

    def __del__(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.mwg_ctx_destroy(self.h)
            self.h = None


_tls = threading.local()


def default_context() -> Context:
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx


@dataclass
class BoundingBox:
    min_scale: int
    corner: tuple
    boxes: tuple
    scaling_factor: float = 1.0


@dataclass
class LegendreBasis:
    order: int
    kind: int = 0


@dataclass
class InterpolatingBasis:
    order: int
    kind: int = 1


class MultiResolutionAnalysis:
    def __init__(self, box: BoundingBox, basis, max_depth: int = 16, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.box = box
        self.basis = basis
        self.order = basis.order
        self.K = basis.order + 1
        self.max_depth = max_depth
        h = C.c_void_p()
        check(lib().mwg_mra_create(self.ctx.h, int(box.min_scale), np.array(box.corner, np.int32),
                                   np.array(box.boxes, np.int32), float(box.scaling_factor), int(basis.order),
                                   int(basis.kind), int(max_depth), C.byref(h)), self.ctx.h)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.mwg_mra_destroy(self.h)
            self.h = None


class FunctionTree:
    def __init__(self, mra: MultiResolutionAnalysis):
        self.mra = mra
        h = C.c_void_p()
        check(lib().mwg_tree_create(mra.h, C.byref(h)), mra.ctx.h)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.mwg_tree_destroy(self.h)
            self.h = None

    def _check(self, rc):
        check(rc, self.mra.ctx.h)

    def getNNodes(self) -> int:
        n = C.c_int64()
        self._check(lib().mwg_tree_nnodes(self.h, C.byref(n)))
        return int(n.value)

    def getSquareNorm(self) -> float:
        v = C.c_double()
        self._check(lib().mwg_tree_sqnorm(self.h, C.byref(v)))
        return float(v.value)

    def getDepth(self) -> int:
        v = C.c_int()
        self._check(lib().mwg_tree_maxdepth(self.h, C.byref(v)))
        return int(v.value)

    def rescale(self, c: float):
        self._check(lib().mwg_tree_rescale(self.h, float(c)))

    def evalf(self, pts) -> np.ndarray:
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pts))
        self._check(lib().mwg_tree_eval(self.h, len(pts), pts.reshape(-1), out))
        return out

    def export(self):
        n = self.getNNodes()
        K3 = self.mra.K ** 3
        depth = np.zeros(n, np.int32)
        lr = np.zeros(3 * n, np.int64)
        coefs = np.zeros(n * 8 * K3)
        self._check(lib().mwg_tree_export(self.h, depth, lr, coefs))
        return depth, lr.reshape(n, 3), coefs.reshape(n, 8, K3)

    def saveTree(self, name: str) -> int:
        """FunctionTree::saveTree: writes "<name>.tree"; returns the file's chunk count."""
        n = C.c_int()
        self._check(lib().mwg_tree_save(self.h, name.encode(), C.byref(n)))
        return n.value

    def loadTree(self, name: str) -> "FunctionTree":
        """FunctionTree::loadTree: reads "<name>.tree" (written for this MRA)."""
        self._check(lib().mwg_tree_load(self.h, name.encode()))
        return self

    def load(self, depth, lr, coefs):
        self._check(lib().mwg_tree_import(self.h, len(depth), np.ascontiguousarray(depth, np.int32),
                                          np.ascontiguousarray(lr, np.int64).reshape(-1),
                                          np.ascontiguousarray(coefs, np.float64).reshape(-1)))
        return self


class _Operator:
    def __init__(self, mra, h):
        self.mra = mra
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.mwg_oper_destroy(self.h)
            self.h = None

    def size(self) -> int:
        n = C.c_int()
        check(lib().mwg_oper_nterms(self.h, C.byref(n)), self.mra.ctx.h)
        return int(n.value)

    def terms(self):
        n = self.size()
        a = np.zeros(n)
        w = np.zeros(n)
        check(lib().mwg_oper_terms(self.h, a, w), self.mra.ctx.h)
        return a, w

    def blocks(self, term: int, depth: int, dlo: int, dhi: int) -> np.ndarray:
        K = self.mra.K
        out = np.zeros((dhi - dlo + 1) * 4 * K * K)
        check(lib().mwg_oper_blocks(self.h, term, depth, dlo, dhi, out), self.mra.ctx.h)
        return out.reshape(dhi - dlo + 1, 4, K, K)


class PoissonOperator(_Operator):
    """Kernel 1/r as a Gaussian sum fitted to prec/10 (SCFDriver.cpp:279)."""

    def __init__(self, mra: MultiResolutionAnalysis, prec: float):
        h = C.c_void_p()
        check(lib().mwg_oper_poisson(mra.h, float(prec), C.byref(h)), mra.ctx.h)
        super().__init__(mra, h)
        self.prec = prec


class HelmholtzOperator(_Operator):
    """Kernel exp(-mu r)/r (HelmholtzVector.cpp:171); the caller applies -1/(2 pi)."""

    def __init__(self, mra: MultiResolutionAnalysis, mu: float, prec: float):
        h = C.c_void_p()
        check(lib().mwg_oper_helmholtz(mra.h, float(mu), float(prec), C.byref(h)), mra.ctx.h)
        super().__init__(mra, h)
        self.mu = mu
        self.prec = prec


def apply(prec: float, out: FunctionTree, oper: _Operator, inp: FunctionTree, max_iter: int = -1,
          flags: int = MWG_APPLY_FP64) -> dict:
    st = ApplyStats()
    check(lib().mwg_apply(float(prec), out.h, oper.h, inp.h, int(max_iter), int(flags), C.byref(st)),
          inp.mra.ctx.h)
    return st.as_dict()


def apply_batch(prec, outs, opers, inps, max_iter: int = -1, flags: int = MWG_APPLY_FP64) -> list:
    n = len(outs)
    if not (len(opers) == n and len(inps) == n):
        raise MwgError("apply_batch: length mismatch")
    precs = np.broadcast_to(np.asarray(prec, np.float64), (n,)).copy()
    arr = C.c_void_p * n
    st = (ApplyStats * n)()
    check(lib().mwg_apply_batch(n, precs, arr(*[o.h for o in outs]), arr(*[o.h for o in opers]),
                                arr(*[i.h for i in inps]), int(max_iter), int(flags), st), inps[0].mra.ctx.h)
    return [s.as_dict() for s in st]


def project(prec: float, tree: FunctionTree, func: GaussSum) -> FunctionTree:
    coef, alpha, center, poly = func.packed()
    check(lib().mwg_project_gauss(float(prec), tree.h, func.n, coef, alpha, center, poly), tree.mra.ctx.h)
    return tree


def project_hydrogen(prec: float, tree: FunctionTree, terms, Z: float = 1.0, origin=(0.0, 0.0, 0.0)):
    """terms: list of (coef, (n,l,m)) or (coef, (n,l,m), (n',l',m')) products of hydrogen orbitals."""
    coef = np.array([t[0] for t in terms], float)
    q = np.zeros((len(terms), 6), np.int32)
    for i, t in enumerate(terms):
        q[i, :3] = t[1]
        if len(t) > 2:
            q[i, 3:] = t[2]
    check(lib().mwg_project_hydrogen(float(prec), tree.h, len(terms), coef, q.reshape(-1), float(Z),
                                     np.array(origin, float)), tree.mra.ctx.h)
    return tree


def multiply(prec: float, out: FunctionTree, coef: float, a: FunctionTree, b: FunctionTree) -> FunctionTree:
    """out = coef * a * b (mrcpp::multiply): union grid of a and b, refined by the split test."""
    check(lib().mwg_multiply(float(prec), out.h, float(coef), a.h, b.h), a.mra.ctx.h)
    return out


def multiply_batch(prec, outs, coefs, as_, bs) -> list:
    """outs[i] = coefs[i] * as_[i] * bs[i] for all i: the same trees as one multiply per
    product, with all products refined together (one host round trip per level)."""
    n = len(outs)
    if not (len(as_) == n and len(bs) == n):
        raise MwgError("multiply_batch: length mismatch")
    if n == 0:
        return outs
    precs = np.broadcast_to(np.asarray(prec, np.float64), (n,)).copy()
    cs = np.broadcast_to(np.asarray(coefs, np.float64), (n,)).copy()
    arr = C.c_void_p * n
    check(lib().mwg_multiply_batch(n, precs, arr(*[o.h for o in outs]), cs, arr(*[a.h for a in as_]),
                                   arr(*[b.h for b in bs])), as_[0].mra.ctx.h)
    return outs


def add(out: FunctionTree, ca: float, a: FunctionTree, cb: float, b: FunctionTree) -> FunctionTree:
    """out = ca * a + cb * b on the union grid (mrcpp::add)."""
    check(lib().mwg_add(out.h, float(ca), a.h, float(cb), b.h), a.mra.ctx.h)
    return out


def dot(a: FunctionTree, b: FunctionTree) -> float:
    v = C.c_double()
    check(lib().mwg_tree_dot(a.h, b.h, C.byref(v)), a.mra.ctx.h)
    return float(v.value)


class Predictor:
    """Learned refinement predictor (2 x 64 tanh MLP, FP32) trained and run by HIP kernels."""

    def __init__(self, ctx: Context | None = None, seed: int = 0):
        self.ctx = ctx or default_context()
        np_, nf = C.c_int(), C.c_int()
        lib().mwg_predictor_sizes(C.byref(np_), C.byref(nf))
        self.nparams, self.nfeatures = np_.value, nf.value
        h = C.c_void_p()
        check(lib().mwg_predictor_create(self.ctx.h, int(seed), C.byref(h)), self.ctx.h)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.mwg_predictor_destroy(self.h)
            self.h = None

    def params(self):
        P = np.zeros(self.nparams, np.float32)
        m = np.zeros(self.nfeatures, np.float32)
        s = np.zeros(self.nfeatures, np.float32)
        check(lib().mwg_predictor_get(self.h, P, m, s), self.ctx.h)
        return P, m, s

    def set_params(self, P, mean, sdev):
        check(lib().mwg_predictor_set(self.h, np.ascontiguousarray(P, np.float32), np.ascontiguousarray(mean, np.float32),
                                      np.ascontiguousarray(sdev, np.float32)), self.ctx.h)

    def train(self, X, y, epochs=200, batch=512, lr=3e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        lb, la = C.c_double(), C.c_double()
        check(lib().mwg_predictor_train(self.h, len(y), X.reshape(-1), y, int(epochs), int(batch), float(lr), int(seed),
                                        C.byref(lb), C.byref(la)), self.ctx.h)
        return lb.value, la.value

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.nparams, np.float32)
        check(lib().mwg_predictor_gradient(self.h, len(y), X.reshape(-1), y, g), self.ctx.h)
        return g

    def infer(self, X):
        X = np.ascontiguousarray(X, np.float32)
        p = np.zeros(len(X), np.float32)
        check(lib().mwg_predictor_infer(self.h, len(X), X.reshape(-1), p), self.ctx.h)
        return p

    def install(self, threshold: float = 0.5):
        """Use this predictor for applies with MWG_APPLY_PREDICT on its context."""
        check(lib().mwg_ctx_set_predictor(self.ctx.h, self.h, float(threshold)), self.ctx.h)


def trace_enable(on: bool = True, ctx: Context | None = None):
    ctx = ctx or default_context()
    check(lib().mwg_trace_enable(ctx.h, int(bool(on))), ctx.h)


def trace_get(ctx: Context | None = None):
    """(features [n, nfeatures], labels [n]) recorded since trace_enable."""
    ctx = ctx or default_context()
    n = C.c_int64()
    check(lib().mwg_trace_size(ctx.h, C.byref(n)), ctx.h)
    nf = C.c_int()
    lib().mwg_predictor_sizes(None, C.byref(nf))
    X = np.zeros(n.value * nf.value, np.float32)
    y = np.zeros(n.value, np.float32)
    if n.value:
        check(lib().mwg_trace_get(ctx.h, X, y), ctx.h)
    return X.reshape(n.value, nf.value), y


def layout_check(boxes, max_depth: int, depth, lr) -> str:
    """The packed-layout check of tree import without a device: '' or the first violation."""
    buf = C.create_string_buffer(256)
    rc = lib().mwg_tree_layout_check(np.ascontiguousarray(boxes, np.int32), int(max_depth), len(depth),
                                     np.ascontiguousarray(depth, np.int32),
                                     np.ascontiguousarray(lr, np.int64).reshape(-1), buf, 256)
    msg = buf.value.decode()
    if rc != 0 and not msg:
        raise MwgError(f"mwg_tree_layout_check: status {rc}")
    return msg


def mra_from(box: dict, order: int, basis: str = "L", max_depth: int = 16, ctx: Context | None = None):
    """MRA from the box dictionaries of ``mrchem_amd.configs`` (n0, corner, boxes, sfac)."""
    b = BoundingBox(box["n0"], box["corner"], box["boxes"], box.get("sfac", 1.0))
    basis_obj = LegendreBasis(order) if basis == "L" else InterpolatingBasis(order)
    return MultiResolutionAnalysis(b, basis_obj, max_depth, ctx)
