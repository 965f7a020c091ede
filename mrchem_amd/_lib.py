"""ctypes binding of libmwgreen.so (the C ABI declared in include/mwg.h).

The library is built in-tree by ``mrchem_amd.build``.  There is no fallback:
if the shared object is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmwgreen.so")

MWG_APPLY_FP64 = 0
MWG_APPLY_MIXED = 1
MWG_APPLY_PREDICT = 2


class ApplyStats(C.Structure):
    _fields_ = [("nodes_out", C.c_int64), ("tuples", C.c_int64), ("tuples_fp64", C.c_int64),
                ("gen_nodes", C.c_int64), ("iters", C.c_int64), ("slices", C.c_int64),
                ("predicted_splits", C.c_int64), ("predicted_wasted", C.c_int64),
                ("t_apply", C.c_double), ("t_contract", C.c_double), ("flops", C.c_double),
                ("device_loop", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OrbitalMeta(C.Structure):
    """FunctionData + OrbitalData of an MRChem .meta file (include/mwg.h)."""
    _fields_ = [("real_size", C.c_int32), ("imag_size", C.c_int32), ("is_shared", C.c_int32),
                ("rank_id", C.c_int32), ("spin", C.c_int32), ("occ", C.c_int32), ("error", C.c_double)]


class TreeFileInfo(C.Structure):
    _fields_ = [("order", C.c_int32), ("basis", C.c_int32), ("root_scale", C.c_int32), ("max_depth", C.c_int32),
                ("corner", C.c_int32 * 3), ("boxes", C.c_int32 * 3), ("scaling_factor", C.c_double),
                ("nnodes", C.c_int64), ("chunk_nodes", C.c_int32), ("nchunks", C.c_int32)]


class MwgError(RuntimeError):
    pass


_LIB = None

# name -> (restype, argtypes); every function returns an int status except mwg_last_error
_SIG = None


def _sig():
    vp, d, i, i64, u = C.c_void_p, C.c_double, C.c_int, C.c_int64, C.c_uint
    fd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C")
    fi = np.ctypeslib.ndpointer(dtype=np.int32, flags="C")
    fl = np.ctypeslib.ndpointer(dtype=np.int64, flags="C")
    ff = np.ctypeslib.ndpointer(dtype=np.float32, flags="C")
    pvp = C.POINTER(C.c_void_p)
    return {
        "mwg_ctx_create": [i, pvp],
        "mwg_ctx_destroy": [vp],
        "mwg_sync": [vp],
        "mwg_ctx_set_tuning": [vp, i, i],
        "mwg_ctx_get_tuning": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "mwg_ctx_trim": [vp, C.POINTER(C.c_int64)],
        "mwg_mra_create": [vp, i, fi, fi, d, i, i, i, pvp],
        "mwg_mra_destroy": [vp],
        "mwg_tree_create": [vp, pvp],
        "mwg_tree_destroy": [vp],
        "mwg_tree_nnodes": [vp, C.POINTER(C.c_int64)],
        "mwg_tree_sqnorm": [vp, C.POINTER(C.c_double)],
        "mwg_tree_maxdepth": [vp, C.POINTER(C.c_int)],
        "mwg_tree_mra": [vp, pvp],
        "mwg_mra_order": [vp, C.POINTER(C.c_int)],
        "mwg_tree_import": [vp, i64, fi, fl, fd],
        "mwg_tree_layout_check": [fi, i, i64, fi, fl, C.c_char_p, i],
        "mwg_tree_export": [vp, fi, fl, fd],
        "mwg_tree_rescale": [vp, d],
        "mwg_tree_dot": [vp, vp, C.POINTER(C.c_double)],
        "mwg_tree_dot_noalloc": [vp, vp, C.POINTER(C.c_double)],
        "mwg_meta_write": [C.c_char_p, C.POINTER(OrbitalMeta), i],
        "mwg_meta_read": [C.c_char_p, C.POINTER(OrbitalMeta), C.POINTER(C.c_int)],
        "mwg_tree_save": [vp, C.c_char_p, C.POINTER(C.c_int)],
        "mwg_tree_load": [vp, C.c_char_p],
        "mwg_tree_file_write": [C.c_char_p, C.POINTER(TreeFileInfo), vp, vp, vp],
        "mwg_tree_file_read": [C.c_char_p, C.POINTER(TreeFileInfo), vp, vp, vp],
        "mwg_tree_eval": [vp, i64, fd, fd],
        "mwg_project_gauss": [d, vp, i, fd, fd, fd, fd],
        "mwg_project_hydrogen": [d, vp, i, fd, fi, d, fd],
        "mwg_oper_poisson": [vp, d, pvp],
        "mwg_oper_helmholtz": [vp, d, d, pvp],
        "mwg_oper_destroy": [vp],
        "mwg_oper_nterms": [vp, C.POINTER(C.c_int)],
        "mwg_oper_terms": [vp, fd, fd],
        "mwg_oper_blocks": [vp, i, i, i64, i64, fd],
        "mwg_multiply": [d, vp, d, vp, vp],
        "mwg_multiply_batch": [i, fd, pvp, fd, pvp, pvp],
        "mwg_add": [vp, d, vp, d, vp],
        "mwg_apply": [d, vp, vp, vp, i, u, C.POINTER(ApplyStats)],
        "mwg_apply_batch": [i, fd, pvp, pvp, pvp, i, u, C.POINTER(ApplyStats)],
        "mwg_predictor_create": [vp, C.c_uint64, pvp],
        "mwg_predictor_destroy": [vp],
        "mwg_predictor_sizes": [C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "mwg_predictor_get": [vp, ff, ff, ff],
        "mwg_predictor_set": [vp, ff, ff, ff],
        "mwg_predictor_train": [vp, i64, ff, ff, i, i, C.c_float, C.c_uint64, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)],
        "mwg_predictor_gradient": [vp, i64, ff, ff, ff],
        "mwg_predictor_infer": [vp, i64, ff, ff],
        "mwg_ctx_set_predictor": [vp, vp, C.c_float],
        "mwg_trace_enable": [vp, i],
        "mwg_trace_size": [vp, C.POINTER(C.c_int64)],
        "mwg_trace_get": [vp, ff, ff],
    }


EXPORTED = ["mwg_ctx_create", "mwg_ctx_destroy", "mwg_last_error", "mwg_sync", "mwg_ctx_set_tuning",
            "mwg_ctx_get_tuning", "mwg_ctx_trim", "mwg_mra_create", "mwg_mra_destroy",
            "mwg_tree_create", "mwg_tree_destroy", "mwg_tree_nnodes", "mwg_tree_sqnorm", "mwg_tree_maxdepth",
            "mwg_tree_mra", "mwg_mra_order",
            "mwg_tree_import", "mwg_tree_layout_check", "mwg_tree_export", "mwg_tree_rescale", "mwg_tree_dot", "mwg_tree_dot_noalloc",
            "mwg_tree_eval", "mwg_meta_write", "mwg_meta_read", "mwg_tree_save", "mwg_tree_load",
            "mwg_tree_file_write", "mwg_tree_file_read",
            "mwg_project_gauss", "mwg_project_hydrogen", "mwg_oper_poisson", "mwg_oper_helmholtz",
            "mwg_oper_destroy", "mwg_oper_nterms", "mwg_oper_terms", "mwg_oper_blocks", "mwg_multiply", "mwg_multiply_batch", "mwg_add", "mwg_apply",
            "mwg_apply_batch", "mwg_predictor_create", "mwg_predictor_destroy", "mwg_predictor_sizes",
            "mwg_predictor_get", "mwg_predictor_set", "mwg_predictor_train", "mwg_predictor_gradient",
            "mwg_predictor_infer", "mwg_ctx_set_predictor", "mwg_trace_enable", "mwg_trace_size", "mwg_trace_get"]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MwgError(f"{LIB_PATH} is not built; run `python -m mrchem_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in _sig().items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.mwg_last_error.restype = C.c_char_p
        L.mwg_last_error.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def check(rc: int, ctx=None):
    if rc != 0:
        msg = lib().mwg_last_error(ctx)
        raise MwgError(f"mwg status {rc}: {msg.decode() if msg else ''}")
