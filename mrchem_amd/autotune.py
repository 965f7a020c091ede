"""Autotuner for the contraction task granularity of the batched apply.

The device apply splits each output node's tuple list into FP64 wave tasks of
``fp64_chunk`` tuples and its MFMA slice list into tasks of ``slice_chunk``
slices (0: sized from each refinement iteration's slice total); partial results of
a node's tasks are summed in FP64 in a fixed order afterwards.  Small tasks spread a node over more waves (better occupancy
for few, heavy nodes); large tasks cut partial-result traffic and reduction work.
The best point depends on k, the precision and the tree shapes, so it is measured:
``autotune`` times the caller's own batch for each candidate (by default the whole
apply, wall clock: small tasks also cost task building and reduction outside the
contraction kernels; or the contraction kernels alone, device clock, ``t_contract``) and
installs the fastest on the context.

Candidates are applied to fresh output trees each time; the results agree to
rounding for every setting (only the FP64 summation grouping changes, on both paths).
"""
from __future__ import annotations

import itertools
import time

from . import mrcpp as G

FP64_CHUNKS = (16, 32, 64, 128, 256)
SLICE_CHUNKS = (0, 16, 32, 64, 128, 256)


def _measure(mra, trees, ops, prec, flags, reps, objective):
    """Smallest (objective, contraction) seconds over reps applies (each minimised on its own)."""
    best_apply, best_contract = float("inf"), float("inf")
    for _ in range(reps):
        outs = [G.FunctionTree(mra) for _ in trees]
        mra.ctx.sync()
        t0 = time.perf_counter()
        st = G.apply_batch(prec, outs, ops, trees, flags=flags)
        mra.ctx.sync()
        best_apply = min(best_apply, time.perf_counter() - t0)
        best_contract = min(best_contract, st[0]["t_contract"])
    return (best_apply if objective == "apply" else best_contract), best_contract


def autotune(mra, trees, ops, prec, flags=G.MWG_APPLY_MIXED, fp64_chunks=FP64_CHUNKS,
             slice_chunks=SLICE_CHUNKS, reps=3, install=True, objective="apply"):
    """Sweep (fp64_chunk, slice_chunk) on ``apply_batch(prec, ., ops, trees, flags)``.

    The FP64 chunk is tuned first with the slice chunk held at its current value, then
    the slice chunk (only when ``flags`` selects the mixed path, which has slices).
    ``objective`` "apply" minimises the apply's wall time, "contraction" the contraction
    kernels' device time.  Returns {"fp64_chunk", "slice_chunk", "objective", "t_best"
    (seconds of the objective), "t_contract" (seconds, the winner's), "table": [(fc, sc,
    t), ...], "seconds"}; installs the winner on ``mra.ctx`` when ``install``.
    """
    if objective not in ("apply", "contraction"):
        raise ValueError("objective must be 'apply' or 'contraction'")
    ctx = mra.ctx
    t0 = time.perf_counter()
    fc0, sc0 = ctx.get_tuning()
    table = []
    # one untimed apply so that operator levels and scratch pools exist
    G.apply_batch(prec, [G.FunctionTree(mra) for _ in trees], ops, trees, flags=flags)
    best = (float("inf"), fc0, sc0, float("inf"))
    for fc in fp64_chunks:
        ctx.set_tuning(fc, sc0)
        t, tc = _measure(mra, trees, ops, prec, flags, reps, objective)
        table.append((fc, sc0, t))
        if t < best[0]:
            best = (t, fc, sc0, tc)
    if flags & G.MWG_APPLY_MIXED:
        for sc in slice_chunks:
            if sc == sc0:
                continue
            ctx.set_tuning(best[1], sc)
            t, tc = _measure(mra, trees, ops, prec, flags, reps, objective)
            table.append((best[1], sc, t))
            if t < best[0]:
                best = (t, best[1], sc, tc)
    t_best, fc, sc, tc_best = best
    ctx.set_tuning(fc, sc) if install else ctx.set_tuning(fc0, sc0)
    return dict(fp64_chunk=fc, slice_chunk=sc, objective=objective, t_best=t_best, t_contract=tc_best,
                table=table, seconds=time.perf_counter() - t0)


def grid(fp64_chunks=FP64_CHUNKS, slice_chunks=SLICE_CHUNKS):
    """Full candidate grid (for exhaustive sweeps)."""
    return list(itertools.product(fp64_chunks, slice_chunks))
