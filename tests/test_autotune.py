"""Contraction task granularity: every setting gives the same tree and coefficients
(to FP64 rounding on the FP64 path and at k = 9; at k = 7 the mixed path's waves
sum their slices in FP32, so there a different grouping changes the result at the
FP32 rounding of the slices routed to FP32, far below prec), invalid settings are refused, and the
autotuner installs one of its candidates."""
import numpy as np
import pytest

from mrchem_amd import configs


@pytest.mark.gpu
def test_tuning_does_not_change_results():
    from mrchem_amd import mrcpp as G
    prec = 1e-5
    mra = G.mra_from(configs.BOX20, 7)
    orbs, nuc, mus = configs.orbitals(5, n_orb=2)
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    ops = [G.HelmholtzOperator(mra, mu, prec) for mu in mus]
    ctx = mra.ctx
    default = ctx.get_tuning()
    ref = None
    try:
        for fc, sc in [(64, 0), (1, 1), (7, 3), (4096, 16), (32, 4096)]:
            ctx.set_tuning(fc, sc)
            for flags in (G.MWG_APPLY_FP64, G.MWG_APPLY_MIXED):
                outs = [G.FunctionTree(mra) for _ in trees]
                G.apply_batch(prec, outs, ops, trees, flags=flags)
                ex = [o.export() for o in outs]
                if ref is None:
                    ref = {}
                if flags not in ref:
                    ref[flags] = ex
                    continue
                for (d0, l0, c0), (d1, l1, c1) in zip(ref[flags], ex):
                    assert np.array_equal(d0, d1) and np.array_equal(l0, l1)
                    tol = 1e-12 if flags == G.MWG_APPLY_FP64 else 1e-3 * prec
                    assert np.abs(c0 - c1).max() <= tol * np.abs(c0).max()
    finally:
        ctx.set_tuning(*default)


@pytest.mark.gpu
def test_invalid_tuning_is_refused():
    from mrchem_amd import mrcpp as G
    ctx = G.default_context()
    before = ctx.get_tuning()
    for fc, sc in [(0, 16), (64, -1), (5000, 16), (64, 4097)]:
        with pytest.raises(G.MwgError):
            ctx.set_tuning(fc, sc)
    assert ctx.get_tuning() == before


@pytest.mark.gpu
def test_autotune_installs_a_candidate():
    from mrchem_amd import mrcpp as G
    from mrchem_amd.autotune import autotune
    prec = 1e-4
    mra = G.mra_from(configs.BOX20, 7)
    orbs, nuc, mus = configs.orbitals(6, n_orb=2)
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    ops = [G.HelmholtzOperator(mra, mu, prec) for mu in mus]
    default = mra.ctx.get_tuning()
    try:
        r = autotune(mra, trees, ops, prec, G.MWG_APPLY_MIXED, fp64_chunks=(32, 64), slice_chunks=(8, 16), reps=1)
        assert (r["fp64_chunk"], r["slice_chunk"]) in {(fc, sc) for fc, sc, _ in r["table"]}
        assert mra.ctx.get_tuning() == (r["fp64_chunk"], r["slice_chunk"])
        assert r["t_best"] == min(t for _, _, t in r["table"])
        assert r["t_contract"] > 0
    finally:
        mra.ctx.set_tuning(*default)
