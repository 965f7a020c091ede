"""GPU parity: the HIP apply (through the C ABI) against the FP64 oracle on the same
seeded inputs, plus closed forms and size-independent properties."""
import math

import numpy as np
import pytest

from mrchem_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    from mrchem_amd import mrcpp
    mrcpp.default_context()
    return mrcpp


@pytest.fixture(scope="module")
def A():
    from oracle import analytic
    return analytic


def _same_tree(t_gpu, t_ref, rtol=1e-12):
    d1, l1, c1 = t_gpu.export()
    d2, l2, c2 = t_ref.export()
    assert len(d1) == len(d2)
    k1 = {(int(d), *map(int, l)): i for i, (d, l) in enumerate(zip(d1, l1))}
    idx = np.array([k1[(int(d), *map(int, l))] for d, l in zip(d2, l2)])
    scale = np.abs(c2).max()
    err = np.abs(c1[idx] - c2).max()
    assert err <= rtol * scale, (err, scale)
    return err / scale


def test_library_exports_every_symbol():
    import ctypes
    from mrchem_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in _lib.EXPORTED:
        assert hasattr(L, name)


@pytest.mark.parametrize("k,basis", [(3, "L"), (4, "L"), (5, "L"), (6, "L"), (7, "L"), (8, "L"), (9, "L"),
                                     (5, "I"), (7, "I"), (9, "I")])
def test_projection_matches_oracle(G, oracle, k, basis):
    g, _ = configs.density3(1)
    prec = 1e-5
    mra = G.mra_from(configs.BOX16, k, basis)
    omra = oracle.MRA(**configs.BOX16, k=k, basis=basis, max_depth=16)
    f = G.project(prec, G.FunctionTree(mra), g)
    of = oracle.Tree(omra).project_gauss(prec, g)
    _same_tree(f, of, 1e-13)
    assert f.getSquareNorm() == pytest.approx(of.sqnorm, rel=1e-13)


@pytest.mark.parametrize("mu", [0.0, 1.1])
def test_operator_blocks_match_oracle(G, oracle, mu):
    mra = G.mra_from(configs.BOX20, 7)
    omra = oracle.MRA(**configs.BOX20, k=7, max_depth=16)
    P = G.HelmholtzOperator(mra, mu, 1e-6) if mu > 0 else G.PoissonOperator(mra, 1e-6)
    OP = oracle.Oper(omra, mu, 1e-6)
    a1, w1 = P.terms()
    a2, w2, _ = OP.terms()
    np.testing.assert_allclose(a1, a2, rtol=1e-14)
    np.testing.assert_allclose(w1, w2, rtol=1e-14)
    for term in range(0, P.size(), 7):
        for depth in (0, 2, 5):
            m = min(3, (2 << depth) - 1)
            b1 = P.blocks(term, depth, -m, m)
            b2 = OP.blocks(term, depth, -m, m)
            assert np.abs(b1 - b2).max() <= 1e-13 * max(np.abs(b2).max(), 1e-300)


def _apply_pair(G, oracle, box, k, prec, g, mu, basis="L"):
    mra = G.mra_from(box, k, basis)
    omra = oracle.MRA(**box, k=k, basis=basis, max_depth=16)
    of = oracle.Tree(omra).project_gauss(prec, g)
    f = G.FunctionTree(mra).load(*of.export())   # the same input tree on both sides
    op = G.HelmholtzOperator(mra, mu, prec) if mu > 0 else G.PoissonOperator(mra, prec)
    oop = oracle.Oper(omra, mu, prec)
    out = G.FunctionTree(mra)
    st = G.apply(prec, out, op, f)
    oout = oracle.Tree(omra)
    ost = oracle.apply(prec, oout, oop, of)
    return out, oout, st, ost


@pytest.mark.parametrize("seed", [0, 1])
def test_poisson_apply_matches_oracle_and_erf(G, oracle, A, seed):
    prec = 1e-4
    g, parts = configs.density3(seed)
    out, oout, st, ost = _apply_pair(G, oracle, configs.BOX16, 5, prec, g, 0.0)
    assert st["tuples"] == ost["tuples"]
    assert st["nodes_out"] == ost["nodes_out"]
    _same_tree(out, oout, 1e-12)
    pts = np.random.default_rng(seed).uniform(-16, 16, (20000, 3))
    err = out.evalf(pts) - A.erf_potential(parts, pts)
    assert math.sqrt((err ** 2).mean() * 32 ** 3) < prec * math.sqrt(out.getSquareNorm())


def test_helmholtz_apply_matches_oracle_and_erfc(G, oracle, A):
    prec = 1e-5
    orbs, nuc, mus = configs.orbitals(0, n_orb=2)
    out, oout, st, ost = _apply_pair(G, oracle, configs.BOX20, 7, prec, orbs[1], mus[1])
    assert st["tuples"] == ost["tuples"]
    assert st["nodes_out"] == ost["nodes_out"]
    _same_tree(out, oout, 1e-12)
    pts = np.random.default_rng(3).uniform(-20, 20, (20000, 3))
    err = out.evalf(pts) - A.potential(orbs[1], pts, mus[1])
    assert math.sqrt((err ** 2).mean() * 40 ** 3) < prec * math.sqrt(out.getSquareNorm())


def test_interpolating_basis_apply_matches_oracle(G, oracle):
    prec = 1e-4
    g, _ = configs.density3(2)
    out, oout, st, ost = _apply_pair(G, oracle, configs.BOX16, 5, prec, g, 0.0, basis="I")
    assert st["tuples"] == ost["tuples"]
    _same_tree(out, oout, 1e-12)


def test_apply_k9(G, oracle):
    prec = 1e-4
    g, _ = configs.density3(3)
    out, oout, st, ost = _apply_pair(G, oracle, configs.BOX16, 9, prec, g, 0.0)
    assert st["tuples"] == ost["tuples"]
    _same_tree(out, oout, 1e-12)


@pytest.mark.parametrize("k", [3, 4, 6, 8])
def test_apply_other_orders_match_oracle(G, oracle, k):
    """The remaining orders of the templated kernels (K = 4, 5, 7, 9): same tuples, grid
    and coefficients as the oracle (Poisson, Legendre)."""
    prec = 1e-4
    g, _ = configs.density3(2)
    out, oout, st, ost = _apply_pair(G, oracle, configs.BOX16, k, prec, g, 0.0)
    assert st["tuples"] == ost["tuples"] and st["nodes_out"] == ost["nodes_out"]
    _same_tree(out, oout, 1e-12)


@pytest.mark.parametrize("basis", ["L", "I"])
def test_helmholtz_apply_k9_matches_oracle(G, oracle, basis):
    """k=9 (the FP64 contraction with two operator rows live) on a bound-state Helmholtz
    kernel, both bases: same tuples, grid and coefficients as the oracle."""
    prec = 1e-4
    orbs, nuc, mus = configs.orbitals(0, n_orb=2)
    out, oout, st, ost = _apply_pair(G, oracle, configs.BOX20, 9, prec, orbs[1], mus[1], basis=basis)
    assert st["tuples"] == ost["tuples"] and st["nodes_out"] == ost["nodes_out"]
    _same_tree(out, oout, 1e-12)


@pytest.mark.parametrize("max_iter,max_depth", [(2, 16), (-1, 3)])
def test_refinement_limits_match_oracle(G, oracle, max_iter, max_depth):
    """mrcpp::apply's maxIter and the MRA's maximum depth cap the output tree the same way
    on both sides (same grid, same tuples, coefficients to 1e-12)."""
    prec = 1e-5
    g, _ = configs.density3(1)
    mra = G.mra_from(configs.BOX16, 5, max_depth=max_depth)
    omra = oracle.MRA(**configs.BOX16, k=5, max_depth=max_depth)
    of = oracle.Tree(omra).project_gauss(prec, g)
    f = G.FunctionTree(mra).load(*of.export())
    out, oout = G.FunctionTree(mra), oracle.Tree(omra)
    st = G.apply(prec, out, G.PoissonOperator(mra, prec), f, max_iter=max_iter)
    ost = oracle.apply(prec, oout, oracle.Oper(omra, 0.0, prec), of, max_iter=max_iter)
    assert st["tuples"] == ost["tuples"] and st["nodes_out"] == ost["nodes_out"]
    assert out.export()[0].max() <= (max_iter if max_iter >= 0 else max_depth)
    _same_tree(out, oout, 1e-12)


def test_non_cubic_root_grid_matches_oracle(G, oracle):
    """A 3 x 2 x 2 root grid (ragged world box, mrenv.cpp's corner/boxes): same grid, same
    tuples and coefficients to 1e-12 as the oracle, also across the root boundaries."""
    prec = 1e-4
    g, _ = configs.density3(1)
    box = dict(n0=-4, corner=(-2, -1, -1), boxes=(3, 2, 2), sfac=1.0)
    out, oout, st, ost = _apply_pair(G, oracle, box, 5, prec, g, 0.0)
    assert st["tuples"] == ost["tuples"] and st["nodes_out"] == ost["nodes_out"]
    _same_tree(out, oout, 1e-12)


def test_batch_equals_single(G, oracle):
    prec = 1e-4
    mra = G.mra_from(configs.BOX16, 5)
    gs = [configs.density3(s)[0] for s in (4, 5, 6)]
    fs = [G.project(prec, G.FunctionTree(mra), g) for g in gs]
    P = G.PoissonOperator(mra, prec)
    H = G.HelmholtzOperator(mra, 0.8, prec)
    ops = [P, H, P]
    singles = []
    for f, op in zip(fs, ops):
        o = G.FunctionTree(mra)
        G.apply(prec, o, op, f)
        singles.append(o)
    outs = [G.FunctionTree(mra) for _ in fs]
    G.apply_batch(prec, outs, ops, fs)
    for a, b in zip(outs, singles):
        d1, l1, c1 = a.export()
        d2, l2, c2 = b.export()
        assert np.array_equal(d1, d2) and np.array_equal(l1, l2)
        assert np.abs(c1 - c2).max() <= 1e-13 * np.abs(c2).max()


def test_large_batch_equals_single(G):
    """A batch of 18 applies (large batches take the split tests of their jobs in
    parallel on the host): the same grids as one apply each, coefficients to FP64 rounding."""
    mra = G.mra_from(configs.BOX16, 5)
    gs = [configs.density3(s)[0] for s in (4, 5, 6)]
    fs = [G.project(1e-4, G.FunctionTree(mra), g) for g in gs]
    P = G.PoissonOperator(mra, 1e-4)
    H = G.HelmholtzOperator(mra, 0.8, 1e-4)
    jobs = [(fs[q % 3], (P, H)[q % 2], (1e-4, 3e-4, 1e-3)[q % 3]) for q in range(18)]
    singles = []
    for f, op, p in jobs:
        o = G.FunctionTree(mra)
        G.apply(p, o, op, f)
        singles.append(o)
    outs = [G.FunctionTree(mra) for _ in jobs]
    G.apply_batch([p for _, _, p in jobs], outs, [op for _, op, _ in jobs], [f for f, _, _ in jobs])
    for a, b in zip(outs, singles):
        d1, l1, c1 = a.export()
        d2, l2, c2 = b.export()
        assert np.array_equal(d1, d2) and np.array_equal(l1, l2)
        assert np.abs(c1 - c2).max() <= 1e-13 * np.abs(c2).max()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_operators_built_together_equal_operators_built_alone(G, mode):
    """Operator levels built for a whole batch at once (all operators and depths in two
    round trips, one allocation per operator with its split-FP16 copies, down to one depth
    below the inputs) give the same applies as levels built for one operator at a time:
    the same grids, coefficients to FP64 rounding (plain) or to the rounding of the FP32
    slice sums (mixed: the order of a node's slices follows how the batch's work list is
    split over workgroups)."""
    prec = 1e-5
    flags = G.MWG_APPLY_MIXED if mode == "mixed" else G.MWG_APPLY_FP64
    mra = G.mra_from(configs.BOX20, 7)
    orbs, nuc, mus = configs.orbitals(7, n_orb=3)
    fs = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    jobs = [(fs[0], mus[0]), (fs[1], mus[1]), (fs[2], mus[2]), (fs[0], 0.0)]

    def make(mu):
        return G.HelmholtzOperator(mra, mu, prec) if mu > 0 else G.PoissonOperator(mra, prec)

    alone = []
    for f, mu in jobs:
        o = G.FunctionTree(mra)
        G.apply_batch(prec, [o], [make(mu)], [f], flags=flags)
        alone.append(o)
    outs = [G.FunctionTree(mra) for _ in jobs]
    G.apply_batch(prec, outs, [make(mu) for _, mu in jobs], [f for f, _ in jobs], flags=flags)
    tol = 1e-13 if mode == "fp64" else 1e-2 * prec
    for a, b in zip(outs, alone):
        d1, l1, c1 = a.export()
        d2, l2, c2 = b.export()
        assert np.array_equal(d1, d2) and np.array_equal(l1, l2)
        assert np.abs(c1 - c2).max() <= tol * np.abs(c2).max()


def test_reference_coulomb_values_on_gpu(G, A):
    """tests/qmoperators/coulomb_operator.cpp:76-82 at its settings, through the C ABI."""
    prec = 1e-3
    mra = G.mra_from(configs.BOX32, 5, "I", max_depth=18)
    orbs = {"1s": (1, 0, 0), "2s": (2, 0, 0), "2px": (2, 1, 0), "2py": (2, 1, 1), "2pz": (2, 1, 2)}
    rho = G.project_hydrogen(prec, G.FunctionTree(mra), [(2.0, q, q) for q in orbs.values()])
    J = G.FunctionTree(mra)
    G.apply(prec / math.sqrt(rho.getSquareNorm()), J, G.PoissonOperator(mra, prec), rho)
    for name, key in [("1s", "1s"), ("2s", "2s"), ("2pz", "2p")]:
        sq = G.project_hydrogen(prec, G.FunctionTree(mra), [(1.0, orbs[name], orbs[name])])
        assert G.dot(sq, J) == pytest.approx(A.REFERENCE_E_P[key], rel=3e-5), key


def test_errors_are_reported(G):
    mra = G.mra_from(configs.BOX16, 5)
    with pytest.raises(G.MwgError):
        G.HelmholtzOperator(mra, -1.0, 1e-4)
    with pytest.raises(G.MwgError):
        G.apply(1e-4, G.FunctionTree(mra), G.PoissonOperator(mra, 1e-4), G.FunctionTree(mra))
    with pytest.raises(G.MwgError):  # blocks below the MRA's deepest scale
        G.PoissonOperator(mra, 1e-4).blocks(0, mra.max_depth + 1, 0, 0)


def test_trim_returns_cached_memory(G):
    """mwg_ctx_trim hands the device blocks the context keeps for reuse back to the device;
    applies afterwards allocate again and give the same trees."""
    prec = 1e-4
    mra = G.mra_from(configs.BOX16, 5)
    f = G.project(prec, G.FunctionTree(mra), configs.density3(4)[0])
    P = G.PoissonOperator(mra, prec)
    o = G.FunctionTree(mra)
    G.apply(prec, o, P, f)
    ref = o.export()
    del o
    assert mra.ctx.trim() > 0
    assert mra.ctx.trim() == 0
    o2 = G.FunctionTree(mra)
    G.apply(prec, o2, P, f)
    d, l, c = o2.export()
    assert np.array_equal(d, ref[0]) and np.array_equal(l, ref[1]) and np.array_equal(c, ref[2])


def _l2_diff(a, b, box_width, n=20000, seed=7):
    pts = np.random.default_rng(seed).uniform(-box_width / 2, box_width / 2, (n, 3))
    d = a.evalf(pts) - b.evalf(pts)
    return math.sqrt((d ** 2).mean() * box_width ** 3)


@pytest.mark.parametrize("i", [0, 1])
def test_mixed_precision_within_prec_of_fp64(G, A, i):
    """Split-FP16 MFMA slices (FP32 sums of <= 64 slices per wave, FP64 beyond; k=7) vs the plain FP64 apply: the
    output differs by less than prec * ||out|| (L2), and both meet the erfc closed form."""
    prec = 1e-6
    orbs, nuc, mus = configs.orbitals(0, n_orb=2)
    mra = G.mra_from(configs.BOX20, 7)
    f = G.project(prec, G.FunctionTree(mra), orbs[i])
    H = G.HelmholtzOperator(mra, mus[i], prec)
    o64 = G.FunctionTree(mra)
    s64 = G.apply(prec, o64, H, f, flags=G.MWG_APPLY_FP64)
    o32 = G.FunctionTree(mra)
    s32 = G.apply(prec, o32, H, f, flags=G.MWG_APPLY_MIXED)
    assert s32["slices"] > 0
    norm = math.sqrt(o64.getSquareNorm())
    assert _l2_diff(o32, o64, 40.0) < prec * norm
    pts = np.random.default_rng(11).uniform(-20, 20, (20000, 3))
    err = o32.evalf(pts) - A.potential(orbs[i], pts, mus[i])
    assert math.sqrt((err ** 2).mean() * 40 ** 3) < prec * norm
    assert abs(s32["nodes_out"] - s64["nodes_out"]) <= 0.05 * s64["nodes_out"]


def test_mixed_precision_interpolating_basis(G):
    """The MFMA path on the Interpolating basis (k=7): same grid as the plain FP64
    apply, difference below prec * ||out|| (L2)."""
    prec = 1e-6
    orbs, nuc, mus = configs.orbitals(0, n_orb=2)
    mra = G.mra_from(configs.BOX20, 7, "I")
    f = G.project(prec, G.FunctionTree(mra), orbs[1])
    H = G.HelmholtzOperator(mra, mus[1], prec)
    o64, o32 = G.FunctionTree(mra), G.FunctionTree(mra)
    s64 = G.apply(prec, o64, H, f, flags=G.MWG_APPLY_FP64)
    s32 = G.apply(prec, o32, H, f, flags=G.MWG_APPLY_MIXED)
    assert s32["slices"] > 0 and s32["nodes_out"] == s64["nodes_out"]
    assert _l2_diff(o32, o64, 40.0) < prec * math.sqrt(o64.getSquareNorm())


def test_mixed_poisson_energy(G, oracle):
    prec = 1e-6
    g, parts = configs.density3(0)
    mra = G.mra_from(configs.BOX16, 7)
    f = G.project(prec, G.FunctionTree(mra), g)
    P = G.PoissonOperator(mra, prec)
    o64, o32 = G.FunctionTree(mra), G.FunctionTree(mra)
    G.apply(prec, o64, P, f)
    G.apply(prec, o32, P, f, flags=G.MWG_APPLY_MIXED)
    e64, e32 = G.dot(f, o64), G.dot(f, o32)
    assert abs(e32 - e64) < prec * abs(e64)


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7, 8, 9])
def test_multiply_matches_oracle(G, oracle, k):
    """mrcpp::multiply on the device vs the oracle on the same input trees."""
    prec = 1e-5
    orbs, nuc, mus = configs.orbitals(1, n_orb=3)
    mra = G.mra_from(configs.BOX20, k)
    omra = oracle.MRA(**configs.BOX20, k=k, max_depth=16)
    oa = oracle.Tree(omra).project_gauss(prec, orbs[0])
    ob = oracle.Tree(omra).project_gauss(prec, orbs[2])
    a = G.FunctionTree(mra).load(*oa.export())
    b = G.FunctionTree(mra).load(*ob.export())
    p = G.multiply(prec, G.FunctionTree(mra), 1.5, a, b)
    op = oracle.multiply(prec, oracle.Tree(omra), 1.5, oa, ob)
    _same_tree(p, op, 1e-12)
    from mrchem_amd import gaussians as GS
    ex = GS.product(orbs[0], orbs[2]).scaled(1.5)
    pts = np.random.default_rng(5).uniform(-3, 3, (2000, 3))
    assert np.abs(p.evalf(pts) - ex(pts)).max() < 100 * prec * np.abs(ex(pts)).max()


@pytest.mark.parametrize("k", [5, 9])
def test_add_is_the_exact_linear_combination(G, oracle, k):
    """mrcpp::add on the device: on the union of the two grids the refinement of the
    coarser tree is exact, so the sum evaluates to ca a(x) + cb b(x) to rounding."""
    prec = 1e-5
    orbs, nuc, mus = configs.orbitals(3, n_orb=2)
    mra = G.mra_from(configs.BOX20, k)
    omra = oracle.MRA(**configs.BOX20, k=k, max_depth=16)
    oa = oracle.Tree(omra).project_gauss(prec, orbs[0])
    ob = oracle.Tree(omra).project_gauss(prec * 10, orbs[1])
    a = G.FunctionTree(mra).load(*oa.export())
    b = G.FunctionTree(mra).load(*ob.export())
    s = G.add(G.FunctionTree(mra), 0.7, a, -1.3, b)
    assert s.getNNodes() >= max(a.getNNodes(), b.getNNodes())
    pts = np.random.default_rng(11).uniform(-6, 6, (3000, 3))
    ex = 0.7 * a.evalf(pts) - 1.3 * b.evalf(pts)
    assert np.abs(s.evalf(pts) - ex).max() <= 1e-12 * np.abs(ex).max()


def test_handles_destroyed_in_any_order(G):
    """A garbage collector may finalise the context and the MRA before the trees made on
    them; every handle keeps what it was made on, so the tree still works and frees last."""
    g, _ = configs.density3(1)
    ctx = G.Context(0)
    mra = G.mra_from(configs.BOX16, 5, ctx=ctx)
    t = G.project(1e-4, G.FunctionTree(mra), g)
    ref = G.project(1e-4, G.FunctionTree(G.mra_from(configs.BOX16, 5)), g)
    n, nrm = t.getNNodes(), t.getSquareNorm()
    t2 = G.FunctionTree(mra)
    ctx.__del__()
    mra.__del__()
    assert t.getNNodes() == n == ref.getNNodes()
    pts = np.random.default_rng(5).uniform(-4, 4, (500, 3))
    assert np.array_equal(t.evalf(pts), ref.evalf(pts))
    G.add(t2, 1.0, t, 1.0, t)
    assert t2.getSquareNorm() == pytest.approx(4.0 * nrm, rel=1e-12)
    del t, t2


@pytest.mark.parametrize("k", [5, 9])
def test_multiply_batch_equals_single_products(G, k):
    """mwg_multiply_batch: every product (pairs with a repeated factor, different
    precisions and coefficients) is bit-identical to its own mwg_multiply."""
    orbs, nuc, mus = configs.orbitals(2, n_orb=3)
    mra = G.mra_from(configs.BOX20, k)
    t = [G.project(1e-5, G.FunctionTree(mra), o) for o in orbs]
    spec = [(1e-5, 1.0, 0, 1), (1e-4, -0.5, 2, 2), (3e-6, 2.0, 1, 2), (1e-5, 1.0, 0, 0)]
    single = [G.multiply(p, G.FunctionTree(mra), c, t[i], t[j]) for p, c, i, j in spec]
    batch = G.multiply_batch([q[0] for q in spec], [G.FunctionTree(mra) for _ in spec], [q[1] for q in spec],
                             [t[q[2]] for q in spec], [t[q[3]] for q in spec])
    for a, b in zip(batch, single):
        (d0, l0, c0), (d1, l1, c1) = a.export(), b.export()
        assert np.array_equal(d0, d1) and np.array_equal(l0, l1) and np.array_equal(c0, c1)
        assert a.getSquareNorm() == b.getSquareNorm()
    with pytest.raises(G.MwgError):
        G.multiply_batch(1e-5, [t[0]], 1.0, [t[0]], [t[1]])
