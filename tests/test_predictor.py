"""Learned refinement predictor: the MLP oracle against finite differences (CPU), the
device forward/backward against the oracle (GPU), and the apply with the predictor
against the plain apply (GPU)."""
import math

import numpy as np
import pytest

from mrchem_amd import configs

NF = 10
NP = 64 * NF + 64 + 64 * 64 + 64 + 64 + 1


def _data(n=300, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, NF)).astype(np.float32) * 3.0 + 1.0
    y = (X[:, 0] + 0.5 * X[:, 5] - 0.2 * X[:, 3] > 1.0).astype(np.float32)
    P = (rng.normal(size=NP) * 0.2).astype(np.float32)
    mean = X.mean(0).astype(np.float32)
    sdev = X.std(0).astype(np.float32)
    return X, y, P, mean, sdev


def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp_ref
    X, y, P, mean, sdev = _data(50)
    P = P.astype(np.float64)
    g = mlp_ref.gradient(P, X, y, mean, sdev)
    rng = np.random.default_rng(1)
    for k in rng.choice(NP, 40, replace=False):
        e = np.zeros(NP)
        h = 1e-6
        e[k] = h
        fd = (mlp_ref.loss(P + e, X, y, mean, sdev) - mlp_ref.loss(P - e, X, y, mean, sdev)) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 + 1e-4 * abs(fd), k


@pytest.mark.gpu
def test_device_gradient_matches_oracle():
    from mrchem_amd import mrcpp as G
    from oracle import mlp_ref
    X, y, P, mean, sdev = _data(600)
    pr = G.Predictor(seed=3)
    pr.set_params(P, mean, sdev)
    gd = pr.gradient(X, y)
    go = mlp_ref.gradient(P.astype(np.float64), X, y, mean, sdev)
    assert np.abs(gd - go).max() <= 1e-5 + 2e-4 * np.abs(go).max()
    pd = pr.infer(X)
    po, _ = mlp_ref.forward(P, X, mean, sdev)
    assert np.abs(pd - po).max() < 1e-5


@pytest.mark.gpu
def test_training_reduces_loss():
    from mrchem_amd import mrcpp as G
    X, y, P, mean, sdev = _data(2000, seed=4)
    pr = G.Predictor(seed=5)
    lb, la = pr.train(X, y, epochs=60, batch=256, lr=3e-3, seed=1)
    assert la < 0.5 * lb
    acc = ((pr.infer(X) > 0.5) == (y > 0.5)).mean()
    assert acc > 0.95


@pytest.mark.gpu
def test_predictor_keeps_the_plain_tree_and_saves_iterations():
    """Train on traces of plain applies, then apply with the predictor: the norm test
    still judges every node (same output tree), fewer refinement iterations."""
    from mrchem_amd import mrcpp as G
    prec = 1e-5
    mra = G.mra_from(configs.BOX20, 7)
    orbs, nuc, mus = configs.orbitals(2, n_orb=4)
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    ops = [G.HelmholtzOperator(mra, mu, prec) for mu in mus]
    G.trace_enable(True)
    plain = [G.FunctionTree(mra) for _ in trees]
    st0 = G.apply_batch(prec, plain, ops, trees)
    X, y = G.trace_get()
    G.trace_enable(False)
    assert len(y) > 100 and 0.05 < y.mean() < 0.95
    pr = G.Predictor(seed=0)
    pr.train(X, y, epochs=150, batch=256, lr=3e-3)
    pr.install(0.5)
    outs = [G.FunctionTree(mra) for _ in trees]
    st1 = G.apply_batch(prec, outs, ops, trees, flags=G.MWG_APPLY_PREDICT)
    assert st1[0]["iters"] < st0[0]["iters"]
    assert st1[0]["predicted_splits"] > 0
    for a, b, s0, s1 in zip(outs, plain, st0, st1):
        assert abs(s1["nodes_out"] - s0["nodes_out"]) <= 0.02 * s0["nodes_out"]
        pts = np.random.default_rng(0).uniform(-20, 20, (5000, 3))
        d = a.evalf(pts) - b.evalf(pts)
        assert math.sqrt((d ** 2).mean() * 40 ** 3) < prec * math.sqrt(b.getSquareNorm())
