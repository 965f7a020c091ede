"""Committed golden fixtures (tests/golden/golden.json, written by
tests/golden/make_golden.py from the FP64 oracle and the closed forms): the seeded
inputs, the kernel fits and small applies.  CPU: the generator and the oracle still
reproduce them, and they agree with the analytic potentials.  GPU: the product
reproduces them."""
import json
import math
import os

import numpy as np
import pytest

from mrchem_amd import configs

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
PTS = np.array(GOLD["points"])


def _rec(g):
    coef, alpha, center, poly = g.packed()
    return coef.tolist(), alpha.tolist(), center.reshape(-1).tolist(), poly.reshape(-1).tolist()


def _same_gauss(g, rec):
    c, a, R, p = _rec(g)
    return (np.array_equal(c, rec["coef"]) and np.array_equal(a, rec["alpha"]) and np.array_equal(R, rec["center"])
            and np.array_equal(p, rec["poly"]))


def test_seeded_inputs_unchanged():
    for rec in GOLD["config1"]:
        g, parts = configs.density3(rec["seed"])
        assert _same_gauss(g, rec["density"])
    c2 = GOLD["config2"]
    orbs, nuc, mus = configs.orbitals(c2["seed"], n_orb=5)
    assert np.array_equal(nuc.reshape(-1), c2["nuclei"])
    assert np.array_equal(mus, c2["mus"])
    for o, rec in zip(orbs, c2["orbitals"]):
        assert np.abs(np.array(_rec(o)[0]) - rec["coef"]).max() <= 1e-13 * np.abs(rec["coef"]).max()


def test_golden_apply_values_match_closed_forms():
    """The stored oracle outputs agree with the stored analytic potentials (config 1:
    erf closed form; config 2: erfc / p-derivative closed form), to the apply's prec."""
    for rec in GOLD["config1"]:
        v, ex = np.array(rec["apply"]["values"]), np.array(rec["apply"]["analytic"])
        assert np.abs(v - ex).max() < 2e-3 * np.abs(ex).max()
    for rec in GOLD["config2"]["applies"]:
        # K * f with K = exp(-mu r)/r; the analytic value is the same convolution in closed form
        v, ex = np.array(rec["apply"]["values"]), np.array(rec["apply"]["analytic"])
        assert np.abs(v - ex).max() < 2e-5 * np.abs(ex).max()


def test_oracle_reproduces_golden(oracle):
    M = oracle
    mra16 = M.MRA(**configs.BOX16, k=5)
    P = M.poisson(mra16, 1e-4)
    a, w, info = P.terms()
    gf = GOLD["fit_poisson_box16_prec1e-4"]
    assert len(a) == gf["n_terms"]
    assert np.allclose(a, gf["a"], rtol=1e-13, atol=0) and np.allclose(w, gf["w"], rtol=1e-13, atol=0)
    for rec in GOLD["config1"]:
        g, parts = configs.density3(rec["seed"])
        t = M.Tree(mra16).project_gauss(1e-4, g)
        assert t.n_nodes == rec["input_nodes"]
        out = M.Tree(mra16)
        st = M.apply(1e-4, out, P, t)
        assert out.n_nodes == rec["apply"]["n_nodes"] and st["tuples"] == rec["apply"]["tuples"]
        assert out.sqnorm == pytest.approx(rec["apply"]["sqnorm"], rel=1e-12)
        assert np.allclose(out.eval(PTS), rec["apply"]["values"], rtol=1e-10, atol=1e-13)


@pytest.mark.gpu
def test_product_reproduces_golden():
    from mrchem_amd import mrcpp as G
    mra16 = G.mra_from(configs.BOX16, 5)
    P = G.PoissonOperator(mra16, 1e-4)
    a, w = P.terms()
    gf = GOLD["fit_poisson_box16_prec1e-4"]
    assert len(a) == gf["n_terms"]
    assert np.allclose(a, gf["a"], rtol=1e-12, atol=0) and np.allclose(w, gf["w"], rtol=1e-12, atol=0)
    for rec in GOLD["config1"]:
        g, parts = configs.density3(rec["seed"])
        t = G.project(1e-4, G.FunctionTree(mra16), g)
        assert t.getNNodes() == rec["input_nodes"]
        out = G.FunctionTree(mra16)
        st = G.apply(1e-4, out, P, t)
        assert out.getNNodes() == rec["apply"]["n_nodes"] and st["tuples"] == rec["apply"]["tuples"]
        assert out.getSquareNorm() == pytest.approx(rec["apply"]["sqnorm"], rel=1e-11)
        assert np.allclose(out.evalf(PTS), rec["apply"]["values"], rtol=1e-10, atol=1e-13)
    mra20 = G.mra_from(configs.BOX20, 7)
    orbs, nuc, mus = configs.orbitals(0, n_orb=5)
    for rec in GOLD["config2"]["applies"]:
        i = rec["orbital"]
        H = G.HelmholtzOperator(mra20, mus[i], 1e-5)
        t = G.project(1e-5, G.FunctionTree(mra20), orbs[i])
        assert t.getNNodes() == rec["input_nodes"]
        out = G.FunctionTree(mra20)
        st = G.apply(1e-5, out, H, t)
        assert out.getNNodes() == rec["apply"]["n_nodes"] and st["tuples"] == rec["apply"]["tuples"]
        assert out.getSquareNorm() == pytest.approx(rec["apply"]["sqnorm"], rel=1e-11)
        v = np.array(rec["apply"]["values"])
        assert np.abs(out.evalf(PTS) - v).max() <= 1e-10 * np.abs(v).max()
