"""Writes tests/golden/golden.json: fixed inputs and expected outputs for the apply path.

Everything here comes from this project's own FP64 oracle (oracle/libmwref.so) and the
closed forms in oracle/analytic.py; nothing is imported or run from the reference.
Contents (one entry per case):
  * inputs: the seeded synthetic densities/orbitals (config 1: 3 normalised Gaussian
    charges, seeds 0..2; config 2: 5 orthonormalised s/p orbitals, seed 0), so that a
    change of the generator is caught;
  * kernel fits: term count, exponents and weights of the Poisson / Helmholtz Gaussian
    sums at the precisions used below;
  * applies: node count, tuple count, square norm and values at fixed points of the
    oracle's output, with the analytic potential at the same points.
Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mrchem_amd import configs  # noqa: E402
from oracle import analytic  # noqa: E402
from oracle import mwref as M  # noqa: E402

POINTS = np.random.default_rng(12345).uniform(-3.0, 3.0, (24, 3))


def gauss_record(g):
    coef, alpha, center, poly = g.packed()
    return dict(coef=coef.tolist(), alpha=alpha.tolist(), center=center.reshape(-1).tolist(),
                poly=poly.reshape(-1).tolist())


def fit_record(op):
    a, w, info = op.terms()
    return dict(n_terms=int(len(a)), a=a.tolist(), w=w.tolist(), err=float(info["err"]))


def apply_record(prec, op, tree, mra, analytic_vals):
    out = M.Tree(mra)
    st = M.apply(prec, out, op, tree)
    return dict(n_nodes=int(out.n_nodes), tuples=int(st["tuples"]), sqnorm=float(out.sqnorm),
                values=out.eval(POINTS).tolist(), analytic=list(map(float, analytic_vals)))


def main():
    M.set_threads(8)
    gold = dict(points=POINTS.tolist(), config1=[], config2={})
    mra16 = M.MRA(**configs.BOX16, k=5)
    P16 = M.poisson(mra16, 1e-4)
    gold["fit_poisson_box16_prec1e-4"] = fit_record(P16)
    for seed in range(3):
        g, parts = configs.density3(seed)
        t = M.Tree(mra16).project_gauss(1e-4, g)
        rec = dict(seed=seed, density=gauss_record(g), input_nodes=int(t.n_nodes), input_sqnorm=float(t.sqnorm))
        rec["apply"] = apply_record(1e-4, P16, t, mra16, analytic.erf_potential(parts, POINTS))
        gold["config1"].append(rec)
    orbs, nuc, mus = configs.orbitals(0, n_orb=5)
    mra20 = M.MRA(**configs.BOX20, k=7, max_depth=16)
    prec = 1e-5
    c2 = dict(seed=0, mus=list(map(float, mus)), nuclei=nuc.reshape(-1).tolist(),
              orbitals=[gauss_record(o) for o in orbs], applies=[])
    for i in (0, 3):
        H = M.helmholtz(mra20, mus[i], prec)
        t = M.Tree(mra20).project_gauss(prec, orbs[i])
        rec = dict(orbital=i, mu=float(mus[i]), fit=fit_record(H), input_nodes=int(t.n_nodes),
                   input_sqnorm=float(t.sqnorm))
        rec["apply"] = apply_record(prec, H, t, mra20, analytic.potential(orbs[i], POINTS, mus[i]))
        c2["applies"].append(rec)
    gold["config2"] = c2
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as fh:
        json.dump(gold, fh, indent=0)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
