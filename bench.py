#!/usr/bin/env python3
"""Benchmark: output nodes/s of the Poisson/Helmholtz apply at prec 1e-6 on one MI355X.

Workload (one GPU): Helmholtz apply (mu ~ U[0.5,2]) to 5 synthetic orbitals, each a
sum of 6 s/p Gaussians on 3 seeded nuclei within 2 bohr (exponents log-U[0.3,30],
coefficients N(0,1), Gram-Schmidt orthonormalised), projected at k=7, prec 1e-6 in
the box [-20,20]^3.  One step = the 5 applies as one batch, inputs resident in HBM,
operators built beforehand (their construction time is reported separately).

Prints one JSON line.  ``value`` is output nodes/s of the fastest variant; the
plain FP64 GPU apply's rate is reported beside it.  ``cpu_baseline`` times the FP64
CPU oracle (OpenMP) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12, "f16": 2.5e15}


def mixed_dtype(k):
    """Arithmetic of the mixed path's slice kernels: split-FP16 operands (hi + lo) on f16 MFMA
    at k = 7 (FP32 sums over a task's slices) and at k = 9 (FP64 sums)."""
    return "f16"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_inputs(G, configs, k, prec, n_orb, seed):
    mra = G.mra_from(configs.BOX20, k, "L", max_depth=16)
    orbs, nuc, mus = configs.orbitals(seed, n_orb=n_orb)
    t0 = time.perf_counter()
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    t_proj = time.perf_counter() - t0
    t0 = time.perf_counter()
    ops = [G.HelmholtzOperator(mra, mu, prec) for mu in mus]
    t_oper = time.perf_counter() - t0
    return mra, orbs, mus, trees, ops, t_proj, t_oper


def _l2diff(G, a, b):
    """||a - b|| exactly: the multiwavelet coefficients of a - b on the union grid (mwg_add)."""
    return math.sqrt(max(G.add(G.FunctionTree(a.mra), 1.0, a, -1.0, b).getSquareNorm(), 0.0))


def max_rel_err(G, outs, refs):
    """Largest ||out - ref|| / ||ref|| over the outputs of a batch (exact coefficient L2)."""
    return max(_l2diff(G, a, b) / math.sqrt(b.getSquareNorm()) for a, b in zip(outs, refs))


def prec_tag(p):
    """1e-06 -> '1e-6' (the precision as the profile file names spell it)."""
    return f"{p:g}".replace("e-0", "e-").replace("e+0", "e+")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter summary
    (profiles/hbm_traffic_<kernel>_<workload>.txt: FETCH_SIZE and WRITE_SIZE runs of this
    benchmark, each scaled by the same counter on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def run_variant(G, mra, trees, ops, prec, flags, steps, warmup):
    n = len(trees)
    for _ in range(warmup):
        outs = [G.FunctionTree(mra) for _ in range(n)]
        G.apply_batch(prec, outs, ops, trees, flags=flags)
    mra.ctx.sync()
    nodes = 0
    flops = 0.0
    t_con = 0.0
    tuples = 0
    tuples64 = 0
    slices = 0
    dloop = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = [G.FunctionTree(mra) for _ in range(n)]
        st = G.apply_batch(prec, outs, ops, trees, flags=flags)
        dloop += int(st[0]["device_loop"])
        nodes += sum(s["nodes_out"] for s in st)
        flops += sum(s["flops"] for s in st)
        tuples += sum(s["tuples"] for s in st)
        tuples64 += sum(s["tuples_fp64"] for s in st)
        slices += sum(s["slices"] for s in st)
        t_con += st[0]["t_contract"]
    mra.ctx.sync()
    dt = time.perf_counter() - t0
    return dict(ms_per_step=1e3 * dt / steps, nodes_per_step=nodes / steps, nodes_per_s=nodes / dt,
                flops_per_step=flops / steps, tuples_per_step=tuples / steps, slices_per_step=slices / steps,
                fp64_tuples_per_step=tuples64 / steps,
                t_contract_per_step=t_con / steps, iters=st[0]["iters"], device_loop_steps=dloop,
                predicted=st[0]["predicted_splits"], wasted=st[0]["predicted_wasted"], outs=outs)


def train_predictor(G, configs, mra, args):
    """Refinement traces from plain applies on other seeded orbitals (Helmholtz, as the
    workload, and Poisson on pair products), then GPU training of the MLP."""
    t0 = time.perf_counter()
    G.trace_enable(True)
    for seed in (101, 102):
        orbs, nuc, mus = configs.orbitals(seed, n_orb=args.orbitals)
        trees = [G.project(args.prec, G.FunctionTree(mra), o) for o in orbs]
        ops = [G.HelmholtzOperator(mra, mu, args.prec) for mu in mus]
        G.apply_batch(args.prec, [G.FunctionTree(mra) for _ in trees], ops, trees)
        pairs = [G.multiply(args.prec, G.FunctionTree(mra), 1.0, trees[i], trees[j])
                 for i in range(len(trees)) for j in range(i, min(i + 2, len(trees)))]
        P = G.PoissonOperator(mra, args.prec)
        G.apply_batch(args.prec, [G.FunctionTree(mra) for _ in pairs], [P] * len(pairs), pairs)
    X, y = G.trace_get()
    G.trace_enable(False)
    pred = G.Predictor(seed=0)
    pred.train(X, y, epochs=150, batch=512, lr=3e-3, seed=0)
    pred.install(args.pred_threshold if args.pred_threshold > 0 else 0.5)
    return pred, len(y), time.perf_counter() - t0


def _cpu_sample(M, mra, orbs, mus, prec, budget_s):
    nodes = 0
    t_apply = 0.0
    done = 0
    for i in range(len(orbs)):
        f = M.Tree(mra).project_gauss(prec, orbs[i])
        H = M.helmholtz(mra, mus[i], prec)
        out = M.Tree(mra)
        st = M.apply(prec, out, H, f)
        nodes += st["nodes_out"]
        t_apply += st["t_total"]
        done += 1
        if t_apply > budget_s:
            break
    return nodes, t_apply, done


def cpu_baseline(orbs, mus, k, prec, budget_s=15.0, budget_1t_s=6.0):
    """FP64 OpenMP oracle on all host threads it is given, and on one thread (the
    reference's published logs are single-threaded)."""
    from mrchem_amd import configs
    from oracle import mwref as M
    threads = M.get_threads()
    mra = M.MRA(**configs.BOX20, k=k, basis="L", max_depth=16)
    nodes, t_apply, done = _cpu_sample(M, mra, orbs, mus, prec, budget_s)
    M.set_threads(1)
    try:
        n1, t1, d1 = _cpu_sample(M, mra, orbs, mus, prec, budget_1t_s)
    finally:
        M.set_threads(threads)
    return dict(value=nodes / t_apply, unit="output nodes/s", cores=threads, kind="port",
                sample=f"FP64 OpenMP oracle (oracle/libmwref.so), Helmholtz apply of {done} of {len(orbs)} "
                       f"orbitals of the same workload, {t_apply:.2f} s of apply time",
                value_1_thread=n1 / t1,
                sample_1_thread=f"same oracle on 1 thread, {d1} orbital(s), {t1:.2f} s of apply time")


def best_flags(res32, res64, G):
    return G.MWG_APPLY_MIXED if res32 is not None and res32["nodes_per_s"] > res64["nodes_per_s"] else G.MWG_APPLY_FP64


def new_operators_cost(G, mra, trees, mus, prec, flags, fixed=()):
    """One step with freshly built Helmholtz operators for the first len(mus) trees (their NS
    levels are computed during the first apply, as after an SCF update of the orbital
    energies) and the operators `fixed` for the rest, then the same step again with the
    operators built."""
    t0 = time.perf_counter()
    ops = [G.HelmholtzOperator(mra, mu, prec) for mu in mus] + list(fixed)
    fit = time.perf_counter() - t0
    times = []
    for _ in range(2):
        outs = [G.FunctionTree(mra) for _ in trees]
        mra.ctx.sync()
        t0 = time.perf_counter()
        G.apply_batch(prec, outs, ops, trees, flags=flags)
        mra.ctx.sync()
        times.append(time.perf_counter() - t0)
    return dict(fit_s=fit, first_apply_ms=1e3 * times[0], warm_apply_ms=1e3 * times[1],
                levels_ms=1e3 * (times[0] - times[1]))


def copy_probe(G, tree):
    """One in-place rescale by 1.0 of a tree's coefficients (k_scale): a streaming kernel
    of known size (reads and writes nnodes * 8 (k+1)^3 doubles) that calibrates HBM
    byte counters against the contraction's."""
    n = tree.getNNodes()
    tree.rescale(1.0)
    tree.mra.ctx.sync()
    K = tree.mra.order + 1
    log(f"[bench] copy probe: k_scale over {n} nodes, {n * 8 * K ** 3 * 8} bytes read and written")


def exchange_run(args, G, configs, steps, warmup):
    """16 orbitals on 8 nuclei (6 bohr cube), k=7, prec 1e-6: all 136 pair products
    phi_i phi_j (device multiply) Poisson-applied as one batch, plain FP64 and mixed;
    per-pair error of the mixed batch against the FP64 batch (exact coefficient L2)."""
    k, prec = 7, args.prec
    mra = G.mra_from(configs.BOX20, k, "L", max_depth=16)
    orbs, nuc, mus = configs.orbitals(args.seed, n_orb=16, n_nuc=8, nuc_box=6.0)
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    pairs = [(i, j) for i in range(16) for j in range(i, 16)]
    t0 = time.perf_counter()
    rho = G.multiply_batch(prec, [G.FunctionTree(mra) for _ in pairs], 1.0, [trees[i] for i, _ in pairs],
                           [trees[j] for _, j in pairs])
    t_mul = time.perf_counter() - t0
    P = G.PoissonOperator(mra, prec)
    ops = [P] * len(rho)
    log(f"[bench] exchange: {len(rho)} pair densities ({sum(r.getNNodes() for r in rho)} nodes) in {t_mul:.3f} s")
    res64 = run_variant(G, mra, rho, ops, prec, G.MWG_APPLY_FP64, steps, warmup)
    res32 = run_variant(G, mra, rho, ops, prec, G.MWG_APPLY_MIXED, steps, warmup)
    errs = [_l2diff(G, a, b) / math.sqrt(b.getSquareNorm()) for a, b in zip(res32["outs"], res64["outs"])]
    log(f"[bench] exchange fp64 {res64['ms_per_step']:.1f} ms/step, mixed {res32['ms_per_step']:.1f} ms/step, "
        f"max per-pair rel L2 err {max(errs):.2e}, device loop {res64['device_loop_steps']}/{steps} (fp64) "
        f"{res32['device_loop_steps']}/{steps} (mixed) steps")
    tc = res32["t_contract_per_step"]
    alg = res64["tuples_per_step"] * 6.0 * (k + 1) ** 4
    return dict(workload=f"exchange pair Poissons, 16 orbitals (136 pairs), k={k}, prec={prec}",
                fp64_nodes_per_s=res64["nodes_per_s"], mixed_nodes_per_s=res32["nodes_per_s"],
                speedup_vs_fp64=res32["nodes_per_s"] / res64["nodes_per_s"],
                fp64_ms_per_step=res64["ms_per_step"], mixed_ms_per_step=res32["ms_per_step"],
                per_pair_rel_l2_err_max=max(errs), per_pair_rel_l2_err_mean=sum(errs) / len(errs),
                output_nodes_per_step=res64["nodes_per_step"], plain_tuples_per_step=res64["tuples_per_step"],
                mixed_contraction_ms_per_step=1e3 * tc,
                mixed_contraction_tflops=alg / tc / 1e12 if tc > 0 else None, multiply_s=t_mul,
                steps=steps, warmup=warmup)


def exchange_main(args, G, configs):
    r = exchange_run(args, G, configs, args.steps, args.warmup)
    best = "mixed" if r["mixed_nodes_per_s"] > r["fp64_nodes_per_s"] and r["per_pair_rel_l2_err_max"] < args.prec \
        else "fp64"
    line = {
        "metric": "output nodes/s for Poisson/Helmholtz apply at prec 1e-6 vs plain FP64 GPU",
        "value": r[f"{best}_nodes_per_s"], "unit": "output nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r[f"{best}_ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": mixed_dtype(7) if best == "mixed" else "f64",
        "data": "synthetic (seeded Gaussian orbitals)", "config": r,
        "roofline": dict(bound="mfma", achieved=r["mixed_contraction_tflops"],
                         peak=PEAK[mixed_dtype(7)] / 1e12, unit="TFLOP/s",
                         frac=(r["mixed_contraction_tflops"] or 0) * 1e12 / PEAK[mixed_dtype(7)],
                         traffic=None),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def predictor_main(args, G, configs):
    """Learned refinement predictor, trained on refinement traces of k=7, prec 1e-6 applies
    (Helmholtz on 5-orbital sets and Poisson on their pair products, seeds 101 and 102),
    evaluated on 8 fresh orbitals (seed 200) at k=9, prec 1e-7: Helmholtz and Poisson
    applies, plain FP64 against FP64 + predictor (refinement iterations, output nodes,
    time, error of the predicted apply against the plain one)."""
    mra7 = G.mra_from(configs.BOX20, 7, "L", max_depth=16)
    targs = argparse.Namespace(**vars(args))
    targs.prec = 1e-6
    targs.orbitals = 5
    pred, ntrain, t_train = train_predictor(G, configs, mra7, targs)
    log(f"[bench] predictor trained on {ntrain} trace samples in {t_train:.1f} s")
    k, prec = 9, 1e-7
    mra = G.mra_from(configs.BOX20, k, "L", max_depth=16)
    orbs, nuc, mus = configs.orbitals(200, n_orb=8)
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    P = G.PoissonOperator(mra, prec)
    cases = {"helmholtz": [G.HelmholtzOperator(mra, mu, prec) for mu in mus], "poisson": [P] * len(trees)}
    report = {}
    best = None
    for name, ops in cases.items():
        res0 = run_variant(G, mra, trees, ops, prec, G.MWG_APPLY_FP64, args.steps, args.warmup)
        cand = {}
        for thr in ([args.pred_threshold] if args.pred_threshold > 0 else [0.5, 0.7, 0.85, 0.95]):
            pred.install(thr)
            cand[thr] = run_variant(G, mra, trees, ops, prec, G.MWG_APPLY_FP64 | G.MWG_APPLY_PREDICT, args.steps,
                                    args.warmup)
            log(f"[bench] {name} threshold {thr}: iters {cand[thr]['iters']}, "
                f"{cand[thr]['ms_per_step']:.1f} ms/step, speculative {cand[thr]['predicted']}, "
                f"discarded {cand[thr]['wasted']}")
        pthr = max(cand, key=lambda t: cand[t]["nodes_per_s"])
        res1 = cand[pthr]
        errs = [_l2diff(G, a, b) / math.sqrt(b.getSquareNorm()) for a, b in zip(res1["outs"], res0["outs"])]
        report[name] = dict(plain_iters=res0["iters"], predict_iters=res1["iters"],
                            plain_nodes_per_s=res0["nodes_per_s"], predict_nodes_per_s=res1["nodes_per_s"],
                            plain_ms=res0["ms_per_step"], predict_ms=res1["ms_per_step"],
                            plain_nodes=res0["nodes_per_step"], predict_nodes=res1["nodes_per_step"],
                            speculative=res1["predicted"], wasted=res1["wasted"], threshold=pthr,
                            threshold_sweep={str(t): dict(ms=r["ms_per_step"], iters=r["iters"],
                                                          speculative=r["predicted"], discarded=r["wasted"])
                                             for t, r in cand.items()},
                            rel_l2_err_max=max(errs))
        log(f"[bench] {name}: iters {res0['iters']} -> {res1['iters']}, {res0['ms_per_step']:.1f} -> "
            f"{res1['ms_per_step']:.1f} ms/step, nodes {res0['nodes_per_step']:.0f} -> {res1['nodes_per_step']:.0f}, "
            f"max rel L2 err {max(errs):.2e}")
        if best is None or name == "helmholtz":
            best = (res0, res1)
    res0, res1 = best
    line = {
        "metric": "output nodes/s for Poisson/Helmholtz apply at prec 1e-6 vs plain FP64 GPU",
        "value": res1["nodes_per_s"], "unit": "output nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res1["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded Gaussian orbitals)",
        "config": {"workload": f"learned predictor: 8 orbitals, k={k}, prec={prec}, Helmholtz (value) and Poisson",
                   "train_samples": ntrain, "train_s": t_train, "cases": report},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def scf_run(args, G, configs, steps, warmup, predict=True, cpu=True):
    """SCF operator step: 32 orbitals on 16 nuclei (10 bohr cube), k=9, prec 1e-6; one step =
    the 32 Helmholtz applies (mu_i = sqrt(-2 eps_i), eps_i ~ U[-2, -0.2]) and the 528 pair
    Poissons phi_i phi_j (i <= j) as one batch: plain FP64, mixed (split-FP16 MFMA) and mixed
    with the learned predictor; every output's error against the FP64 batch (exact L2); the
    CPU oracle timed on two of the step's Helmholtz applies."""
    k, prec = 9, args.prec
    mra = G.mra_from(configs.BOX20, k, "L", max_depth=16)
    orbs, nuc, _ = configs.orbitals(args.seed, n_orb=32, n_nuc=16, nuc_box=10.0)
    rng = configs.SplitMix64(args.seed + 1000)
    eps = [-rng.uniform(0.2, 2.0) for _ in range(32)]
    mus = [math.sqrt(-2.0 * e) for e in eps]
    t0 = time.perf_counter()
    trees = [G.project(prec, G.FunctionTree(mra), o) for o in orbs]
    pairs = [(i, j) for i in range(32) for j in range(i, 32)]
    rho = G.multiply_batch(prec, [G.FunctionTree(mra) for _ in pairs], 1.0, [trees[i] for i, _ in pairs],
                           [trees[j] for _, j in pairs])
    t_in = time.perf_counter() - t0
    t0 = time.perf_counter()
    P = G.PoissonOperator(mra, prec)
    Hs = [G.HelmholtzOperator(mra, mu, prec) for mu in mus]
    t_op = time.perf_counter() - t0
    inps = trees + rho
    ops = Hs + [P] * len(rho)
    log(f"[bench] scf step: {len(inps)} applies, input nodes {sum(t.getNNodes() for t in inps)}, "
        f"inputs {t_in:.2f} s, operators {t_op:.2f} s")
    res64 = run_variant(G, mra, inps, ops, prec, G.MWG_APPLY_FP64, steps, warmup)
    K = k + 1
    alg = res64["tuples_per_step"] * 6.0 * K ** 4
    log(f"[bench] scf fp64: {res64['ms_per_step']:.0f} ms/step, {res64['nodes_per_s']:.0f} nodes/s, "
        f"iters {res64['iters']}, {res64['tuples_per_step']:.0f} tuples, contraction "
        f"{1e3 * res64['t_contract_per_step']:.0f} ms/step, device loop {res64['device_loop_steps']}/{steps} steps")
    res32 = run_variant(G, mra, inps, ops, prec, G.MWG_APPLY_MIXED, steps, warmup)
    err32 = max_rel_err(G, res32["outs"], res64["outs"])
    res32["outs"] = None
    log(f"[bench] scf mixed: {res32['ms_per_step']:.0f} ms/step, {res32['nodes_per_s']:.0f} nodes/s, "
        f"contraction {1e3 * res32['t_contract_per_step']:.0f} ms/step, max rel L2 err vs fp64 {err32:.2e}, "
        f"device loop {res32['device_loop_steps']}/{steps} steps")
    new_ops = new_operators_cost(G, mra, inps, mus, prec, G.MWG_APPLY_MIXED, fixed=[P] * len(rho))
    log(f"[bench] scf new Helmholtz operators (mixed): first step {new_ops['first_apply_ms']:.0f} ms, then "
        f"{new_ops['warm_apply_ms']:.0f} ms (kernel fits {1e3 * new_ops['fit_s']:.0f} ms)")
    resP, errP = None, None
    if predict:
        pred, ntrain, t_train = train_predictor(G, configs, G.mra_from(configs.BOX20, 7, "L", max_depth=16),
                                                argparse.Namespace(**dict(vars(args), prec=1e-6, orbitals=5)))
        pred.install(args.pred_threshold if args.pred_threshold > 0 else 0.85)
        resP = run_variant(G, mra, inps, ops, prec, G.MWG_APPLY_MIXED | G.MWG_APPLY_PREDICT, steps, warmup)
        errP = max_rel_err(G, resP["outs"], res64["outs"])
        resP["outs"] = None
        log(f"[bench] scf mixed+predictor: {resP['ms_per_step']:.0f} ms/step, {resP['nodes_per_s']:.0f} nodes/s, "
            f"iters {resP['iters']}, speculative {resP['predicted']}, discarded {resP['wasted']}, "
            f"max rel L2 err vs fp64 {errP:.2e}")
    cpu_r = None
    if cpu:
        from oracle import mwref as M
        omra = M.MRA(**configs.BOX20, k=k, basis="L", max_depth=16)
        nodes, t_apply = 0, 0.0
        for i in range(2):
            f = M.Tree(omra).project_gauss(prec, orbs[i])
            st = M.apply(prec, M.Tree(omra), M.helmholtz(omra, mus[i], prec), f)
            nodes += st["nodes_out"]
            t_apply += st["t_total"]
        cpu_r = dict(value=nodes / t_apply, unit="output nodes/s", cores=M.get_threads(), kind="port",
                     sample=f"FP64 OpenMP oracle, 2 of the 32 Helmholtz applies of the step, {t_apply:.1f} s")
    res64["outs"] = None
    tc64, tc32 = res64["t_contract_per_step"], res32["t_contract_per_step"]
    return dict(workload=f"SCF operator step: 32 Helmholtz + 528 pair Poissons, k={k}, prec={prec}",
                fp64_nodes_per_s=res64["nodes_per_s"], fp64_ms_per_step=res64["ms_per_step"],
                mixed_nodes_per_s=res32["nodes_per_s"], mixed_ms_per_step=res32["ms_per_step"],
                mixed_rel_l2_err_vs_fp64=err32, speedup_mixed_vs_fp64=res32["nodes_per_s"] / res64["nodes_per_s"],
                predict_nodes_per_s=resP["nodes_per_s"] if resP else None,
                predict_ms_per_step=resP["ms_per_step"] if resP else None,
                predict_iters=resP["iters"] if resP else None, fp64_iters=res64["iters"],
                predict_rel_l2_err_vs_fp64=errP,
                output_nodes_per_step=res64["nodes_per_step"], inputs_s=t_in, operators_s=t_op,
                new_operators=new_ops,
                plain_tuples_per_step=res64["tuples_per_step"],
                fp64_contraction_ms_per_step=1e3 * tc64, mixed_contraction_ms_per_step=1e3 * tc32,
                fp64_contraction_tflops=alg / tc64 / 1e12 if tc64 > 0 else None,
                mixed_contraction_tflops=alg / tc32 / 1e12 if tc32 > 0 else None,
                cpu_baseline=cpu_r, steps=steps, warmup=warmup)


def scf_main(args, G, configs):
    r = scf_run(args, G, configs, args.steps, args.warmup, predict=not args.no_predict, cpu=not args.no_cpu)
    cands = [("fp64", r["fp64_nodes_per_s"], 0.0), ("mixed", r["mixed_nodes_per_s"], r["mixed_rel_l2_err_vs_fp64"])]
    if r["predict_nodes_per_s"]:
        cands.append(("predict", r["predict_nodes_per_s"], r["predict_rel_l2_err_vs_fp64"]))
    best = max((c for c in cands if c[2] < args.prec), key=lambda c: c[1])
    tfl = r["fp64_contraction_tflops"] if best[0] == "fp64" else r["mixed_contraction_tflops"]
    peak = PEAK["f64"] if best[0] == "fp64" else PEAK[mixed_dtype(9)]
    line = {
        "metric": "output nodes/s for Poisson/Helmholtz apply at prec 1e-6 vs plain FP64 GPU",
        "value": best[1], "unit": "output nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r[f"{best[0]}_ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if best[0] == "fp64" else mixed_dtype(9),
        "data": "synthetic (seeded Gaussian orbitals)", "config": {k: v for k, v in r.items() if k != "cpu_baseline"},
        "roofline": {"bound": "mfma", "achieved": tfl, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": (tfl or 0) * 1e12 / peak, "traffic": None},
        "cpu_baseline": r["cpu_baseline"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--k", type=int, default=7)
    ap.add_argument("--prec", type=float, default=1e-6)
    ap.add_argument("--orbitals", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-predict", action="store_true", help="skip the learned-predictor variant")
    ap.add_argument("--pred-threshold", type=float, default=0.0,
                    help="predictor decision threshold (0: the fastest of 0.5, 0.7, 0.85, 0.95)")
    ap.add_argument("--workload", default="helmholtz", choices=["helmholtz", "exchange", "predictor", "scf"])
    ap.add_argument("--no-autotune", action="store_true", help="keep the default contraction task sizes")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after timing, rescale one output tree in place (HBM counter calibration)")
    ap.add_argument("--no-exchange", action="store_true",
                    help="default workload only: skip the exchange (136 pair Poissons) results")
    ap.add_argument("--no-scf", action="store_true",
                    help="default workload only: skip the SCF operator step (k=9) results")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the batched applies would shard by orbital/pair across GPUs)", file=sys.stderr)
        sys.exit(2)

    from mrchem_amd import configs
    from mrchem_amd import mrcpp as G

    if args.workload == "exchange":
        return exchange_main(args, G, configs)
    if args.workload == "predictor":
        return predictor_main(args, G, configs)
    if args.workload == "scf":
        return scf_main(args, G, configs)

    log(f"[bench] building inputs: {args.orbitals} orbitals, k={args.k}, prec={args.prec}")
    mra, orbs, mus, trees, ops, t_proj, t_oper = build_inputs(G, configs, args.k, args.prec, args.orbitals,
                                                              args.seed)
    log(f"[bench] input nodes {[t.getNNodes() for t in trees]}, projection {t_proj:.3f} s, operators {t_oper:.3f} s")

    from mrchem_amd.autotune import autotune
    tuning = {}
    default_tuning = mra.ctx.get_tuning()  # (the context is shared by every workload below)

    def tune(name, flags):
        if args.no_autotune:
            return
        r = autotune(mra, trees, ops, args.prec, flags)
        tuning[name] = dict(fp64_chunk=r["fp64_chunk"], slice_chunk=r["slice_chunk"])
        log(f"[bench] autotune {name}: fp64_chunk {r['fp64_chunk']}, slice_chunk {r['slice_chunk']}, "
            f"apply {1e3 * r['t_best']:.2f} ms, contraction {1e3 * r['t_contract']:.2f} ms "
            f"({len(r['table'])} candidates, {r['seconds']:.1f} s)")

    tune("fp64", G.MWG_APPLY_FP64)
    res64 = run_variant(G, mra, trees, ops, args.prec, G.MWG_APPLY_FP64, args.steps, args.warmup)
    log(f"[bench] fp64: {res64['ms_per_step']:.2f} ms/step, {res64['nodes_per_s']:.1f} nodes/s, "
        f"contraction {1e3 * res64['t_contract_per_step']:.2f} ms/step, "
        f"device loop {res64['device_loop_steps']}/{args.steps} steps")
    best, dtype = res64, "f64"
    res32 = None
    mixed_ok = args.k in (7, 9)
    if mixed_ok:
        tune("mixed", G.MWG_APPLY_MIXED)
        res32 = run_variant(G, mra, trees, ops, args.prec, G.MWG_APPLY_MIXED, args.steps, args.warmup)
        err = max_rel_err(G, res32["outs"], res64["outs"])
        res32["rel_err_vs_fp64"] = err
        log(f"[bench] mixed: {res32['ms_per_step']:.2f} ms/step, {res32['nodes_per_s']:.1f} nodes/s, "
            f"contraction {1e3 * res32['t_contract_per_step']:.2f} ms/step, rel L2 err vs fp64 {err:.2e}, "
            f"device loop {res32['device_loop_steps']}/{args.steps} steps")
        if res32["nodes_per_s"] > res64["nodes_per_s"] and err < args.prec:
            best, dtype = res32, mixed_dtype(args.k)
    # an SCF iteration brings new mu values: fresh operators build their NS levels during
    # their first apply (the timed steps above reuse built operators)
    new_ops = new_operators_cost(G, mra, trees, mus, args.prec, best_flags(res32, res64, G))
    log(f"[bench] new operators: first apply {new_ops['first_apply_ms']:.2f} ms, then "
        f"{new_ops['warm_apply_ms']:.2f} ms (kernel fit {1e3 * new_ops['fit_s']:.1f} ms)")
    resP = None
    if not args.no_predict:
        pred, ntrain, t_train = train_predictor(G, configs, mra, args)
        flags = (G.MWG_APPLY_MIXED if mixed_ok else G.MWG_APPLY_FP64) | G.MWG_APPLY_PREDICT
        # decision threshold of the predictor: the given one, or the fastest of a few
        sweep = {}
        for thr in ([args.pred_threshold] if args.pred_threshold > 0 else [0.5, 0.7, 0.85, 0.95]):
            pred.install(thr)
            r = run_variant(G, mra, trees, ops, args.prec, flags, args.steps, args.warmup)
            sweep[thr] = r
            log(f"[bench] mixed+predictor (threshold {thr}): {r['ms_per_step']:.2f} ms/step, iters {r['iters']}, "
                f"speculative {r['predicted']}, discarded {r['wasted']}")
        pthr = max(sweep, key=lambda t: sweep[t]["nodes_per_s"])
        resP = sweep[pthr]
        errP = max_rel_err(G, resP["outs"], res64["outs"])
        resP["rel_err_vs_fp64"] = errP
        resP["train_samples"] = ntrain
        resP["train_s"] = t_train
        resP["threshold"] = pthr
        resP["threshold_sweep"] = {str(t): dict(ms_per_step=r["ms_per_step"], iters=r["iters"],
                                                speculative=r["predicted"], discarded=r["wasted"])
                                   for t, r in sweep.items()}
        log(f"[bench] mixed+predictor: {resP['ms_per_step']:.2f} ms/step, {resP['nodes_per_s']:.1f} nodes/s, "
            f"iters {resP['iters']} (fp64 {res64['iters']}), threshold {pthr}, rel L2 err vs fp64 {errP:.2e}, "
            f"trained on {ntrain} trace samples in {t_train:.1f} s")
        if resP["nodes_per_s"] > best["nodes_per_s"] and errP < args.prec:
            best, dtype = resP, (mixed_dtype(args.k) if mixed_ok else "f64")

    # algorithmic flops: the plain FP64 method's surviving tuples x 6 (k+1)^4 (the same work for
    # every variant); issued flops count what the variant's kernels actually execute
    K = args.k + 1
    F = res64["tuples_per_step"]
    alg = F * 6.0 * K ** 4
    tc = best["t_contract_per_step"]
    achieved = alg / tc if tc > 0 else 0.0
    kern = f"k_contract_mfma{K}" if dtype != "f64" else "k_contract_fp64"
    per_launch = measured_traffic(kern, f"helmholtz{args.orbitals}_k{args.k}_prec{prec_tag(args.prec)}")
    # algorithmic bytes of a step: every input and output node block read or written once (FP64)
    in_nodes = sum(t.getNNodes() for t in trees)
    alg_bytes = 8.0 * K ** 3 * 8 * (in_nodes + best["nodes_per_step"])
    roofline = dict(bound="mfma", achieved=achieved / 1e12, peak=PEAK[dtype] / 1e12, unit="TFLOP/s",
                    frac=achieved / PEAK[dtype],
                    traffic=per_launch,
                    kernel=(kern + " + k_contract_fp64 + k_reduce_partials") if dtype != "f64"
                    else "k_contract_fp64 + k_reduce_partials",
                    plain_tuples_per_step=F, alg_flops_per_tuple=6.0 * K ** 4,
                    issued_tflops=(best["flops_per_step"] / tc / 1e12) if tc > 0 else None,
                    contraction_ms_per_step=1e3 * tc,
                    launches_per_step=best["iters"],
                    traffic_per_step=per_launch * best["iters"] if per_launch else None,
                    alg_bytes_per_step=alg_bytes,
                    traffic_over_alg_bytes=(per_launch * best["iters"] / alg_bytes) if per_launch else None)

    if args.copy_probe:
        copy_probe(G, res64["outs"][0])

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU oracle baseline ...")
        cpu = cpu_baseline(orbs, mus, args.k, args.prec)
        log(f"[bench] cpu: {cpu['value']:.1f} nodes/s on {cpu['cores']} threads")

    sub_steps, sub_warm = min(args.steps, 3), 1
    # the exchange step is k = 7 like the tuned workload; the SCF step (k = 9) runs with the
    # context's default task sizes
    exch = None if args.no_exchange else exchange_run(args, G, configs, sub_steps, sub_warm)
    mra.ctx.set_tuning(*default_tuning)
    scf = None if args.no_scf else scf_run(args, G, configs, min(args.steps, 2), sub_warm,
                                           predict=not args.no_predict, cpu=not args.no_cpu)

    line = {
        "metric": "output nodes/s for Poisson/Helmholtz apply at prec 1e-6 vs plain FP64 GPU",
        "value": best["nodes_per_s"],
        "unit": "output nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded Gaussian orbitals)",
        "config": {
            "workload": f"Helmholtz apply, {args.orbitals} orbitals, k={args.k}, prec={args.prec}, box [-20,20]^3",
            "orbitals": args.orbitals, "k": args.k, "prec": args.prec, "seed": args.seed,
            "output_nodes_per_step": best["nodes_per_step"], "tuples_per_step": best["tuples_per_step"],
            "plain_tuples_per_step": F, "input_nodes": in_nodes,
            "fp64_nodes_per_s": res64["nodes_per_s"], "speedup_vs_fp64": best["nodes_per_s"] / res64["nodes_per_s"],
            "fp64_ms_per_step": res64["ms_per_step"],
            "device_loop_steps": {"fp64": res64["device_loop_steps"],
                                  "mixed": res32["device_loop_steps"] if res32 else None,
                                  "predict": resP["device_loop_steps"] if resP else None},
            "mixed_nodes_per_s": res32["nodes_per_s"] if res32 else None,
            "mixed_rel_l2_err_vs_fp64": res32["rel_err_vs_fp64"] if res32 else None,
            "mixed_tuples_per_step": res32["tuples_per_step"] if res32 else None,
            "mixed_fp64_routed_tuples_per_step": res32["fp64_tuples_per_step"] if res32 else None,
            "mixed_slices_per_step": res32["slices_per_step"] if res32 else None,
            "predict_nodes_per_s": resP["nodes_per_s"] if resP else None,
            "predict_rel_l2_err_vs_fp64": resP["rel_err_vs_fp64"] if resP else None,
            "predict_iters": resP["iters"] if resP else None, "fp64_iters": res64["iters"],
            "predict_speculative_nodes": resP["predicted"] if resP else None,
            "predict_speculative_wasted": resP["wasted"] if resP else None,
            "predict_train_samples": resP["train_samples"] if resP else None,
            "predict_threshold": resP["threshold"] if resP else None,
            "predict_threshold_sweep": resP["threshold_sweep"] if resP else None,
            "operator_build_s": t_oper, "projection_s": t_proj,
            "new_operators": new_ops,
            "contraction_ms_per_step": 1e3 * best["t_contract_per_step"],
            "tuning": tuning or None,
            "exchange": exch,
            "scf": scf,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
