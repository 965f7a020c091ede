"""MRChem caller mirrors (mrchem_amd/mrchem.py) over the device apply."""
import math

import numpy as np
import pytest

from mrchem_amd import configs


@pytest.fixture(scope="module")
def G():
    from mrchem_amd import mrcpp
    mrcpp.default_context()
    return mrcpp


@pytest.fixture(scope="module")
def A():
    from oracle import analytic
    return analytic


def test_helmholtz_vector_clamps_and_mu():
    from mrchem_amd import mrchem as MC

    class FakeMRA:
        pass
    hv = MC.HelmholtzVector(1e-4, [-0.5, 0.3, -2.0], FakeMRA())
    assert list(hv.lambdas) == [-0.5, -0.5, -2.0]          # lambda > 0 -> -0.5 (HelmholtzVector.cpp:53-55)
    assert hv.mu(0) == pytest.approx(1.0)                   # mu = sqrt(-2 lambda) (:169)
    assert hv.mu(2) == pytest.approx(2.0)


def test_spin_factors():
    from mrchem_amd import mrchem as MC
    O = MC.Orbital
    a, b, p = O(None, 1.0, "a"), O(None, 1.0, "b"), O(None, 2.0, "p")
    assert MC.spin_factor(a, a) == 1.0 and MC.spin_factor(a, b) == 0.0 and MC.spin_factor(a, p) == 1.0
    assert MC.spin_factor(p, a) == 0.5 and MC.spin_factor(p, p) == 1.0


@pytest.mark.gpu
def test_helmholtz_vector_against_closed_form(G, A):
    from mrchem_amd import mrchem as MC
    prec = 1e-5
    orbs, nuc, mus = configs.orbitals(3, n_orb=3)
    mra = G.mra_from(configs.BOX20, 7)
    Phi = [MC.Orbital(G.project(prec, G.FunctionTree(mra), o)) for o in orbs]
    lambdas = [-0.5 * m * m for m in mus]
    out = MC.HelmholtzVector(prec, lambdas, mra)(Phi)
    pts = np.random.default_rng(0).uniform(-20, 20, (10000, 3))
    for i, o in enumerate(out):
        ex = -1.0 / (2.0 * math.pi) * A.potential(orbs[i], pts, mus[i])
        err = o.tree.evalf(pts) - ex
        assert math.sqrt((err ** 2).mean() * 40 ** 3) < prec * math.sqrt(o.tree.getSquareNorm())


@pytest.mark.gpu
def test_coulomb_potential_reference_flow(G, A):
    """tests/qmoperators/coulomb_operator.cpp flow: project hydrogen orbitals (prec 1e-3,
    Interpolating k=5, [-32,32]^3), rho by multiply, J = P[rho] at prec/||rho||,
    <phi|J phi> by multiply and dot; compared with the reference's E_P values."""
    from mrchem_amd import mrchem as MC
    prec = 1e-3
    mra = G.mra_from(configs.BOX32, 5, "I", max_depth=18)
    qn = {"1s": (1, 0, 0), "2s": (2, 0, 0), "2px": (2, 1, 0), "2py": (2, 1, 1), "2pz": (2, 1, 2)}
    Phi = {k: MC.Orbital(G.project_hydrogen(prec, G.FunctionTree(mra), [(1.0, q)])) for k, q in qn.items()}
    P = G.PoissonOperator(mra, prec)
    J = MC.CoulombPotential(P, list(Phi.values()), mra)
    J.setup(prec)
    for name, key in [("1s", "1s"), ("2s", "2s"), ("2pz", "2p")]:
        Jphi = J(Phi[name], prec)
        v = G.dot(Phi[name].tree, Jphi.tree)
        assert v == pytest.approx(A.REFERENCE_E_P[key], rel=5e-3), name


@pytest.mark.gpu
def test_exchange_potential_matches_oracle(G, oracle):
    """Pair Poissons of the exchange setup (batched on the device) against the oracle's
    multiply + apply on the same orbital trees."""
    from mrchem_amd import mrchem as MC
    prec = 1e-4
    orbs, nuc, mus = configs.orbitals(4, n_orb=3)
    mra = G.mra_from(configs.BOX20, 5)
    omra = oracle.MRA(**configs.BOX20, k=5, max_depth=16)
    otrees = [oracle.Tree(omra).project_gauss(prec, o) for o in orbs]
    Phi = [MC.Orbital(G.FunctionTree(mra).load(*t.export())) for t in otrees]
    K = MC.ExchangePotential(G.PoissonOperator(mra, prec), Phi, mra, screen=False)
    ex = K.setup(prec)
    assert len(K.stats) == 6
    OP = oracle.poisson(omra, prec)
    for (i, j), V in K.pair_potentials.items():
        rho = oracle.multiply(min(prec, 1e-1), oracle.Tree(omra), 1.0, otrees[i], otrees[j])
        ov = oracle.Tree(omra)
        oracle.apply(min(prec, 1e-1), ov, OP, rho)
        pts = np.random.default_rng(i * 7 + j).uniform(-4, 4, (2000, 3))
        a, b = V.evalf(pts), ov.eval(pts)
        assert np.abs(a - b).max() < 1e-10 * max(np.abs(b).max(), 1e-30)
    # K is Hermitian: <phi_a|K phi_b> = <phi_b|K phi_a>
    for a in range(3):
        for b in range(a + 1, 3):
            kab = G.dot(Phi[a].tree, ex[b])
            kba = G.dot(Phi[b].tree, ex[a])
            assert abs(kab - kba) < 10 * prec * max(abs(kab), 1e-3)
