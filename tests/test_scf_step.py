"""Orbital vector algebra and the SCF Helmholtz step (mrchem_amd/mrchem.py), mirroring
orbital_utils.cpp (add, rotate, calc_lowdin_matrix, orthonormalize) and
OrbitalOptimizer::optimize (OrbitalOptimizer.cpp:148-170).

Checks: matrix algebra on the CPU; on the GPU, rotations through their overlaps
(S' = A S A^T, exact bilinearity), Loewdin orthonormalization (S = I), and the hydrogen
1s / 2s fixed point of the step with a Gaussian-smoothed nucleus."""
import math

import numpy as np
import pytest

from mrchem_amd import configs


def test_hermitian_matrix_pow():
    from mrchem_amd import mrchem as MC
    rng = np.random.default_rng(0)
    B = rng.standard_normal((5, 5))
    S = B @ B.T + 5 * np.eye(5)
    X = MC.hermitian_matrix_pow(S, -0.5)
    assert np.allclose(X @ S @ X, np.eye(5), atol=1e-12)
    assert np.allclose(X, X.T, atol=1e-14)
    # an eigenvalue below MachineZero is dropped, not inverted
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    A = Q @ np.diag([4.0, 1.0, 1e-16]) @ Q.T
    Y = MC.hermitian_matrix_pow(A, -0.5)
    assert np.allclose(Y, Q @ np.diag([0.5, 1.0, 0.0]) @ Q.T, atol=1e-12)


def test_rotate_refuses_a_wrong_matrix():
    from mrchem_amd import mrchem as MC
    from mrchem_amd.mrcpp import MwgError
    with pytest.raises(MwgError):
        MC.rotate(np.eye(3), [MC.Orbital(None), MC.Orbital(None)])
    with pytest.raises(MwgError):
        MC.add(1.0, [MC.Orbital(None)], 1.0, [])


@pytest.fixture(scope="module")
def G():
    from mrchem_amd import mrcpp
    mrcpp.default_context()
    return mrcpp


@pytest.mark.gpu
def test_rotate_and_orthonormalize(G):
    from mrchem_amd import mrchem as MC
    prec = 1e-5
    orbs, _, _ = configs.orbitals(4, n_orb=3)
    mra = G.mra_from(configs.BOX20, 7)
    Phi = [MC.Orbital(G.project(prec, G.FunctionTree(mra), o)) for o in orbs]
    S = MC.calc_overlap_matrix(Phi)
    A = np.array([[1.0, 0.3, -0.2], [0.1, 0.9, 0.4], [-0.3, 0.2, 1.2]])
    Chi = MC.rotate(A, Phi)
    S2 = MC.calc_overlap_matrix(Chi)
    assert np.abs(S2 - A @ S @ A.T).max() < 1e-12 * np.abs(S2).max()
    Ort, U = MC.orthonormalize(prec, Chi)
    assert np.abs(MC.calc_overlap_matrix(Ort) - np.eye(3)).max() < 1e-12
    assert np.abs(U @ S2 @ U - np.eye(3)).max() < 1e-12
    # vector add: (Chi + Phi) - Phi = Chi, through the norms of the difference
    back = MC.add(1.0, MC.add(1.0, Chi, 1.0, Phi), -1.0, Chi)
    d = MC.add(1.0, back, -1.0, Phi)
    assert MC.get_norms(d).max() < 1e-12 * MC.get_norms(Phi).max()


def _nucleus(G, mra, prec, Z=1.0, a=1e4):
    """V = -Z erf(sqrt(a) r) / r: the Poisson potential of a normalised Gaussian charge."""
    from mrchem_amd.gaussians import normalized_s
    rho = G.project(prec, G.FunctionTree(mra), normalized_s(Z, a, (0.0, 0.0, 0.0)))
    V = G.FunctionTree(mra)
    G.apply(prec, V, G.PoissonOperator(mra, prec), rho)
    V.rescale(-1.0)
    return V


@pytest.mark.gpu
def test_helmholtz_step_hydrogen_fixed_point(G):
    """1s and 2s of hydrogen with F = diag(-1/2, -1/8): the step returns them (up to the
    nucleus smoothing: the orbital moves by about 1e-4), keeps F to that level, and reports
    the update norms."""
    from mrchem_amd import mrchem as MC
    prec = 1e-5
    mra = G.mra_from(configs.BOX20, 7, max_depth=18)
    Phi = [MC.Orbital(G.project_hydrogen(prec, G.FunctionTree(mra), [(1.0, (1, 0, 0))]), 1.0, "a"),
           MC.Orbital(G.project_hydrogen(prec, G.FunctionTree(mra), [(1.0, (2, 0, 0))]), 1.0, "a")]
    V = _nucleus(G, mra, prec * 0.1)
    F = np.diag([-0.5, -0.125])
    Phi1, F1, err = MC.helmholtz_step(prec, F, Phi, V, mra)
    print("hydrogen step: update norms", err, "F", np.diag(F1))
    assert err.max() < 5e-4
    # F <- U F U^T with U = S^-1/2 of the updated orbitals: their norms move by the update
    assert np.abs(F1 - F).max() < 1e-3 * np.abs(F).max()
    assert np.abs(MC.calc_overlap_matrix(Phi1) - np.eye(2)).max() < 1e-3
    for p, q in zip(Phi1, Phi):  # (the projected inputs are normalised to the projection prec)
        assert G.dot(p.tree, q.tree) > 1.0 - 1e-4
