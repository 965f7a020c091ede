"""The allocation-free dot (mwg_tree_dot_noalloc) against mwg_tree_dot through the C ABI:
two calls in a row, the second on smaller trees (stale scratch must not leak into it), and
the empty-tree and self-dot cases.  Agreement to FP64 rounding of the summation order."""
import ctypes as C
import math

import pytest

from mrchem_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    from mrchem_amd import mrcpp
    mrcpp.default_context()
    return mrcpp


def _dots(G, a, b):
    from mrchem_amd._lib import lib
    v0, v1 = C.c_double(), C.c_double()
    assert lib().mwg_tree_dot(a.h, b.h, C.byref(v0)) == 0
    assert lib().mwg_tree_dot_noalloc(a.h, b.h, C.byref(v1)) == 0
    return v0.value, v1.value


def test_noalloc_dot_matches_dot_on_two_calls(G):
    orbs, _, _ = configs.orbitals(0, n_orb=3)
    big = G.mra_from(configs.BOX20, 7, "L")
    a = G.project(1e-6, G.FunctionTree(big), orbs[0])
    b = G.project(1e-6, G.FunctionTree(big), orbs[1])
    small = G.mra_from(configs.BOX20, 5, "I")
    c = G.project(1e-3, G.FunctionTree(small), orbs[2])
    d = G.project(1e-3, G.FunctionTree(small), orbs[0])
    assert c.getNNodes() < a.getNNodes()
    print(f"noalloc dot: nodes {a.getNNodes()}, {b.getNNodes()} then {c.getNNodes()}, {d.getNNodes()}", flush=True)
    for x, y in ((a, b), (c, d)):
        ref, got = _dots(G, x, y)
        scale = math.sqrt(x.getSquareNorm() * y.getSquareNorm())
        assert abs(got - ref) <= 1e-12 * scale, (got, ref)
    # self dot: the squared norm
    ref, got = _dots(G, c, c)
    assert abs(got - c.getSquareNorm()) <= 1e-12 * c.getSquareNorm()
    assert abs(got - ref) <= 1e-12 * c.getSquareNorm()


def test_noalloc_dot_of_an_empty_tree_is_zero(G):
    orbs, _, _ = configs.orbitals(0, n_orb=1)
    mra = G.mra_from(configs.BOX20, 5, "L")
    a = G.project(1e-3, G.FunctionTree(mra), orbs[0])
    e = G.FunctionTree(mra)
    assert _dots(G, a, e) == (0.0, 0.0)
    assert _dots(G, e, a) == (0.0, 0.0)
