"""MRChem orbital files (f3): the .meta layout against the reference's stored file
(tests/h2_pol_lda/orbitals/phi_0.meta, kept as tests/golden/h2_pol_lda_phi_0.meta) and
the tree file format, on the CPU through the C ABI (no device); device trees, orbital
vectors and refusals on the GPU."""
import os
import struct

import numpy as np
import pytest

from mrchem_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_META = os.path.join(HERE, "golden", "h2_pol_lda_phi_0.meta")

needs_lib = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libmwgreen.so not built")


def _layout(boxes, levels, rng):
    """A valid import layout: the roots (x fastest), then every node of each level split."""
    depth, lr = [], []
    for z in range(boxes[2]):
        for y in range(boxes[1]):
            for x in range(boxes[0]):
                depth.append(0)
                lr.append((x, y, z))
    frontier = list(range(len(depth)))
    for _ in range(levels):
        nxt = []
        for p in frontier:
            for a in range(8):
                x, y, z = lr[p]
                depth.append(depth[p] + 1)
                lr.append((2 * x + (a & 1), 2 * y + ((a >> 1) & 1), 2 * z + (a >> 2)))
                nxt.append(len(depth) - 1)
        frontier = nxt
    return np.array(depth, np.int32), np.array(lr, np.int64)


@needs_lib
def test_meta_matches_the_reference_file(tmp_path):
    from mrchem_amd import orbital_io as io
    m = io.read_meta(GOLDEN_META)
    assert (m.real_size, m.imag_size, m.is_shared, m.rank_id, m.spin, m.occ, m.error, m.orbital) == \
        (6, 0, False, 0, io.SPIN_PAIRED, 2, 1.0, True)
    out = str(tmp_path / "phi_0.meta")
    io.write_meta(out, m)
    a, b = open(GOLDEN_META, "rb").read(), open(out, "rb").read()
    assert len(b) == 36
    pad = set(range(9, 12)) | set(range(24, 28))  # alignment padding of the two structs
    assert all(a[i] == b[i] for i in range(36) if i not in pad)
    assert all(b[i] == 0 for i in pad)
    # the layout itself: int, int, bool (+3), int, int, int (+4), double
    assert struct.unpack("<ii?3xiii4xd", b) == (6, 0, False, 0, 0, 2, 1.0)


@needs_lib
def test_density_meta_and_refusals(tmp_path):
    from mrchem_amd import orbital_io as io
    p = str(tmp_path / "rho.meta")
    io.write_meta(p, io.Meta(real_size=3, imag_size=1, is_shared=True, orbital=False))
    assert os.path.getsize(p) == 12
    m = io.read_meta(p)
    assert (m.real_size, m.imag_size, m.is_shared, m.orbital) == (3, 1, True, False)
    bad = tmp_path / "bad.meta"
    bad.write_bytes(b"\0" * 20)
    with pytest.raises(_lib.MwgError):
        io.read_meta(str(bad))
    bad.write_bytes(struct.pack("<ii?3x", 1, 0, True)[:8] + b"\x02\0\0\0")  # is_shared = 2
    with pytest.raises(_lib.MwgError):
        io.read_meta(str(bad))
    with pytest.raises(_lib.MwgError):
        io.read_meta(str(tmp_path / "missing.meta"))


@needs_lib
@pytest.mark.parametrize("levels,k", [(1, 3), (2, 3), (1, 7)])
def test_tree_file_round_trip(tmp_path, levels, k):
    from mrchem_amd import orbital_io as io
    rng = np.random.default_rng(levels * 10 + k)
    boxes = (4, 4, 4) if levels == 2 else (2, 1, 3)
    depth, lr = _layout(boxes, levels, rng)
    coefs = rng.standard_normal((len(depth), 8 * (k + 1) ** 3))
    info = dict(order=k, basis=1, root_scale=-4, max_depth=12, corner=(-2, -1, -3), boxes=boxes,
                scaling_factor=0.75)
    name = str(tmp_path / f"t{levels}_{k}")
    io.tree_file_write(name, info, depth, lr, coefs)
    got, d2, l2, c2 = io.tree_file_read(name)
    for key, v in info.items():
        assert got[key] == (tuple(v) if isinstance(v, tuple) else v), key
    assert got["nnodes"] == len(depth) and got["chunk_nodes"] == 1024
    assert got["nchunks"] == (len(depth) + 1023) // 1024
    assert (d2 == depth).all() and (l2 == lr).all() and (c2 == coefs).all()
    size = os.path.getsize(name + ".tree")
    assert size == 96 + len(depth) * (32 + 8 * 8 * (k + 1) ** 3)


@needs_lib
def test_tree_file_refusals(tmp_path):
    from mrchem_amd import orbital_io as io
    rng = np.random.default_rng(1)
    depth, lr = _layout((2, 2, 2), 1, rng)
    coefs = rng.standard_normal((len(depth), 8 * 64))
    info = dict(order=3, basis=0, root_scale=0, max_depth=8, corner=(0, 0, 0), boxes=(2, 2, 2), scaling_factor=1.0)
    # a child before its parent is not the import layout
    with pytest.raises(_lib.MwgError):
        io.tree_file_write(str(tmp_path / "x"), info, depth[::-1].copy(), lr[::-1].copy(), coefs)
    name = str(tmp_path / "t")
    io.tree_file_write(name, info, depth, lr, coefs)
    raw = open(name + ".tree", "rb").read()
    for cut, what in ((raw[:50], "truncated header"), (raw[:-8], "truncated records"), (raw + b"\0", "trailing"),
                      (b"X" + raw[1:], "magic")):
        open(name + ".tree", "wb").write(cut)
        with pytest.raises(_lib.MwgError):
            io.tree_file_read(name)


@pytest.mark.gpu
def test_orbitals_save_and_load_on_device(tmp_path):
    from mrchem_amd import configs
    from mrchem_amd import mrcpp as G
    from mrchem_amd import orbital_io as io
    G.default_context()
    orbs, _, _ = configs.orbitals(0, n_orb=3)
    mra = G.mra_from(configs.BOX20, 5, "L")
    trees = [G.project(1e-4, G.FunctionTree(mra), o) for o in orbs]
    Phi = [io.Orbital(re=trees[0], spin=io.SPIN_ALPHA, occ=1, error=1e-3),
           io.Orbital(re=trees[1], im=trees[2], spin=io.SPIN_BETA, occ=1, error=2e-3),
           io.Orbital(re=trees[2], occ=2, rank_id=1)]
    prefix = str(tmp_path / "phi")
    io.save_orbitals(Phi, prefix)
    got = io.load_orbitals(prefix, mra)
    assert len(got) == 3
    for a, b in zip(Phi, got):
        assert (a.spin, a.occ, a.rank_id, a.error) == (b.spin, b.occ, b.rank_id, b.error)
        for part in ("re", "im"):
            x, y = getattr(a, part), getattr(b, part)
            assert (x is None) == (y is None)
            if x is not None:
                for u, v in zip(x.export(), y.export()):
                    assert np.array_equal(u, v)
                assert y.getSquareNorm() == pytest.approx(x.getSquareNorm(), rel=1e-14)
    assert io.read_meta(prefix + "_1.meta").imag_size == 1
    assert len(io.load_orbitals(prefix, mra, n_orbs=2)) == 2
    # a file written for another MRA is refused
    with pytest.raises(_lib.MwgError):
        G.FunctionTree(G.mra_from(configs.BOX20, 7, "L")).loadTree(prefix + "_0_re")
    # a loaded orbital applies exactly like the one saved
    P = G.PoissonOperator(mra, 1e-4)
    o1, o2 = G.FunctionTree(mra), G.FunctionTree(mra)
    G.apply(1e-4, o1, P, trees[0])
    G.apply(1e-4, o2, P, got[0].re)
    for u, v in zip(o1.export(), o2.export()):
        assert np.array_equal(u, v)
