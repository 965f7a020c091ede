"""CPU checks of the drop-in boundary: the C ABI library loads, exports every symbol
declared in include/mwg.h, and the mrcpp facade compiles against it.  No GPU calls."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mwg.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(mwg_\w+)\(", text, flags=re.M)))


def test_header_declarations_match_binding():
    from mrchem_amd import _lib
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_loads_and_exports_every_declared_symbol():
    from mrchem_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmwgreen.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    _lib.lib()  # signatures bind


def test_mrcpp_facade_compiles(tmp_path):
    from mrchem_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmwgreen.so not built")
    exe = tmp_path / "facade_example"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cxx", "facade_example.cpp"), "-L", os.path.dirname(_lib.LIB_PATH),
           "-lmwgreen", "-o", str(exe)]
    subprocess.run(cmd, check=True)
    assert exe.exists()


def test_cxx_poisson_caller_compiles(tmp_path):
    """The C++ caller of the widened facade (project, multiply, add, grids, files) builds."""
    from mrchem_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmwgreen.so not built")
    exe = tmp_path / "poisson_caller"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cxx", "poisson_caller.cpp"), "-L", os.path.dirname(_lib.LIB_PATH),
           "-lmwgreen", "-o", str(exe)]
    subprocess.run(cmd, check=True)
    assert exe.exists()


@pytest.mark.gpu
def test_cxx_poisson_caller_runs(tmp_path):
    """A C++ program over the mrcpp facade: a real Poisson apply against the erf closed form,
    and the tree utilities (tests/cxx/poisson_caller.cpp prints one line per check)."""
    from mrchem_amd import _lib
    exe = tmp_path / "poisson_caller"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "poisson_caller.cpp"), "-L", libdir, "-lmwgreen",
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "poisson caller: ok" in r.stdout


def test_bench_refuses_multi_gpu():
    r = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_configs_are_seeded_and_orthonormal():
    from mrchem_amd import configs, gaussians as G
    a1, n1, m1 = configs.orbitals(0)
    a2, n2, m2 = configs.orbitals(0)
    assert (n1 == n2).all() and m1 == m2
    S = [[G.overlap(a, b) for b in a1] for a in a1]
    import numpy as np
    assert np.abs(np.array(S) - np.eye(len(a1))).max() < 1e-12
    assert all(0.5 <= m <= 2.0 for m in m1)
    g, parts = configs.density3(0)
    assert abs(G.integral(g) - sum(q for q, a, R in parts)) < 1e-12
    for q, a, R in parts:
        assert 1.0 <= a <= 50.0 and 0.5 <= q <= 2.0 and np.all(np.abs(R) <= 2.0)


def test_gaussian_product_theorem():
    import numpy as np
    from mrchem_amd import gaussians as G
    a = G.primitive("px", 0.7, (0.1, -0.3, 0.2), 1.3) + G.primitive("s", 2.0, (0.5, 0.0, -0.1))
    b = G.primitive("pz", 1.1, (-0.4, 0.2, 0.3)) + G.primitive("py", 0.3, (0.0, 0.0, 0.0), -0.6)
    p = G.product(a, b)
    pts = np.random.default_rng(0).uniform(-2, 2, (200, 3))
    assert np.abs(p(pts) - a(pts) * b(pts)).max() < 1e-13


def test_tree_import_refuses_malformed_layouts(oracle):
    """The packed-layout contract of mwg_tree_import (include/mwg.h), checked on the host:
    a valid export passes; a node too deep, a translation outside the box, a duplicate,
    a node before its parent, an incomplete or reordered sibling group and roots out of
    order are each refused with a message naming the node."""
    import numpy as np
    from mrchem_amd import configs
    from mrchem_amd import mrcpp as G
    omra = oracle.MRA(**configs.BOX16, k=3, max_depth=6)
    t = oracle.Tree(omra).project_gauss(1e-3, configs.density3(0)[0])
    depth, lr, _ = t.export()
    boxes = configs.BOX16["boxes"]
    assert G.layout_check(boxes, 6, depth, lr) == ""
    assert depth.max() >= 2

    def bad(d, l, what, max_depth=6):
        msg = G.layout_check(boxes, max_depth, d, l)
        assert what in msg, (what, msg)

    first = int(np.argmax(depth == depth.max()))  # first node of a deepest group
    bad(depth, lr, "depth outside", max_depth=int(depth.max()) - 1)
    l2 = lr.copy()
    l2[first, 0] = (boxes[0] << depth[first]) + 2
    bad(depth, l2, "translation outside")
    d2, l2 = depth.copy(), lr.copy()
    l2[first + 1] = l2[first]
    bad(d2, l2, "duplicate")
    perm = np.arange(len(depth))
    perm[[first, first + 1]] = perm[[first + 1, first]]
    bad(depth[perm], lr[perm], "child order")
    bad(depth[:-3], lr[:-3], "complete sibling groups")
    nr = int(np.prod(boxes))
    perm = np.arange(len(depth))
    perm[[0, 1]] = perm[[1, 0]]
    bad(depth[perm], lr[perm], "roots must come first")
    g = nr + 8 * ((first - nr) // 8)   # move the deepest group before everything below the roots
    order = list(range(nr)) + list(range(g, g + 8)) + [i for i in range(nr, len(depth)) if not g <= i < g + 8]
    bad(depth[order], lr[order], "before its parent")
