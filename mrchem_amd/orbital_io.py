"""MRChem orbital and density files over the C ABI (include/mwg.h).

Mirrors ``Orbital::saveOrbital`` / ``loadOrbital`` (src/qmfunctions/Orbital.cpp:118-190),
``Density::saveDensity`` / ``loadDensity`` (src/qmfunctions/Density.cpp:55-125) and
``orbital::save_orbitals`` / ``load_orbitals`` (src/qmfunctions/orbital_utils.cpp:255-300):
``<prefix>.meta`` holds FunctionData (+ OrbitalData for an orbital), the real and
imaginary parts go to ``<prefix>_re.tree`` / ``<prefix>_im.tree``, and an orbital vector
is stored as ``<file>_<suffix><i>``; loading stops at the first index without a file.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from ._lib import MwgError, OrbitalMeta, TreeFileInfo, lib

SPIN_PAIRED, SPIN_ALPHA, SPIN_BETA = 0, 1, 2  # src/mrchem.h:41


def _check(rc: int):
    if rc != 0:
        raise MwgError(f"mwg error {rc}: {lib().mwg_last_error(None).decode()}")


@dataclass
class Meta:
    real_size: int = 0
    imag_size: int = 0
    is_shared: bool = False
    rank_id: int = 0
    spin: int = SPIN_PAIRED
    occ: int = 0
    error: float = 1.0
    orbital: bool = True


def write_meta(path: str, m: Meta) -> None:
    c = OrbitalMeta(m.real_size, m.imag_size, int(m.is_shared), m.rank_id, m.spin, m.occ, m.error)
    _check(lib().mwg_meta_write(path.encode(), C.byref(c), int(m.orbital)))


def read_meta(path: str) -> Meta:
    c = OrbitalMeta()
    orb = C.c_int()
    _check(lib().mwg_meta_read(path.encode(), C.byref(c), C.byref(orb)))
    return Meta(c.real_size, c.imag_size, bool(c.is_shared), c.rank_id, c.spin, c.occ, c.error, bool(orb.value))


def tree_file_write(name: str, info: dict, depth, lr, coefs) -> TreeFileInfo:
    """Host arrays (import layout) to "<name>.tree" without a device."""
    depth = np.ascontiguousarray(depth, np.int32)
    lr = np.ascontiguousarray(lr, np.int64).reshape(-1)
    coefs = np.ascontiguousarray(coefs, np.float64).reshape(-1)
    t = TreeFileInfo()
    for k in ("order", "basis", "root_scale", "max_depth", "scaling_factor"):
        setattr(t, k, info[k])
    t.corner[:] = list(info["corner"])
    t.boxes[:] = list(info["boxes"])
    t.nnodes = len(depth)
    _check(lib().mwg_tree_file_write(name.encode(), C.byref(t), depth.ctypes.data, lr.ctypes.data,
                                     coefs.ctypes.data))
    return t


def tree_file_read(name: str):
    """"<name>.tree" to (info, depth, lr, coefs) without a device."""
    t = TreeFileInfo()
    _check(lib().mwg_tree_file_read(name.encode(), C.byref(t), None, None, None))
    n, K = t.nnodes, t.order + 1
    depth = np.zeros(n, np.int32)
    lr = np.zeros((n, 3), np.int64)
    coefs = np.zeros((n, 8 * K ** 3), np.float64)
    _check(lib().mwg_tree_file_read(name.encode(), C.byref(t), depth.ctypes.data, lr.ctypes.data,
                                    coefs.ctypes.data))
    info = dict(order=t.order, basis=t.basis, root_scale=t.root_scale, max_depth=t.max_depth,
                corner=tuple(t.corner), boxes=tuple(t.boxes), scaling_factor=t.scaling_factor,
                nnodes=t.nnodes, chunk_nodes=t.chunk_nodes, nchunks=t.nchunks)
    return info, depth, lr, coefs


@dataclass
class Orbital:
    """An MRChem orbital: real and imaginary FunctionTrees (either may be absent) and its
    OrbitalData (Orbital.h:51-56)."""
    re: Optional[object] = None
    im: Optional[object] = None
    spin: int = SPIN_PAIRED
    occ: int = 2
    rank_id: int = 0
    error: float = 1.0


def save_orbital(prefix: str, phi: Orbital) -> None:
    """Orbital::saveOrbital (Orbital.cpp:118-146)."""
    m = Meta(spin=phi.spin, occ=phi.occ, rank_id=phi.rank_id, error=phi.error, orbital=True)
    if phi.re is not None:
        m.real_size = phi.re.saveTree(prefix + "_re")
    if phi.im is not None:
        m.imag_size = phi.im.saveTree(prefix + "_im")
    write_meta(prefix + ".meta", m)


def load_orbital(prefix: str, mra) -> Orbital:
    """Orbital::loadOrbital (Orbital.cpp:156-190): the parts the .meta marks present."""
    from .mrcpp import FunctionTree
    m = read_meta(prefix + ".meta")
    phi = Orbital(spin=m.spin, occ=m.occ, rank_id=m.rank_id, error=m.error)
    if m.real_size > 0:
        phi.re = FunctionTree(mra).loadTree(prefix + "_re")
    if m.imag_size > 0:
        phi.im = FunctionTree(mra).loadTree(prefix + "_im")
    return phi


def save_density(prefix: str, re, im=None) -> None:
    """Density::saveDensity (Density.cpp:55-83): a 12-byte .meta."""
    m = Meta(orbital=False)
    if re is not None:
        m.real_size = re.saveTree(prefix + "_re")
    if im is not None:
        m.imag_size = im.saveTree(prefix + "_im")
    write_meta(prefix + ".meta", m)


def load_density(prefix: str, mra):
    from .mrcpp import FunctionTree
    m = read_meta(prefix + ".meta")
    re = FunctionTree(mra).loadTree(prefix + "_re") if m.real_size > 0 else None
    im = FunctionTree(mra).loadTree(prefix + "_im") if m.imag_size > 0 else None
    return re, im


def save_orbitals(Phi: List[Orbital], file: str, suffix: str = "", n_orbs: int = -1) -> None:
    """orbital::save_orbitals (orbital_utils.cpp:255-270): "<file>_<suffix><i>"."""
    if n_orbs < 0:
        n_orbs = len(Phi)
    if n_orbs > len(Phi):
        raise MwgError("Index out of bounds")
    for i in range(n_orbs):
        save_orbital(f"{file}_{suffix}{i}", Phi[i])


def load_orbitals(file: str, mra, suffix: str = "", n_orbs: int = -1) -> List[Orbital]:
    """orbital::load_orbitals (orbital_utils.cpp:281-300): until the first index without
    a .meta file (or n_orbs), the orbitals with a part present."""
    Phi = []
    i = 0
    while n_orbs <= 0 or i < n_orbs:
        prefix = f"{file}_{suffix}{i}"
        if not os.path.exists(prefix + ".meta"):
            break
        phi = load_orbital(prefix, mra)
        if phi.re is None and phi.im is None:
            break
        Phi.append(phi)
        i += 1
    return Phi
