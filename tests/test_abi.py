"""The C-ABI library loads and exports every entry point its header
declares; host-side logic of the product (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT
from fast_amd import mesh as M


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(fgpu_[A-Za-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    from fast_amd import fgpu
    L = fgpu.lib()
    names = _declared()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(L, n), n
    # and the Python mirror binds each of them
    assert names == set(fgpu.EXPORTS)


def test_dof_numbering_follows_reference_semantics():
    # ctrl.cpp:2004-2104: active if used by an element, supports -> 0,
    # node-major 1-based numbering
    m = M.hex8_block(2, 1, 1)
    assert m.dofmap.shape == (m.nnode, 3)
    fixed_nodes = np.isclose(m.coords[:, 0], 0.0)
    assert np.all(m.dofmap[fixed_nodes] == 0)
    free = m.dofmap[~fixed_nodes].ravel()
    assert np.array_equal(np.sort(free), np.arange(1, m.neq + 1))
    assert np.all(np.diff(m.dofmap[m.dofmap > 0]) > 0)


def test_bsp_block_connectivity_and_extraction():
    m = M.bsp_block(ne=(3, 2, 2), p=3, perturb=0.0)
    assert m.nn == 64 and m.nelem == 12
    # interior extraction operator of uniform cubic knots (closed form)
    C = M.bezier_extraction(3, M.open_uniform_knots(3, 5))
    mid = np.array([[1 / 6, 0, 0, 0], [2 / 3, 2 / 3, 1 / 3, 1 / 6],
                    [1 / 6, 1 / 3, 2 / 3, 2 / 3], [0, 0, 0, 1 / 6]])
    assert np.allclose(C[2], mid)
    # every column sums to one (partition of unity in Bernstein form)
    assert np.allclose(C.sum(axis=1), 1.0)
    # element e's control points: i fastest then j then k
    e = 1 + 3 * 1  # span (1, 1, 0)
    nb = m.nbas
    expect = [(1 + i) + nb[0] * ((1 + j) + nb[1] * k)
              for k in range(4) for j in range(4) for i in range(4)]
    assert list(m.conn[e]) == expect
    # control points at Greville abscissae reproduce the identity map
    assert np.allclose(m.coords[:nb[0], 0], M.greville(3, m.knots[0]))


def test_load_vector_of_face_traction_sums_to_total_force():
    m = M.hex8_block(4, 3, 2, size=(2.0, 1.5, 0.5), traction=(-1.0, 0.0, 0.0))
    assert np.isclose(m.fext[:, 0].sum(), -1.0 * 1.5 * 0.5)
    f = m.load_vector()
    assert np.isclose(f.sum(), -0.75)
