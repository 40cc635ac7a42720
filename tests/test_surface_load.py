"""The oracle's surface-load integration (orc_surface_load, restating
cElemLoad::ExtForce over cFaceUnifLoad / cFaceUnifLoadIGA faces) pinned by
closed forms, and the neutral reader's surface groups and prescribed
displacements.

Beyond these, the reference program's own runs pin it: the decks with face
loads (TestB8, BeamB8Face: brick faces in global and local axes;
ThinCylinderDeg5, CurvedBeamDeg4: IGA faces, local and global) are solved
with the oracle's load vector and match the reference's printed
displacements (tests/test_neutral.py)."""
import os

import numpy as np
import pytest

from fast_amd import mesh as M
from fast_amd import neutral
from fast_amd.mesh import SurfaceGroup
from test_neutral import DECKS


def free_mesh(coords, cpw=None):
    """A mesh over the given points whose every direction is an equation
    (one dummy brick on the first eight points), so that load vectors hold
    every node's total."""
    coords = np.asarray(coords, dtype=np.float64)
    n = len(coords)
    conn = np.arange(8, dtype=np.int32)[None, :] % n
    dofmap = np.arange(1, 3 * n + 1, dtype=np.int32).reshape(n, 3)
    return M.Mesh(coords=coords, conn=conn, shape="hex8", nip=(2, 2, 2), form="linear",
                  mat=np.array([[1.0, 0.3]]), fixed=np.zeros((n, 3), bool), dofmap=dofmap,
                  cpw=cpw)


def test_parallelogram_face_corner_shares(oracle_mod):
    """A flat parallelogram face: detJ is constant, so every corner takes a
    quarter of q * area, in global axes; in local axes the w component acts
    along the unit normal (x_r x x_s)."""
    a, b = np.array([2.0, 0.5, 0.1]), np.array([0.3, 1.5, -0.2])
    x0 = np.array([1.0, -1.0, 0.5])
    pts = [x0, x0 + a, x0 + a + b, x0 + b]        # n0 n1 n2 n3 of the face
    m = free_mesh(pts)
    area = np.linalg.norm(np.cross(a, b))
    q = np.array([3.0, -2.0, 7.0])
    g = SurfaceGroup(deg=(1, 1), nip=2, nodes=[[0, 1, 3, 2]], q=[q])
    f = m.scatter(oracle_mod.OracleMesh(m).surface_load([g]))
    assert np.allclose(f, np.tile(q * area / 4, (4, 1)), rtol=1e-14, atol=0)
    # local: u = a/|a|, w = u x b/|b|, v = w x u; a pure w load is a pressure
    # along (a x b)/(|a||b|) per unit area
    gl = SurfaceGroup(deg=(1, 1), nip=3, nodes=[[0, 1, 3, 2]], q=[[0.0, 0.0, 5.0]], local=[1])
    f = m.scatter(oracle_mod.OracleMesh(m).surface_load([gl]))
    w = np.cross(a / np.linalg.norm(a), b / np.linalg.norm(b))
    assert np.allclose(f.sum(axis=0), 5.0 * area * w, rtol=1e-13)


def test_bezier_face_shares_follow_the_first_parameter(oracle_mod):
    """A degree-(2, 2) Bezier face x(s, t) = (s, t (1 + s), 0) on [0, 1]^2
    (identity extraction): dA = (1 + s) ds dt, so node (i, j) takes
    q * (1/3) * (1/3 + (i + 1)/12) -- exact under 3 x 3 Gauss points.  The
    shares differ along the first parameter only, which pins the node
    order (first parameter fastest)."""
    pts = [(i / 2.0, (j / 2.0) * (1.0 + i / 2.0), 0.0) for j in range(3) for i in range(3)]
    m = free_mesh(pts)
    g = SurfaceGroup(deg=(2, 2), nip=3, nodes=[list(range(9))], q=[[0.0, 0.0, 6.0]],
                     ext=[np.concatenate([np.eye(3).ravel(), np.eye(3).ravel()])])
    f = m.scatter(oracle_mod.OracleMesh(m).surface_load([g]))
    want = np.array([6.0 / 3.0 * (1.0 / 3.0 + (i + 1) / 12.0) for j in range(3) for i in range(3)])
    assert np.allclose(f[:, 2], want, rtol=1e-14, atol=0)
    assert not f[:, :2].any()


def quarter_cylinder(R=2.0, L=3.0):
    """Rational quadratic quarter circle of radius R (weights 1, 1/sqrt 2, 1)
    times a straight segment of length L along z: a degree-(2, 1) face."""
    arc = [(R, 0.0), (R, R), (0.0, R)]
    w = [1.0, np.sqrt(0.5), 1.0]
    pts, cpw = [], []
    for z in (0.0, L):
        for (x, y), wi in zip(arc, w):
            pts.append((x, y, z))
            cpw.append(wi)
    return pts, np.array(cpw)


def test_rational_quarter_cylinder_pressure(oracle_mod):
    """A pressure along the local w axis of the exact quarter cylinder: the
    integrand x_s x x_t is a rational function, and 12 x 12 Gauss points give
    the closed forms: total R L (1, 1, 0) for unit pressure, and the area
    pi R L / 2 for a global unit load."""
    R, L = 2.0, 3.0
    pts, cpw = quarter_cylinder(R, L)
    m = free_mesh(pts, cpw=cpw)
    ext = np.concatenate([np.eye(3).ravel(), np.eye(2).ravel()])
    g = SurfaceGroup(deg=(2, 1), nip=12, nodes=[list(range(6))], q=[[0.0, 0.0, 1.0]],
                     ext=[ext], local=[1])
    f = m.scatter(oracle_mod.OracleMesh(m).surface_load([g]))
    assert np.allclose(f.sum(axis=0), [R * L, R * L, 0.0], rtol=1e-11, atol=1e-11)
    g2 = SurfaceGroup(deg=(2, 1), nip=12, nodes=[list(range(6))], q=[[0.0, 1.0, 0.0]],
                      ext=[ext])
    f = m.scatter(oracle_mod.OracleMesh(m).surface_load([g2]))
    assert f[:, 1].sum() == pytest.approx(np.pi * R * L / 2, rel=1e-11)
    # the load is symmetric about the arc's middle: end columns mirror
    assert f[0, 1] == pytest.approx(f[2, 1], rel=1e-13)


def test_deck_surface_groups():
    d = neutral.read(os.path.join(DECKS, "BeamB8Face.dat"))
    assert len(d.surfaces) == 1 and not d.load.any()
    g = d.surfaces[0]
    assert g.deg == (1, 1) and g.ext is None and g.nface == 8 and g.local.all()
    assert np.allclose(g.q, [[0.02, -0.01, -0.3]] * 8)
    d = neutral.read(os.path.join(DECKS, "ThinCylinderDeg5.dat"))
    g = d.surfaces[0]
    p = d.mesh.deg
    assert g.nface == 80 and g.nip == max(d.mesh.nip) and g.local.all()
    assert g.nodes.shape[1] == (g.deg[0] + 1) * (g.deg[1] + 1)
    assert g.ext.shape == (80, (g.deg[0] + 1) ** 2 + (g.deg[1] + 1) ** 2)
    assert set(g.deg) <= set(p)


PRESCRIBED = """%LOAD.CASE.PRESCRIBED.DOF
4
2 1 0.001
3 1 0.001
6 1 0.001
7 1 0.001

%END"""


def prescribed_deck(tmp_path):
    """TestB8 with the end face pulled by a prescribed ux = 1 mm instead of
    the face load."""
    text = open(os.path.join(DECKS, "TestB8.dat")).read()
    head = text[:text.index("%LOAD.CASE.AREA.UNIFORM")]
    path = tmp_path / "TestB8Presc.dat"
    path.write_text(head + PRESCRIBED + "\n")
    return str(path)


def test_prescribed_dofs_read(tmp_path):
    d = neutral.read(prescribed_deck(tmp_path))
    eq, val = d.prescribed
    assert np.array_equal(val, [0.001] * 4) and not d.surfaces and not d.load.any()
    idx = {int(l): i for i, l in enumerate(d.labels)}
    assert np.array_equal(eq, [d.mesh.dofmap[idx[n], 0] for n in (2, 3, 6, 7)])


def uniaxial_residual(d, u):
    """Largest deviation of u from the uniaxial state u = (2e-4 x,
    -nu 2e-4 y, -nu 2e-4 z) of the bar, up to a rotation about the x axis:
    the deck's supports (node 5 fixed, nodes 1, 4 and 8 in x) leave that
    rigid mode free, and a solver settles it by rounding."""
    X = d.mesh.coords
    want = np.stack([2e-4 * X[:, 0], -0.3 * 2e-4 * X[:, 1], -0.3 * 2e-4 * X[:, 2]], axis=1)
    mode = np.stack([np.zeros(len(X)), -X[:, 2], X[:, 1]], axis=1)
    r = (u - want).ravel()
    theta = r @ mode.ravel() / (mode.ravel() @ mode.ravel())
    return np.abs(r - theta * mode.ravel()).max()


def test_prescribed_penalty_solution_closed_form(oracle_mod, tmp_path):
    """The penalty method of the reference (K_ii += 1e30, f_i += 1e30 v,
    ctrl/ctrl.cpp:958-974, 698-713) on the bar: a prescribed end
    displacement of 1 mm over 5 m gives the uniaxial state exactly (the
    brick reproduces linear fields)."""
    d = neutral.read(prescribed_deck(tmp_path))
    om = oracle_mod.OracleMesh(d.mesh)
    rowptr, col = om.pattern()
    K, ok = om.assemble_K()
    assert ok
    eq, val = d.prescribed
    f = np.zeros(d.mesh.neq)
    for e, v in zip(eq, val):
        r = e - 1
        K[rowptr[r] + np.searchsorted(col[rowptr[r]:rowptr[r + 1]], r)] += 1e30
        f[r] += 1e30 * v
    ue, ok = oracle_mod.skyline_solve(rowptr, col, K, f)
    assert ok
    assert uniaxial_residual(d, d.mesh.scatter(ue)) <= 1e-12 * 1e-3
