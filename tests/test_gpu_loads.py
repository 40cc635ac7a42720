"""External loads on the device: uniform surface loads (fgpu_add_surface_loads /
fgpu_surface_load) against the oracle's restatement of cElemLoad::ExtForce,
and the prescribed-equation penalty (fgpu_set_penalty) in the stiffness, the
internal force and the drivers.

Surface loads are FP64 on both sides with the same point order; they agree
to 1e-13 of the largest entry (the device may contract products into fused
multiply-adds).  The penalty adds its values after the element
contributions: the assembled diagonal is the plain one plus the values,
bit for bit."""
import os

import numpy as np
import pytest

from fast_amd import mesh as M
from fast_amd.mesh import SurfaceGroup
from test_neutral import DECKS
from test_surface_load import prescribed_deck, quarter_cylinder, uniaxial_residual

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def same(a, b, tol=1e-13):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", ["TestB8", "BeamB8Face", "ThinCylinderDeg5", "CurvedBeamDeg4"])
def test_deck_surface_loads_vs_oracle(fg, oracle_mod, name):
    from fast_amd import neutral
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    g = fg.GpuModel(d.mesh)
    for s in d.surfaces:
        g.add_surface_loads(s)
    f = g.surface_load()
    same(f, oracle_mod.OracleMesh(d.mesh).surface_load(d.surfaces))
    g.clear_surface_loads()
    assert not g.surface_load().any()
    g.close()


def warped_bsp_faces(seed=5):
    """A cubic-quadratic B-spline block with random weights and a warped
    bottom face; two groups of faces of different degrees that share
    nodes: the bottom (z = 0: parameters x, y, degree (3, 2)) and the side
    (x = 0: parameters y, z, degree (2, 2)), mixed local flags."""
    rng = np.random.default_rng(seed)
    ne, p = (4, 3, 2), (3, 2, 2)
    m = M.bsp_block(ne=ne, p=p, perturb=0.1, seed=seed, clamp_x0=False)
    nb = m.nbas
    k = np.arange(m.nnode) // (nb[0] * nb[1])
    i = np.arange(m.nnode) % nb[0]
    m.coords[k == 0, 2] += rng.uniform(-0.03, 0.03, int((k == 0).sum()))
    m.coords[i == 0, 0] += rng.uniform(-0.03, 0.03, int((i == 0).sum()))
    m.cpw = rng.uniform(0.7, 1.3, m.nnode)
    q = [d + 1 for d in p]
    groups = []
    # bottom: element (sr, ss, 0); its extraction rows of x then y
    nodes, ext = [], []
    for ss in range(ne[1]):
        for sr in range(ne[0]):
            e = sr + ne[0] * ss
            nodes.append([(sr + a) + nb[0] * (ss + b) for b in range(q[1]) for a in range(q[0])])
            ext.append(m.ext[e][:q[0] ** 2 + q[1] ** 2])
    n = len(nodes)
    groups.append(SurfaceGroup(deg=(p[0], p[1]), nip=5, nodes=nodes, ext=ext,
                               q=rng.uniform(-1, 1, (n, 3)), local=rng.integers(0, 2, n)))
    # side x = 0: element (0, ss, st); rows of y then z
    nodes, ext = [], []
    o = q[0] ** 2
    for st in range(ne[2]):
        for ss in range(ne[1]):
            e = ne[0] * (ss + ne[1] * st)
            nodes.append([nb[0] * ((ss + a) + nb[1] * (st + b))
                          for b in range(q[2]) for a in range(q[1])])
            ext.append(m.ext[e][o:o + q[1] ** 2 + q[2] ** 2])
    n = len(nodes)
    groups.append(SurfaceGroup(deg=(p[1], p[2]), nip=4, nodes=nodes, ext=ext,
                               q=rng.uniform(-1, 1, (n, 3)), local=rng.integers(0, 2, n)))
    return m, groups


def test_rational_face_groups_vs_oracle(fg, oracle_mod):
    m, groups = warped_bsp_faces()
    g = fg.GpuModel(m)
    for s in groups:
        g.add_surface_loads(s)
    om = oracle_mod.OracleMesh(m)
    same(g.surface_load(), om.surface_load(groups))
    # groups keep their order and add: the second alone, then both again
    g.clear_surface_loads()
    g.add_surface_loads(groups[1])
    same(g.surface_load(), om.surface_load(groups[1:]))
    g.close()


def test_quarter_cylinder_closed_form(fg):
    """The exact quarter cylinder's pressure total R L (1, 1, 0); the face
    is the bottom of a one-element degree-(2, 1, 1) solid whose control
    points carry the circle's weights (the device reads the weights of
    B-spline meshes)."""
    R, L = 2.0, 3.0
    pts, cpw = quarter_cylinder(R, L)
    pts = np.asarray(pts)
    eye = [np.eye(3).ravel(), np.eye(2).ravel(), np.eye(2).ravel()]
    mb = M.Mesh(coords=np.concatenate([pts, pts * [0.5, 0.5, 1.0]]),
                conn=np.arange(12)[None, :], shape="bsp", nip=(3, 2, 2), form="linear",
                mat=np.array([[1.0, 0.3]]), fixed=np.zeros((12, 3), bool), deg=(2, 1, 1),
                ext=np.concatenate(eye)[None, :], cpw=np.concatenate([cpw, cpw]))
    g = fg.GpuModel(mb)
    g.add_surface_loads(SurfaceGroup(deg=(2, 1), nip=12, nodes=[list(range(6))],
                                     q=[[0.0, 0.0, 1.0]],
                                     ext=[np.concatenate(eye[:2])], local=[1]))
    f = mb.scatter(g.surface_load())
    assert np.allclose(f.sum(axis=0), [R * L, R * L, 0.0], rtol=1e-11, atol=1e-11)
    g.close()


def test_surface_load_arguments_rejected(fg):
    m = M.hex8_block(2, 2, 2)
    g = fg.GpuModel(m)
    base = dict(deg=(1, 1), nodes=[[0, 1, 3, 4]], q=[[0.0, 0.0, 1.0]])
    with pytest.raises(fg.FgpuError, match="Gauss points"):
        g.add_surface_loads(SurfaceGroup(nip=17, **base))
    with pytest.raises(fg.FgpuError):
        g.add_surface_loads(SurfaceGroup(nip=2, deg=(1, 1), nodes=[[0, 1, 3, m.nnode]],
                                         q=[[0.0, 0.0, 1.0]]))
    with pytest.raises(fg.FgpuError):
        g.add_surface_loads(SurfaceGroup(nip=2, deg=(8, 1), nodes=[list(range(18))],
                                         q=[[0.0, 0.0, 1.0]]))
    assert not g.surface_load().any()  # nothing was added
    g.close()


@pytest.mark.parametrize("kind", ["hex8", "bsp"])
def test_penalty_in_stiffness_and_internal_force(fg, kind):
    """The penalty values land on the diagonal after the element
    contributions (repeated equations add up) and the internal force takes
    v * u at the prescribed equations."""
    m = (M.hex8_block(4, 3, 3, perturb=0.2) if kind == "hex8"
         else M.bsp_block(ne=(3, 3, 2), p=3))
    rng = np.random.default_rng(1)
    g = fg.GpuModel(m)
    g.stiffness()
    rowptr, col, K0 = g.csr()
    u = rng.standard_normal((m.nnode, 3)) * 1e-3
    f0 = g.int_force(u)
    eq = rng.choice(np.arange(1, m.neq + 1), 12, replace=False)
    eq = np.concatenate([eq, eq[:3]])  # three equations given twice
    v = rng.uniform(1.0, 1e3, len(eq))
    g.set_penalty(eq, v)
    g.stiffness()
    _, _, K1 = g.csr()
    tot = {}
    for e, x in zip(eq, v):
        tot[e] = tot[e] + x if e in tot else x
    want = K0.copy()
    for e, x in tot.items():
        r = e - 1
        want[rowptr[r] + np.searchsorted(col[rowptr[r]:rowptr[r + 1]], r)] += x
    assert np.array_equal(K1, want)
    f1 = g.int_force(u)
    un = u.ravel()
    node_dir = {int(m.dofmap.ravel()[k]): k for k in np.flatnonzero(m.dofmap.ravel())}
    want = f0.copy()
    for e, x in tot.items():
        want[e - 1] += x * un[node_dir[e]]
    same(f1, want, tol=1e-15)
    g.set_penalty([], [])
    g.stiffness()
    assert np.array_equal(g.csr()[2], K0)
    g.close()


def test_prescribed_deck_closed_form(fg, tmp_path):
    """The bar pulled by a prescribed end displacement through the deck
    runner (penalty 1e30 on the device, direct solve): the uniaxial state."""
    from fast_amd import deck
    d, un, _ = deck.run(prescribed_deck(tmp_path), stresses=False)
    assert uniaxial_residual(d, un) <= 1e-9 * 1e-3


def test_prescribed_newton_free_norm(fg):
    """Load-control Newton with prescribed equations: the convergence test
    runs over the free equations (cControl::Convergence), so a
    displacement-driven total-Lagrangian block converges, its prescribed
    equations hold their values and its free equations are in balance."""
    from fast_amd import deck
    m = M.hex8_block(6, 2, 2, size=(3.0, 1.0, 1.0), form="tl_svk", E=100.0, nu=0.3)
    xmax = np.isclose(m.coords[:, 0], 3.0)
    eq = m.dofmap[xmax, 0]
    P = deck.PENALTY_STIFFNESS
    g = fg.GpuModel(m)
    g.set_penalty(eq, np.full(len(eq), P))
    f = np.zeros(m.neq)
    f[eq - 1] = P * 0.3  # 10 % stretch over the load path
    us, st, err = g.newton(f, steps=3, factor=1.0 / 3.0, maxit=30, tol=1e-8, direct=True)
    assert err == "" and all(s["converged"] for s in st), (err, st)
    u = m.scatter(us[-1])
    assert np.allclose(u[xmax, 0], 0.3, rtol=1e-12)
    g.set_penalty([], [])
    r = g.int_force(u)
    free = np.ones(m.neq, bool)
    free[eq - 1] = False
    assert np.abs(r[free]).max() <= 1e-7 * np.abs(r[~free]).max()
    g.close()
