"""20-node bricks on the device (k_hex20.hip, the 8-node face kernel of
k_load.hip) against the FP64 oracle, all in FP64: the assembled stiffness
(linear, St.Venant-Kirchhoff and neo-Hooke on 2x2x2 and 3x3x3 Gauss
points; pattern bit-exact, values 1e-12 relative; the element matrices as
read back in FP32, 1e-6), the internal force, body force, stresses and
serendipity face loads (1e-12 relative), and load-control Newton on a
total-Lagrangian cantilever (the same iteration counts, states to the
Newton tolerance)."""
import numpy as np
import pytest

from fast_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def nrel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


MATS = {"linear": (3.0, 0.25), "tl_svk": (2.6, 0.3), "tl_neohooke": (1.0, 1.5)}


def block(form, nip, seed=3):
    return M.hex20_block(3, 2, 2, size=(1.2, 0.8, 0.7), form=form, nip=nip, perturb=0.3,
                         seed=seed, mat=np.tile(MATS[form], (12, 1)))


def deformation(m, seed=4):
    X = m.coords
    u = np.stack([0.03 * X[:, 0] * X[:, 2], 0.02 * X[:, 1] ** 2, -0.04 * X[:, 0] ** 2], 1)
    return u + np.random.default_rng(seed).uniform(-0.01, 0.01, u.shape)


@pytest.mark.parametrize("nip", [2, 3])
@pytest.mark.parametrize("form", ["linear", "tl_svk", "tl_neohooke"])
def test_hex20_stiffness_and_internal_force_vs_oracle(fg, oracle_mod, form, nip):
    m = block(form, nip)
    u = deformation(m)
    g = fg.GpuModel(m)
    om = oracle_mod.OracleMesh(m)
    g.stiffness(None if form == "linear" else u)
    Kg = g.element_matrices(0, m.nelem).astype(np.float64)
    worst = 0.0
    for e in range(m.nelem):
        Ko, ok = om.elem_stiff(e, None if form == "linear" else u)
        assert ok
        worst = max(worst, nrel(Kg[e], Ko))
    print("hex20 %s nip %d: worst element |dK|/|K| %.2e" % (form, nip, worst))
    assert worst <= 1e-6
    rowptr, col, val = g.csr()
    r0, c0 = om.pattern()
    assert np.array_equal(rowptr, r0) and np.array_equal(col, c0)
    v0, ok = om.assemble_K(None if form == "linear" else u)
    print("assembled |dK|/|K| %.2e" % nrel(val, v0))
    assert ok and nrel(val, v0) <= 1e-12
    fi = g.int_force(u)
    f0, ok = om.assemble_fint(u)
    assert ok and nrel(fi, f0) <= 1e-12
    g.close()


@pytest.mark.parametrize("form", ["linear", "tl_svk", "tl_neohooke"])
def test_hex20_stresses_and_body_force_vs_oracle(fg, oracle_mod, form):
    m = block(form, 3, seed=7)
    u = deformation(m, seed=8)
    g = fg.GpuModel(m)
    om = oracle_mod.OracleMesh(m)
    si, sn, _ = g.stresses(u)
    s0, n0, _, ok = om.stresses(u)
    assert ok
    assert np.abs(si - s0).max() <= 1e-12 * np.abs(s0).max()
    assert np.abs(sn - n0).max() <= 1e-12 * np.abs(n0).max()
    b = np.random.default_rng(0).uniform(-1, 1, (m.nelem, 3))
    fb = g.body_force(b)
    f0, ok = om.assemble_body(b)
    assert ok and np.abs(fb - f0).max() <= 1e-13 * np.abs(f0).max()
    g.close()


@pytest.mark.parametrize("nip", [2, 3])
def test_hex20_face_loads_vs_oracle(fg, oracle_mod, nip):
    """Serendipity faces of the distorted block's x = 1.2 end and top, in
    global and in local axes, against the oracle, and the flat-face closed
    form (-1/12 of qA per corner, 1/3 per mid-edge node)."""
    m = block("linear", nip, seed=9)
    x = m.coords
    faces = []
    for cn in m.conn:
        for fc in M.HEX20_FACES:
            nd = [int(cn[a]) for a in fc]
            if np.allclose(x[nd, 0], 1.2) or np.allclose(x[nd, 2], 0.7):
                faces.append(nd)
    faces = np.asarray(faces)
    rng = np.random.default_rng(2)
    q = rng.uniform(-1, 1, (len(faces), 3))
    loc = rng.integers(0, 2, len(faces)).astype(np.int32)
    grp = M.SurfaceGroup(deg=(2, 2), nip=nip, nodes=faces, q=q, local=loc, shape="q8")
    g = fg.GpuModel(m)
    g.add_surface_loads(grp)
    f = g.surface_load()
    f0 = oracle_mod.OracleMesh(m).surface_load([grp])
    assert np.abs(f - f0).max() <= 1e-13 * np.abs(f0).max()
    g.close()
    mc = M.hex20_block(1, 1, 1, size=(2.0, 3.0, 1.5), clamp_x0=False)
    face = [int(mc.conn[0][a]) for a in M.HEX20_FACES[0]]
    gc = fg.GpuModel(mc)
    gc.add_surface_loads(M.SurfaceGroup(deg=(2, 2), nip=nip, nodes=np.array([face]),
                                        q=[[0.0, 0.0, 1.0]], shape="q8"))
    fc = mc.scatter(gc.surface_load())
    for k, n in enumerate(face):
        assert fc[n, 2] == pytest.approx(6.0 * (-1.0 / 12.0 if k % 2 == 0 else 1.0 / 3.0),
                                         rel=1e-13)
    gc.close()


def test_hex20_newton_vs_oracle(fg, oracle_mod):
    """Load-control Newton (device driver, direct solves) on a
    St.Venant-Kirchhoff cantilever of 20-node bricks under a tip load,
    against the oracle's restatement of ctrlnr.cpp: the same iteration
    counts and states to the Newton tolerance."""
    m = M.hex20_block(4, 1, 1, size=(4.0, 1.0, 1.0), form="tl_svk",
                      mat=np.tile((1000.0, 0.3), (4, 1)))
    tip = np.isclose(m.coords[:, 0], 4.0)
    fext = np.zeros((m.nnode, 3))
    fext[tip, 2] = -2.0 / tip.sum()
    q = m.gather(fext)
    om = oracle_mod.OracleMesh(m)
    us0, its0 = oracle_mod.newton_raphson(om, q, 3, 1.0, 20, 1e-10)[:2]
    g = fg.GpuModel(m)
    us, st, err = g.newton(q, steps=3, factor=1.0, maxit=20, tol=1e-10, direct=True)
    assert not err and len(us) == 3
    for k in range(3):
        assert nrel(us[k], us0[k]) <= 1e-8, k
    print("hex20 Newton iterations", list(its0), [s["iters"] for s in st])
    assert [s["iters"] for s in st] == list(its0)
    g.close()
