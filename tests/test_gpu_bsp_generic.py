"""B-spline solids outside the tuned cubic kernels (k_bspg.hip: any degree up
to 7 per direction, any Gauss order) against the CPU FP64 oracle; these
kernels compute and keep FP64 throughout: element matrices and the
assembled matrix (1e-12 normwise), internal and body forces (1e-11 /
1e-13), stresses at the points and at the nodes (1e-12), and a
total-Lagrangian Newton solve."""
import numpy as np
import pytest

from fast_amd import mesh as M

pytestmark = pytest.mark.gpu

MATS = {"linear": (2.0, 0.3), "tl_svk": (2.0, 0.3), "tl_neohooke": (1.0, 1.5)}
SHAPES = [((2, 2, 2), None), ((4, 2, 3), None), ((1, 1, 1), 2), ((5, 2, 2), None), ((3, 3, 3), 5)]


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def nrel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(p, nip, form, ne=(3, 2, 2), seed=3):
    nel = int(np.prod(ne))
    return M.bsp_block(ne=ne, p=p, nip=nip, perturb=0.08, seed=seed, form=form,
                       mat=np.tile(MATS[form], (nel, 1)))


def state(m, amp=0.03):
    X = m.coords
    u = amp * np.stack([np.sin(1.3 * X[:, 0] + X[:, 1]), X[:, 0] * X[:, 2] - 0.5 * X[:, 1],
                        np.cos(X[:, 0]) * X[:, 1] ** 2], 1)
    u[m.fixed] = 0
    return u


@pytest.mark.parametrize("form", list(MATS))
@pytest.mark.parametrize("p,nip", SHAPES)
def test_generic_kernels_vs_oracle(fg, oracle_mod, p, nip, form):
    m = make(p, nip, form)
    u = state(m)
    uk = None if form == "linear" else u
    g = fg.GpuModel(m)
    om = oracle_mod.OracleMesh(m)
    g.stiffness(uk)
    Ke = g.element_matrices(0, m.nelem)
    for e in (0, m.nelem - 1):
        K0, ok = om.elem_stiff(e, uk)
        assert ok and nrel(Ke[e].astype(np.float64), K0) <= 1e-6, (p, nip, form, e)  # (FP32 export)
    val = g.csr()[2]
    v0, ok = om.assemble_K(uk)
    assert ok and nrel(val, v0) <= 1e-12
    f = g.int_force(u)
    f0, ok = om.assemble_fint(u)
    assert ok and nrel(f, f0) <= 1e-11
    b = np.random.default_rng(0).uniform(-1, 1, (m.nelem, 3))
    fb = g.body_force(b)
    fb0, ok = om.assemble_body(b)
    assert ok and np.abs(fb - fb0).max() <= 1e-13 * np.abs(fb0).max()
    si, sn, sa = g.stresses(u, nodal=True, average=True)
    r_si, r_sn, r_sa, ok = om.stresses(u)
    sc = np.abs(r_si).max()
    assert np.abs(si - r_si).max() <= 1e-12 * sc
    assert np.abs(sn - r_sn).max() <= 1e-12 * sc and np.abs(sa - r_sa).max() <= 1e-12 * sc
    if form != "tl_neohooke":
        g.set_stress_nodes(False)
        _, sn2, _ = g.stresses(u)
        r_sn2, ok = om.node_stresses(u)
        assert ok and np.abs(sn2 - r_sn2).max() <= 1e-12 * np.abs(r_sn2).max()
    g.close()


def test_node_stresses_of_tuned_cubic_solids(fg, oracle_mod):
    """Evaluation at the nodes also serves the tuned (3, 3, 3) and (3, 3, 2)
    elements (their extraction rows are stored padded)."""
    for p in (3, (3, 3, 2)):
        m = make(p, None, "linear")
        u = state(m)
        g = fg.GpuModel(m)
        g.set_stress_nodes(False)
        _, sn, _ = g.stresses(u)
        r_sn, ok = oracle_mod.OracleMesh(m).node_stresses(u)
        assert ok and np.abs(sn - r_sn).max() <= 1e-12 * np.abs(r_sn).max()
        g.close()


@pytest.mark.parametrize("form", ["tl_svk", "tl_neohooke"])
def test_generic_newton_vs_oracle(fg, oracle_mod, form):
    m = make((4, 2, 2), None, form, ne=(4, 1, 1))
    tip = np.isclose(m.coords[:, 0], m.coords[:, 0].max())
    fn = np.zeros((m.nnode, 3))
    fn[tip, 2] = -0.002 / tip.sum()
    q = np.zeros(m.neq)
    d = m.dofmap
    for a in range(m.nnode):
        for k in range(3):
            if d[a, k] > 0:
                q[d[a, k] - 1] += fn[a, k]
    us0, its0 = oracle_mod.newton_raphson(oracle_mod.OracleMesh(m), q, 3, 1.0 / 3, 30, 1e-10)[:2]
    g = fg.GpuModel(m)
    us, st, err = g.newton(q, steps=3, factor=1.0 / 3, maxit=30, tol=1e-10, pcg_tol=1e-12)
    assert err == "", err
    assert nrel(us[-1], us0[-1]) <= 1e-8
    g.close()
