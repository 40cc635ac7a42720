"""Parity of the HIP path (through the C-ABI) with the CPU oracle, the
reference's golden fixtures, and size-independent properties at full size.

Tolerances (written per test):
* element / global stiffness: FP32 element matrices vs the FP64 oracle,
  normwise relative error |K_gpu - K_64|_F / |K_64|_F <= 1e-6;
* internal force: FP64 throughout, relative error <= 1e-11;
* assembly pattern and DOF maps: bit-exact (integer);
* PCG on identical FP64 matrices: same iteration count as the oracle's
  restatement of the reference PCG, solution to 1e-10.
"""
import numpy as np
import pytest

from conftest import load_golden, mesh_from_golden
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


def elem_check(fg, oracle_mod, mesh, u=None, tol=1e-6, elems=None):
    g = fg.GpuModel(mesh)
    g.stiffness(u)
    om = oracle_mod.OracleMesh(mesh)
    elems = range(mesh.nelem) if elems is None else elems
    Kg = g.element_matrices(0, mesh.nelem)
    worst = 0.0
    for e in elems:
        Ko, ok = om.elem_stiff(e, u)
        err = nrel(Kg[e].astype(np.float64), Ko)
        worst = max(worst, err)
        assert err <= tol, (e, err)
    return worst, g, om


@pytest.mark.parametrize("form", ["linear", "tl_svk"])
def test_hex8_element_matrices_vs_reference_fixture(fg, oracle_mod, form):
    d = load_golden("elem_hex8_%s.npz" % form)
    m = mesh_from_golden(d)
    g = fg.GpuModel(m)
    u = d["u"] if form != "linear" else None
    g.stiffness(u)
    Kg = g.element_matrices(0, m.nelem).astype(np.float64)
    for e in range(m.nelem):
        assert nrel(Kg[e], d["K"][e]) <= 1e-6


@pytest.mark.parametrize("form,matp", [("linear", (200e3, 0.3)), ("tl_svk", (2.6, 0.3)),
                                        ("tl_neohooke", (1.0, 1.5))])
@pytest.mark.parametrize("nip", [1, 2, 3])
def test_hex8_element_matrices_vs_oracle(fg, oracle_mod, form, matp, nip):
    m = M.hex8_block(3, 2, 2, form=form, nip=nip, perturb=0.3, seed=nip,
                     mat=np.tile(matp, (12, 1)))
    u = None
    if form != "linear":
        u = np.random.default_rng(3).uniform(-0.05, 0.05, (m.nnode, 3))
    elem_check(fg, oracle_mod, m, u, tol=1e-6 if nip > 1 else 2e-6)


@pytest.mark.parametrize("form", ["linear", "tl_svk"])
def test_bsp_element_matrices_vs_reference_fixture(fg, oracle_mod, form):
    d = load_golden("elem_bsp3_%s.npz" % form)
    m = mesh_from_golden(d)
    g = fg.GpuModel(m)
    g.stiffness(d["u"] if form != "linear" else None)
    Kg = g.element_matrices(0, m.nelem).astype(np.float64)
    assert nrel(Kg[0], d["K0"]) <= 1e-6
    for e in range(m.nelem):
        assert nrel(Kg[e] @ d["v"], d["Kv"][e]) <= 2e-6
        assert nrel(np.diag(Kg[e]), d["Kdiag"][e]) <= 1e-6


@pytest.mark.parametrize("deg", [3, (3, 3, 2)])
@pytest.mark.parametrize("form,matp", [("tl_svk", (2.6, 0.3)), ("tl_neohooke", (1.0, 1.5))])
def test_bsp_total_lagrangian_vs_oracle(fg, oracle_mod, form, matp, deg):
    """B-spline total-Lagrangian tangent (vector-ALU kernel) and internal
    force against the FP64 oracle on a perturbed, deformed patch."""
    m = M.bsp_block(ne=(3, 2, 2), p=deg, size=(1.0, 0.7, 0.5), form=form, perturb=0.08, seed=5,
                    mat=np.tile(matp, (12, 1)))
    X = m.coords
    u = np.stack([0.04 * X[:, 0] * X[:, 2], 0.03 * X[:, 1] ** 2, -0.05 * X[:, 0] ** 2], 1)
    u += np.random.default_rng(4).uniform(-0.01, 0.01, u.shape)
    worst, g, om = elem_check(fg, oracle_mod, m, u, tol=1e-6)
    print("bsp %s worst element rel err %.2e" % (form, worst))
    fi = g.int_force(u)
    f0, ok = om.assemble_fint(u)
    assert ok
    assert nrel(fi, f0) <= 1e-11


@pytest.mark.parametrize("deg", [3, (3, 3, 2)])
def test_bsp_element_matrices_vs_oracle_perturbed_patch(fg, oracle_mod, deg):
    m = M.bsp_block(ne=(4, 3, 2), p=deg, size=(1.0, 0.7, 0.4), E=2.6, nu=0.3, perturb=0.08,
                    seed=21)
    worst, g, om = elem_check(fg, oracle_mod, m, tol=1e-6)
    print("bsp", deg, "worst element rel err", worst)
    X = m.coords
    u = np.stack([0.01 * X[:, 0] * X[:, 2], 0.02 * X[:, 1], -0.01 * X[:, 0] ** 2], 1)
    f0, ok = om.assemble_fint(u)
    assert nrel(g.int_force(u), f0) <= 1e-11


@pytest.mark.parametrize("kind", ["hex8_linear", "hex8_tl", "hex8_nh", "bsp", "bsp332"])
def test_assembly_pattern_bitexact_and_values(fg, oracle_mod, kind):
    rng = np.random.default_rng(8)
    if kind in ("bsp", "bsp332"):
        m = M.bsp_block(ne=(3, 3, 2), p=3 if kind == "bsp" else (3, 3, 2), perturb=0.05, seed=2)
        u = None
    else:
        form = {"hex8_linear": "linear", "hex8_tl": "tl_svk", "hex8_nh": "tl_neohooke"}[kind]
        mat = (1.0, 1.5) if form == "tl_neohooke" else (2.6, 0.3)
        m = M.hex8_block(5, 4, 3, form=form, perturb=0.25, seed=4, mat=np.tile(mat, (60, 1)))
        u = None if form == "linear" else rng.uniform(-0.05, 0.05, (m.nnode, 3))
    g = fg.GpuModel(m)
    g.stiffness(u)
    rowptr, col, val = g.csr()
    om = oracle_mod.OracleMesh(m)
    r0, c0 = om.pattern()
    assert np.array_equal(rowptr, r0)
    assert np.array_equal(col, c0)
    v0, ok = om.assemble_K(u)
    assert nrel(val, v0) <= 1e-6
    # the assembled matrix is exactly symmetric (mirrored blocks)
    A = om.to_scipy(val)
    assert (A - A.T).count_nonzero() == 0


@pytest.mark.parametrize("form,matp", [("linear", (3.0, 0.25)), ("tl_svk", (2.6, 0.3)),
                                        ("tl_neohooke", (1.0, 1.5))])
def test_internal_force_vs_oracle_fp64(fg, oracle_mod, form, matp):
    m = M.hex8_block(4, 3, 3, form=form, perturb=0.3, seed=5, mat=np.tile(matp, (36, 1)))
    u = np.random.default_rng(6).uniform(-0.06, 0.06, (m.nnode, 3))
    g = fg.GpuModel(m).int_force(u)
    g0, ok = oracle_mod.OracleMesh(m).assemble_fint(u)
    assert ok
    assert nrel(g, g0) <= 1e-11


@pytest.mark.parametrize("form", ["linear", "tl_svk"])
def test_internal_force_bsp_vs_reference_fixture(fg, oracle_mod, form):
    d = load_golden("elem_bsp3_%s.npz" % form)
    m = mesh_from_golden(d)
    gg = fg.GpuModel(m).int_force(d["u"])
    g0 = np.zeros(m.neq)
    for e in range(m.nelem):
        dofs = d["edof"][e]
        np.add.at(g0, dofs[dofs > 0] - 1, d["g"][e][dofs > 0])
    assert nrel(gg, g0) <= 1e-11


def test_pcg_same_iterations_as_reference_restatement(fg, oracle_mod):
    m = M.hex8_block(6, 4, 4, perturb=0.2, seed=3, traction=(0, 0, -1.0))
    om = oracle_mod.OracleMesh(m)
    v0, _ = om.assemble_K()
    rowptr, col = om.pattern()
    f = m.load_vector()
    u0, it0, rr0, conv0 = oracle_mod.pcg(rowptr, col, v0, f, 10000, 1e-10)
    g = fg.GpuModel(m)
    g.set_csr_values(v0)
    x, it, rr, conv = g.solve(f, tol=1e-10, maxit=10000)
    assert conv and conv0
    assert abs(it - it0) <= 1, (it, it0)
    assert nrel(x, u0) <= 1e-9
    # warm start from the solution converges at once
    x2, it2, rr2, conv2 = g.solve(f, x0=u0, tol=1e-8)
    assert conv2 and it2 <= 1


def test_block_jacobi_batch_solve(fg):
    """The batch solve with node-block Jacobi: every mesh converges to the
    point-Jacobi solution within the tolerance, in no more iterations in
    total, and each mesh matches its own single-mesh block-Jacobi solve."""
    m, prm = M.cube_batch(7, nmin=2, nmax=8, seed=4)
    g = fg.GpuModel(m)
    g.stiffness()
    f = m.load_vector()
    x0, it0, _, n0 = g.solve_batch(f, tol=1e-10)
    g.set_preconditioner(1)
    x1, it1, _, n1 = g.solve_batch(f, tol=1e-10)
    print("batch PCG iterations: point", it0.sum(), "node blocks", it1.sum())
    assert n0 == 7 and n1 == 7
    assert it1.sum() < it0.sum()
    assert nrel(x1, x0) <= 1e-8
    bp = m.batch_nodeptr
    for k in (0, 6):
        mk = M.hex8_cube(int(prm["n"][k]), nu=float(prm["nu"][k]),
                         clamp_face=int(prm["clamp_face"][k]), load_face=int(prm["load_face"][k]))
        gk = fg.GpuModel(mk)
        gk.stiffness()
        gk.set_preconditioner(1)
        xk, itk, _, conv = gk.solve(mk.load_vector(), tol=1e-10)
        assert conv and abs(int(it1[k]) - itk) <= 1, (k, it1[k], itk)
        assert nrel(m.scatter(x1)[bp[k]:bp[k + 1]], mk.scatter(xk)) <= 1e-9


def test_block_jacobi_pcg(fg, oracle_mod):
    """Node-block Jacobi (fgpu_set_preconditioner 1) solves the oracle's
    system to the same tolerance as point Jacobi in fewer iterations, eager
    and from the captured iteration graphs, and switching back gives point
    Jacobi's solve bit for bit."""
    m = M.hex8_block(12, 8, 8, perturb=0.2, seed=3, traction=(0, 0, -1.0))
    om = oracle_mod.OracleMesh(m)
    v0, _ = om.assemble_K()
    g = fg.GpuModel(m)
    g.set_csr_values(v0)
    f = m.load_vector()
    x0, it0, rr0, c0 = g.solve(f, tol=1e-10, maxit=20000)
    g.set_preconditioner(1)
    x1, it1, rr1, c1 = g.solve(f, tol=1e-10, maxit=20000)
    x1b, it1b, _, _ = g.solve(f, tol=1e-10, maxit=20000)
    print("PCG iterations: point Jacobi", it0, "node blocks", it1)
    assert c0 and c1 and it1 < it0
    assert it1b == it1 and np.array_equal(x1b, x1)
    A = om.to_scipy(v0)
    assert np.linalg.norm(A @ x1 - f) <= 1e-9 * np.linalg.norm(f)
    assert nrel(x1, x0) <= 1e-8
    g.set_preconditioner(0)
    x2, it2, _, _ = g.solve(f, tol=1e-10, maxit=20000)
    assert it2 == it0 and np.array_equal(x2, x0)
    with pytest.raises(fg.FgpuError):
        g.set_preconditioner(2)


def test_linear_static_cfg1_vs_reference_golden(fg, oracle_mod):
    d = load_golden("cfg1_linear_u.npz")
    m = mesh_from_golden(d)
    g = fg.GpuModel(m)
    g.stiffness()
    x, it, rr, conv = g.solve(m.load_vector(), tol=1e-10)
    assert conv
    err = nrel(m.scatter(x), d["u"])
    print("cfg1 u rel err vs reference", err, "iters", it)
    assert err <= 1e-5


def test_newton_raphson_tl_vs_reference_golden(fg, oracle_mod):
    d = load_golden("nr_tl_hex8_4x4x4.npz")
    m = mesh_from_golden(d)
    g = fg.GpuModel(m)
    q = m.load_vector()
    u = np.zeros(m.neq)
    fint = g.int_force(m.scatter(u))
    lam = 0.0
    for step in range(5):
        lam += 0.2
        f = lam * q
        r = f - fint
        for it in range(30):
            g.stiffness(m.scatter(u))
            du, _, _, conv = g.solve(r, tol=1e-12)
            u = u + du
            fint = g.int_force(m.scatter(u))
            r = f - fint
            if np.linalg.norm(r) / max(1.0, np.linalg.norm(f)) < 1e-9:
                break
        err = nrel(m.scatter(u), d["u_steps"][step])
        assert err <= 1e-8, (step, err)


def test_full_size_cfg3_properties(fg):
    """64^3 hex8 (neo-Hookean) at a deformed state: K_t symmetric, and
    f_int(u + h v) - f_int(u - h v) = 2 h K_t v to FP32 accuracy."""
    m = M.hex8_block(64, 64, 64, form="tl_neohooke", mat=np.tile((1.0, 1.5), (64 ** 3, 1)))
    rng = np.random.default_rng(0)
    X = m.coords
    u = np.stack([0.02 * X[:, 0] * X[:, 2], 0.01 * X[:, 1] ** 2, -0.03 * X[:, 0] ** 2], 1)
    u[m.dofmap[:, 0] == 0] = 0.0
    g = fg.GpuModel(m)
    g.stiffness(u)
    rowptr, col, val = g.csr()
    assert len(val) == 63_695_790
    v = rng.standard_normal(m.neq)
    h = 1e-6
    vp = m.scatter(v)
    d = (g.int_force(u + h * vp) - g.int_force(u - h * vp)) / (2 * h)
    import scipy.sparse as sp
    K = sp.csr_matrix((val, col, rowptr), shape=(m.neq, m.neq))
    Kv = K @ v
    assert nrel(Kv, d) <= 1e-5
    assert abs(K - K.T).max() == 0


def test_newton_driver_on_device_vs_reference_golden(fg, oracle_mod):
    """fgpu_newton (the ctrlnr.cpp loop run natively on the device)."""
    d = load_golden("nr_tl_hex8_4x4x4.npz")
    m = mesh_from_golden(d)
    g = fg.GpuModel(m)
    us, st, err = g.newton(m.load_vector(), steps=5, factor=0.2, maxit=30, tol=1e-9,
                           pcg_tol=1e-12)
    assert len(us) == 5, err
    for s in range(5):
        assert nrel(m.scatter(us[s]), d["u_steps"][s]) <= 1e-8
        assert st[s]["converged"]


def test_batched_solve_equals_independent_solves(fg, oracle_mod):
    """A batch of uneven cubes solved by one per-mesh PCG equals each mesh
    solved alone (same iteration count, same iterate), and the oracle's
    direct solution of each mesh."""
    m, prm = M.cube_batch(7, nmin=2, nmax=6, seed=3)
    g = fg.GpuModel(m)
    g.stiffness()
    f = m.load_vector()
    x, its, rr, nconv = g.solve_batch(f, tol=1e-10)
    assert nconv == 7
    bp = m.batch_nodeptr
    for k in range(7):
        mk = M.hex8_cube(int(prm["n"][k]), nu=float(prm["nu"][k]),
                         clamp_face=int(prm["clamp_face"][k]), load_face=int(prm["load_face"][k]))
        gk = fg.GpuModel(mk)
        gk.stiffness()
        xk, itk, rrk, conv = gk.solve(mk.load_vector(), tol=1e-10)
        assert conv
        eq = m.dofmap[bp[k]:bp[k + 1]]
        xb = m.scatter(x)[bp[k]:bp[k + 1]]
        assert abs(int(its[k]) - itk) <= 1, (k, its[k], itk)
        assert nrel(xb, mk.scatter(xk)) <= 1e-9
        om = oracle_mod.OracleMesh(mk)
        ud = oracle_mod.linear_static(om, mk.load_vector())
        # FP32 element matrices: solution error <= cond(K) * 1e-7
        assert nrel(mk.scatter(xk), mk.scatter(ud)) <= 1e-5
        assert eq.max() > 0


def test_graph_replays_equal_eager_launches(fg, oracle_mod):
    """PCG chunks replayed from a captured hipGraph give bit-identical
    iterates to eagerly launched kernels, across repeated solves."""
    import os
    m = M.hex8_block(6, 5, 4, perturb=0.2, seed=9, traction=(0, 0, -1.0))
    f = m.load_vector()
    ga = fg.GpuModel(m)
    ga.stiffness()
    os.environ["FGPU_NO_GRAPH"] = "1"
    try:
        gb = fg.GpuModel(m)
    finally:
        del os.environ["FGPU_NO_GRAPH"]
    gb.stiffness()
    for tol in (1e-6, 1e-10, 1e-8):
        xa, ia, _, _ = ga.solve(f, tol=tol)
        xb, ib, _, _ = gb.solve(f, tol=tol)
        assert ia == ib
        assert np.array_equal(xa, xb)


def test_linear_static_driver(fg, oracle_mod):
    d = load_golden("cfg1_linear_u.npz")
    m = mesh_from_golden(d)
    g = fg.GpuModel(m)
    u, it, rr = g.linear_static(m.load_vector(), tol=1e-10)
    # FP64 residual refinement: the FP32 element matrices limit nothing
    assert rr <= 1e-10
    assert nrel(m.scatter(u), d["u"]) <= 1e-8


@pytest.mark.parametrize("overlap", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["hex8_nh", "hex8_linear", "bsp_svk", "bsp332_linear",
                                  "bsp_linear"])
def test_stiffness_beside_internal_force_identical_bits(fg, kind, overlap):
    """fgpu_assemble_K_fint in every order, forced (0: K then f_int on one
    stream; 1: f_int on the second stream beside the assembly; 2: its element
    vectors beside the assembly, the gather after it; 3: its element vectors
    first, the gather beside the element stiffness; for the linear
    tuned B-spline kernels of degree (3,3,3) and (3,3,2) finished from the
    per-point records the stiffness kernel writes) gives the matrix and the
    force of fgpu_assemble_K + fgpu_int_force, bit for bit, and reports a
    failed stress update of either side."""
    if kind == "hex8_nh":
        m = M.hex8_block(5, 4, 3, form="tl_neohooke", perturb=0.2, seed=4,
                         mat=np.tile((1.0, 1.5), (60, 1)))
    elif kind == "hex8_linear":
        m = M.hex8_block(5, 4, 3, perturb=0.2, seed=6)
    elif kind == "bsp_svk":
        m = M.bsp_block(ne=(3, 3, 2), p=3, perturb=0.05, seed=2, form="tl_svk",
                        mat=np.tile((2.6, 0.3), (18, 1)))
    elif kind == "bsp_linear":
        m = M.bsp_block(ne=(3, 3, 2), p=3, perturb=0.05, seed=5)
    else:
        m = M.bsp_block(ne=(3, 2, 2), p=(3, 3, 2), perturb=0.05, seed=3)
    X = m.coords
    u = np.stack([0.02 * X[:, 0] * X[:, 2], 0.01 * X[:, 1] ** 2, -0.03 * X[:, 0] ** 2], 1)
    u[m.fixed] = 0
    g = fg.GpuModel(m)
    g.set_overlap(overlap)
    g.stiffness(u)
    K1 = g.csr()[2].copy()
    f1 = g.int_force(u)
    f2 = g.stiffness_and_int_force(u)
    assert "f_int beside" not in g.tuning_report()  # the forced order, untimed
    K2 = g.csr()[2]
    assert np.array_equal(K1, K2) and np.array_equal(f1, f2)
    if kind == "hex8_nh":
        bad = u.copy()
        bad[:, 0] = -X[:, 0]  # F_xx = 0: det C = 0 everywhere
        with pytest.raises(fg.FgpuError, match="stress update failed"):
            g.stiffness_and_int_force(bad)
    g.close()


@pytest.mark.parametrize("kind", ["hex8_nh", "hex8_linear", "bsp"])
def test_launch_configurations_identical_bits(fg, oracle_mod, kind):
    """Every launch configuration the autotuner may pick (hex8 elements per
    workgroup, or one element per lane; assembly loads in batches of 1, 4, 8 or 16, with and without the
    XCD-contiguous workgroup order) gives the same global matrix, bit for bit,
    on neo-Hookean and linear bricks and on a cubic B-spline patch, and it
    matches the oracle to the FP32 element-matrix tolerance."""
    if kind == "bsp":
        m = M.bsp_block(ne=(4, 3, 3), p=3, perturb=0.05, seed=2)
        u = None
    elif kind == "hex8_linear":
        m = M.hex8_block(7, 6, 5, perturb=0.2, seed=8, E=2.0, nu=0.35)
        u = None
    else:
        n = 6
        m = M.hex8_block(n, n, n, form="tl_neohooke", perturb=0.1, seed=3,
                         mat=np.tile((1.0, 1.5), (n ** 3, 1)))
        X = m.coords
        u = np.stack([0.02 * X[:, 0] * X[:, 2], 0.01 * X[:, 1] ** 2, -0.03 * X[:, 0] ** 2], 1)
        u[m.fixed] = 0
    g = fg.GpuModel(m)
    g.stiffness(u)
    assert g.tuning_report()  # the autotuner ran on the first assembly
    ref = g.csr()[2].copy()
    seen = 0
    for te in ([0, 1, 4, 8, 16, 32, 64] if kind != "bsp" else [0]):
        for mode in range(9):
            try:
                g.set_launch(te, mode)
            except fg.FgpuError:
                continue
            g.stiffness(u)
            val = g.csr()[2]
            assert np.array_equal(val, ref), (te, mode)
            seen += 1
    assert seen >= 2
    om = oracle_mod.OracleMesh(m)
    Ko, ok = om.assemble_K(u)
    assert nrel(ref, Ko) <= 1e-6


@pytest.mark.parametrize("kind", ["cubic_renumbered", "p332_svk", "quadratic_f64"])
def test_row_assembly_identical_bits(fg, kind):
    """The row assembly (one wave per node row, mode 8) gives the global
    matrix of the per-block kernel bit for bit: on a mesh whose control
    points are renumbered at random (so that element blocks are read
    transposed as well), on the cubic-cubic-quadratic total-Lagrangian
    kernel, and on FP64 element matrices (degree 2)."""
    import dataclasses
    if kind == "cubic_renumbered":
        m = M.bsp_block(ne=(4, 3, 3), p=3, perturb=0.05, seed=12)
        perm = np.random.default_rng(1).permutation(m.nnode).astype(np.int32)
        coords = np.empty_like(m.coords)
        coords[perm] = m.coords
        fixed = np.empty_like(m.fixed)
        fixed[perm] = m.fixed
        cpw = None
        if m.cpw is not None:
            cpw = np.empty_like(m.cpw)
            cpw[perm] = m.cpw
        m = dataclasses.replace(m, coords=coords, conn=perm[m.conn], fixed=fixed, cpw=cpw,
                                dofmap=None, fext=None)
        u = None
    elif kind == "p332_svk":
        m = M.bsp_block(ne=(3, 3, 3), p=(3, 3, 2), perturb=0.05, seed=13, form="tl_svk",
                        mat=np.tile((2.6, 0.3), (27, 1)))
        X = m.coords
        u = np.stack([0.02 * X[:, 0] * X[:, 2], 0.01 * X[:, 1] ** 2, -0.03 * X[:, 0] ** 2], 1)
        u[m.fixed] = 0
    else:
        m = M.bsp_block(ne=(4, 3, 3), p=2, perturb=0.05, seed=14)
        u = None
    g = fg.GpuModel(m)
    g.set_launch(0, 2)
    g.stiffness(u)
    ref = g.csr()[2].copy()
    g.set_launch(0, 8)
    g.stiffness(u)
    assert np.array_equal(g.csr()[2], ref)
    g.close()


@pytest.mark.parametrize("form,matp,deg", [("tl_svk", (2.6, 0.3), 3),
                                           ("tl_neohooke", (1.0, 1.5), (3, 3, 2))])
def test_newton_bspline_vs_oracle(fg, oracle_mod, form, matp, deg):
    """Load-control Newton on a B-spline solid cantilever (total Lagrangian)
    against the oracle's restatement of ctrlnr.cpp on the same model: same
    converged states per step (the device tangent is FP32, the residual FP64,
    so the converged states agree to the Newton tolerance)."""
    m = M.bsp_block(ne=(4, 2, 2), p=deg, size=(2.0, 0.5, 0.5), form=form, perturb=0.05, seed=3,
                    mat=np.tile(matp, (16, 1)))
    tip = np.isclose(m.coords[:, 0], 2.0)
    fn = np.zeros((m.nnode, 3))
    fn[tip, 2] = -0.002 / tip.sum()
    q = np.zeros(m.neq)
    dm = m.dofmap
    q[dm[tip, 2] - 1] = fn[tip, 2]
    om = oracle_mod.OracleMesh(m)
    u_ref, it_ref = oracle_mod.newton_raphson(om, q, 3, 1.0, 30, 1e-10)
    assert len(u_ref) == 3
    g = fg.GpuModel(m)
    us, st, err = g.newton(q, steps=3, factor=1.0, maxit=30, tol=1e-10, pcg_tol=1e-12)
    assert err == "" and len(us) == 3, err
    for s in range(3):
        assert nrel(us[s], u_ref[s]) <= 1e-8, (s, nrel(us[s], u_ref[s]))
        assert st[s]["converged"]
    print("Newton iterations oracle", it_ref, "device", [x["iters"] for x in st])
