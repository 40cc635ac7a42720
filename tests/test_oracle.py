"""The CPU oracle pinned against the reference's outputs and against the
mathematics (closed forms, invariants, brute force).  No GPU needed."""
import os
import numpy as np
import pytest

from conftest import load_golden, mesh_from_golden
from fast_amd import mesh as M


# ---------------------------------------------------------------- quadrature

@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_exactness(oracle_mod, n):
    r, w = oracle_mod.gauss(n)
    for k in range(2 * n):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        # table values carry 15 decimals, so exactness holds to ~1e-14
        assert abs(np.sum(w * r ** k) - exact) < 5e-14


def test_gauss_table_values(oracle_mod):
    # decimal table values printed in the reference's Gauss table
    r, w = oracle_mod.gauss(2)
    assert r[1] == 0.577350269189626 and w[0] == 1.0
    r, w = oracle_mod.gauss(4)
    assert r[0] == -0.861136311594053 and w[1] == 0.652145154862546


# ---------------------------------------------------------------- shapes

def test_hex8_shape_invariants(oracle_mod):
    rng = np.random.default_rng(0)
    corners = np.array([[-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                        [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]])
    for a, c in enumerate(corners):
        N, _ = oracle_mod.hex8_shape(*c)
        assert np.allclose(N, np.eye(8)[a])
    for _ in range(10):
        p = rng.uniform(-1, 1, 3)
        N, dN = oracle_mod.hex8_shape(*p)
        assert abs(N.sum() - 1) < 1e-15
        assert np.abs(dN.sum(axis=0)).max() < 1e-15
        h = 1e-6
        for d in range(3):
            pp, pm = p.copy(), p.copy()
            pp[d] += h
            pm[d] -= h
            fd = (oracle_mod.hex8_shape(*pp)[0] - oracle_mod.hex8_shape(*pm)[0]) / (2 * h)
            assert np.allclose(fd, dN[:, d], atol=1e-9)


def test_bezier_extraction_matches_cox_de_boor(oracle_mod):
    from scipy.interpolate import BSpline
    rng = np.random.default_rng(1)
    for p, ns in ((2, 5), (3, 7), (4, 3)):
        U = M.open_uniform_knots(p, ns)
        Cx = oracle_mod.bezier_extraction(p, U)
        assert np.array_equal(Cx, M.bezier_extraction(p, U))
        nb = len(U) - p - 1
        for s in range(ns):
            a, b = s / ns, (s + 1) / ns
            for _ in range(4):
                t = rng.uniform(-1, 1)
                x = a + (t + 1) / 2 * (b - a)
                Bern = oracle_mod.bernstein(p, (t + 1) / 2)
                N = Cx[s] @ Bern
                ref = np.array([BSpline(U, np.eye(nb)[i], p)(x) for i in range(s, s + p + 1)])
                assert np.allclose(N, ref, atol=1e-13)
    # a non-uniform vector with a repeated interior knot
    U = np.array([0, 0, 0, 0, 0.2, 0.5, 0.5, 0.7, 1, 1, 1, 1.0])
    Cx = oracle_mod.bezier_extraction(3, U)
    assert np.allclose(Cx, M.bezier_extraction(3, U), atol=1e-15)
    spans = [(0, 0.2), (0.2, 0.5), (0.5, 0.7), (0.7, 1.0)]
    first = [0, 1, 2, 4]  # first basis of each span (knot index - p)
    first = [k - 3 for k in (3, 4, 6, 7)]
    nb = len(U) - 4
    for s, (a, b) in enumerate(spans):
        t = 0.3
        x = a + (t + 1) / 2 * (b - a)
        N = Cx[s] @ oracle_mod.bernstein(3, (t + 1) / 2)
        ref = np.array([BSpline(U, np.eye(nb)[i], 3)(x) for i in range(first[s], first[s] + 4)])
        assert np.allclose(N, ref, atol=1e-13)


# ---------------------------------------------------------------- elements

@pytest.mark.parametrize("form", ["linear", "tl_svk"])
def test_hex8_elements_bitexact_vs_reference(oracle_mod, form):
    d = load_golden("elem_hex8_%s.npz" % form)
    m = mesh_from_golden(d)
    om = oracle_mod.OracleMesh(m)
    assert np.array_equal(m.dofmap, d["dofs"])
    for e in range(m.nelem):
        K, ok = om.elem_stiff(e, d["u"] if form != "linear" else None)
        assert ok
        np.testing.assert_array_equal(K, d["K"][e])
        g, ok = om.elem_intforce(e, d["u"])
        np.testing.assert_array_equal(g, d["g"][e])
        assert np.array_equal(m.dofmap[m.conn[e]].ravel(), d["edof"][e])


@pytest.mark.parametrize("form", ["linear", "tl_svk"])
def test_bsp_elements_vs_reference(oracle_mod, form):
    d = load_golden("elem_bsp3_%s.npz" % form)
    m = mesh_from_golden(d)
    om = oracle_mod.OracleMesh(m)
    assert np.array_equal(m.dofmap, d["dofs"])
    u = d["u"] if form != "linear" else None
    for e in range(m.nelem):
        K, ok = om.elem_stiff(e, u)
        if e == 0:
            np.testing.assert_array_equal(K, d["K0"])
        np.testing.assert_array_equal(np.diag(K), d["Kdiag"][e])
        np.testing.assert_allclose(K @ d["v"], d["Kv"][e], rtol=1e-12,
                                   atol=1e-12 * np.abs(K).max())
        g, _ = om.elem_intforce(e, d["u"])
        np.testing.assert_array_equal(g, d["g"][e])


def test_element_invariants(oracle_mod):
    """Symmetry, rigid-body null space, patch test, scaling -- for hex8 and
    cubic B-spline elements (closed-form properties of B^T C B)."""
    for m in (M.hex8_block(2, 2, 2, E=3.0, nu=0.25, perturb=0.3, seed=3),
              M.bsp_block(ne=(2, 2, 2), p=3, E=3.0, nu=0.25, perturb=0.2, seed=3)):
        om = oracle_mod.OracleMesh(m)
        for e in (0, m.nelem - 1):
            K, _ = om.elem_stiff(e)
            X = m.coords[m.conn[e]]
            assert np.abs(K - K.T).max() <= 1e-13 * np.abs(K).max()
            R = []
            for c in range(3):
                t = np.zeros_like(X)
                t[:, c] = 1
                R.append(t.ravel())
            for (i, j) in ((0, 1), (1, 2), (2, 0)):
                r = np.zeros_like(X)
                r[:, i] = -X[:, j]
                r[:, j] = X[:, i]
                R.append(r.ravel())
            R = np.array(R).T
            assert np.abs(K @ R).max() <= 1e-12 * np.abs(K).max() * np.abs(R).max()
            ev = np.linalg.eigvalsh(K)
            assert np.sum(ev < 1e-9 * ev.max()) == 6
            # constant-strain patch: f_int = K u for linear forms, and the
            # strain energy of a uniform strain field equals V * e^T C e
            eps = np.array([1e-3, -2e-3, 5e-4, 1e-3, 0, 2e-4])
            Lg = np.array([[eps[0], eps[3] / 2, eps[4] / 2],
                           [eps[3] / 2, eps[1], eps[5] / 2],
                           [eps[4] / 2, eps[5] / 2, eps[2]]])
            u = (X @ Lg.T).ravel()
            un = np.zeros((m.nnode, 3))
            un[m.conn[e]] = u.reshape(-1, 3)
            g, _ = om.elem_intforce(e, un)
            assert np.allclose(g, K @ u, rtol=1e-10, atol=1e-14)


def test_neohooke_tangent_is_derivative_of_fint(oracle_mod):
    """No reference oracle exists for neo-Hookean: pin K_t = d f_int / d u
    by central differences, and the small-strain limit against the
    reference's linear model (mu=1, lambda=1.5 <=> E=2.6, nu=0.3)."""
    m = M.hex8_block(1, 1, 1, form="tl_neohooke", perturb=0.0,
                     mat=np.array([[1.0, 1.5]]))
    m.coords[:] += np.random.default_rng(2).uniform(-0.05, 0.05, m.coords.shape)
    om = oracle_mod.OracleMesh(m)
    rng = np.random.default_rng(4)
    u = rng.uniform(-0.08, 0.08, (m.nnode, 3))
    K, ok = om.elem_stiff(0, u)
    assert ok
    h = 1e-6
    Kfd = np.zeros_like(K)
    for j in range(24):
        up, um = u.copy(), u.copy()
        node = m.conn[0][j // 3]  # element-local dof j -> node, direction
        up[node, j % 3] += h
        um[node, j % 3] -= h
        Kfd[:, j] = (om.elem_intforce(0, up)[0] - om.elem_intforce(0, um)[0]) / (2 * h)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()
    # small strain limit
    lin = M.hex8_block(1, 1, 1, E=2.6, nu=0.3)
    lin.coords[:] = m.coords
    K0, _ = oracle_mod.OracleMesh(lin).elem_stiff(0)
    Kn, _ = om.elem_stiff(0, np.zeros((m.nnode, 3)))
    assert np.abs(K0 - Kn).max() < 1e-12 * np.abs(K0).max()


def test_tl_svk_tangent_is_derivative_of_fint(oracle_mod):
    m = M.bsp_block(ne=(1, 1, 1), p=2, E=2.6, nu=0.3, form="tl_svk", perturb=0.0)
    om = oracle_mod.OracleMesh(m)
    u = np.random.default_rng(5).uniform(-0.05, 0.05, (m.nnode, 3))
    K, _ = om.elem_stiff(0, u)
    h = 1e-6
    nd = K.shape[0]
    Kfd = np.zeros_like(K)
    for j in range(nd):
        up, um = u.copy(), u.copy()
        node = m.conn[0][j // 3]
        up[node, j % 3] += h
        um[node, j % 3] -= h
        Kfd[:, j] = (om.elem_intforce(0, up)[0] - om.elem_intforce(0, um)[0]) / (2 * h)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()


# ---------------------------------------------------------------- assembly

def test_assembly_equals_dense_brute_force(oracle_mod):
    m = M.hex8_block(3, 2, 2, perturb=0.2, seed=9)
    om = oracle_mod.OracleMesh(m)
    val, ok = om.assemble_K()
    Kd = np.zeros((m.neq, m.neq))
    for e in range(m.nelem):
        Ke, _ = om.elem_stiff(e)
        dofs = m.dofmap[m.conn[e]].ravel()
        for i in range(24):
            for j in range(24):
                if dofs[i] and dofs[j]:
                    Kd[dofs[i] - 1, dofs[j] - 1] += Ke[i, j]
    A = om.to_scipy(val).toarray()
    assert np.allclose(A, Kd, rtol=0, atol=1e-12 * np.abs(Kd).max())
    assert np.array_equal(A != 0, Kd != 0) or np.all((A != 0) >= (Kd != 0))


# ---------------------------------------------------------------- solvers

def test_linear_static_cfg1_vs_reference(oracle_mod):
    d = load_golden("cfg1_linear_u.npz")
    m = mesh_from_golden(d)
    assert m.neq == 3630
    om = oracle_mod.OracleMesh(m)
    u = oracle_mod.linear_static(om, m.load_vector())
    un = m.scatter(u)
    np.testing.assert_allclose(un, d["u"], rtol=1e-12, atol=1e-12 * np.abs(d["u"]).max())


def test_bsp_linear_static_vs_reference(oracle_mod):
    d = load_golden("bsp3_linear_u.npz")
    m = mesh_from_golden(d)
    om = oracle_mod.OracleMesh(m)
    u = oracle_mod.linear_static(om, m.load_vector())
    np.testing.assert_allclose(m.scatter(u), d["u"], rtol=1e-11,
                               atol=1e-11 * np.abs(d["u"]).max())


def test_newton_raphson_tl_vs_reference(oracle_mod):
    d = load_golden("nr_tl_hex8_4x4x4.npz")
    m = mesh_from_golden(d)
    om = oracle_mod.OracleMesh(m)
    us, its = oracle_mod.newton_raphson(om, m.load_vector(), 5, 0.2, 30, 1e-9)
    assert len(us) == 5
    for s in range(5):
        np.testing.assert_allclose(m.scatter(us[s]), d["u_steps"][s], rtol=1e-11,
                                   atol=1e-11 * np.abs(d["u_steps"]).max())
    # the reference log prints one "Iter:" line per Newton iteration
    ref_its = []
    for k in d["iters"]:
        if k == 1:
            ref_its.append(1)
        else:
            ref_its[-1] = int(k)
    assert its == ref_its


def test_pcg_matches_textbook_and_direct(oracle_mod):
    m = M.hex8_block(4, 3, 3, perturb=0.2, seed=2, traction=(0, 0, -1.0))
    om = oracle_mod.OracleMesh(m)
    val, _ = om.assemble_K()
    rowptr, col = om.pattern()
    f = m.load_vector()
    u, it, rr, conv = oracle_mod.pcg(rowptr, col, val, f, 5000, 1e-10)
    assert conv
    A = om.to_scipy(val)
    import scipy.sparse.linalg as sla
    ud = sla.spsolve(A.tocsc(), f)
    assert np.linalg.norm(u - ud) <= 1e-8 * np.linalg.norm(ud)
    # textbook Jacobi-PCG, x0 = 0, same stopping rule -> same iteration count
    d = A.diagonal()
    x = np.zeros_like(f)
    r = f.copy()
    z = r / d
    p = z.copy()
    rz = z @ r
    k = 0
    for k in range(1, 5001):
        w = A @ p
        a = rz / (p @ w)
        x += a * p
        r -= a * w
        z = r / d
        if np.linalg.norm(r) / np.linalg.norm(f) < 1e-10:
            break
        rz1 = z @ r
        p = z + (rz1 / rz) * p
        rz = rz1
    assert k == it
    us, ok = oracle_mod.skyline_solve(rowptr, col, val, f)
    assert ok and np.linalg.norm(us - ud) <= 1e-10 * np.linalg.norm(ud)


def test_bsp_mixed_degree_element_invariants(oracle_mod):
    """Degree (3, 3, 2) B-spline solid (the reference's Scordelis-Lo roof
    uses it): symmetric K_e, six rigid-body modes, constant-strain patch."""
    m = M.bsp_block(ne=(2, 2, 1), p=(3, 3, 2), E=2.0, nu=0.3, perturb=0.1, seed=7)
    assert m.nn == 48
    om = oracle_mod.OracleMesh(m)
    X = m.coords
    for e in range(m.nelem):
        K, ok = om.elem_stiff(e)
        assert ok and np.allclose(K, K.T, rtol=0, atol=1e-12 * np.abs(K).max())
        Xe = X[m.conn[e]]
        R = []
        for d in range(3):
            t = np.zeros_like(Xe)
            t[:, d] = 1.0
            R.append(t.ravel())
        for a, b in ((0, 1), (0, 2), (1, 2)):
            r = np.zeros_like(Xe)
            r[:, a], r[:, b] = -Xe[:, b], Xe[:, a]
            R.append(r.ravel())
        for r in R:
            assert np.linalg.norm(K @ r) <= 1e-9 * np.abs(K).max() * np.linalg.norm(r)
    # linear field u = G x is reproduced: f_int equals the boundary traction
    # of the constant stress, so interior control points see zero force
    G = np.array([[1e-3, 2e-4, 0.0], [0.0, -5e-4, 1e-4], [3e-4, 0.0, 2e-4]])
    u = X @ G.T
    g, ok = om.assemble_fint(u)
    nb = m.nbas
    i = np.arange(m.nnode) % nb[0]
    j = (np.arange(m.nnode) // nb[0]) % nb[1]
    k = np.arange(m.nnode) // (nb[0] * nb[1])
    inner = (i > 0) & (i < nb[0] - 1) & (j > 0) & (j < nb[1] - 1) & (k > 0) & (k < nb[2] - 1)
    fn = m.scatter(g)
    assert inner.sum() == 0 or np.abs(fn[inner]).max() <= 1e-12 * np.abs(fn).max()


def _stress_meshes(M, form, mat):
    return [M.hex8_block(3, 2, 2, size=(1.0, 0.6, 0.5), perturb=0.25, seed=3, form=form,
                         mat=np.tile(mat, (12, 1))),
            M.hex8_block(2, 2, 1, perturb=0.2, seed=4, form=form, nip=3, mat=np.tile(mat, (4, 1))),
            M.bsp_block(ne=(2, 2, 2), p=3, perturb=0.1, seed=2, form=form, mat=np.tile(mat, (8, 1))),
            M.bsp_block(ne=(2, 1, 2), p=(3, 3, 2), perturb=0.1, seed=5, form=form,
                        mat=np.tile(mat, (4, 1)))]


def test_stress_extrapolation_matrix(oracle_mod):
    """CompTRMatrix: a least-squares linear fit, so rows sum to one, linear
    fields at the points come out exactly at the nodes, and a single point
    along r gives the mean."""
    from fast_amd import mesh as M
    rng = np.random.default_rng(0)
    for m in _stress_meshes(M, "linear", (1.0, 0.3)):
        om = oracle_mod.OracleMesh(m)
        TR = om.stress_extrap()
        pts = oracle_mod.brick_points(m.nip)[0].reshape(-1, 3)
        assert TR.shape == (m.nn, len(pts))
        assert np.allclose(TR.sum(axis=1), 1.0, rtol=0, atol=1e-14)
        if m.shape == "hex8":
            nodes = np.array([[-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                              [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]], float)
        else:
            g = [np.linspace(-1, 1, p + 1) for p in m.deg]
            nodes = np.array([[r, s, t] for t in g[2] for s in g[1] for r in g[0]])
        c = rng.standard_normal(4)
        f = c[0] + pts @ c[1:]
        assert np.allclose(TR @ f, c[0] + nodes @ c[1:], rtol=0, atol=1e-13)
    m = M.hex8_block(2, 1, 1, nip=1)
    assert np.allclose(oracle_mod.OracleMesh(m).stress_extrap(), 1.0)


def _homogeneous(m, A):
    return (m.coords - m.coords.mean(axis=0)) @ A.T


@pytest.mark.parametrize("form,mat", [("linear", (2.0, 0.3)), ("tl_svk", (2.0, 0.3)),
                                      ("tl_neohooke", (1.0, 1.5))])
def test_stress_of_homogeneous_deformation(oracle_mod, form, mat):
    """u = (F - I) X reproduces F exactly on every element shape, so every
    point carries the closed-form stress: sigma = lambda tr(eps) I + 2 mu eps
    (linear), S = lambda tr(E) I + 2 mu E with E = (F^T F - I) / 2 (SVK),
    S = mu (I - C^-1) + lambda ln J C^-1 (neo-Hooke); the element-nodal and
    averaged values equal it too.  A rigid rotation leaves S = 0 for the
    total-Lagrangian forms."""
    from fast_amd import mesh as M
    rng = np.random.default_rng(1)
    H = 0.08 * rng.standard_normal((3, 3))
    F = np.eye(3) + H
    if form == "linear":
        E, nu = mat
        lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
        eps = 0.5 * (H + H.T)
        S = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    elif form == "tl_svk":
        E, nu = mat
        lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
        Eg = 0.5 * (F.T @ F - np.eye(3))
        S = lam * np.trace(Eg) * np.eye(3) + 2 * mu * Eg
    else:
        mu, lam = mat
        Cm = F.T @ F
        Ci = np.linalg.inv(Cm)
        S = mu * (np.eye(3) - Ci) + lam * 0.5 * np.log(np.linalg.det(Cm)) * Ci
    voigt = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    th = 0.4
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    for m in _stress_meshes(M, form, mat):
        om = oracle_mod.OracleMesh(m)
        si, sn, sa, ok = om.stresses(_homogeneous(m, H))
        assert ok
        assert np.abs(si - voigt).max() <= 1e-12 * np.abs(voigt).max()
        assert np.abs(sn - voigt).max() <= 1e-12 * np.abs(voigt).max()
        assert np.abs(sa - voigt).max() <= 1e-12 * np.abs(voigt).max()
        if form != "linear":
            si, _, _, ok = om.stresses(_homogeneous(m, Q - np.eye(3)))
            assert ok and np.abs(si).max() <= 1e-12


@pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_probe")),
    reason="the reference build (oracle/_ref) is not present")
def test_reference_cpu_baseline_runs():
    """bench.py's CPU baseline runs the reference's own assembly loops
    (ref_probe's timing mode) on a small sample of each step workload, the
    oracle on the same sample beside it, and reports their rates; the
    batch's sample is its first cubes, with their own edges and faces."""
    import bench
    res = bench.reference_baseline(M.bsp_block(ne=(2, 2, 2), p=3, perturb=0.05, seed=0),
                                   "a 2x2x2 patch")
    assert res["kind"] == "reference" and res["cores"] == 1
    assert res["value"] > 0 and "2x2x2 patch" in res["sample"]
    assert res["stiffness_s"] > 0 and res["int_force_s"] >= 0
    _, prm = M.cube_batch(8, nmin=4, nmax=6, seed=3)
    sample, what = bench.reference_sample("batch", prm, 200)
    assert sample.nelem >= 200 and sample.nelem == sum(int(n) ** 3 for n in
                                                       prm["n"][:len(sample.batch_nodeptr) - 1])
    res = bench.cpu_baselines("batch", prm, target=200)
    assert res["kind"] == "reference" and res["port"]["kind"] == "port"
    assert res["host"]["nproc"] >= 1
    assert bench.cpu_baselines("hex8", None) is None  # neo-Hookean: no reference model
