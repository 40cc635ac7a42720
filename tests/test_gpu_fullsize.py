"""Parity at the benchmark's full sizes: the cubic B-spline patch 32x32x8
(8,192 elements, 4x4x4 Gauss) and the batch of 1024 uneven hex8 cubes
(10.33 M elements, 33.6 M equations), each against the FP64 oracle.

Per element the FP32 element matrices are held to 1e-6 (normwise, also in
K_e u_e relative to |K_e|_2 |u_e|); the FP64 internal force to 1e-11 of the
oracle's; solutions through size-independent properties: the patch's
refined linear solve reproduces a chosen displacement through the oracle's
FP64 matrix, and each sampled cube of the batch, solved by the per-mesh PCG
on the FP32-element matrix to 1e-8, is within a normwise backward error of
1e-6 of its own FP64 system (the FP32 rounding, independent of the cube's
conditioning)."""
import os

import numpy as np
import pytest

from fast_amd import mesh as M

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def smooth(X):
    return 1e-3 * np.stack([np.sin(3 * X[:, 0]) * X[:, 2], np.cos(2 * X[:, 1]) * X[:, 0],
                            X[:, 0] * X[:, 1]], 1)


def element_checks(g, om, m, u, idx):
    worst_f = worst_ku = 0.0
    for e in idx:
        K0, ok = om.elem_stiff(int(e), None)
        assert ok
        Ke = g.element_matrices(int(e), 1)[0].astype(np.float64)
        ue = u[m.conn[e]].ravel()
        worst_f = max(worst_f, np.linalg.norm(Ke - K0) / np.linalg.norm(K0))
        worst_ku = max(worst_ku, np.linalg.norm((Ke - K0) @ ue) /
                       (np.linalg.norm(K0, 2) * np.linalg.norm(ue)))
    print("element matrices: worst |dK|_F/|K|_F %.2e, worst |dK u|/(|K| |u|) %.2e"
          % (worst_f, worst_ku))
    assert worst_f <= 1e-6 and worst_ku <= 1e-6


def test_cubic_patch_32x32x8(fg, oracle_mod):
    m = M.bsp_block(ne=(32, 32, 8), p=3, perturb=0.05, seed=0, E=1.0, nu=0.3)
    u = smooth(m.coords)
    u[m.fixed] = 0.0
    g = fg.GpuModel(m)
    om = oracle_mod.OracleMesh(m)
    gd = g.stiffness_and_int_force(u)
    element_checks(g, om, m, u, np.linspace(0, m.nelem - 1, 48).astype(int))
    g0, ok = om.assemble_fint(u, nthreads=THREADS)
    assert ok
    rel = np.linalg.norm(gd - g0) / np.linalg.norm(g0)
    print("f_int: %.2e" % rel)
    assert rel <= 1e-11
    # the refined solve reproduces v from f = K64 v (K64: the oracle's FP64
    # matrix, the system the device's FP64 residuals define)
    import scipy.sparse as sp
    rowptr, col = om.pattern()
    K64, ok = om.assemble_K(nthreads=THREADS)
    assert ok
    A = sp.csr_matrix((K64, col.astype(np.int64), rowptr), shape=(m.neq, m.neq))
    v = np.zeros((m.nnode, 3))
    v[:] = smooth(m.coords + 0.3)
    v = v[m.dofmap > 0]
    ve = np.zeros(m.neq)
    ve[m.dofmap[m.dofmap > 0] - 1] = v
    x, it, rr = g.linear_static(A @ ve, tol=1e-10)
    err = np.linalg.norm(x - ve) / np.linalg.norm(ve)
    print("linear static: %d PCG iterations, relres %.1e, |u - v|/|v| %.2e" % (it, rr, err))
    assert rr <= 1e-10 and err <= 1e-7
    g.close()


def test_batch_1024_cubes(fg, oracle_mod):
    m, prm = M.cube_batch(1024, nmin=4, nmax=32, seed=0)
    assert m.nelem == 10_330_882
    u = smooth(m.coords)
    u[m.fixed] = 0.0
    g = fg.GpuModel(m)
    gd = g.stiffness_and_int_force(u)
    om = oracle_mod.OracleMesh(m)
    element_checks(g, om, m, u, np.linspace(0, m.nelem - 1, 64).astype(int))
    # f_int of the first meshes (independent: their equations get nothing
    # from the others) against the oracle's assembly of those elements
    nptr = np.asarray(m.batch_nodeptr)
    mesh_of = np.searchsorted(nptr, m.conn[:, 0], side="right") - 1
    k = int(mesh_of[300_000])
    ne = int(np.searchsorted(mesh_of, k))
    neq = int(m.dofmap[:nptr[k]].max())
    g0, ok = om.assemble_fint(u, nthreads=THREADS, elem_range=(0, ne))
    assert ok
    rel = np.linalg.norm(gd[:neq] - g0[:neq]) / np.linalg.norm(g0[:neq])
    print("f_int of the first %d meshes (%d elements): %.2e" % (k, ne, rel))
    assert rel <= 1e-11
    # the whole batch solved by the per-mesh PCG; sampled cubes (the
    # smallest, a middle one, the largest) checked against their own FP64
    # systems assembled by the oracle
    f = m.load_vector()
    x, its, rr, nconv = g.solve_batch(f, tol=1e-8)
    assert nconv == 1024
    xn = m.scatter(x)
    ns = np.asarray(prm["n"])
    order = np.argsort(ns, kind="stable")
    import scipy.sparse as sp
    for kk in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        mk = M.hex8_cube(int(ns[kk]), nu=float(prm["nu"][kk]), clamp_face=int(prm["clamp_face"][kk]),
                         load_face=int(prm["load_face"][kk]))
        omk = oracle_mod.OracleMesh(mk)
        rp, cl = omk.pattern()
        Kk, ok = omk.assemble_K(nthreads=THREADS)
        assert ok
        A = sp.csr_matrix((Kk, cl.astype(np.int64), rp), shape=(mk.neq, mk.neq))
        xk = np.zeros(mk.neq)
        sub = xn[nptr[kk]:nptr[kk + 1]]
        live = mk.dofmap > 0
        xk[mk.dofmap[live] - 1] = sub[live]
        fk = mk.load_vector()
        r = fk - A @ xk
        res = np.linalg.norm(r) / np.linalg.norm(fk)
        # normwise backward error (infinity norms): x solves a system within
        # this relative distance of the FP64 one -- the FP32 element matrices'
        # rounding, not the conditioning
        anorm = abs(A).sum(axis=1).max()
        eta = np.abs(r).max() / (anorm * np.abs(xk).max() + np.abs(fk).max())
        print("cube %d (n = %d): %d PCG iterations, FP64 residual %.2e, backward error %.2e"
              % (kk, ns[kk], its[kk], res, eta))
        assert eta <= 1e-6 and rr[kk] <= 1e-8
    g.close()
