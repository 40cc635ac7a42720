"""Device direct solve (fgpu_factor / fgpu_factor_solve / fgpu_direct_solve,
LDL^T in band storage) against the oracle's restatement of the reference's
skyline Crout solver on the same assembled matrix, the inertia of a shifted
matrix, singular-matrix detection, and the linear-static / Newton drivers
with the direct solver against the reference program's deck solutions.

Both factorizations are FP64 on the same matrix; the solutions agree to
1e-10 of the largest entry (orderings of the sums differ).  The drivers
refine with FP64 residuals, so they meet the reference to the tolerance of
the PCG drivers (1e-9)."""
import os

import numpy as np
import pytest

from conftest import load_golden
from fast_amd import mesh as M
from test_neutral import DECKS, assert_same_displacements

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def half_band(rowptr, col):
    bw = 0
    for i in range(len(rowptr) - 1):
        c = col[rowptr[i]:rowptr[i + 1]]
        bw = max(bw, i - c.min())
    return bw


CASES = {
    "hex8_10cubed": lambda: M.hex8_block(10, 10, 10, perturb=0.2, seed=1, traction=(0, 0, -1.0)),
    "hex8_tiny": lambda: M.hex8_block(2, 1, 1, perturb=0.1, seed=2, traction=(0, 0, -1.0)),
    "hex8_slab": lambda: M.hex8_block(3, 7, 2, perturb=0.2, seed=3, traction=(0, 0, -1.0)),
    "bsp": lambda: M.bsp_block(ne=(6, 3, 2), p=3, perturb=0.1, seed=4),
    "bsp332": lambda: M.bsp_block(ne=(4, 3, 2), p=(3, 3, 2), perturb=0.1, seed=5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_direct_solve_vs_oracle_skyline(fg, oracle_mod, case):
    m = CASES[case]()
    g = fg.GpuModel(m)
    g.stiffness()
    rowptr, col, val = g.csr()
    rng = np.random.default_rng(0)
    b = rng.standard_normal(m.neq)
    x = g.direct_solve(b)
    x0, ok = oracle_mod.skyline_solve(rowptr, col, val, b)
    assert ok
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max(), case
    bw, neg, nbytes = g.factor_info()
    assert bw == half_band(rowptr, col) and neg == 0
    assert nbytes == m.neq * (bw + 1) * 8
    # the factors serve further right-hand sides until the matrix changes
    b2 = rng.standard_normal(m.neq)
    x2 = g.factor_solve(b2)
    x20, _ = oracle_mod.skyline_solve(rowptr, col, val, b2)
    assert np.abs(x2 - x20).max() <= 1e-10 * np.abs(x20).max()
    g.stiffness()
    with pytest.raises(fg.FgpuError, match="no factorization"):
        g.factor_solve(b2)
    g.close()


def test_negative_pivots_count_the_eigenvalues_below_a_shift(fg):
    """Sylvester's law of inertia: the LDL^T of K - sigma I has as many
    negative pivots as K has eigenvalues below sigma (the reference reads
    this count with cSymSkylMatrix::GetNgtPvt)."""
    m = M.hex8_block(3, 2, 2, perturb=0.2, seed=6)
    g = fg.GpuModel(m)
    g.stiffness()
    rowptr, col, val = g.csr()
    K = np.zeros((m.neq, m.neq))
    for i in range(m.neq):
        K[i, col[rowptr[i]:rowptr[i + 1]]] = val[rowptr[i]:rowptr[i + 1]]
    ev = np.linalg.eigvalsh(K)
    sigma = 0.5 * (ev[6] + ev[7])
    g.add_diag(np.arange(1, m.neq + 1), np.full(m.neq, -sigma))
    g.factor()
    assert g.factor_info()[1] == 7
    g.close()


def test_singular_matrix_is_reported(fg, oracle_mod):
    """A free-floating block keeps its six rigid-body modes: with the FP64
    stiffness of the oracle loaded into the context, a pivot falls below
    1e-12 and the solve reports a singular matrix (CroutSolver's test).  (The
    matrix assembled from FP32 element matrices keeps those modes only to
    FP32 accuracy, so its pivots stay above the absolute threshold.)"""
    m = M.hex8_block(3, 2, 2, clamp_x0=False)
    g = fg.GpuModel(m)
    g.stiffness()
    om = oracle_mod.OracleMesh(m)
    rowptr, col, _ = g.csr()
    r0, c0 = om.pattern()
    assert np.array_equal(rowptr, r0) and np.array_equal(col, c0)
    K64, ok = om.assemble_K()
    g.set_csr_values(K64)
    with pytest.raises(fg.FgpuError, match="singular matrix"):
        g.factor()
    g.close()


@pytest.mark.parametrize("name", ["ScordelisLoDeg3", "BeamB8"])
def test_linear_static_direct_vs_reference_run(fg, name):
    from fast_amd import deck, neutral
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    g = fg.GpuModel(d.mesh)
    deck.prepare(d, g)
    f = deck.load_vector(d, g)
    # (the FP64 residual of the Scordelis-Lo shell bottoms out near 1.6e-10)
    u, k, rr = g.linear_static_direct(f, tol=1e-9)
    assert rr <= 1e-9 and k <= 30, (k, rr)
    print(name, "refinements", k, "relres", rr)
    ref = load_golden("deck_%s.npz" % name)
    assert_same_displacements(d, d.mesh.scatter(u), ref["u"], name, tol=1e-9)
    g.close()


@pytest.mark.parametrize("name", ["BeamB8TL", "ScordelisLoDeg3TL"])
def test_newton_direct_vs_reference_run(fg, oracle_mod, name):
    from fast_amd import deck, neutral
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    g = fg.GpuModel(d.mesh)
    deck.prepare(d, g)
    f = deck.load_vector(d, g)
    us, st, err = g.newton(f, steps=d.steps, factor=d.factor, maxit=d.maxit, tol=d.tol,
                           direct=True)
    assert err == "" and all(s["converged"] for s in st), err
    ref = load_golden("deck_%s.npz" % name)
    assert_same_displacements(d, d.mesh.scatter(us[-1]), ref["u"], name, tol=1e-8)
    print(name, "Newton iterations per step:", [s["iters"] for s in st])
    g.close()
