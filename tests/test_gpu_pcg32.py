"""PCG with the FP32 copy of the assembled matrix (fgpu_set_matrix_precision
32): the plain solve solves the rounded matrix (agreement with the FP64-matrix
solve to the FP32 rounding times the condition number, 1e-4 here), and the
drivers, whose residuals are FP64, reach the FP64-matrix results to their
tolerances.  Switching back re-captures the iteration graph: FP64 results
are then bit-identical to a context that never switched."""
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


def cube(n=8):
    return M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                        perturb=0.1, seed=2, traction=(0.0, 0.0, -0.2))


def nrel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_plain_solve(fg):
    m = cube()
    q = m.load_vector()
    g = fg.GpuModel(m)
    g.stiffness()
    x64, it64, rr64, ok64 = g.solve(q, tol=1e-10, maxit=20000)
    g.set_matrix_precision(32)
    x32, it32, rr32, ok32 = g.solve(q, tol=1e-10, maxit=20000)
    assert ok64 and ok32 and rr32 <= 1e-10
    assert nrel(x32, x64) <= 1e-4
    assert abs(it32 - it64) <= 0.1 * it64
    g.set_matrix_precision(64)
    x64b, it64b, _, _ = g.solve(q, tol=1e-10, maxit=20000)
    assert np.array_equal(x64b, x64) and it64b == it64
    g.close()


def test_drivers(fg):
    m = cube()
    q = m.load_vector()
    kw = dict(steps=3, factor=1.0 / 3, maxit=30, tol=1e-10, pcg_tol=1.0, adaptive_pcg=2)
    g = fg.GpuModel(m)
    u64, s64, e64 = g.newton(q, **kw)
    g.set_matrix_precision(32)
    u32, s32, e32 = g.newton(q, **kw)
    assert e64 == "" and e32 == ""
    assert nrel(u32[-1], u64[-1]) <= 1e-8
    lin = M.hex8_block(8, 8, 8, perturb=0.1, seed=2, traction=(0.0, 0.0, -0.2))
    gl = fg.GpuModel(lin)
    ql = lin.load_vector()
    l64, _, r64 = gl.linear_static(ql, tol=1e-11)
    gl.set_matrix_precision(32)
    l32, _, r32 = gl.linear_static(ql, tol=1e-11)
    assert r32 <= 1e-11 and nrel(l32, l64) <= 1e-9
    gl.close()
    print("Newton iterations FP64 matrix", [s["iters"] for s in s64], "FP32 copy",
          [s["iters"] for s in s32], "PCG", sum(s["pcg_iters"] for s in s64),
          sum(s["pcg_iters"] for s in s32))
    g.close()
