"""fgpu_int_force_dev_noalloc: the internal force with its element vectors in
the library's fixed device scratch instead of the context's buffer.  Called
through the C-ABI on two contexts in a row, the second mesh smaller than the
first (both share the device's scratch), it gives fgpu_int_force_dev's
result bit for bit and the oracle's to 1e-11; meshes whose element vectors
exceed the scratch, and shapes outside the tuned kernels, are refused."""
import ctypes as C

import numpy as np
import pytest

from fast_amd import mesh as M

pytestmark = pytest.mark.gpu


def _field(m, scale):
    X = m.coords
    u = scale * np.stack([X[:, 0] * X[:, 2], X[:, 1] ** 2, -X[:, 0] ** 2], 1)
    u[m.fixed] = 0.0
    return u


def _meshes(kind):
    if kind == "hex8_nh":
        mk = lambda n: M.hex8_block(n, n - 1, n - 2, form="tl_neohooke", perturb=0.2, seed=n,
                                    mat=np.tile((1.0, 1.5), (n * (n - 1) * (n - 2), 1)))
        return mk(7), mk(4), 0.02
    if kind == "hex8_linear":
        return M.hex8_block(9, 8, 7, perturb=0.2, seed=1), M.hex8_block(3, 2, 2), 1e-3
    return (M.bsp_block(ne=(4, 3, 3), p=3, perturb=0.05, seed=2),
            M.bsp_block(ne=(2, 2, 1), p=(3, 3, 2), perturb=0.05, seed=3), 1e-3)


@pytest.mark.parametrize("kind", ["hex8_nh", "hex8_linear", "bsp"])
def test_noalloc_int_force_matches_entry_point(kind, oracle_mod):
    import torch
    from fast_amd import fgpu
    L = fgpu.lib()
    big, small, scale = _meshes(kind)
    assert small.nelem < big.nelem
    for m in (big, small):  # two calls in a row, the second smaller
        g = fgpu.GpuModel(m)
        u = _field(m, scale)
        du = torch.from_numpy(np.ascontiguousarray(u)).cuda()
        f_ref = torch.full((g.neq,), np.nan, dtype=torch.float64, device="cuda")
        f_new = torch.full((g.neq,), np.nan, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert L.fgpu_int_force_dev(g.h, du.data_ptr(), f_ref.data_ptr()) == 1
        assert L.fgpu_int_force_dev_noalloc(g.h, du.data_ptr(), f_new.data_ptr()) == 1, \
            L.fgpu_last_error(g.h).decode()
        a, b = f_ref.cpu().numpy(), f_new.cpu().numpy()
        assert np.array_equal(a, b)
        f0, ok = oracle_mod.OracleMesh(m).assemble_fint(u)
        assert ok and np.linalg.norm(b - f0) <= 1e-11 * np.linalg.norm(f0)
        g.close()


def test_noalloc_int_force_refuses_what_does_not_fit():
    import torch
    from fast_amd import fgpu
    L = fgpu.lib()
    # 72 x 70 x 70 bricks: 352,800 element vectors of 192 B exceed 64 MiB
    m = M.hex8_block(72, 70, 70)
    g = fgpu.GpuModel(m)
    du = torch.zeros(3 * m.nnode, dtype=torch.float64, device="cuda")
    f = torch.zeros(g.neq, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert L.fgpu_int_force_dev_noalloc(g.h, du.data_ptr(), f.data_ptr()) == 0
    assert "fixed scratch" in L.fgpu_last_error(g.h).decode()
    g.close()
    # a quadratic B-spline solid runs the FP64 kernels, whose point scratch
    # is allocated: refused
    q = M.bsp_block(ne=(2, 2, 2), p=2, perturb=0.0)
    gq = fgpu.GpuModel(q)
    du = torch.zeros(3 * q.nnode, dtype=torch.float64, device="cuda")
    f = torch.zeros(gq.neq, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert L.fgpu_int_force_dev_noalloc(gq.h, du.data_ptr(), f.data_ptr()) == 0
    assert "only" in L.fgpu_last_error(gq.h).decode()
    gq.close()
