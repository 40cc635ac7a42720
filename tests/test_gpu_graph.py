"""Newton iterations replayed from captured hipGraphs (fgpu_nr_params.graph)
against the eager launches of the same driver: identical bits in every
step's displacement, identical Newton and PCG iteration counts.

The graph path runs the eager path's kernels in the eager path's order (the
first iteration of a call is eager and does the lazy set-up); it differs only
in how the host waits: one read per iteration and per PCG chunk."""
import os

import numpy as np
import pytest

from fast_amd import mesh as M
from test_neutral import DECKS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def both(g, q, **kw):
    out = []
    for graph in (False, True, True):  # the second graph run replays a fresh capture
        us, st, err = g.newton(q, graph=graph, **kw)
        assert err == "" and all(s["converged"] for s in st), err
        out.append((us, st))
    (u0, s0) = out[0]
    for us, st in out[1:]:
        assert np.array_equal(us, u0)
        assert [s["iters"] for s in st] == [s["iters"] for s in s0]
        assert [s["pcg_iters"] for s in st] == [s["pcg_iters"] for s in s0]
        assert [s["err"] for s in st] == [s["err"] for s in s0]
    return s0


@pytest.mark.parametrize("form", ["tl_svk", "tl_neohooke"])
def test_hex8_pcg(fg, form):
    m = M.hex8_block(6, 3, 3, size=(2.0, 1.0, 1.0), perturb=0.15, seed=4, form=form,
                     traction=(0.0, 0.0, -0.004))
    g = fg.GpuModel(m)
    st = both(g, m.load_vector(), steps=3, factor=1.0, maxit=30, tol=1e-9, pcg_tol=1e-10)
    assert sum(s["iters"] for s in st) >= 6  # the captured path did run
    g.close()


def test_hex8_pcg_fp32_matrix_adaptive(fg):
    m = M.hex8_block(5, 4, 3, perturb=0.1, seed=7, form="tl_svk", traction=(0.0, 0.002, -0.004))
    g = fg.GpuModel(m)
    g.set_matrix_precision(32)
    both(g, m.load_vector(), steps=2, factor=1.0, maxit=30, tol=1e-10, pcg_tol=1.0,
         adaptive_pcg=2)
    g.close()


@pytest.mark.parametrize("direct", [False, True])
def test_bspline_tl(fg, direct):
    m = M.bsp_block(ne=(4, 2, 2), p=3, form="tl_svk", perturb=0.05, seed=3)
    q = np.zeros(m.neq)
    tip = np.isclose(m.coords[:, 0], 1.0)
    q[m.dofmap[tip, 2] - 1] = -2e-4
    g = fg.GpuModel(m)
    both(g, q, steps=3, factor=1.0, maxit=30, tol=1e-9, pcg_tol=1e-11, direct=direct)
    g.close()


def test_generic_bspline_deck_direct(fg):
    """The reference's CurvedBeamDeg4 (degree 4 solids, total-Lagrangian,
    an end-face load) on its first steps with the direct solver."""
    from fast_amd import deck, neutral
    d = neutral.read(os.path.join(DECKS, "CurvedBeamDeg4.dat"))
    g = fg.GpuModel(d.mesh)
    deck.prepare(d, g)
    f = deck.load_vector(d, g)
    both(g, f, steps=6, factor=d.factor, maxit=d.maxit, tol=d.tol, direct=True)
    g.close()


def test_prescribed_penalty_newton(fg):
    from fast_amd import deck
    m = M.hex8_block(6, 2, 2, size=(3.0, 1.0, 1.0), form="tl_svk", E=100.0, nu=0.3)
    xmax = np.isclose(m.coords[:, 0], 3.0)
    eq = m.dofmap[xmax, 0]
    g = fg.GpuModel(m)
    g.set_penalty(eq, np.full(len(eq), deck.PENALTY_STIFFNESS))
    f = np.zeros(m.neq)
    f[eq - 1] = deck.PENALTY_STIFFNESS * 0.3
    both(g, f, steps=3, factor=1.0 / 3.0, maxit=30, tol=1e-8, direct=True)
    g.close()
