"""Device stresses (fgpu_stresses: integration points, element nodes, node
averages) against the CPU FP64 oracle on random smooth states, and against
what the reference program printed for its own displacements on the decks
of tests/golden/ (deck_*.npz).  Both sides compute in FP64 from the same
nodal state, so the bar is 1e-12 of the largest stress against the oracle;
against the reference's printed values it is 1e-11, since the gradients of
the curved Scordelis-Lo shell sum 48 terms with cancellation in an order of
their own on the device (the oracle, in the reference's order, matches
those values to the last bit)."""
import os

import numpy as np
import pytest

from conftest import load_golden
from fast_amd import mesh as M
from test_neutral import ALL, DECKS, reference_state
from test_oracle import _stress_meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def smooth_state(m, amp):
    X = m.coords
    return amp * np.stack([np.sin(1.3 * X[:, 0] + X[:, 1]), X[:, 0] * X[:, 2] - 0.5 * X[:, 1],
                           np.cos(X[:, 0]) * X[:, 1] ** 2], 1)


@pytest.mark.parametrize("form,mat", [("linear", (2.0, 0.3)), ("tl_svk", (2.0, 0.3)),
                                      ("tl_neohooke", (1.0, 1.5))])
def test_stresses_vs_oracle(fg, oracle_mod, form, mat):
    for m in _stress_meshes(M, form, mat):
        u = smooth_state(m, 0.05)
        g = fg.GpuModel(m)
        si, sn, sa = g.stresses(u, nodal=True, average=True)
        r_si, r_sn, r_sa, ok = oracle_mod.OracleMesh(m).stresses(u)
        assert ok
        scale = np.abs(r_si).max()
        tag = (m.shape, m.deg, m.nip)
        assert np.abs(si - r_si).max() <= 1e-12 * scale, tag
        assert np.abs(sn - r_sn).max() <= 1e-12 * scale, tag
        assert np.abs(sa - r_sa).max() <= 1e-12 * scale, tag
        # the optional outputs are optional
        si2, sn2, sa2 = g.stresses(u, nodal=False, average=False)
        assert sn2 is None and sa2 is None and np.array_equal(si2, si)
        si3, sn3, sa3 = g.stresses(u, nodal=False, average=True)
        assert np.array_equal(sa3, sa)
        g.close()


@pytest.mark.parametrize("name", ALL)
def test_stresses_vs_reference_run(fg, name):
    from fast_amd import neutral
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    ref = load_golden("deck_%s.npz" % name)
    g = fg.GpuModel(d.mesh)
    if not d.stress_extrapolation:
        g.set_stress_nodes(False)
    si, sn, _ = g.stresses(reference_state(d, ref))
    scale = np.abs(ref["sig_ip"]).max()
    assert np.abs(si - ref["sig_ip"]).max() <= 1e-11 * scale
    assert np.abs(sn - ref["sig_node"]).max() <= 1e-11 * scale
    g.close()


def test_failed_stress_update_is_reported(fg):
    """A state with det C = 0 (here F_xx = 0 everywhere) fails the neo-Hooke
    stress update."""
    m = M.hex8_block(2, 2, 2, form="tl_neohooke", mat=np.tile((1.0, 1.5), (8, 1)))
    u = np.zeros((m.nnode, 3))
    u[:, 0] = -m.coords[:, 0]
    g = fg.GpuModel(m)
    with pytest.raises(fg.FgpuError, match="stress update failed"):
        g.stresses(u)
    g.close()
