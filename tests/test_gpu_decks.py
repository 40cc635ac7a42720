"""The reference's own example decks through the device path
(fast_amd/deck.py: neutral-file reader, device body forces, device solve)
against the displacements and stresses the reference program printed for
them (tests/golden/deck_*.npz), and the orientation rules of the element
kernels.

Tolerance: the reference solves with a direct (skyline) solver; the device
path with Jacobi-PCG to 1e-12 and FP32 element matrices (1e-7 relative per
element): displacements agree to 1e-6 of the largest."""
import os

import numpy as np
import pytest

from conftest import load_golden
from fast_amd import mesh as M
from test_neutral import ALL, DECKS, assert_same_displacements

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


@pytest.mark.parametrize("solver", ["direct", "pcg"])
@pytest.mark.parametrize("name", ALL)
def test_deck_matches_reference_run(fg, tmp_path, name, solver):
    """Displacements and stresses of the last step (linear static, or load
    control for the TL decks), with the device direct solver (the decks'
    default skyline matrix) and with PCG, against the reference program's
    output; the written .pos file reads back the same."""
    from fast_amd import deck, neutral
    out = str(tmp_path / (name + ".pos"))
    d, un, sig = deck.run(os.path.join(DECKS, name + ".dat"), out=out, solver=solver)
    ref = load_golden("deck_%s.npz" % name)
    assert_same_displacements(d, un, ref["u"], name, tol=1e-6)
    # (TestB8's free rigid rotation carries no small-strain stress)
    scale = np.abs(ref["sig_ip"]).max()
    assert np.abs(sig[0] - ref["sig_ip"]).max() <= 1e-5 * scale
    assert np.abs(sig[1] - ref["sig_node"]).max() <= 1e-5 * scale
    steps = neutral.read_pos(out)
    st = steps[-1]
    assert np.array_equal(st["labels"], d.labels) and np.array_equal(st["u"], un)
    assert np.array_equal(np.stack(st["sig_ip"]), sig[0])
    assert np.array_equal(np.stack(st["sig_node"]), sig[1])
    # every load step is printed, as the reference prints them
    assert len(steps) == int(ref["nsteps"])
    for k, s in enumerate(steps):
        assert abs(s["factor"] - ref["factors"][k]) <= 1e-12 * abs(ref["factors"][k])
        assert_same_displacements(d, s["u"], ref["u_steps"][k], name, tol=1e-6)


def test_body_force_vs_oracle(fg, oracle_mod):
    for m in (M.hex8_block(3, 2, 2, size=(1.0, 0.7, 0.4), perturb=0.2, seed=1),
              M.bsp_block(ne=(3, 2, 2), p=3, perturb=0.1, seed=2),
              M.bsp_block(ne=(2, 2, 1), p=(3, 3, 2), perturb=0.1, seed=2)):
        b = np.random.default_rng(0).uniform(-1, 1, (m.nelem, 3))
        f = fg.GpuModel(m).body_force(b)
        f0, ok = oracle_mod.OracleMesh(m).assemble_body(b)
        assert ok and np.abs(f - f0).max() <= 1e-13 * np.abs(f0).max()


def _mirror(conn):
    """Reference local order with the t = +1 and t = -1 faces swapped:
    a left-handed parametrization of the same brick."""
    return np.ascontiguousarray(conn[:, [4, 5, 6, 7, 0, 1, 2, 3]])


def test_left_handed_mesh_gives_the_same_solution(fg):
    m = M.hex8_block(4, 3, 3, perturb=0.2, seed=5, traction=(0.0, 0.0, -1.0))
    f = m.load_vector()
    u0, _, _ = fg.GpuModel(m).linear_static(f, tol=1e-12)
    ml = M.hex8_block(4, 3, 3, perturb=0.2, seed=5, traction=(0.0, 0.0, -1.0))
    ml.conn = _mirror(ml.conn)
    u1, _, _ = fg.GpuModel(ml).linear_static(f, tol=1e-12)
    assert np.abs(u1 - u0).max() <= 1e-8 * np.abs(u0).max()


def test_inverted_element_is_reported(fg):
    m = M.hex8_block(4, 3, 3, perturb=0.1, seed=5)
    m.conn[7] = m.conn[7][[4, 5, 6, 7, 0, 1, 2, 3]]
    g = fg.GpuModel(m)
    with pytest.raises(fg.FgpuError, match="inverted elements"):
        g.stiffness()
    with pytest.raises(fg.FgpuError, match="inverted elements"):
        g.int_force(np.zeros((m.nnode, 3)))


def test_pinched_semicylinder_matches_reference_run(fg, tmp_path):
    """PinSemiCylDeg4 (256 total-Lagrangian NURBS solids of degree (4, 4, 2),
    a point load, 'lcm' steps of 0.025; TARGET.ITERATIONS read and unused, as
    the reference's load control does) for the 40 steps the fixture holds,
    direct solver: at each of the reference program's 20 printed steps the
    displacement field to 1e-6 of the largest and the loaded node's w to
    1e-6 relative; the last step's stresses of the first 32 elements."""
    from fast_amd import deck, neutral
    text = open(os.path.join(DECKS, "PinSemiCylDeg4.dat")).read()
    k0 = text.index("%HEADER.ANALYSIS.MAXIMUM.STEPS")
    k1 = text.index("%", k0 + 1)
    text = text[:k0] + "%HEADER.ANALYSIS.MAXIMUM.STEPS\n40\n\n" + text[k1:]
    path = tmp_path / "PinSemiCylDeg4.dat"
    path.write_text(text)
    out = str(tmp_path / "psc.pos")
    d, un, sig = deck.run(str(path), out=out, solver="direct")
    ref = load_golden("deck_PinSemiCylDeg4.npz")
    steps = neutral.read_pos(out)
    assert len(steps) == 40
    node = int(np.flatnonzero(d.labels == 400)[0])
    kref = int(np.flatnonzero(ref["labels"] == 400)[0])
    for i in range(20):
        s = steps[2 * i + 1]
        assert abs(s["factor"] - ref["factors"][i]) <= 1e-12
        assert_same_displacements(d, s["u"], ref["u_steps"][i], "PinSemiCylDeg4", tol=1e-6)
        w, w0 = s["u"][node, 1], ref["u_steps"][i][kref, 1]
        assert abs(w - w0) <= 1e-6 * abs(w0), (i, w, w0)
    ne = len(ref["sig_ip"])
    scale = np.abs(ref["sig_ip"]).max()
    assert np.abs(sig[0][:ne] - ref["sig_ip"]).max() <= 1e-5 * scale
    assert np.abs(sig[1][:ne] - ref["sig_node"]).max() <= 1e-5 * scale
