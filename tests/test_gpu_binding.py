"""The FAST-side binding (fast_amd/binding/ctrlgpu.cpp) inside the reference's
own set-up: oracle/_ref/ref_gpu reads each deck with the reference's reader
and model classes, hands the model to the device through GpuInit (nodes,
VecDof equations, element incidences, control-point weights and extraction
operators read from the reference's objects), and runs the deck's linear
static or load-control analysis with the binding's device stiffness,
internal force and solve.  Its displacements must match the reference
program's CPU run of the same deck (tests/golden/deck_*.npz) to 1e-6 of the
largest (FP32 element matrices, the reference solves in FP64).

Not run: the Scordelis-Lo decks.  Their patch is parametrized left-handed;
the device integrates with |det J| while the reference's load routines
(cControl::ExtLoadVector, which the binding leaves to the reference) use the
signed det J for the body load, so the binding would pair a positive
stiffness with a negated load (the product's own deck runner forms the body
load on the device and matches them: tests/test_gpu_decks.py)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from test_neutral import DECKS, assert_same_displacements

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_gpu")
CASES = [("TestB8", "direct"), ("BeamB8", "pcg"), ("BeamB8Face", "direct"),
         ("BeamB8TL", "pcg"), ("PinCylDeg4", "direct"), ("ThinCylinderDeg5", "direct"),
         ("CurvedBeamDeg4", "direct"), ("BeamB20", "direct"), ("BeamB20TL", "direct")]


def read_steps(path):
    steps, cur = [], None
    for ln in open(path):
        f = ln.split()
        if f[0] == "step":
            cur = {"factor": float(f[3]), "labels": [], "u": []}
            steps.append(cur)
        else:
            cur["labels"].append(int(f[0]))
            cur["u"].append([float(x) for x in f[1:4]])
    for s in steps:
        s["labels"] = np.array(s["labels"])
        s["u"] = np.array(s["u"])
    return steps


@pytest.mark.skipif(not os.path.exists(EXE), reason="the reference build (oracle/_ref) is "
                    "not present: it is compiled where the reference sources are")
@pytest.mark.parametrize("name,solver", CASES)
def test_binding_runs_reference_deck(tmp_path, name, solver):
    from fast_amd import neutral
    import shutil
    shutil.copy(os.path.join(DECKS, name + ".dat"), tmp_path / (name + ".dat"))
    out = tmp_path / "u.txt"
    res = subprocess.run([EXE, str(tmp_path / name), str(out), solver], capture_output=True,
                         text=True, timeout=600, cwd=str(tmp_path))
    assert res.returncode == 0, res.stderr[-2000:]
    steps = read_steps(str(out))
    ref = load_golden("deck_%s.npz" % name)
    last = steps[-1]
    assert np.array_equal(last["labels"], ref["labels"])
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    assert_same_displacements(d, last["u"], ref["u"], name, tol=1e-6)
    if "u_steps" in ref and len(steps) > 1:  # every printed step of a load path
        by = {round(s["factor"], 12): s["u"] for s in steps}
        for lam, u in zip(ref["factors"], ref["u_steps"]):
            got = by[round(float(lam), 12)]
            assert np.abs(got - u).max() <= 1e-6 * np.abs(ref["u"]).max()
