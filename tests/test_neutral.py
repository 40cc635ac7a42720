"""The neutral-file reader (fast_amd/neutral.py) on the reference's own decks:
the parsed model, its load vector and the oracle's solution of it against the
nodal displacements the reference program printed for the same deck
(tests/golden/deck_*.npz, written by tests/golden/gen_decks.py)."""
import os

import numpy as np
import pytest

from conftest import load_golden
from fast_amd import neutral

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def test_testb8_model_and_face_load():
    d = neutral.read(os.path.join(DECKS, "TestB8.dat"))
    m = d.mesh
    assert m.shape == "hex8" and m.nelem == 1 and m.nnode == 8 and m.neq == 18
    assert m.nip == (2, 2, 2) and np.allclose(m.mat, [[200e9, 0.3]])
    # 100 MPa on the 1 m x 1 m end face: a quarter on each corner, along x
    fn = m.scatter(d.load)
    assert np.allclose(fn[[1, 2, 5, 6], 0], 25e6, rtol=1e-14)
    assert np.abs(fn).sum() == pytest.approx(100e6, rel=1e-14)


def test_scordelis_model():
    d = neutral.read(os.path.join(DECKS, "ScordelisLoDeg3.dat"))
    m = d.mesh
    assert m.shape == "bsp" and m.deg == (3, 3, 2) and m.nn == 48 and m.nelem == 64
    assert m.nnode == 363 and m.nip == (4, 4, 4)
    assert np.allclose(d.body, [[0.0, 0.0, -360.0]] * 64)
    # every element's extraction operators reproduce a partition of unity
    assert np.allclose(m.ext.reshape(64, -1)[:, :16].reshape(64, 4, 4).sum(axis=1).sum(axis=1), 4.0)


def test_solver_order_numbering():
    """NODE.SOLVER.ORDER lists the rank of every node (in label order); the
    equations follow increasing rank."""
    d = neutral.read(os.path.join(DECKS, "ScordelisLoDeg3.dat"))
    dm = d.mesh.dofmap
    toks = open(os.path.join(DECKS, "ScordelisLoDeg3.dat")).read().split("%NODE.SOLVER.ORDER")[1]
    rank = np.asarray(toks.split("%")[0].split()[1:], dtype=int)
    assert sorted(rank) == list(range(1, 364))
    first = int(np.argmin(rank))  # rank 1: node 121, supported in x and y
    assert d.labels[first] == 121 and list(dm[first]) == [0, 0, 1]
    # the equation numbers grow with the rank of their node
    eq_min = np.where(dm > 0, dm, 10 ** 9).min(axis=1)
    live = eq_min < 10 ** 9
    r = rank[live]
    assert np.all(np.diff(eq_min[live][np.argsort(r)]) > 0)
    assert dm.max() == d.mesh.neq


@pytest.mark.parametrize("name", ["TestB8", "ScordelisLoDeg3"])
def test_oracle_solution_matches_reference_run(oracle_mod, name):
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    om = oracle_mod.OracleMesh(d.mesh)
    f = d.load.copy()
    if d.body is not None:
        fb, ok = om.assemble_body(d.body)
        assert ok
        f += fb
    u = oracle_mod.linear_static(om, f)
    ref = load_golden("deck_%s.npz" % name)
    assert np.array_equal(ref["labels"], d.labels)
    un = d.mesh.scatter(u)
    assert_same_displacements(d, un, ref["u"], name)


def assert_same_displacements(d, un, ref, name, tol=1e-10):
    """Displacements agree to tol (relative to the largest).  TestB8's
    supports leave the rotation about the x axis through node 5 free (its
    stiffness matrix is singular), so there the difference may be such a
    rotation, which is fitted and removed first; the axial displacement of
    the loaded end is sigma L / E = 2.5 mm."""
    diff = un - ref
    if name == "TestB8":
        X = d.mesh.coords - d.mesh.coords[4]
        rot = np.stack([np.zeros(len(X)), -X[:, 2], X[:, 1]], 1).ravel()
        theta = rot @ diff.ravel() / (rot @ rot)
        diff = diff - theta * rot.reshape(-1, 3)
        assert np.allclose(un[[1, 2, 5, 6], 0], 100e6 * 5.0 / 200e9, rtol=1e-12)
    err = np.abs(diff).max() / np.abs(ref).max()
    assert err <= tol, err
