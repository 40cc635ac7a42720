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


def test_testb8_model_and_face_load(oracle_mod):
    d = neutral.read(os.path.join(DECKS, "TestB8.dat"))
    m = d.mesh
    assert m.shape == "hex8" and m.nelem == 1 and m.nnode == 8 and m.neq == 18
    assert m.nip == (2, 2, 2) and np.allclose(m.mat, [[200e9, 0.3]])
    assert not d.load.any() and len(d.surfaces) == 1
    g = d.surfaces[0]
    assert g.deg == (1, 1) and g.nip == 2 and g.nodes.shape == (1, 4) and g.ext is None
    # 100 MPa on the 1 m x 1 m end face: a quarter on each corner, along x
    fn = m.scatter(oracle_mod.OracleMesh(m).surface_load(d.surfaces))
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


ALL = ["TestB8", "ScordelisLoDeg3", "PinCylDeg4", "ThinCylinderDeg5", "CurvedBeamDeg4", "BeamB8",
       "BeamB8TL", "BeamB8Face",
       "ScordelisLoDeg3TL", "BeamB20", "BeamB20TL"]


def oracle_solution(oracle_mod, d):
    """Nodal displacements of the deck's last step by the oracle."""
    om = oracle_mod.OracleMesh(d.mesh)
    f = d.load.copy()
    if d.surfaces:
        f += om.surface_load(d.surfaces)
    if d.body is not None:
        fb, ok = om.assemble_body(d.body)
        assert ok
        f += fb
    if d.analysis == "linear":
        u = oracle_mod.linear_static(om, f)
    else:
        us, its = oracle_mod.newton_raphson(om, f, d.steps, d.factor, d.maxit, d.tol)[:2]
        assert len(us) == d.steps
        u = us[-1]
    return om, d.mesh.scatter(u)


def reference_state(d, ref):
    """The reference's printed displacements in the reader's node order."""
    idx = {int(l): i for i, l in enumerate(d.labels)}
    u = np.zeros((d.mesh.nnode, 3))
    for l, v in zip(ref["labels"], ref["u"]):
        u[idx[int(l)]] = v
    return u


@pytest.mark.parametrize("name", ALL)
def test_oracle_solution_matches_reference_run(oracle_mod, name):
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    ref = load_golden("deck_%s.npz" % name)
    assert np.array_equal(ref["labels"], d.labels)
    assert int(ref["nsteps"]) == d.steps
    om, un = oracle_solution(oracle_mod, d)
    # ThinCylinderDeg5: both direct solutions leave a relative residual of
    # 7e-12 in the same FP64 system (they are each accurate to ~1e-9 there);
    # BeamB20: condition number 4e9 (tests/test_hex20.py)
    tol = (1e-8 if name == "ThinCylinderDeg5" else 1e-7 if name == "BeamB20"
           else 1e-9 if "TL" in name or "Curved" in name else 1e-10)
    assert_same_displacements(d, un, ref["u"], name, tol=tol)


@pytest.mark.parametrize("name", ALL)
def test_oracle_stresses_match_reference_run(oracle_mod, name):
    """Integration-point and element-nodal stresses of the reference's own
    displacements (cElmParam[TL]::IntPntStress, NodalStress with the
    default extrapolation, or PntStress at the nodes where the deck turns
    the extrapolation off) against what the reference printed for them."""
    d = neutral.read(os.path.join(DECKS, name + ".dat"))
    ref = load_golden("deck_%s.npz" % name)
    assert np.array_equal(ref["elem_labels"], d.elem_labels)
    om = oracle_mod.OracleMesh(d.mesh)
    u = reference_state(d, ref)
    si, sn, sa, ok = om.stresses(u)
    assert ok
    if not d.stress_extrapolation:  # PinCylDeg4: evaluated at the nodes
        sn, ok = om.node_stresses(u)
        assert ok
    scale = np.abs(ref["sig_ip"]).max()
    # (degree 5: the oracle's basis derivatives round differently at 1e-12;
    # 20-node bricks: the serendipity functions are formed in another order)
    tol = 1e-11 if name in ("ThinCylinderDeg5", "BeamB20", "BeamB20TL") else 1e-13
    assert np.abs(si - ref["sig_ip"]).max() <= tol * scale
    assert np.abs(sn - ref["sig_node"]).max() <= tol * scale


def test_pos_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    labels = np.arange(1, 9)
    u = rng.standard_normal((8, 3))
    si, sn, sa = rng.standard_normal((2, 8, 6)), rng.standard_normal((2, 8, 6)), \
        rng.standard_normal((8, 6))
    out = str(tmp_path / "a.pos")
    neutral.write_pos(out, labels, u, 0.5, elem_labels=np.array([4, 7]), sig_ip=si,
                      sig_node=sn, sig_avg=sa)
    st = neutral.read_pos(out)
    assert len(st) == 1 and st[0]["factor"] == 0.5
    assert np.array_equal(st[0]["labels"], labels) and np.array_equal(st[0]["u"], u)
    assert np.array_equal(st[0]["elem_labels"], [4, 7])
    assert np.array_equal(np.stack(st[0]["sig_ip"]), si)
    assert np.array_equal(np.stack(st[0]["sig_node"]), sn)
    assert np.array_equal(st[0]["sig_avg"], sa)


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


# the pinched semi-cylinder's header (dat/iga/PinSemiCylDeg4.dat:10-29): load
# factor and the "FAST" column of the displacement w at the loading point
PSC_HEADER = [(0.05, -5.39093), (0.10, -15.7282), (0.15, -27.1612), (0.20, -37.3189),
              (0.25, -48.8163), (0.30, -66.7192), (0.35, -94.8713), (0.40, -113.724),
              (0.45, -124.549), (0.50, -132.390), (0.55, -138.921), (0.60, -144.041),
              (0.65, -148.793), (0.70, -152.909), (0.75, -156.634), (0.80, -160.468),
              (0.85, -163.393), (0.90, -166.205), (0.95, -169.427), (1.00, -171.766)]


def test_pinched_semicylinder_fixture_and_header():
    """PinSemiCylDeg4 as the reference program built from its shipped
    sources runs it (tests/golden/gen_decks.py: 40 load-control steps of
    0.025, every second printed): the printed load factors are the header's,
    and the displacement w (y) of the loaded node 400 follows the header's
    FAST column to within 0.75% -- not to the digits printed: the shipped
    program differs from its own header by 0.03% to 0.7%, so the header was
    made with another version.  The device is held to the program's output
    (tests/test_gpu_decks.py)."""
    d = neutral.read(os.path.join(DECKS, "PinSemiCylDeg4.dat"))
    assert d.mesh.shape == "bsp" and d.mesh.nelem == 256 and d.mesh.deg == (4, 4, 2)
    assert d.analysis == "newton" and d.steps == 100 and d.factor == 0.025
    ref = load_golden("deck_PinSemiCylDeg4.npz")
    assert int(ref["nsteps"]) == 40 and len(ref["factors"]) == 20
    k = int(np.flatnonzero(ref["labels"] == 400)[0])
    dev = []
    for i, (lf, w) in enumerate(PSC_HEADER):
        assert abs(ref["factors"][i] - lf) <= 1e-12
        dev.append(abs(ref["u_steps"][i][k, 1] - w) / abs(w))
    print("reference program vs header: %.2e .. %.2e" % (min(dev), max(dev)))
    assert max(dev) <= 7.5e-3 and min(dev) >= 1e-4


def test_pinched_semicylinder_stresses(oracle_mod):
    """The oracle's integration-point and element-nodal stresses of the
    reference's last printed state, on the first 32 elements the fixture
    keeps (degree (4, 4, 2) total-Lagrangian NURBS solids, 5x5x5 points)."""
    d = neutral.read(os.path.join(DECKS, "PinSemiCylDeg4.dat"))
    ref = load_golden("deck_PinSemiCylDeg4.npz")
    om = oracle_mod.OracleMesh(d.mesh)
    u = reference_state(d, ref)
    si, sn, sa, ok = om.stresses(u)
    assert ok
    ne = len(ref["sig_ip"])
    scale = np.abs(ref["sig_ip"]).max()
    assert np.abs(si[:ne] - ref["sig_ip"]).max() <= 1e-11 * scale
    assert np.abs(sn[:ne] - ref["sig_node"]).max() <= 1e-11 * scale
