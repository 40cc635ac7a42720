"""Pins of the CPU FP64 graph-network oracle (oracle/gnn_oracle.py): the
hand-written backward against central finite differences of the forward, the
mean aggregation against a brute-force neighbour search, and the parameter
layout against the counts of the network definition."""
import numpy as np
import pytest

from fast_amd import mesh as M
from oracle import gnn_oracle as G


def _case(seed=0, n=3):
    m = M.hex8_block(n, n, n, form="tl_neohooke", traction=(0.0, 0.0, -1.0))
    rng = np.random.default_rng(seed)
    fmask = (~m.fixed).astype(np.uint8) @ np.array([1, 2, 4], dtype=np.uint8)
    lmask = (m.fext != 0).astype(np.uint8) @ np.array([1, 2, 4], dtype=np.uint8)
    u = 0.05 * rng.standard_normal((m.nnode, 3))
    u[m.fixed] = 0
    ut = u + 0.01 * rng.standard_normal((m.nnode, 3))
    ut[m.fixed] = 0
    uh = 0.5 * u + 0.001 * rng.standard_normal((m.nnode, 3))
    uh[m.fixed] = 0
    X = G.features(m.coords, u, 0.1, 0.02, uh, 0.05, fmask, lmask, 0.05, 0.3, 0.5)
    Mg = G.graph(m.conn, m.nnode)
    T = G.target(u, ut, 0.02, 0.5, fmask)
    P = 0.3 * rng.standard_normal(G.layout()["total"])
    return m, X, Mg, T, fmask, P


def test_layout_counts():
    L = G.layout()
    # 32x64 + 64, 4 x (2 x 64x64 + 64), 64x64 + 64, 64x16 + 16
    assert L["total"] == 32 * 64 + 64 + 4 * (2 * 64 * 64 + 64) + 64 * 64 + 64 + 64 * 16 + 16
    P = np.arange(L["total"], dtype=np.float64)
    assert np.array_equal(G.pack(G.unpack(P)), P)
    W = G.unpack(P)
    o = L["ws2"][0]
    assert W["ws2"][3, 5] == o + 3 * 64 + 5  # row-major W[k][n]


def test_graph_mean_matches_brute_force():
    m = M.hex8_block(3, 2, 2)
    Mg = G.graph(m.conn, m.nnode).toarray()
    for i in range(m.nnode):
        nb = set()
        for e in range(m.nelem):
            if i in m.conn[e]:
                nb.update(int(v) for v in m.conn[e])
        row = np.zeros(m.nnode)
        row[sorted(nb)] = 1.0 / len(nb)
        assert np.allclose(Mg[i], row, rtol=0, atol=1e-15)
    # interior nodes of a structured hex mesh see 27 nodes
    assert np.isclose(Mg.max(), 1.0 / 8) and np.isclose(Mg[Mg > 0].min(), 1.0 / 27)


@pytest.mark.parametrize("seed,beta", [(0, 0.0), (1, 0.0), (2, 30.0)])
def test_backward_matches_finite_differences(seed, beta):
    m, X, Mg, T, fmask, P = _case(seed)
    neq = int((~m.fixed).sum())
    l0, g, _ = G.loss_grad(P, X, Mg, T, fmask, neq, beta)
    rng = np.random.default_rng(100 + seed)
    L = G.layout()
    # single coordinates from every block, plus random directions
    idx = [L[k][0] + int(rng.integers(L[k][1])) for k in L if k != "total"]
    eps = 1e-6
    for k in idx:
        e = np.zeros_like(P)
        e[k] = eps
        lp = G.loss_grad(P + e, X, Mg, T, fmask, neq, beta)[0]
        lm = G.loss_grad(P - e, X, Mg, T, fmask, neq, beta)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (k, fd, g[k])
    for _ in range(3):
        d = rng.standard_normal(P.size)
        d /= np.linalg.norm(d)
        lp = G.loss_grad(P + eps * d, X, Mg, T, fmask, neq, beta)[0]
        lm = G.loss_grad(P - eps * d, X, Mg, T, fmask, neq, beta)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd) + 1e-10


def test_target_and_prediction_are_inverse():
    """A network whose correction y equals t - v reproduces the increment:
    pins the skip term v and the scaling of the target and the prediction."""
    m, X, Mg, T, fmask, P = _case(3)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((m.nnode, 3))
    ut = u + rng.standard_normal((m.nnode, 3))
    uh = rng.standard_normal((m.nnode, 3))
    lam, dlam, lamh, osc = 0.2, 0.04, 0.15, 2.0
    X = G.features(m.coords, u, lam, dlam, uh, lamh, fmask, fmask, 1.0, 1.0, osc)
    T = G.target(u, ut, dlam, osc, fmask)
    y = np.zeros((m.nnode, G.GF))
    y[:, :3] = T[:, :3] - X[:, G.SKIP:G.SKIP + 3]
    l, _ = G.loss(y, X, T, fmask, m.neq)
    assert l < 1e-28
    du = np.zeros((m.nnode, 3))
    for c in range(3):
        free = ((fmask >> c) & 1).astype(bool)
        du[free, c] = dlam * osc * (y[free, c] + X[free, G.SKIP + c])
    free3 = ~m.fixed
    assert np.allclose(du[free3], (ut - u)[free3], rtol=1e-13, atol=1e-13)
    assert np.all(du[m.fixed] == 0)
    # the slope features
    assert np.allclose(X[:, 6:9], u / (lam * osc))
    assert np.allclose(X[:, 9:12], (u - uh) / ((lam - lamh) * osc))
    X0 = G.features(m.coords, u, 0.0, dlam, None, 0.0, fmask, fmask, 1.0, 1.0, osc)
    assert np.all(X0[:, 6:12] == 0)
    Xn = G.features(m.coords, u, lam, dlam, None, 0.0, fmask, fmask, 1.0, 1.0, osc)
    assert np.array_equal(Xn[:, 6:9], Xn[:, 9:12])  # no history: v = s


def test_prediction_with_zero_network_is_the_secant():
    m, X, Mg, T, fmask, P = _case(4)
    rng = np.random.default_rng(2)
    u = rng.standard_normal((m.nnode, 3))
    uh = rng.standard_normal((m.nnode, 3))
    u[m.fixed] = 0
    uh[m.fixed] = 0
    lam, dlam, lamh, osc = 0.1, 0.05, 0.07, 3.0
    X = G.features(m.coords, u, lam, dlam, uh, lamh, fmask, fmask, 1.0, 1.0, osc)
    du = G.predict(np.zeros_like(P), X, Mg, dlam, osc, fmask)
    assert np.allclose(du, (u - uh) * dlam / (lam - lamh), rtol=1e-12, atol=1e-14)
    X = G.features(m.coords, u, lam, dlam, None, 0.0, fmask, fmask, 1.0, 1.0, osc)
    du = G.predict(np.zeros_like(P), X, Mg, dlam, osc, fmask)
    assert np.allclose(du, u * dlam / lam, rtol=1e-12, atol=1e-14)
