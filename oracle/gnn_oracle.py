"""CPU FP64 restatement of the mesh graph network that warm-starts the Newton
solves (fast_amd/csrc/k_gnn.hip).  Test infrastructure only: tests use it to
check the device forward, loss and hand-written backward.

The network has no counterpart in the reference (its solve path is
cCtrlNR::Solver, ctrl/ctrlnr.cpp:47-267, with no predictor); this module is
its definition in plain numpy:

  graph      nodes of the mesh, neighbours = nodes sharing an element, self
             included; mean aggregation over the neighbour list
  sample     state u at lam, earlier state u_h at lam_h (the unloaded state
             when there is none), step dlam
  features   32 per node: p = (x - lo) / extent (3), u / u_scale (3),
             s = u / (lam out_scale) (3), v = (u - u_h) / ((lam - lam_h)
             out_scale) (3), lam / lam_scale, dlam / lam_scale,
             (lam - lam_h) / lam_scale, free-direction bit mask / 7, load bit
             mask / 7, 1, then zeros; s and v are 0 at lam = 0
  h0         relu(X W_in + b_in)
  layer l    h <- h + relu(h W_s,l + mean_N(h) W_n,l + b_l),  l = 0..3
  decoder    d = relu(h W_d1 + b_d1);  y = M M (d W_d2 + b_d2), M the
             mean over the neighbours (two smoothing passes; M has unit row
             sums, so this equals (M M d) W_d2 + b_d2, the form used here)
  loss       (sum over free directions of e^2 + beta sum |e - M e|^2) / neq,
             e = v + y - t on free directions (0 elsewhere),
             t = (u_target - u) / (dlam out_scale)
  prediction du = dlam out_scale (v + y[:, :3]) on free directions

Parameters are one flat vector; every weight W (fan_in x fan_out) is stored
row-major, W[k][n] at k * fan_out + n.  The backward pass here is written out
by hand and is pinned against central finite differences of ``forward`` in
tests/test_gnn_oracle.py.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

GH = 64
GFI = 32
GF = 16
GL = 4
GS = 2  # smoothing passes of the decoder


def layout():
    """Offsets of every parameter block in the flat vector."""
    o = 0
    L = {}

    def take(name, n):
        nonlocal o
        L[name] = (o, n)
        o += n

    take("w_in", GFI * GH)
    take("b_in", GH)
    for l in range(GL):
        take("ws%d" % l, GH * GH)
        take("wn%d" % l, GH * GH)
        take("bl%d" % l, GH)
    take("w_d1", GH * GH)
    take("b_d1", GH)
    take("w_d2", GH * GF)
    take("b_d2", GF)
    L["total"] = o
    return L


SHAPES = {"w_in": (GFI, GH), "w_d1": (GH, GH), "w_d2": (GH, GF)}


def _shape(name):
    if name in SHAPES:
        return SHAPES[name]
    if name.startswith("ws") or name.startswith("wn"):
        return (GH, GH)
    return None


def unpack(P):
    L = layout()
    out = {}
    for k, v in L.items():
        if k == "total":
            continue
        o, n = v
        s = _shape(k)
        out[k] = P[o:o + n].reshape(s) if s else P[o:o + n]
    return out


def pack(d):
    L = layout()
    P = np.zeros(L["total"])
    for k, (o, n) in ((k, v) for k, v in L.items() if k != "total"):
        P[o:o + n] = np.asarray(d[k]).ravel()
    return P


def graph(conn, nnode):
    """Row-normalised adjacency M (mean over the neighbours, self included)."""
    conn = np.asarray(conn)
    ne, nn = conn.shape
    rows = np.repeat(conn, nn, axis=1).ravel()
    cols = np.tile(conn, (1, nn)).ravel()
    A = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(nnode, nnode))
    A.data[:] = 1.0
    A.sum_duplicates()
    A.data[:] = 1.0
    deg = np.asarray(A.sum(axis=1)).ravel()
    return sp.diags(1.0 / deg) @ A


def box(coords):
    """Bounding box as the device keeps it (float32 lower corner, float32
    inverse extent)."""
    lo = coords.min(axis=0)
    hi = coords.max(axis=0)
    inv = np.where(hi > lo, 1.0 / np.where(hi > lo, hi - lo, 1.0), 1.0)
    return lo.astype(np.float32).astype(np.float64), inv.astype(np.float32).astype(np.float64)


SKIP = 9  # feature offset of v


def features(coords, u, lam, dlam, uh, lamh, fmask, lmask, u_scale, lam_scale, out_scale):
    """coords, u, uh: [nnode, 3] (uh None: the unloaded state); fmask/lmask:
    per-node direction bit masks."""
    if uh is None:
        uh, lamh = np.zeros_like(u), 0.0
    lo, inv = box(coords)
    n = coords.shape[0]
    X = np.zeros((n, GFI))
    X[:, 0:3] = (coords - lo) * inv
    X[:, 3:6] = u / u_scale
    X[:, 6:9] = u / (lam * out_scale) if lam != 0 else 0.0
    X[:, 9:12] = (u - uh) / ((lam - lamh) * out_scale) if lam > lamh else 0.0
    X[:, 12] = lam / lam_scale
    X[:, 13] = dlam / lam_scale
    X[:, 14] = (lam - lamh) / lam_scale
    X[:, 15] = fmask / 7.0
    X[:, 16] = lmask / 7.0
    X[:, 17] = 1.0
    return X


def relu(x):
    return np.maximum(x, 0.0)


def forward(P, X, M):
    """Returns y [n, 16] and the cache the backward pass needs."""
    W = unpack(P)
    h = relu(X @ W["w_in"] + W["b_in"])
    cache = {"X": X, "h": [h], "ag": [], "a": []}
    for l in range(GL):
        ag = M @ h
        a = relu(h @ W["ws%d" % l] + ag @ W["wn%d" % l] + W["bl%d" % l])
        h = h + a
        cache["ag"].append(ag)
        cache["a"].append(a)
        cache["h"].append(h)
    d1 = relu(h @ W["w_d1"] + W["b_d1"])
    e = d1
    for _ in range(GS):
        e = M @ e
    y = e @ W["w_d2"] + W["b_d2"]
    cache["d1"] = d1
    cache["e"] = e
    return y, cache


def target(u, ut, dlam, out_scale, fmask):
    T = np.zeros((u.shape[0], GF))
    for c in range(3):
        free = ((fmask >> c) & 1).astype(bool)
        T[free, c] = (ut[free, c] - u[free, c]) / (dlam * out_scale)
    return T


def loss(y, X, T, fmask, neq, M=None, beta=0.0):
    """(sum_free e^2 + beta sum_nodes |e - M e|^2) / neq and d loss / d y."""
    e = np.zeros_like(y)
    for c in range(3):
        free = ((fmask >> c) & 1).astype(bool)
        e[free, c] = y[free, c] + X[free, SKIP + c] - T[free, c]
    val = (e * e).sum()
    dy = 2.0 * e
    if beta:
        d = e - M @ e
        val += beta * (d * d).sum()
        g = 2.0 * beta * (d - M.T @ d)
        for c in range(3):
            free = ((fmask >> c) & 1).astype(bool)
            dy[free, c] += g[free, c]
    return float(val / neq), dy / neq


def backward(P, M, cache, dy):
    """Gradient of the loss with respect to the flat parameter vector."""
    W = unpack(P)
    G = {}
    d1 = cache["d1"]
    h = cache["h"]
    G["w_d2"] = cache["e"].T @ dy
    G["b_d2"] = dy.sum(axis=0)
    de = dy @ W["w_d2"].T
    for _ in range(GS):
        de = M.T @ de
    dz = de * (d1 > 0)
    G["w_d1"] = h[GL].T @ dz
    G["b_d1"] = dz.sum(axis=0)
    dh = dz @ W["w_d1"].T
    for l in range(GL - 1, -1, -1):
        dzl = dh * (cache["a"][l] > 0)
        G["ws%d" % l] = h[l].T @ dzl
        G["wn%d" % l] = cache["ag"][l].T @ dzl
        G["bl%d" % l] = dzl.sum(axis=0)
        dh = dh + dzl @ W["ws%d" % l].T + M.T @ (dzl @ W["wn%d" % l].T)
    dz0 = dh * (h[0] > 0)
    G["w_in"] = cache["X"].T @ dz0
    G["b_in"] = dz0.sum(axis=0)
    return pack(G)


def loss_grad(P, X, M, T, fmask, neq, beta=0.0):
    y, cache = forward(P, X, M)
    l, dy = loss(y, X, T, fmask, neq, M, beta)
    return l, backward(P, M, cache, dy), y


def predict(P, X, M, dlam, out_scale, fmask):
    y, _ = forward(P, X, M)
    du = np.zeros((X.shape[0], 3))
    for c in range(3):
        free = ((fmask >> c) & 1).astype(bool)
        du[free, c] = dlam * out_scale * (y[free, c] + X[free, SKIP + c])
    return du
