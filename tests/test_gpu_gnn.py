"""The learned warm start on the device (fgpu_gnn_*), through the C-ABI:
forward, loss and hand-written backward against the CPU FP64 oracle
(oracle/gnn_oracle.py), device training, and what the trained network gains
in the Newton driver over the same driver without it.

Tolerances: the dense layers take FP16 operands (unit roundoff 2^-11 =
4.9e-4) with FP32 accumulation through seven products in a row, so the loss,
the gradient and the prediction are held to 2e-2 normwise relative error
against FP64; the Newton solves with and without the network must reach the
same converged states to 1e-6.
"""
import numpy as np
import pytest

from fast_amd import mesh as M
from oracle import gnn_oracle as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from fast_amd import fgpu
    fgpu.lib()
    return fgpu


def nrel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def masks(m):
    bits = np.array([1, 2, 4], dtype=np.uint8)
    fmask = ((~m.fixed).astype(np.uint8) * bits).sum(1).astype(np.uint8)
    lmask = ((m.fext != 0).astype(np.uint8) * bits).sum(1).astype(np.uint8)
    return fmask, lmask


def to_eq(m, u):
    out = np.zeros(m.neq)
    d = m.dofmap.ravel()
    k = d > 0
    out[d[k] - 1] = u.ravel()[k]
    return out


def cube(n):
    return M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                        traction=(0.0, 0.0, -1.0))


def fields(m, seed):
    rng = np.random.default_rng(seed)
    X = m.coords
    u = 0.03 * np.stack([np.sin(2 * X[:, 0] + X[:, 1]), X[:, 0] * X[:, 2], -X[:, 0] ** 2], 1)
    ut = u + 0.004 * np.stack([X[:, 0], X[:, 1] * X[:, 0], -2 * X[:, 0] ** 2], 1)
    ut += 1e-4 * rng.standard_normal(ut.shape)
    u[m.fixed] = 0
    ut[m.fixed] = 0
    return u, ut


SC = (0.05, 0.3, 0.2)  # u_scale, lam_scale, out_scale


@pytest.mark.parametrize("case,beta", [("history", 0.0), ("no_history", 0.0), ("unloaded", 0.0),
                                       ("history", 20.0)])
def test_loss_and_gradient_match_oracle(fg, case, beta):
    import torch
    m = cube(6)
    fmask, lmask = masks(m)
    model = fg.GpuModel(m)
    net = fg.Gnn(model, lmask, seed=3)
    net.set_scales(*SC)
    net.set_loss(beta)
    u, ut = fields(m, 0)
    uh = 0.7 * u
    lam, dlam, lamh = 0.12, 0.02, 0.08
    if case == "unloaded":
        u[:] = 0
        lam = 0.0
    du = torch.from_numpy(to_eq(m, u)).cuda()
    dut = torch.from_numpy(to_eq(m, ut)).cuda()
    duh = torch.from_numpy(to_eq(m, uh)).cuda()
    torch.cuda.synchronize()
    hist = case == "history"
    loss, grad = net.loss_grad(None if case == "unloaded" else du.data_ptr(), lam, dlam,
                               dut.data_ptr(), duh.data_ptr() if hist else None, lamh)
    P = net.params().astype(np.float64)
    X = G.features(m.coords, u, lam, dlam, uh if hist else None, lamh, fmask, lmask,
                   SC[0], SC[1], SC[2])
    Mg = G.graph(m.conn, m.nnode)
    T = G.target(u, ut, dlam, SC[2], fmask)
    l_ref, g_ref, _ = G.loss_grad(P, X, Mg, T, fmask, m.neq, beta)
    assert abs(loss - l_ref) <= 2e-2 * l_ref, (loss, l_ref)
    assert nrel(grad.astype(np.float64), g_ref) <= 2e-2, nrel(grad, g_ref)
    # every parameter block receives its gradient (a dropped term shows here)
    L = G.layout()
    for k, (o, n) in ((k, v) for k, v in L.items() if k != "total"):
        gb, rb = grad[o:o + n].astype(np.float64), g_ref[o:o + n]
        if np.linalg.norm(rb) > 1e-6 * np.linalg.norm(g_ref):
            assert nrel(gb, rb) <= 5e-2, (k, nrel(gb, rb))


def test_prediction_matches_oracle(fg):
    import torch
    m = cube(5)
    fmask, lmask = masks(m)
    model = fg.GpuModel(m)
    net = fg.Gnn(model, lmask, seed=7)
    net.set_scales(*SC)
    # parameters from the oracle side: the layouts must agree
    rng = np.random.default_rng(1)
    P = (0.15 * rng.standard_normal(G.layout()["total"])).astype(np.float32)
    net.set_params(P)
    assert np.array_equal(net.params(), P)
    u, _ = fields(m, 1)
    uh = 0.8 * u
    lam, dlam, lamh = 0.2, 0.03, 0.16
    du = torch.from_numpy(to_eq(m, u)).cuda()
    duh = torch.from_numpy(to_eq(m, uh)).cuda()
    out = torch.zeros(m.neq, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    net.predict(du.data_ptr(), lam, dlam, out.data_ptr(), duh.data_ptr(), lamh)
    X = G.features(m.coords, u, lam, dlam, uh, lamh, fmask, lmask, SC[0], SC[1], SC[2])
    ref = G.predict(P.astype(np.float64), X, G.graph(m.conn, m.nnode), dlam, SC[2], fmask)
    assert nrel(out.cpu().numpy(), to_eq(m, ref)) <= 2e-2


def test_training_and_warm_start_gain(fg):
    """Continuation data from the device Newton driver, device training, then
    the same load path solved with and without the trained predictor."""
    import torch
    m = cube(8)
    fmask, lmask = masks(m)
    model = fg.GpuModel(m)
    q = m.load_vector()
    rng = np.random.default_rng(0)
    lams = np.sort(rng.uniform(0.0, 0.3, 200))
    states = torch.zeros((len(lams), m.neq), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    us, st, err = model.newton(q, steps=len(lams), lambdas=lams, maxit=20, tol=1e-8,
                               pcg_tol=1.0, adaptive_pcg=2, secant=True,
                               u_steps_dev=states.data_ptr(), keep_host=False)
    assert err == "" and all(s["converged"] for s in st), err
    umax = float(states[-1].abs().max())
    net = fg.Gnn(model, lmask, seed=0)
    net.set_scales(umax, 0.3, umax / 0.3)
    hist = net.train(states.data_ptr(), lams, nsteps=1500, maxgap=64, lr=2e-3, seed=0)
    assert np.all(np.isfinite(hist))
    assert hist[-100:].mean() < 0.05 * hist[:20].mean(), (hist[:20].mean(), hist[-100:].mean())
    # held-out load path: 10 equal increments to 0.3
    kw = dict(steps=10, factor=0.03, maxit=20, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2)
    u0, s0, e0 = model.newton(q, **kw)
    u1, s1, e1 = model.newton(q, predict=net, **kw)
    assert e0 == "" and e1 == "", (e0, e1)
    for a, b in zip(s0, s1):
        print("step %d plain: iters %d pcg %d err0 %.2e | net: iters %d pcg %d err0 %.2e"
              % (a["step"], a["iters"], a["pcg_iters"], a["err_start"], b["iters"],
                 b["pcg_iters"], b["err_start"]))
    print("loss %.3e -> %.3e" % (hist[:20].mean(), hist[-100:].mean()))
    assert nrel(u1[-1], u0[-1]) <= 1e-6
    assert all(s["predicted"] == 1 for s in s1)
    # the network's start saves PCG iterations over the plain start
    assert sum(s["pcg_iters"] for s in s1) < sum(s["pcg_iters"] for s in s0)


def test_newton_driver_options(fg):
    """Load-factor lists, restarts from a device state, the secant predictor
    and the rejection of a prediction that raises the residual all reach the
    converged states of the plain fixed-increment run."""
    import torch
    m = cube(5)
    model = fg.GpuModel(m)
    q = m.load_vector()
    base = dict(maxit=30, tol=1e-10, pcg_tol=1e-12)
    uA, sA, eA = model.newton(q, steps=4, factor=0.05, **base)
    assert eA == ""
    uB, sB, eB = model.newton(q, steps=4, lambdas=[0.05, 0.1, 0.15, 0.2], secant=True,
                              maxit=30, tol=1e-10, pcg_tol=1.0, adaptive_pcg=2)
    assert eB == "" and nrel(uB[-1], uA[-1]) <= 1e-7
    assert [s["predicted"] for s in sB] == [0, 1, 1, 1]
    states = torch.zeros((2, m.neq), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    uC, sC, eC = model.newton(q, steps=2, factor=0.05, u_steps_dev=states.data_ptr(), **base)
    assert np.allclose(states.cpu().numpy(), uC, rtol=0, atol=0)
    uD, sD, eD = model.newton(q, steps=2, factor=0.05, u0_dev=states[1].data_ptr(), lam0=0.1,
                              **base)
    assert eD == "" and nrel(uD[-1], uA[-1]) <= 1e-8
    assert abs(sD[-1]["lam"] - 0.2) < 1e-15
    for p in (1, 2, 3, 4):  # polynomial extrapolation through the kept states
        uX, sX, eX = model.newton(q, steps=4, factor=0.05, extrap=p, maxit=30, tol=1e-10,
                                  pcg_tol=1.0, adaptive_pcg=2)
        assert eX == "" and nrel(uX[-1], uA[-1]) <= 1e-7, (p, eX)
        assert [s["predicted"] for s in sX] == [0, 1, 1, 1]
    ones = torch.ones(m.neq, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    L = fg.lib()

    def bad(step, u_dev, lam_prev, lam, uh_dev, lam_h, du_dev):  # du = K 1: far off
        return bool(L.fgpu_spmv_dev(model.h, ones.data_ptr(), du_dev))

    uE, sE, eE = model.newton(q, steps=4, factor=0.05, predict=bad, **base)
    assert eE == "" and nrel(uE[-1], uA[-1]) <= 1e-8
    assert all(s["predicted"] == -1 for s in sE)


def test_extrapolation_predictors_save_work(fg):
    """The load path of the warm-start benchmark on a 16^3 block: starting
    each step from the polynomial through the last converged states reaches
    the plain run's states with fewer Newton and PCG iterations; the order-4
    polynomial beats the secant, which beats the plain start."""
    m = cube(16)
    model = fg.GpuModel(m)
    q = m.load_vector()
    kw = dict(steps=10, factor=0.03, maxit=30, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2)
    runs = {}
    for name, extra in (("plain", {}), ("secant", {"secant": True}), ("extrap2", {"extrap": 2}),
                        ("extrap4", {"extrap": 4})):
        u, st, err = model.newton(q, **kw, **extra)
        assert err == "", (name, err)
        runs[name] = (u[-1], sum(s["iters"] for s in st), sum(s["pcg_iters"] for s in st))
        print(name, runs[name][1:])
    for name in ("secant", "extrap2", "extrap4"):
        assert nrel(runs[name][0], runs["plain"][0]) <= 1e-6
    assert runs["extrap4"][2] < runs["secant"][2] < runs["plain"][2]
    assert runs["extrap4"][1] <= runs["secant"][1] <= runs["plain"][1]


def test_block_jacobi_newton(fg):
    """Node-block Jacobi PCG inside the load-control Newton driver: the same
    converged states as point Jacobi (the reference's preconditioner) with
    fewer PCG iterations on the 16^3 neo-Hookean block."""
    m = cube(16)
    model = fg.GpuModel(m)
    q = m.load_vector()
    kw = dict(steps=6, factor=0.03, maxit=30, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2, extrap=4)
    out = {}
    for kind in (0, 1):
        model.set_preconditioner(kind)
        u, st, err = model.newton(q, **kw)
        assert err == "", (kind, err)
        out[kind] = (u[-1], sum(s["iters"] for s in st), sum(s["pcg_iters"] for s in st))
        print("preconditioner", kind, out[kind][1:])
    assert nrel(out[1][0], out[0][0]) <= 1e-6
    assert out[1][2] < out[0][2]
    assert out[1][1] <= out[0][1] + 2
