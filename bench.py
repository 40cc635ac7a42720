"""Benchmark: element integration + assembly throughput on one MI355X.

Metric: elements/s over one step = global stiffness (element K_e on MFMA +
deterministic assembly into the FP64 CSR) + internal force f_int (FP64), with
inputs resident in HBM.  Default workload: cubic B-spline solid patch
32x32x8 elements (open uniform knots, control points on the unit box moved by
seeded U(-0.05, 0.05) h), 4x4x4 Gauss, E = 1, nu = 0.3.
``--workload hex8`` runs the 64^3 neo-Hookean hex8 block instead.

``--workload warmstart`` runs the learned warm start on the same 64^3
neo-Hookean block: solver-generated training data (2000 load levels
U(0, 0.3) of a unit traction, seed 0, solved by continuation), device training
of the graph network, then one load path (10 equal steps to 0.3) solved to
|r| / max(1, |f|) <= 1e-8 plain, with the secant predictor and with the
network; the value is the wall-time ratio network / plain.

Prints ONE JSON line (rank 0).  Multi-GPU runs are not part of this build.
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_MFMA = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0       # GB/s


def make_workload(name):
    from fast_amd import mesh as M
    if name == "iga":
        m = M.bsp_block(ne=(32, 32, 8), p=3, perturb=0.05, seed=0, E=1.0, nu=0.3)
        cfg = {"workload": "iga_cubic_bspline_32x32x8", "elements": m.nelem, "nqp": 64,
               "ndof_elem": 192, "neq": int(m.neq), "E": 1.0, "nu": 0.3, "gauss": "4x4x4",
               "perturb": "U(-0.05,0.05)*h seed 0"}
        X = m.coords
        u = 1e-3 * np.stack([np.sin(3 * X[:, 0]) * X[:, 2], np.cos(2 * X[:, 1]) * X[:, 0],
                             X[:, 0] * X[:, 1]], 1)
    elif name == "batch":
        m, prm = M.cube_batch(1024, nmin=4, nmax=32, seed=0)
        cfg = {"workload": "hex8_batch_1024_cubes", "meshes": 1024, "elements": m.nelem,
               "nqp": 8, "ndof_elem": 24, "neq": int(m.neq), "edge_n": "U{4..32} seed 0",
               "nu": "U(0.2,0.45)", "E": 1.0, "gauss": "2x2x2"}
        X = m.coords
        u = 1e-3 * np.stack([np.sin(3 * X[:, 0]) * X[:, 2], np.cos(2 * X[:, 1]) * X[:, 0],
                             X[:, 0] * X[:, 1]], 1)
    else:
        n = 64
        m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                         traction=(0.0, 0.0, -0.2))
        cfg = {"workload": "hex8_neohooke_64x64x64", "elements": m.nelem, "nqp": 8,
               "ndof_elem": 24, "neq": int(m.neq), "mu": 1.0, "lambda": 1.5, "gauss": "2x2x2"}
        X = m.coords
        u = np.stack([0.02 * X[:, 0] * X[:, 2], 0.01 * X[:, 1] ** 2, -0.03 * X[:, 0] ** 2], 1)
    u[m.fixed] = 0.0
    return m, cfg, u


# algorithmic flops of the element-stiffness kernel per element
def kernel_flops(m):
    if m.shape == "bsp":
        # the nine sums P_ij[a][b] = sum_q c_q G_i[q][a] G_j[q][b] for every
        # stored node block a <= b (2080), 64 points, multiply + add
        return 2080 * 9 * 64 * 2
    return None


# what the kernel issues on MFMA: 10 upper 16x16 tile pairs x 9 products,
# 16x16 outputs, K = 64 (full tiles on the diagonal)
BSP_EXECUTED_FLOPS = 10 * 9 * 16 * 16 * 64 * 2


def graph_counts(m):
    """Stored node blocks (all pairs of nodes sharing an element) and upper
    node blocks (A <= B) of the global block matrix."""
    import scipy.sparse as sp
    ne, nn = m.conn.shape
    rows = np.repeat(m.conn, nn, axis=1).ravel()
    cols = np.tile(m.conn, (1, nn)).ravel()
    G = sp.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(m.nnode, m.nnode))
    G.sum_duplicates()
    nbf = int(G.nnz)
    return nbf, (nbf + m.nnode) // 2


def assembly_bytes(m):
    """Algorithmic HBM bytes of one assembly: every stored element block read
    once (9 FP32 + its 4-byte contribution code), per upper node block its
    segment pointer (8 B) and four 4-byte indices (rows, positions), every
    stored global block written once (9 FP64)."""
    nsym = m.nn * (m.nn + 1) // 2
    nbf, nblk = graph_counts(m)
    return m.nelem * nsym * (36 + 4) + nblk * (8 + 16) + nbf * 72


def measured_traffic(workload):
    """HBM bytes per assembly launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    import re
    path = os.path.join(ROOT, "profiles", "assembly_counters_%s.md" % workload)
    if not os.path.exists(path):
        return None
    mt = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(mt.group(1)) if mt else None


def element_roofline(m, tk):
    fl = kernel_flops(m)
    if fl:
        achieved = fl * m.nelem / (tk * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP32_MFMA,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_MFMA, 4), "traffic": None,
                "kernel": "bsp_K_kernel", "flops_per_elem": fl,
                "executed_frac": round(BSP_EXECUTED_FLOPS * m.nelem / (tk * 1e-3) / 1e12
                                       / PEAK_FP32_MFMA, 4),
                "kernel_ms": round(float(tk), 4)}
    # hex8: element kernel bound by its FP32 element-matrix stores
    bytes_per = 36 * 9 * 4
    achieved = bytes_per * m.nelem / (tk * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4), "traffic": None, "kernel": "hex8_K_kernel",
            "bytes_per_elem": bytes_per, "kernel_ms": round(float(tk), 4)}


def assembly_roofline(m, ta, cfg, report=""):
    b = assembly_bytes(m)
    achieved = b / (ta * 1e-3) / 1e9
    # the autotuner's pick: mode 8 is the row kernel (one wave per node row)
    import re
    mt = re.search(r"picked assembly mode (\d+)", report)
    kernel = "assemble_rows_kernel" if mt and int(mt.group(1)) == 8 else "assemble_kernel"
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4), "traffic": measured_traffic(cfg["workload"]),
            "kernel": kernel, "algorithmic_bytes": int(b),
            "kernel_ms": round(float(ta), 4)}


def cpu_baseline(m, u, budget_s=12.0):
    """The oracle (C restatement of the reference loops) on a bounded sample of
    the same workload: K_e + assembly + f_int for the first elements."""
    from oracle import oracle as O
    om = O.OracleMesh(m)
    om.pattern()
    ncores = os.cpu_count() or 1
    res = {}
    for nt in (1, ncores):
        ne = 8
        while True:
            t0 = time.perf_counter()
            om.assemble_K(u, nthreads=nt, elem_range=(0, ne))
            om.assemble_fint(u, nthreads=nt, elem_range=(0, ne))
            dt = time.perf_counter() - t0
            if dt > budget_s / 4 or ne >= m.nelem:
                break
            ne = min(m.nelem, ne * 4)
        res[nt] = (ne / dt, ne, dt)
    return res, ncores


def reference_baseline(cfg, ne=(8, 8, 8)):
    """The reference's own code on the CPU: oracle/_ref/ref_probe (linked from
    the reference's object files, built from its sources as it ships: -O3, one
    thread) times cControl::StiffnessMatrix (element StiffMat + AddGlobMat into
    its sparse system matrix, ctrl/ctrl.cpp:886-975) and
    cControl::IntForceVector (ctrl/ctrl.cpp:719-813) on a smaller patch of the
    same discretization.  None where the build or the workload's model is
    missing from the reference."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_probe")
    if not os.path.exists(exe) or not cfg["workload"].startswith("iga_cubic_bspline"):
        return None
    from fast_amd import mesh as M
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from gen_golden import write_deck
    sample = M.bsp_block(ne=ne, p=3, perturb=0.05, seed=0)
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, "sample")
        write_deck(base + ".dat", sample, matrix="sparse")
        try:
            out = subprocess.run([exe, base, "-", "--time", "1"], capture_output=True, text=True,
                                 timeout=600, check=True).stdout
        except (subprocess.SubprocessError, OSError) as err:
            print("bench: reference baseline failed (%s); the oracle's stands" % err,
                  file=sys.stderr)
            return None
    f = out.split()
    nelem, tk, tf = int(f[1]), float(f[3]), float(f[5])
    return {"value": round(nelem / (tk + tf), 2), "unit": "elements/s", "cores": 1,
            "kind": "reference",
            "sample": "the reference's cControl::StiffnessMatrix + IntForceVector (its own "
                      "code, built from its sources, one thread as shipped) on a %dx%dx%d cubic "
                      "B-spline patch (%d elements, 4x4x4 Gauss, same generator), sparse system "
                      "matrix: %.2f s + %.2f s" % (ne + (nelem, tk, tf)),
            "stiffness_s": tk, "int_force_s": tf}


def accuracy(model, m, u, nsample=32):
    """Relative error of K_e u_e (FP32 path) against the FP64 oracle on a
    sample of elements, normwise and forward."""
    from oracle import oracle as O
    om = O.OracleMesh(m)
    idx = np.linspace(0, m.nelem - 1, nsample).astype(int)
    num = den = 0.0
    worst_fw = 0.0
    for e in idx:
        K0, _ = om.elem_stiff(int(e), u if m.form != "linear" else None)
        Ke = model.element_matrices(int(e), 1)[0].astype(np.float64)
        ue = u[m.conn[e]].ravel()
        d = (Ke - K0) @ ue
        num += d @ d
        den += (np.linalg.norm(K0) * np.linalg.norm(ue)) ** 2
        worst_fw = max(worst_fw, np.linalg.norm(d) / max(np.linalg.norm(K0 @ ue), 1e-300))
    return float(np.sqrt(num / den)), float(worst_fw)


def run_warmstart(args):
    """Learned warm start: data generation, training, timed solves."""
    import torch
    from fast_amd import mesh as M, fgpu
    n = 64
    m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                     traction=(0.0, 0.0, -1.0))
    model = fgpu.GpuModel(m)
    q = m.load_vector()
    bits = np.array([1, 2, 4], dtype=np.uint8)
    lmask = ((m.fext != 0).astype(np.uint8) * bits).sum(1).astype(np.uint8)
    lams = np.sort(np.random.default_rng(0).uniform(0.0, 0.3, args.nstates))
    print("bench: warm start, neq %d, generating %d states" % (m.neq, args.nstates),
          file=sys.stderr, flush=True)
    states = torch.zeros((args.nstates, m.neq), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chunk = 250  # continuation in chunks, restarting from the last state
    for c0 in range(0, args.nstates, chunk):
        c1 = min(args.nstates, c0 + chunk)
        restart = {} if c0 == 0 else {"u0_dev": states[c0 - 1].data_ptr(), "lam0": lams[c0 - 1]}
        _, st, err = model.newton(q, steps=c1 - c0, lambdas=lams[c0:c1], maxit=20, tol=1e-8,
                                  pcg_tol=1.0, adaptive_pcg=2, secant=True,
                                  u_steps_dev=states[c0].data_ptr(), keep_host=False, **restart)
        if err:
            raise RuntimeError("data generation: " + err)
        print("bench: states %d/%d (%.1f s)" % (c1, args.nstates, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    t_data = time.perf_counter() - t0
    umax = float(states[-1].abs().max())
    net = fgpu.Gnn(model, lmask, seed=0)
    net.set_scales(umax, 0.3, umax / 0.3)
    net.set_loss(args.beta)
    print("bench: data %.1f s; training %d steps" % (t_data, args.train_steps), file=sys.stderr,
          flush=True)
    os.environ["FGPU_GNN_VERBOSE"] = "1"
    t0 = time.perf_counter()
    hist = net.train(states.data_ptr(), lams, nsteps=args.train_steps, maxgap=400, lr=1e-3,
                     seed=0, origin_frac=0.1)
    t_train = time.perf_counter() - t0
    del states
    torch.cuda.empty_cache()
    kw = dict(steps=10, factor=0.03, maxit=30, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2)
    res = {}
    for rep in range(2):  # the first pass warms up; the second is timed
        for name, extra in (("plain", {}), ("secant", {"secant": True}), ("network", {"predict": net})):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, s, e = model.newton(q, keep_host=False, **kw, **extra)
            wall = time.perf_counter() - t0
            if e:
                raise RuntimeError("%s: %s" % (name, e))
            res[name] = {"wall_s": round(wall, 4), "newton_iters": sum(x["iters"] for x in s),
                         "pcg_iters": sum(x["pcg_iters"] for x in s),
                         "predictions_used": sum(x["predicted"] == 1 for x in s)}
            print("bench: %s %s" % (name, res[name]), file=sys.stderr, flush=True)
    ratio = res["network"]["wall_s"] / res["plain"]["wall_s"]
    k = max(1, len(hist) // 10)
    out = {
        "metric": "solve wall time to 1e-8, NN-warm / plain (10 load steps to 0.3)",
        "value": round(ratio, 4), "unit": "ratio", "n_gpus": 1, "steps": 10, "warmup": 10,
        "ms_per_step": round(res["network"]["wall_s"] * 100, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solve / fp16-mfma network",
        "data": "synthetic (seeded generated mesh and load levels)",
        "config": {"workload": "warmstart_hex8_neohooke_64x64x64", "neq": int(m.neq),
                   "train_states": args.nstates, "train_steps": args.train_steps,
                   "loss_beta": args.beta, "network": "4 message-passing layers, hidden 64"},
        "warm_start": {"solves": res, "secant_over_plain": round(res["secant"]["wall_s"] /
                                                                 res["plain"]["wall_s"], 4),
                       "datagen_s": round(t_data, 2), "train_s": round(t_train, 2),
                       "train_ms_per_step": round(1e3 * t_train / args.train_steps, 3),
                       "loss_curve": [float("%.3e" % hist[i:i + k].mean())
                                      for i in range(0, len(hist), k)]},
    }
    print(json.dumps(out))


def run_newton(args):
    """The 64^3 neo-Hookean cantilever under load control: 20 increments to
    the tip traction 0.2, Newton to |r| / |f| <= 1e-8 with inexact PCG
    corrections (FP32 element tangents, FP64 f_int and PCG)."""
    import torch
    from fast_amd import mesh as M, fgpu
    n = 64
    m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                     traction=(0.0, 0.0, -0.2))
    model = fgpu.GpuModel(m)
    q = m.load_vector()
    kw = dict(steps=20, factor=0.05, maxit=30, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2,
              keep_host=False)
    print("bench: newton, neq %d, 20 increments" % m.neq, file=sys.stderr, flush=True)
    model.newton(q, **dict(kw, steps=1))  # warm-up: tuning, first launches
    res = {}
    for name, extra in (("plain", {}), ("secant", {"secant": True}),
                        ("plain_mat32", {}), ("secant_mat32", {"secant": True})):
        model.set_matrix_precision(32 if name.endswith("mat32") else 64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        us, st, err = model.newton(q, **kw, **extra)
        wall = time.perf_counter() - t0
        if err:
            raise RuntimeError("%s: %s" % (name, err))
        res[name] = {"wall_s": round(wall, 4), "newton_iters": sum(x["iters"] for x in st),
                     "pcg_iters": sum(x["pcg_iters"] for x in st),
                     "ms_stiffness": round(sum(x["ms_stiffness"] for x in st), 2),
                     "ms_fint": round(sum(x["ms_fint"] for x in st), 2),
                     "ms_solve": round(sum(x["ms_solve"] for x in st), 2),
                     "final_relres": float("%.3e" % st[-1]["err"])}
        print("bench: %s %s" % (name, res[name]), file=sys.stderr, flush=True)
    # headline: plain Newton (no predictor), PCG on the FP32 copy of the
    # matrix; the FP64-matrix runs reach the same states (same iterations)
    out = {
        "metric": "load-control Newton wall time, 20 increments to traction 0.2",
        "value": res["plain_mat32"]["wall_s"], "unit": "s", "n_gpus": 1, "steps": 20,
        "warmup": 1, "ms_per_step": round(res["plain_mat32"]["wall_s"] * 1e3 / 20, 2),
        "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32-mfma/f64",
        "data": "synthetic (generated mesh)",
        "config": {"workload": "newton_hex8_neohooke_64x64x64", "neq": int(m.neq),
                   "mu": 1.0, "lambda": 1.5, "traction": 0.2, "increments": 20, "tol": 1e-8,
                   "pcg_matrix": "fp32 copy of the fp64 assembly"},
        "newton": res,
    }
    print(json.dumps(out))


def run_copy(args):
    """Device-to-device copy of a 1 GiB FP64 buffer: the practical HBM
    bandwidth, and the calibration of the FETCH_SIZE / WRITE_SIZE counters
    (each copy reads and writes exactly 1 GiB)."""
    import torch
    n = (1 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    gbs = 2 * (1 << 30) / dt / 1e9
    print(json.dumps({"metric": "device copy bandwidth (read + write)", "value": round(gbs, 1),
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(dt * 1e3, 4), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic", "config": {"workload": "copy_1GiB"},
                      "bytes_per_copy": 2 * (1 << 30), "frac_of_peak": round(gbs / PEAK_HBM, 4)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="iga",
                    choices=["iga", "hex8", "batch", "warmstart", "newton", "copy"])
    ap.add_argument("--nstates", type=int, default=2000, help="warmstart: training states")
    ap.add_argument("--train-steps", type=int, default=4000, help="warmstart: Adam steps")
    ap.add_argument("--beta", type=float, default=30.0, help="warmstart: roughness weight")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve", action="store_true", help="also time a plain PCG solve to 1e-8")
    ap.add_argument("--serial", action="store_true",
                    help="K, then f_int (default: f_int beside K on a second stream)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (multi-GPU runs are not part of "
              "it); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    from fast_amd import fgpu
    if args.workload == "warmstart":
        run_warmstart(args)
        return
    if args.workload == "copy":
        run_copy(args)
        return
    if args.workload == "newton":
        run_newton(args)
        return

    m, cfg, u = make_workload(args.workload)
    print("bench: workload %s nelem %d neq %d" % (cfg["workload"], m.nelem, m.neq),
          file=sys.stderr, flush=True)
    model = fgpu.GpuModel(m)
    L = fgpu.lib()
    model.stiffness(u)                     # uploads u into the context
    d_u = model.device_buffer(2)
    g = torch.zeros(model.neq, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the library runs on its own non-blocking stream

    def step():
        if args.serial:
            if not L.fgpu_assemble_K_dev(model.h, d_u):
                raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
            if not L.fgpu_int_force_dev(model.h, d_u, g.data_ptr()):
                raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
        # K (element + assembly) and f_int of the state, f_int on the
        # context's second stream beside the stiffness
        elif not L.fgpu_assemble_K_fint_dev(model.h, d_u, g.data_ptr()):
            raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
        return tuple(L.fgpu_last_time_ms(model.h, k) for k in range(3))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        times.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    times = np.array(times)
    print(model.tuning_report(), file=sys.stderr, flush=True)
    tk, ta, tf = times.mean(axis=0)
    value = m.nelem * args.steps / wall
    rooflines = {"element_K": element_roofline(m, tk),
                 "assembly": assembly_roofline(m, ta, cfg, model.tuning_report())}
    # the dominant kernel of the step (element K or assembly) carries the
    # headline roofline; the other is reported beside it
    dom = "assembly" if ta >= tk else "element_K"
    roof = dict(rooflines[dom])
    roof["other"] = rooflines["element_K" if dom == "assembly" else "assembly"]
    # the context picks the faster order on its first combined call: with f_int
    # beside the assembly, the assembly's launches share the GPU with it and
    # its durations (and so the roofline) are those of the shared run
    overlap = False
    if not args.serial:
        mm = re.search(r"K then f_int: ([0-9.]+) ms; f_int beside the assembly: ([0-9.]+) ms",
                       model.tuning_report())
        overlap = bool(mm) and float(mm.group(2)) < float(mm.group(1))
    if overlap and dom == "assembly":
        roof["shares_gpu_with"] = "f_int (second stream)"
    nrm, fwd = accuracy(model, m, u)
    solve = None
    if args.solve:
        # plain Jacobi-PCG on the tangent at u, right-hand side = reference load
        f = m.load_vector() if m.fext is not None else model.int_force(u)
        b = torch.from_numpy(f).cuda()
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        it = fgpu.C.c_int(0)
        rr = fgpu.C.c_double(0.0)
        model.stiffness(u)
        solve = {}
        # the PCG on the FP64 matrix, then on its FP32 copy (FP64 vectors and
        # dot products either way)
        for bits in (64, 32):
            model.set_matrix_precision(bits)
            x.zero_()
            torch.cuda.synchronize()
            key = "fp%d_matrix" % bits
            if m.batch_nodeptr is not None:
                nb = len(m.batch_nodeptr) - 1
                its = np.zeros(nb, dtype=np.int32)
                rrs = np.zeros(nb)
                nconv = L.fgpu_solve_batch_dev(model.h, b.data_ptr(), x.data_ptr(), 100000, 1e-8,
                                               its.ctypes.data, rrs.ctypes.data)
                ms = L.fgpu_last_time_ms(model.h, 3)
                solve[key] = {"rtol": 1e-8, "meshes": nb, "converged": int(nconv),
                              "iters_max": int(its.max()), "iters_mean": float(its.mean()),
                              "relres_max": float(rrs.max()), "ms": round(ms, 3)}
            else:
                conv = L.fgpu_solve_dev(model.h, b.data_ptr(), x.data_ptr(), None, 100000, 1e-8,
                                        fgpu.C.byref(it), fgpu.C.byref(rr))
                ms = L.fgpu_last_time_ms(model.h, 3)
                solve[key] = {"rtol": 1e-8, "iters": it.value, "relres": rr.value,
                              "converged": bool(conv), "ms": round(ms, 3),
                              "ms_per_iter": round(ms / max(it.value, 1), 4)}
        model.set_matrix_precision(64)
    cpu = None
    if not args.no_cpu:
        res, ncores = cpu_baseline(m, u)
        v, ne, dt = res[ncores]
        v1 = res[1]
        port = {"value": round(v, 2), "unit": "elements/s", "cores": ncores, "kind": "port",
                "sample": "oracle K_e+assembly+f_int over the first %d elements (%.1f s) of the "
                          "same workload" % (ne, dt),
                "single_thread": round(v1[0], 2)}
        # the reference's own code where it builds and has the workload's model;
        # the oracle (a restatement of its loops, threaded) beside it
        cpu = reference_baseline(cfg)
        if cpu is None:
            cpu = port
        else:
            cpu["port"] = port
    out = {
        "metric": "elements/s (K+f_int+assembly)",
        "value": round(value, 1), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32-mfma/f64",
        "data": "synthetic (seeded generated mesh)", "config": cfg,
        "phase_ms": {"element_K": round(float(tk), 4), "assembly": round(float(ta), 4),
                     "f_int": round(float(tf), 4)},
        "f_int_beside_assembly": overlap,
        "accuracy": {"Ku_normwise_rel": nrm, "Ku_forward_rel_max": fwd},
        "roofline": roof, "cpu_baseline": cpu, "solve_plain_pcg": solve,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
