"""Benchmark: element integration + assembly throughput on one MI355X.

Metric: elements/s over one step = global stiffness (element K_e + the
deterministic assembly into the FP64 block matrix) + internal force f_int
(FP64), with inputs resident in HBM.

Default workload (the largest configuration that fits one GPU): 1024
independent hex8 elastic unit cubes, edge n ~ U{4..32}, nu ~ U(0.2, 0.45),
E = 1, one clamped face and pressure 1 on another, all from seed 0 (10.33 M
elements, 33.6 M equations), 2x2x2 Gauss, one launch per phase for the whole
batch.  Beside the headline the same JSON line carries, each measured in this
run: the batch's per-mesh PCG solve to 1e-8, the 10^3 hex8 linear-static
solve (configuration 1), the cubic B-spline
32x32x8 patch step, the 64^3 neo-Hookean hex8 step
and its 20-increment Newton solve, the learned warm start's solve-time ratios
(plain / secant / polynomial extrapolation of orders 2 and 4 / network), and the relative errors of K.u, f_int and u
against the FP64 oracle for every configuration.

``--workload iga|iga_svk|hex8|batch`` runs one step workload alone (iga_svk:
the cubic patch with total-Lagrangian St.Venant-Kirchhoff elements);
``warmstart``, ``newton`` and ``copy`` run those measurements alone.

Prints ONE JSON line (rank 0).  Multi-GPU runs are not part of this build.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_MFMA = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_FP64 = 78.6        # TFLOP/s, MI355X FP64 vector
PEAK_HBM = 8000.0       # GB/s

STEP_WORKLOADS = {"batch": "hex8_batch_1024_cubes", "iga": "iga_cubic_bspline_32x32x8",
                  "hex8": "hex8_neohooke_64x64x64",
                  "iga_svk": "iga_cubic_bspline_32x32x8_tl_svk"}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def smooth_field(X):
    return 1e-3 * np.stack([np.sin(3 * X[:, 0]) * X[:, 2], np.cos(2 * X[:, 1]) * X[:, 0],
                            X[:, 0] * X[:, 1]], 1)


def make_workload(name):
    from fast_amd import mesh as M
    prm = None
    if name in ("iga", "iga_svk"):
        form = "tl_svk" if name == "iga_svk" else "linear"
        m = M.bsp_block(ne=(32, 32, 8), p=3, perturb=0.05, seed=0, E=1.0, nu=0.3, form=form)
        cfg = {"workload": STEP_WORKLOADS[name], "elements": m.nelem, "nqp": 64,
               "ndof_elem": 192, "neq": int(m.neq), "E": 1.0, "nu": 0.3, "gauss": "4x4x4",
               "perturb": "U(-0.05,0.05)*h seed 0", "form": form}
        u = smooth_field(m.coords)
    elif name == "batch":
        m, prm = M.cube_batch(1024, nmin=4, nmax=32, seed=0)
        cfg = {"workload": STEP_WORKLOADS[name], "meshes": 1024, "elements": m.nelem,
               "nqp": 8, "ndof_elem": 24, "neq": int(m.neq), "edge_n": "U{4..32} seed 0",
               "nu": "U(0.2,0.45)", "E": 1.0, "gauss": "2x2x2"}
        u = smooth_field(m.coords)
    else:
        n = 64
        m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                         traction=(0.0, 0.0, -0.2))
        cfg = {"workload": STEP_WORKLOADS[name], "elements": m.nelem, "nqp": 8,
               "ndof_elem": 24, "neq": int(m.neq), "mu": 1.0, "lambda": 1.5, "gauss": "2x2x2"}
        X = m.coords
        u = np.stack([0.02 * X[:, 0] * X[:, 2], 0.01 * X[:, 1] ** 2, -0.03 * X[:, 0] ** 2], 1)
    u[m.fixed] = 0.0
    return m, cfg, u, prm


# ---------------------------------------------------------------- rooflines
# weighted products per stored block and point of the B-spline stiffness
# kernels: linear 9 (P_ij = sum_q c_q G_i[q][a] G_j[q][b]); total Lagrangian
# St.Venant-Kirchhoff 30 (9 of F g, 6 x 3 of (g.g) F F^T, 3 of g^T S g) and
# neo-Hooke 21 (9 + 9 of F^-T g, 3 of g'^T F F^T g')
BSP_CHAINS = {"linear": 9, "tl_svk": 30, "tl_neohooke": 21}


def kernel_flops(m):
    """Algorithmic flops of the B-spline element-stiffness kernel: its
    weighted sums for every stored node block a <= b (2080), 64 points,
    multiply + add."""
    if m.shape == "bsp":
        return 2080 * BSP_CHAINS[m.form] * 64 * 2
    return None


def bsp_executed_flops(m):
    """What the B-spline kernel issues on MFMA: 10 upper 16x16 tile pairs x
    its chains, 16x16 outputs, K = 64 (full tiles on the diagonal)."""
    return 10 * BSP_CHAINS[m.form] * 16 * 16 * 64 * 2


def picked_mode(report):
    mt = re.search(r"picked assembly mode (\d+)", report or "")
    return int(mt.group(1)) if mt else -1


def assembly_bytes(m, nbf, nblk):
    """Algorithmic HBM bytes of one assembly: every stored element block read
    once (9 FP32 + its 4-byte contribution code), per upper node block its
    4-byte segment pointer, its 1-byte direction mask and its two 4-byte
    positions, every stored global block written once (9 FP64)."""
    nsym = m.nn * (m.nn + 1) // 2
    return m.nelem * nsym * (36 + 4) + nblk * (4 + 1 + 8) + nbf * 72


def measured_traffic(workload, kernel):
    """HBM bytes per assembly launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated), when that
    summary measured the same assembly kernel."""
    path = os.path.join(ROOT, "profiles", "assembly_counters_%s.md" % workload)
    if not os.path.exists(path):
        return None
    text = open(path).read()
    mk = re.search(r"Dominant kernel of the step: (?:void )?(\w+)", text)
    if not mk or mk.group(1) != kernel:
        return None
    mt = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(mt.group(1)) if mt else None


def element_roofline(m, tk, report=""):
    fl = kernel_flops(m)
    if fl:
        achieved = fl * m.nelem / (tk * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP32_MFMA,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_MFMA, 4), "traffic": None,
                "kernel": "bsp_K_kernel" if m.form == "linear" else "bsp_Ktm_kernel",
                "flops_per_elem": fl,
                "executed_frac": round(bsp_executed_flops(m) * m.nelem / (tk * 1e-3) / 1e12
                                       / PEAK_FP32_MFMA, 4),
                "kernel_ms": round(float(tk), 4)}
    # hex8: element kernel bound by its FP32 element-matrix stores
    bytes_per = 36 * 9 * 4
    achieved = bytes_per * m.nelem / (tk * 1e-3) / 1e9
    lane = re.search(r"picked hex8_K elements/workgroup 1\b", report or "")
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4), "traffic": None,
            "kernel": "hex8_K_lane_kernel" if lane else "hex8_K_kernel",
            "bytes_per_elem": bytes_per, "kernel_ms": round(float(tk), 4)}


def assembly_roofline(m, ta, cfg, counts, report=""):
    b = assembly_bytes(m, *counts)
    achieved = b / (ta * 1e-3) / 1e9
    kernel = "assemble_rows_kernel" if picked_mode(report) == 8 else "assemble_kernel"
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4),
            "traffic": measured_traffic(cfg["workload"], kernel),
            "kernel": kernel, "algorithmic_bytes": int(b), "kernel_ms": round(float(ta), 4)}


def fint_roofline(m, tf):
    """f_int (element vectors + gather): HBM-bound; algorithmic bytes = per
    element its connectivity (4 B per node), the nodal coordinates and
    displacements it reads (2 x 24 B per node) and its element vector written
    and read back (2 x 24 B per node); per node its adjacency (4 B per
    element-node) and its 24-byte result."""
    per_elem = m.nn * (4 + 48 + 48 + 4)
    b = m.nelem * per_elem + m.nnode * 24
    achieved = b / (tf * 1e-3) / 1e9
    rec = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
           "frac": round(achieved / PEAK_HBM, 4), "algorithmic_bytes": int(b),
           "kernel_ms": round(float(tf), 4)}
    if m.shape == "bsp":
        # B-spline f_int is FP64 arithmetic: per point grad u = sum_a u_a (x) g_a
        # (9 multiply-adds per node), then the forces f_a = sum_q c sigma g_a by
        # sum factorization (per component 4 + 3 + 2 contractions of 4 points
        # over the 4 x 4 x 4 tensor grid); the basis gradients' evaluation not
        # counted
        nq = int(np.prod(m.nip))
        grid = 4 * 4 * (m.nn // 16)
        fl = m.nelem * 2 * (nq * m.nn * 9 + 3 * (4 + 3 + 2) * 4 * grid)
        t = fl / (tf * 1e-3) / 1e12
        rec = {"bound": "fp64", "achieved": round(t, 2), "peak": PEAK_FP64, "unit": "TFLOP/s",
               "frac": round(t / PEAK_FP64, 4), "flops_per_elem": fl // m.nelem,
               "hbm_GBs": round(achieved, 1), "algorithmic_bytes": int(b),
               "kernel_ms": round(float(tf), 4)}
    return rec


# ---------------------------------------------------------------- CPU baselines
def host_info():
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def port_threads():
    n = os.cpu_count() or 1
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def port_baseline(sample, what):
    """The oracle (C restatement of the reference loops) on the reference's
    sample: K_e + assembly into its CSR + f_int of every element, one thread
    and all the host threads the run may use."""
    from oracle import oracle as O
    om = O.OracleMesh(sample)
    om.pattern()
    u = smooth_field(sample.coords)
    u[sample.fixed] = 0.0
    ncores = port_threads()
    res = {}
    for nt in (1, ncores):
        t0 = time.perf_counter()
        om.assemble_K(u, nthreads=nt)
        om.assemble_fint(u, nthreads=nt)
        dt = time.perf_counter() - t0
        res[nt] = (sample.nelem / dt, dt)
    v, dt = res[ncores]
    return {"value": round(v, 2), "unit": "elements/s", "cores": ncores, "kind": "port",
            "sample": "oracle K_e + assembly + f_int on %s (%.2f s)" % (what, dt),
            "single_thread": round(res[1][0], 2)}


def reference_sample(name, prm, target):
    """A smaller instance of the workload for the reference's CPU run: the
    first cubes of the batch (same edges, Poisson ratios and faces) up to
    about ``target`` elements; an 8x8x8 patch of the cubic B-spline
    discretization."""
    from fast_amd import mesh as M
    if name == "batch":
        cubes, tot = [], 0
        for k in range(len(prm["n"])):
            n = int(prm["n"][k])
            cubes.append(M.hex8_cube(n, E=1.0, nu=float(prm["nu"][k]),
                                     clamp_face=int(prm["clamp_face"][k]),
                                     load_face=int(prm["load_face"][k])))
            tot += n ** 3
            if tot >= target:
                break
        return M.union(cubes), "the first %d cubes of the batch (%d hex8 elements, linear " \
                               "BRICK8, 2x2x2 Gauss)" % (len(cubes), tot)
    if name == "iga":
        s = M.bsp_block(ne=(8, 8, 8), p=3, perturb=0.05, seed=0)
        return s, "an 8x8x8 cubic B-spline patch of the same generator (512 elements, 4x4x4 Gauss)"
    return None, None


def cpu_baselines(name, prm, target=300000):
    """The reference's own loops on a sample of the workload, with the
    oracle on the same sample beside them; the oracle alone where the
    reference build is absent."""
    sample, what = reference_sample(name, prm, target)
    if sample is None:
        return None
    port = port_baseline(sample, what)
    ref = reference_baseline(sample, what)
    if ref is None:
        out = port
    else:
        out = ref
        out["port"] = port
    out["host"] = host_info()
    return out


def reference_baseline(sample, what):
    """The reference's own code on the CPU: oracle/_ref/ref_probe (linked from
    the reference's object files, built from its sources as it ships: -O3, one
    thread) times cControl::StiffnessMatrix (element StiffMat + AddGlobMat into
    its sparse system matrix, ctrl/ctrl.cpp:886-975) and
    cControl::IntForceVector (ctrl/ctrl.cpp:719-813) on a smaller instance of
    the same workload.  None where the build is missing (it is compiled in
    the container that has the reference sources) or the workload's model is
    not the reference's (neo-Hookean)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_probe")
    if not os.path.exists(exe):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from gen_golden import write_deck
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, "sample")
        write_deck(base + ".dat", sample, matrix="sparse")
        try:
            out = subprocess.run([exe, base, "-", "--time", "1"], capture_output=True, text=True,
                                 timeout=300, check=True).stdout
        except (subprocess.SubprocessError, OSError) as err:
            log("reference baseline failed (%s); the oracle's stands" % err)
            return None
    f = out.split()
    nelem, tk, tf = int(f[1]), float(f[3]), float(f[5])
    return {"value": round(nelem / (tk + tf), 2), "unit": "elements/s", "cores": 1,
            "kind": "reference",
            "sample": "the reference's cControl::StiffnessMatrix + IntForceVector (its own code, "
                      "built from its sources, one thread as shipped) on %s, sparse system "
                      "matrix: %.2f s + %.2f s" % (what, tk, tf),
            "stiffness_s": tk, "int_force_s": tf}


# ---------------------------------------------------------------- accuracy
def accuracy_Ku(model, m, u, nsample=32):
    """Relative error of K_e u_e (FP32 element matrices) against the FP64
    oracle on a sample of elements: normwise over the sample
    (sqrt(sum |dK_e u_e|^2 / sum (|K_e|_F |u_e|)^2)), per element normwise
    (worst |dK_e u_e| / (|K_e|_2 |u_e|)), and forward (worst
    |dK_e u_e| / |K_e u_e|, which cancellation in K_e u_e inflates)."""
    from oracle import oracle as O
    om = O.OracleMesh(m)
    idx = np.linspace(0, m.nelem - 1, nsample).astype(int)
    num = den = 0.0
    worst_fw = worst_el = 0.0
    for e in idx:
        K0, _ = om.elem_stiff(int(e), u if m.form != "linear" else None)
        Ke = model.element_matrices(int(e), 1)[0].astype(np.float64)
        ue = u[m.conn[e]].ravel()
        d = (Ke - K0) @ ue
        num += d @ d
        den += (np.linalg.norm(K0) * np.linalg.norm(ue)) ** 2
        worst_fw = max(worst_fw, np.linalg.norm(d) / max(np.linalg.norm(K0 @ ue), 1e-300))
        worst_el = max(worst_el, np.linalg.norm(d) /
                       max(np.linalg.norm(K0, 2) * np.linalg.norm(ue), 1e-300))
    return {"Ku_normwise_rel": float(np.sqrt(num / den)),
            "Ku_elementwise_normwise_max": float(worst_el),
            "Ku_forward_rel_max": float(worst_fw)}


def accuracy_fint(g_dev, m, u, target=300000):
    """Relative error of the device's global f_int against the oracle's FP64
    assembly (all host threads): the whole mesh, or for a batch the first
    meshes up to about ``target`` elements (meshes are independent, so their
    equations receive nothing from the others)."""
    from oracle import oracle as O
    ne, neq = m.nelem, m.neq
    if m.batch_nodeptr is not None and m.nelem > target:
        nptr = np.asarray(m.batch_nodeptr)
        # mesh of each element (elements are stored mesh by mesh)
        mesh_of = np.searchsorted(nptr, m.conn[:, 0], side="right") - 1
        k = max(1, int(mesh_of[target]))  # meshes 0..k-1 lie inside the sample
        ne = int(np.searchsorted(mesh_of, k))
        neq = int(m.dofmap[:nptr[k]].max())
    g0, _ = O.OracleMesh(m).assemble_fint(u, nthreads=port_threads(), elem_range=(0, ne))
    d = g_dev[:neq] - g0[:neq]
    return float(np.linalg.norm(d) / max(np.linalg.norm(g0[:neq]), 1e-300))


def accuracy_u_batch(prm, ks=(None,)):
    """Linear static solutions of single cubes of the batch on the device
    (FP32 element matrices, FP64 residual refinement to 1e-10) against the
    oracle's FP64 skyline solve: the smallest and a middle-sized cube."""
    from fast_amd import fgpu, mesh as M
    from oracle import oracle as O
    ns = np.asarray(prm["n"])
    order = np.argsort(ns, kind="stable")
    pick = [int(order[0]), int(order[len(order) // 4])]
    worst = 0.0
    info = []
    for k in pick:
        c = M.hex8_cube(int(ns[k]), E=1.0, nu=float(prm["nu"][k]),
                        clamp_face=int(prm["clamp_face"][k]), load_face=int(prm["load_face"][k]))
        f = c.load_vector()
        g = fgpu.GpuModel(c)
        ug, _, _ = g.linear_static(f, tol=1e-10)
        g.close()
        u0 = O.linear_static(O.OracleMesh(c), f, solver="skyline")
        e = float(np.linalg.norm(ug - u0) / np.linalg.norm(u0))
        worst = max(worst, e)
        info.append({"cube": k, "n": int(ns[k]), "u_rel": e})
    return worst, info


def accuracy_u_newton():
    """Config-3 model at 10^3: the device Newton (FP32 element tangents, FP64
    residuals and PCG) against the oracle's FP64 Newton, 20 increments to the
    tip traction 0.2, both to |r| / max(1, |f|) <= 1e-9."""
    from fast_amd import fgpu, mesh as M
    from oracle import oracle as O
    n = 10
    m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                     traction=(0.0, 0.0, -0.2))
    q = m.load_vector()
    g = fgpu.GpuModel(m)
    us, st, err = g.newton(q, steps=20, factor=0.05, maxit=30, tol=1e-9, pcg_tol=1e-12)
    g.close()
    if err:
        raise RuntimeError("accuracy newton: " + err)
    uo, _ = O.newton_raphson(O.OracleMesh(m), q, steps=20, factor=0.05, maxit=30, tol=1e-9,
                             solver="pcg", pcg_tol=1e-12)
    return float(np.linalg.norm(us[-1] - uo[-1]) / np.linalg.norm(uo[-1]))


# ---------------------------------------------------------------- batch solve
def batch_solve_record(model, m, tol=1e-8):
    """Configuration 4's solve: every mesh of the batch by its own Jacobi-PCG
    to |r| / |b| <= tol inside shared launches (fgpu_solve_batch_dev, device
    vectors), on the FP64 matrix and on its FP32 copy, then on the FP32 copy
    with node-block Jacobi."""
    import ctypes
    import torch
    from fast_amd import fgpu
    L = fgpu.lib()
    nb = L.fgpu_nbatch(model.h)
    b = torch.from_numpy(m.load_vector()).cuda()
    x = torch.zeros_like(b)
    out = {"tol": tol, "meshes": int(nb)}
    for bits, pc in ((64, 0), (32, 0), (32, 1)):
        model.set_matrix_precision(bits)
        model.set_preconditioner(pc)
        it = np.zeros(nb, dtype=np.int32)
        rr = np.zeros(nb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = L.fgpu_solve_batch_dev(model.h, b.data_ptr(), x.data_ptr(), m.neq, tol,
                                   it.ctypes.data_as(ctypes.c_void_p),
                                   rr.ctypes.data_as(ctypes.c_void_p))
        wall = time.perf_counter() - t0
        key = "fp%d_matrix" % bits + ("_node_blocks" if pc else "")
        out[key] = {
            "wall_s": round(wall, 4), "meshes_converged": int(n),
            "pcg_iters_max": int(it.max()), "pcg_iters_mean": round(float(it.mean()), 1),
            "relres_max": float("%.3e" % rr.max())}
        log("batch solve, %s: %s" % (key, out[key]))
    model.set_matrix_precision(64)
    model.set_preconditioner(0)
    del b, x
    return out


# ---------------------------------------------------------------- step bench
def run_step(name, steps, warmup, serial=False, with_accuracy=True):
    """K + f_int + assembly of one state, ``steps`` timed passes; returns the
    bench record (without the CPU baselines)."""
    import torch
    from fast_amd import fgpu
    m, cfg, u, prm = make_workload(name)
    log("workload %s nelem %d neq %d" % (cfg["workload"], m.nelem, m.neq))
    model = fgpu.GpuModel(m)
    L = fgpu.lib()
    model.stiffness(u)                     # uploads u into the context
    d_u = model.device_buffer(2)
    g = torch.zeros(model.neq, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the library runs on its own non-blocking stream

    def step():
        if serial:
            if not L.fgpu_assemble_K_dev(model.h, d_u):
                raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
            if not L.fgpu_int_force_dev(model.h, d_u, g.data_ptr()):
                raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
        # K (element + assembly) and f_int of the state in one call; the
        # context picks the order (f_int beside the assembly or after it)
        elif not L.fgpu_assemble_K_fint_dev(model.h, d_u, g.data_ptr()):
            raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
        return tuple(L.fgpu_last_time_ms(model.h, k) for k in range(3))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for _ in range(steps):
        times.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    times = np.array(times)
    # f_int alone (element vectors + gather, nothing beside it), for its
    # roofline; the step above may run it beside the assembly
    tfa = []
    for _ in range(max(3, min(steps, 10))):
        if not L.fgpu_int_force_dev(model.h, d_u, g.data_ptr()):
            raise fgpu.FgpuError(L.fgpu_last_error(model.h).decode())
        tfa.append(L.fgpu_last_time_ms(model.h, 2))
    tf_alone = float(np.mean(tfa))
    report = model.tuning_report()
    log(report.strip().replace("\n", " | "))
    tk, ta, tf = times.mean(axis=0)
    counts = model.block_counts()
    rooflines = {"element_K": element_roofline(m, tk, report),
                 "assembly": assembly_roofline(m, ta, cfg, counts, report)}
    dom = "assembly" if ta >= tk else "element_K"
    roof = dict(rooflines[dom])
    roof["other"] = rooflines["element_K" if dom == "assembly" else "assembly"]
    overlap = None
    if not serial:
        mm = re.search(r"K then f_int: ([0-9.]+) ms; f_int beside the assembly: ([0-9.]+) ms; "
                       r"f_int element vectors beside the assembly, gather after: ([0-9.]+) ms; "
                       r"f_int element vectors first, gather beside the element K: ([0-9.]+) ms",
                       report)
        if mm:
            t = [float(mm.group(k)) for k in (1, 2, 3, 4)]
            overlap = ["serial", "f_int beside the assembly",
                       "f_int element vectors beside the assembly, gather after",
                       "f_int element vectors first, gather beside the element K"][t.index(min(t))]
    if overlap not in (None, "serial") and dom == "assembly" and "first" not in overlap:
        roof["shares_gpu_with"] = "f_int (second stream)"
    rec = {
        "metric": "elements/s (K+f_int+assembly)",
        "value": round(m.nelem * steps / wall, 1), "unit": "elements/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 element matrices (MFMA for B-spline) / fp64 assembly, f_int",
        "data": "synthetic (seeded generated mesh)", "config": cfg,
        "phase_ms": {"element_K": round(float(tk), 4), "assembly": round(float(ta), 4),
                     "f_int": round(float(tf), 4), "f_int_alone": round(tf_alone, 4)},
        "step_order": overlap or "serial",
        "roofline": roof,
        "f_int_roofline": fint_roofline(m, tf_alone),
    }
    if overlap not in (None, "serial"):
        rec["phase_ms"]["note"] = ("f_int: its span in the step, parts of it beside other "
                                   "kernels on the second stream; f_int_alone: its kernels "
                                   "with the GPU to themselves (the f_int roofline's time)")
    if name == "batch":
        rec["solve"] = batch_solve_record(model, m)
    if with_accuracy:
        acc = accuracy_Ku(model, m, u)
        gh = g.cpu().numpy()
        acc["f_int_rel"] = accuracy_fint(gh, m, u)
        rec["accuracy"] = acc
    model.close()
    del g
    torch.cuda.empty_cache()
    return rec, m, u, prm


# ---------------------------------------------------------------- warm start
def warmstart_record(nstates=2000, train_steps=4000, beta=30.0):
    """Learned warm start on the 64^3 neo-Hookean block: solver-generated
    training data (load levels U(0, 0.3) of a unit traction, seed 0, solved
    by continuation), device training of the graph network, then one load
    path (10 equal steps to 0.3) solved to |r| / max(1, |f|) <= 1e-8 plain,
    with the secant predictor, with the polynomial extrapolation through the
    last 3 and 5 converged states, and with the network (second pass timed)."""
    import torch
    from fast_amd import mesh as M, fgpu
    n = 64
    m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                     traction=(0.0, 0.0, -1.0))
    model = fgpu.GpuModel(m)
    q = m.load_vector()
    bits = np.array([1, 2, 4], dtype=np.uint8)
    lmask = ((m.fext != 0).astype(np.uint8) * bits).sum(1).astype(np.uint8)
    lams = np.sort(np.random.default_rng(0).uniform(0.0, 0.3, nstates))
    log("warm start, neq %d, generating %d states" % (m.neq, nstates))
    states = torch.zeros((nstates, m.neq), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chunk = 250  # continuation in chunks, restarting from the last state
    for c0 in range(0, nstates, chunk):
        c1 = min(nstates, c0 + chunk)
        restart = {} if c0 == 0 else {"u0_dev": states[c0 - 1].data_ptr(), "lam0": lams[c0 - 1]}
        _, st, err = model.newton(q, steps=c1 - c0, lambdas=lams[c0:c1], maxit=20, tol=1e-8,
                                  pcg_tol=1.0, adaptive_pcg=2, secant=True,
                                  u_steps_dev=states[c0].data_ptr(), keep_host=False, **restart)
        if err:
            raise RuntimeError("data generation: " + err)
        log("states %d/%d (%.1f s)" % (c1, nstates, time.perf_counter() - t0))
    t_data = time.perf_counter() - t0
    umax = float(states[-1].abs().max())
    net = fgpu.Gnn(model, lmask, seed=0)
    net.set_scales(umax, 0.3, umax / 0.3)
    net.set_loss(beta)
    log("data %.1f s; training %d steps" % (t_data, train_steps))
    t0 = time.perf_counter()
    hist = net.train(states.data_ptr(), lams, nsteps=train_steps, maxgap=400, lr=1e-3,
                     seed=0, origin_frac=0.1)
    t_train = time.perf_counter() - t0
    del states
    torch.cuda.empty_cache()
    kw = dict(steps=10, factor=0.03, maxit=30, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2)
    res = {}
    last = {}
    ust = torch.zeros((10, m.neq), dtype=torch.float64, device="cuda")
    for rep in range(2):  # the first pass warms up; the second is timed
        for name, extra in (("plain", {}), ("secant", {"secant": True}),
                            ("extrap2", {"extrap": 2}), ("extrap4", {"extrap": 4}),
                            ("network", {"predict": net})):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, s, e = model.newton(q, keep_host=False, u_steps_dev=ust.data_ptr(), **kw, **extra)
            wall = time.perf_counter() - t0
            if e:
                raise RuntimeError("%s: %s" % (name, e))
            res[name] = {"wall_s": round(wall, 4), "newton_iters": sum(x["iters"] for x in s),
                         "pcg_iters": sum(x["pcg_iters"] for x in s),
                         "predictions_used": sum(x["predicted"] == 1 for x in s)}
            last[name] = ust[-1].clone()
            if rep:
                log("%s %s" % (name, res[name]))
    for name in res:  # every start converges to the same state (Newton tolerance 1e-8)
        res[name]["u_rel_vs_plain"] = float("%.3e" % float(
            torch.linalg.norm(last[name] - last["plain"]) / torch.linalg.norm(last["plain"])))
    del ust, last
    k = max(1, len(hist) // 10)
    net.close()
    model.close()
    return {
        "network_over_plain": round(res["network"]["wall_s"] / res["plain"]["wall_s"], 4),
        "secant_over_plain": round(res["secant"]["wall_s"] / res["plain"]["wall_s"], 4),
        "network_over_secant": round(res["network"]["wall_s"] / res["secant"]["wall_s"], 4),
        "extrap2_over_plain": round(res["extrap2"]["wall_s"] / res["plain"]["wall_s"], 4),
        "extrap4_over_plain": round(res["extrap4"]["wall_s"] / res["plain"]["wall_s"], 4),
        "solves": res, "neq": int(m.neq), "train_states": nstates, "train_steps": train_steps,
        "loss_beta": beta, "datagen_s": round(t_data, 2), "train_s": round(t_train, 2),
        "train_ms_per_step": round(1e3 * t_train / train_steps, 3),
        "loss_curve": [float("%.3e" % hist[i:i + k].mean()) for i in range(0, len(hist), k)],
    }


def run_warmstart(args):
    ws = warmstart_record(args.nstates, args.train_steps, args.beta)
    out = {
        "metric": "solve wall time to 1e-8, NN-warm / plain (10 load steps to 0.3)",
        "value": ws["network_over_plain"], "unit": "ratio", "n_gpus": 1, "steps": 10,
        "warmup": 10, "ms_per_step": round(ws["solves"]["network"]["wall_s"] * 100, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 solve / fp16-mfma network",
        "data": "synthetic (seeded generated mesh and load levels)",
        "config": {"workload": "warmstart_hex8_neohooke_64x64x64", "neq": ws["neq"],
                   "train_states": args.nstates, "train_steps": args.train_steps,
                   "loss_beta": args.beta, "network": "4 message-passing layers, hidden 64"},
        "warm_start": ws,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------- config 1
def config1_record():
    """Configuration 1: unit cube of 10^3 hex8 (E = 200e3, nu = 0.3), face x = 0
    clamped, pressure 1.0 on face x = 1.  The device's linear-static solves
    (Jacobi-PCG with FP64 residual refinement to 1e-10, and the band LDL^T with
    refinement) against the oracle's FP64 skyline Crout solve of the same
    system, timed on the host too."""
    from fast_amd import fgpu, mesh as M
    from oracle import oracle as O
    m = M.hex8_cube(10, E=200e3, nu=0.3, clamp_face=0, load_face=1, pressure=1.0)
    f = m.load_vector()
    g = fgpu.GpuModel(m)
    g.linear_static(f, tol=1e-10)  # first calls: tuning, band set-up
    g.linear_static_direct(f)
    t0 = time.perf_counter()
    up, it, rr = g.linear_static(f, tol=1e-10)
    tp = time.perf_counter() - t0
    t0 = time.perf_counter()
    ud, k, rrd = g.linear_static_direct(f)
    td = time.perf_counter() - t0
    g.close()
    t0 = time.perf_counter()
    u0 = O.linear_static(O.OracleMesh(m), f, solver="skyline")
    to = time.perf_counter() - t0
    n0 = float(np.linalg.norm(u0))
    return {"neq": int(m.neq), "elements": int(m.nelem),
            "pcg": {"wall_s": round(tp, 4), "pcg_iters": int(it), "relres": float("%.3e" % rr),
                    "u_rel_vs_fp64": float("%.3e" % (np.linalg.norm(up - u0) / n0))},
            "direct": {"wall_s": round(td, 4), "refinements": int(k),
                       "relres": float("%.3e" % rrd),
                       "u_rel_vs_fp64": float("%.3e" % (np.linalg.norm(ud - u0) / n0))},
            "oracle_skyline_s": round(to, 4)}


# ---------------------------------------------------------------- Newton
def newton_record(variants=("plain_mat32", "secant_mat32", "extrap4_mat32", "extrap4_mat32_bj")):
    """The 64^3 neo-Hookean cantilever under load control: 20 increments to
    the tip traction 0.2, Newton to |r| / max(1, |f|) <= 1e-8 with inexact
    PCG corrections (FP32 element tangents, FP64 f_int and PCG vectors).
    Suffix _bj: node-block Jacobi PCG instead of the reference's point
    Jacobi."""
    import torch
    from fast_amd import mesh as M, fgpu
    n = 64
    m = M.hex8_block(n, n, n, form="tl_neohooke", mat=np.tile((1.0, 1.5), (n ** 3, 1)),
                     traction=(0.0, 0.0, -0.2))
    model = fgpu.GpuModel(m)
    q = m.load_vector()
    kw = dict(steps=20, factor=0.05, maxit=30, tol=1e-8, pcg_tol=1.0, adaptive_pcg=2,
              keep_host=False)
    log("newton, neq %d, 20 increments" % m.neq)
    model.newton(q, **dict(kw, steps=1))  # warm-up: tuning, first launches
    res = {}
    for name in variants:
        extra = {"secant": True} if name.startswith("secant") else {}
        if name.startswith("extrap"):
            extra = {"extrap": int(name[6])}
        model.set_matrix_precision(32 if "mat32" in name else 64)
        model.set_preconditioner(1 if name.endswith("_bj") else 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st, err = model.newton(q, **kw, **extra)
        wall = time.perf_counter() - t0
        if err:
            raise RuntimeError("%s: %s" % (name, err))
        res[name] = {"wall_s": round(wall, 4), "newton_iters": sum(x["iters"] for x in st),
                     "pcg_iters": sum(x["pcg_iters"] for x in st),
                     "ms_stiffness": round(sum(x["ms_stiffness"] for x in st), 2),
                     "ms_fint": round(sum(x["ms_fint"] for x in st), 2),
                     "ms_solve": round(sum(x["ms_solve"] for x in st), 2),
                     "final_relres": float("%.3e" % st[-1]["err"])}
        log("%s %s" % (name, res[name]))
    model.close()
    return {"neq": int(m.neq), "increments": 20, "traction": 0.2, "tol": 1e-8, "runs": res}


def run_newton(args):
    rec = newton_record(("plain", "secant", "extrap4", "plain_mat32", "secant_mat32",
                         "extrap4_mat32", "extrap4_mat32_bj"))
    res = rec["runs"]
    out = {
        "metric": "load-control Newton wall time, 20 increments to traction 0.2",
        "value": res["plain_mat32"]["wall_s"], "unit": "s", "n_gpus": 1, "steps": 20,
        "warmup": 1, "ms_per_step": round(res["plain_mat32"]["wall_s"] * 1e3 / 20, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 element tangents / fp64 f_int and PCG",
        "data": "synthetic (generated mesh)",
        "config": {"workload": "newton_hex8_neohooke_64x64x64", "neq": rec["neq"],
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

    def rate(f, nbytes):  # GB/s of f over args.steps runs
        f()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            f()
        torch.cuda.synchronize()
        return round(nbytes * args.steps / (time.perf_counter() - t1) / 1e9, 1)

    # write-only (fill) and read-only (sum) rates of an 8 GiB buffer
    big = torch.empty((8 << 30) // 8, dtype=torch.float64, device="cuda")
    big.fill_(1.0)
    fill = rate(lambda: big.fill_(2.0), 8 << 30)
    read = rate(lambda: big.sum(), 8 << 30)
    del big
    print(json.dumps({"metric": "device copy bandwidth (read + write)", "value": round(gbs, 1),
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(dt * 1e3, 4), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic", "config": {"workload": "copy_1GiB"},
                      "bytes_per_copy": 2 * (1 << 30), "frac_of_peak": round(gbs / PEAK_HBM, 4),
                      "write_only_fill_8GiB_GBs": fill, "read_only_sum_8GiB_GBs": read}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batch",
                    choices=["iga", "iga_svk", "hex8", "batch", "warmstart", "newton", "copy"])
    ap.add_argument("--nstates", type=int, default=2000, help="warmstart: training states")
    ap.add_argument("--train-steps", type=int, default=4000, help="warmstart: Adam steps")
    ap.add_argument("--beta", type=float, default=30.0, help="warmstart: roughness weight")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the other configurations' measurements")
    ap.add_argument("--serial", action="store_true",
                    help="K, then f_int (default: the context's timed choice)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (multi-GPU runs are not part of "
              "it); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.workload == "warmstart":
        run_warmstart(args)
        return
    if args.workload == "copy":
        run_copy(args)
        return
    if args.workload == "newton":
        run_newton(args)
        return

    t_start = time.perf_counter()
    rec, m, u, prm = run_step(args.workload, args.steps, args.warmup, args.serial)
    if args.workload == "batch":
        worst, info = accuracy_u_batch(prm)
        rec["accuracy"]["u_rel"] = worst
        rec["accuracy"]["u_cases"] = info
    rec["cpu_baseline"] = None if args.no_cpu else cpu_baselines(args.workload, prm)
    del m, u
    if args.workload == "batch" and not args.no_extras:
        extras = {}
        for name in ("iga", "hex8"):
            r, _, _, _ = run_step(name, 10, 2)
            extras[r["config"]["workload"]] = {
                "elements_per_s": r["value"], "ms_per_step": r["ms_per_step"],
                "phase_ms": r["phase_ms"], "step_order": r["step_order"],
                "roofline": r["roofline"], "f_int_roofline": r["f_int_roofline"],
                "accuracy": r["accuracy"]}
        extras["linear_static_hex8_10x10x10"] = config1_record()
        nr = newton_record()
        nr["u_rel_vs_fp64_newton_10cubed"] = accuracy_u_newton()
        extras["newton_hex8_neohooke_64x64x64"] = nr
        extras["warmstart_hex8_neohooke_64x64x64"] = warmstart_record()
        rec["configs"] = extras
    rec["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
