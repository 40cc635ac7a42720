"""Python host mirror of the C-ABI in include/fgpu.h (ctypes).

The class ``GpuModel`` exposes the reference's operator seam for this path:
``stiffness`` (cControl::StiffnessMatrix, ctrl/ctrl.cpp:886-975),
``int_force`` (cControl::IntForceVector, ctrl/ctrl.cpp:719-813) and ``solve``
(cSysMatrix::Solve with PCGSolver, mlib/sysmat.cpp:1061-1125), with the same
1 = ok / 0 = failure convention surfaced as ``FgpuError``.

The native library must be built (``python -m fast_amd.build``); there is no
fallback: importing without it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .mesh import FORMS, SHAPES, Mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfgpu.so")
_LIB = None


class FgpuError(RuntimeError):
    pass


class FgpuMesh(C.Structure):
    _fields_ = [("nnode", C.c_int), ("nelem", C.c_int), ("shape", C.c_int),
                ("nn", C.c_int), ("deg", C.c_int * 3), ("nip", C.c_int * 3),
                ("form", C.c_int), ("coords", C.c_void_p), ("conn", C.c_void_p),
                ("dofmap", C.c_void_p), ("ext", C.c_void_p), ("cpw", C.c_void_p),
                ("mat", C.c_void_p), ("nbatch", C.c_int), ("batch_nodeptr", C.c_void_p)]


PREDICT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double,
                         C.c_void_p, C.c_double, C.c_void_p)


class FgpuSurface(C.Structure):
    _fields_ = [("nface", C.c_int), ("deg", C.c_int * 2), ("nip", C.c_int),
                ("nodes", C.c_void_p), ("ext", C.c_void_p), ("local", C.c_void_p),
                ("q", C.c_void_p), ("shape", C.c_int)]


class NrParams(C.Structure):
    _fields_ = [("steps", C.c_int), ("factor", C.c_double), ("maxit", C.c_int),
                ("tol", C.c_double), ("pcg_maxit", C.c_int), ("pcg_tol", C.c_double),
                ("adaptive_pcg", C.c_int), ("predict", PREDICT_FN), ("predict_user", C.c_void_p),
                ("lambdas", C.c_void_p), ("u0_dev", C.c_void_p), ("lam0", C.c_double),
                ("u_steps_dev", C.c_void_p), ("secant", C.c_int), ("galerkin", C.c_int),
                ("direct", C.c_int), ("graph", C.c_int), ("extrap", C.c_int)]


class NrStat(C.Structure):
    _fields_ = [("step", C.c_int), ("iters", C.c_int), ("pcg_iters", C.c_int),
                ("converged", C.c_int), ("predicted", C.c_int), ("lam", C.c_double),
                ("err", C.c_double), ("err_start", C.c_double),
                ("ms_stiffness", C.c_double), ("ms_fint", C.c_double),
                ("ms_solve", C.c_double), ("ms_total", C.c_double)]


EXPORTS = {
    "fgpu_linear_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                     C.c_void_p, C.c_void_p]),
    "fgpu_newton": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_create": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_destroy": (None, [C.c_void_p]),
    "fgpu_last_error": (C.c_char_p, [C.c_void_p]),
    "fgpu_neq": (C.c_int, [C.c_void_p]),
    "fgpu_nnz": (C.c_longlong, [C.c_void_p]),
    "fgpu_block_counts": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_assemble_K": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_assemble_K_dev": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_int_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_int_force_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_int_force_dev_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_add_diag": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "fgpu_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                             C.c_double, C.c_void_p, C.c_void_p]),
    "fgpu_solve_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_double, C.c_void_p, C.c_void_p]),
    "fgpu_spmv_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_solve_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                   C.c_void_p, C.c_void_p]),
    "fgpu_solve_batch_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_double, C.c_void_p, C.c_void_p]),
    "fgpu_nbatch": (C.c_int, [C.c_void_p]),
    "fgpu_get_csr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_set_csr_values": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_get_element_matrices": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "fgpu_body_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_body_force_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_assemble_K_fint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_assemble_K_fint_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_set_matrix_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_set_preconditioner": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_factor": (C.c_int, [C.c_void_p]),
    "fgpu_linear_static_direct": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                            C.c_void_p, C.c_void_p]),
    "fgpu_factor_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_factor_solve_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_direct_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_direct_solve_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_factor_info": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fgpu_set_stress_nodes": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_num_points": (C.c_int, [C.c_void_p]),
    "fgpu_stresses": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4),
    "fgpu_stresses_dev": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4),
    "fgpu_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fgpu_set_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_add_surface_loads": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_clear_surface_loads": (C.c_int, [C.c_void_p]),
    "fgpu_surface_load": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_surface_load_dev": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_set_penalty": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "fgpu_tuning_report": (C.c_char_p, [C.c_void_p]),
    "fgpu_last_time_ms": (C.c_double, [C.c_void_p, C.c_int]),
    "fgpu_device_buffer": (C.c_void_p, [C.c_void_p, C.c_int]),
    "fgpu_gnn_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint]),
    "fgpu_gnn_destroy": (None, [C.c_void_p]),
    "fgpu_gnn_num_params": (C.c_int, [C.c_void_p]),
    "fgpu_gnn_set_scales": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "fgpu_gnn_set_loss": (C.c_int, [C.c_void_p, C.c_double]),
    "fgpu_gnn_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_gnn_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgpu_gnn_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                     C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "fgpu_gnn_predict": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                   C.c_void_p, C.c_double, C.c_void_p]),
    "fgpu_gnn_train": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                 C.c_double, C.c_float, C.c_uint, C.c_void_p]),
    "fgpu_gnn_newton_predictor": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_double,
                                            C.c_double, C.c_void_p, C.c_double, C.c_void_p]),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise FgpuError("libfgpu.so is not built (run: python -m fast_amd.build)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


class GpuModel:
    """One fgpu context: a mesh resident on the current HIP device."""

    def __init__(self, mesh: Mesh):
        self.mesh = mesh
        L = lib()
        self._keep = []

        def c(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        m = FgpuMesh()
        m.nnode, m.nelem = mesh.nnode, mesh.nelem
        m.shape = SHAPES[mesh.shape]
        m.nn = mesh.nn
        for d in range(3):
            m.deg[d] = int(mesh.deg[d])
            m.nip[d] = int(mesh.nip[d])
        m.form = FORMS[mesh.form]
        m.coords = _p(c(mesh.coords, np.float64))
        m.conn = _p(c(mesh.conn, np.int32))
        m.dofmap = _p(c(mesh.dofmap, np.int32))
        m.mat = _p(c(mesh.mat, np.float64))
        m.ext = _p(c(mesh.ext, np.float64)) if mesh.ext is not None else None
        m.cpw = _p(c(mesh.cpw, np.float64)) if mesh.cpw is not None else None
        if mesh.batch_nodeptr is not None:
            m.nbatch = len(mesh.batch_nodeptr) - 1
            m.batch_nodeptr = _p(c(mesh.batch_nodeptr, np.int64))
        self._m = m
        h = C.c_void_p()
        if not L.fgpu_create(C.byref(h), C.byref(m)):
            raise FgpuError("fgpu_create failed (see stderr)")
        self.h = h
        self.neq = L.fgpu_neq(h)
        self.nnz = L.fgpu_nnz(h)

    def close(self):
        if getattr(self, "h", None):
            lib().fgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, ok, what):
        if not ok:
            raise FgpuError("%s: %s" % (what, lib().fgpu_last_error(self.h).decode()))

    # ---- the operator seam
    def stiffness(self, u_nodal=None):
        u = None if u_nodal is None else np.ascontiguousarray(u_nodal, dtype=np.float64)
        self._check(lib().fgpu_assemble_K(self.h, _p(u)), "fgpu_assemble_K")

    def stiffness_and_int_force(self, u_nodal):
        """K of the state (kept in the context) and its internal force, the
        internal force computed beside the stiffness on a second stream."""
        u = np.ascontiguousarray(u_nodal, dtype=np.float64)
        g = np.zeros(self.neq)
        self._check(lib().fgpu_assemble_K_fint(self.h, _p(u), _p(g)), "fgpu_assemble_K_fint")
        return g

    def int_force(self, u_nodal):
        u = np.ascontiguousarray(u_nodal, dtype=np.float64)
        g = np.zeros(self.neq)
        self._check(lib().fgpu_int_force(self.h, _p(u), _p(g)), "fgpu_int_force")
        return g

    def add_diag(self, eq, v):
        eq = np.ascontiguousarray(eq, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._check(lib().fgpu_add_diag(self.h, len(eq), _p(eq), _p(v)), "fgpu_add_diag")

    def solve(self, b, x0=None, maxit=None, tol=1e-8, raise_on_fail=False):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.neq)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        it = C.c_int(0)
        rr = C.c_double(0.0)
        conv = lib().fgpu_solve(self.h, _p(b), _p(x), _p(x0),
                                self.neq if maxit is None else int(maxit), tol,
                                C.byref(it), C.byref(rr))
        if raise_on_fail:
            self._check(conv, "fgpu_solve")
        return x, it.value, rr.value, bool(conv)

    def solve_batch(self, b, maxit=None, tol=1e-8):
        """Per-mesh Jacobi-PCG of a batch context.  Returns x, per-mesh
        iteration counts, per-mesh relative residuals, number converged."""
        nb = lib().fgpu_nbatch(self.h)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.neq)
        it = np.zeros(nb, dtype=np.int32)
        rr = np.zeros(nb)
        n = lib().fgpu_solve_batch(self.h, _p(b), _p(x), self.neq if maxit is None else int(maxit),
                                   tol, _p(it), _p(rr))
        return x, it, rr, n

    def csr(self):
        rowptr = np.zeros(self.neq + 1, dtype=np.int64)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        self._check(lib().fgpu_get_csr(self.h, _p(rowptr), _p(col), _p(val)), "fgpu_get_csr")
        return rowptr, col, val

    def linear_static(self, f, maxit=None, tol=1e-8):
        """cLinStat::Solver on the device (ctrl/ctrllins.cpp:46-99)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        u = np.zeros(self.neq)
        it = C.c_int(0)
        rr = C.c_double(0.0)
        ok = lib().fgpu_linear_static(self.h, _p(f), _p(u), 0 if maxit is None else int(maxit),
                                      tol, C.byref(it), C.byref(rr))
        self._check(ok, "fgpu_linear_static")
        return u, it.value, rr.value

    def linear_static_direct(self, f, tol=1e-10):
        """cLinStat::Solver with the device LDL^T: (u, refinements, relres)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        u = np.zeros(self.neq)
        k = C.c_int(0)
        rr = C.c_double(0.0)
        ok = lib().fgpu_linear_static_direct(self.h, _p(f), _p(u), tol, C.byref(k), C.byref(rr))
        self._check(ok, "fgpu_linear_static_direct")
        return u, k.value, rr.value

    def newton(self, q, steps, factor=0.0, maxit=20, tol=1e-5, pcg_tol=1e-5, pcg_maxit=0,
               adaptive_pcg=0, predict=None, lambdas=None, u0_dev=None, lam0=0.0,
               u_steps_dev=None, secant=False, galerkin=False, direct=False, keep_host=True,
               graph=True, extrap=0):
        """cCtrlNR::Solver on the device (ctrl/ctrlnr.cpp:47-267).

        ``predict`` is a Python callable ``(step, u_dev_ptr, lam_prev, lam,
        u_hist_dev_ptr, lam_hist, du_dev_ptr) -> bool`` or a ``Gnn`` (its native predictor is passed
        without a Python hop).  ``lambdas`` replaces the fixed increment
        ``factor`` by a list of load factors; ``u0_dev``/``lam0`` start from
        a converged state; ``u_steps_dev`` (device, steps x neq) receives the
        converged displacement of every step; ``direct`` solves each
        correction with the device LDL^T instead of PCG; ``graph`` replays
        each Newton iteration from captured hipGraphs (bit-identical to the
        eager launches); ``extrap`` = p starts each step from the polynomial
        in lambda through the last p + 1 converged states (p <= 4)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        prm = NrParams()
        prm.steps, prm.factor, prm.maxit, prm.tol = steps, factor, maxit, tol
        prm.pcg_maxit, prm.pcg_tol, prm.adaptive_pcg = pcg_maxit, pcg_tol, int(adaptive_pcg)
        prm.secant = int(secant)
        prm.galerkin = int(galerkin)
        prm.direct = int(direct)
        prm.graph = int(graph)
        prm.extrap = int(extrap)
        cb = None
        if isinstance(predict, Gnn):
            prm.predict = C.cast(lib().fgpu_gnn_newton_predictor, PREDICT_FN)
            prm.predict_user = predict.h
        elif predict is not None:
            def _cb(user, step, u_dev, lam_prev, lam, uh_dev, lam_h, du_dev):
                try:
                    return 1 if predict(step, u_dev, lam_prev, lam, uh_dev, lam_h, du_dev) else 0
                except Exception as exc:  # never let an exception cross the ABI
                    print("predictor failed:", exc)
                    return 0
            cb = PREDICT_FN(_cb)
            prm.predict = cb
        lam_arr = None
        if lambdas is not None:
            lam_arr = np.ascontiguousarray(lambdas, dtype=np.float64)
            assert lam_arr.shape == (steps,)
            prm.lambdas = lam_arr.ctypes.data
        prm.u0_dev = u0_dev
        prm.lam0 = lam0
        prm.u_steps_dev = u_steps_dev
        us = np.zeros((steps, self.neq)) if keep_host else None
        stats = (NrStat * steps)()
        done = lib().fgpu_newton(self.h, _p(q), C.byref(prm), _p(us), stats)
        st = [{k: getattr(stats[i], k) for k, _ in NrStat._fields_} for i in range(steps)]
        err = lib().fgpu_last_error(self.h).decode() if done < steps else ""
        if us is not None:
            us = us[:max(done, 0)]
        return us, st[:max(done, 0) + (1 if done < steps else 0)], err

    def set_csr_values(self, val):
        val = np.ascontiguousarray(val, dtype=np.float64)
        assert val.shape == (self.nnz,)
        self._check(lib().fgpu_set_csr_values(self.h, _p(val)), "fgpu_set_csr_values")

    def element_matrices(self, e0, ne):
        nd = 3 * self.mesh.nn
        K = np.zeros((ne, nd, nd), dtype=np.float32)
        self._check(lib().fgpu_get_element_matrices(self.h, e0, ne, _p(K)),
                    "fgpu_get_element_matrices")
        return K

    def body_force(self, b_elem):
        """Load vector of uniform body forces b_elem [nelem, 3]."""
        b = np.ascontiguousarray(b_elem, dtype=np.float64)
        assert b.shape == (self.mesh.nelem, 3)
        f = np.zeros(self.neq)
        self._check(lib().fgpu_body_force(self.h, _p(b), _p(f)), "fgpu_body_force")
        return f

    def set_matrix_precision(self, bits):
        """32: the PCG multiplies by an FP32 copy of the matrix; 64: FP64."""
        self._check(lib().fgpu_set_matrix_precision(self.h, int(bits)),
                    "fgpu_set_matrix_precision")

    def set_preconditioner(self, kind):
        """PCG preconditioner: 0 point Jacobi (the reference's), 1 the inverse
        of each node's 3x3 diagonal block."""
        self._check(lib().fgpu_set_preconditioner(self.h, int(kind)), "fgpu_set_preconditioner")

    def direct_solve(self, b):
        """Solve K x = b with the device LDL^T of the current matrix (factorized
        on first use after an assembly)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.neq)
        self._check(lib().fgpu_direct_solve(self.h, _p(b), _p(x)), "fgpu_direct_solve")
        return x

    def factor(self):
        self._check(lib().fgpu_factor(self.h), "fgpu_factor")

    def factor_solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.neq)
        self._check(lib().fgpu_factor_solve(self.h, _p(b), _p(x)), "fgpu_factor_solve")
        return x

    def factor_info(self):
        """(half bandwidth, negative pivots of the last factorization or -1,
        band bytes)."""
        bw, neg, nb = C.c_int(0), C.c_int(0), C.c_longlong(0)
        self._check(lib().fgpu_factor_info(self.h, C.byref(bw), C.byref(neg), C.byref(nb)),
                    "fgpu_factor_info")
        return bw.value, neg.value, nb.value

    def set_stress_nodes(self, extrapolate):
        """True: nodal stresses by extrapolation from the points (default);
        False: evaluated at the nodes (B-spline solids)."""
        self._check(lib().fgpu_set_stress_nodes(self.h, int(bool(extrapolate))),
                    "fgpu_set_stress_nodes")

    def stresses(self, u_nodal, nodal=True, average=False):
        """Stresses of the nodal state u [nnode, 3] (Voigt xx yy zz xy xz yz):
        (integration points [nelem, nq, 6], element nodes [nelem, nn, 6] or
        None, node averages [nnode, 6] or None)."""
        m = self.mesh
        u = np.ascontiguousarray(u_nodal, dtype=np.float64)
        assert u.shape == (m.nnode, 3)
        nq = lib().fgpu_num_points(self.h)
        si = np.zeros((m.nelem, nq, 6))
        sn = np.zeros((m.nelem, m.nn, 6)) if nodal else None
        sa = np.zeros((m.nnode, 6)) if average else None
        self._check(lib().fgpu_stresses(self.h, _p(u), _p(si), _p(sn), _p(sa)), "fgpu_stresses")
        return si, sn, sa

    def set_launch(self, hex8_te=0, asm_mode=-1):
        """Fix the launch configuration (replaces the autotuner)."""
        self._check(lib().fgpu_set_launch(self.h, hex8_te, asm_mode), "fgpu_set_launch")

    def add_surface_loads(self, group):
        """Append a group of surface loads (a SurfaceGroup of fast_amd.mesh:
        deg, nip, nodes [nface, (p1+1)(p2+1)] (8 for "q8" faces), ext or
        None, local, q, shape)."""
        nodes = np.ascontiguousarray(group.nodes, dtype=np.int32)
        ext = None if group.ext is None else np.ascontiguousarray(group.ext, dtype=np.float64)
        local = (None if group.local is None
                 else np.ascontiguousarray(group.local, dtype=np.int32))
        q = np.ascontiguousarray(group.q, dtype=np.float64)
        s = FgpuSurface()
        s.nface = len(nodes)
        s.deg[0], s.deg[1] = int(group.deg[0]), int(group.deg[1])
        s.nip = int(group.nip)
        s.nodes, s.ext, s.local, s.q = _p(nodes), _p(ext), _p(local), _p(q)
        s.shape = 1 if getattr(group, "shape", "tensor") == "q8" else 0
        self._check(lib().fgpu_add_surface_loads(self.h, C.byref(s)), "fgpu_add_surface_loads")

    def clear_surface_loads(self):
        self._check(lib().fgpu_clear_surface_loads(self.h), "fgpu_clear_surface_loads")

    def surface_load(self):
        """Equation vector of the context's surface loads (device integration)."""
        f = np.zeros(self.neq)
        self._check(lib().fgpu_surface_load(self.h, _p(f)), "fgpu_surface_load")
        return f

    def set_penalty(self, eq, v):
        """Prescribed equations (1-based) and their penalty (fgpu_set_penalty)."""
        eq = np.ascontiguousarray(eq, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._check(lib().fgpu_set_penalty(self.h, len(eq), _p(eq), _p(v)), "fgpu_set_penalty")

    def set_overlap(self, mode):
        """Order of stiffness_and_int_force: 0 serial, 1 f_int beside the
        assembly (fused point records where the kernels have them), 2 only
        f_int's element vectors beside the assembly and its gather after it,
        3 f_int's element vectors first and its gather beside the element
        stiffness, -1 timed on the first call."""
        self._check(lib().fgpu_set_overlap(self.h, mode), "fgpu_set_overlap")

    def block_counts(self):
        """(stored node blocks, upper node blocks) of the global matrix."""
        a, b = C.c_longlong(0), C.c_longlong(0)
        self._check(lib().fgpu_block_counts(self.h, C.byref(a), C.byref(b)), "fgpu_block_counts")
        return a.value, b.value

    def tuning_report(self):
        return lib().fgpu_tuning_report(self.h).decode()

    def last_times(self):
        return [lib().fgpu_last_time_ms(self.h, k) for k in range(4)]

    def device_buffer(self, which):
        return lib().fgpu_device_buffer(self.h, which)


class Gnn:
    """Learned warm start of a load step (fgpu_gnn_*): a 4-layer residual
    message-passing network, hidden width 64, trained on the device with
    Adam on converged states of a continuation run of the same mesh."""

    def __init__(self, model: GpuModel, load_mask=None, seed=0):
        self.model = model
        lm = None
        if load_mask is not None:
            lm = np.ascontiguousarray(load_mask, dtype=np.uint8)
            assert lm.shape == (model.mesh.nnode,)
        h = C.c_void_p()
        if not lib().fgpu_gnn_create(C.byref(h), model.h, _p(lm), seed):
            raise FgpuError("fgpu_gnn_create: %s" % lib().fgpu_last_error(model.h).decode())
        self.h = h
        self.nparam = lib().fgpu_gnn_num_params(h)

    def _check(self, ok, what):
        if not ok:
            raise FgpuError("%s: %s" % (what, lib().fgpu_last_error(self.model.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            lib().fgpu_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scales(self, u_scale, lam_scale, out_scale):
        self._check(lib().fgpu_gnn_set_scales(self.h, u_scale, lam_scale, out_scale),
                    "fgpu_gnn_set_scales")

    def set_loss(self, beta):
        self._check(lib().fgpu_gnn_set_loss(self.h, beta), "fgpu_gnn_set_loss")

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._check(lib().fgpu_gnn_get_params(self.h, _p(p)), "fgpu_gnn_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.nparam,)
        self._check(lib().fgpu_gnn_set_params(self.h, _p(p)), "fgpu_gnn_set_params")

    def loss_grad(self, u_dev, lam, dlam, ut_dev, uh_dev=None, lam_h=0.0, grad=True):
        loss = C.c_double(0.0)
        g = np.zeros(self.nparam, dtype=np.float32) if grad else None
        self._check(lib().fgpu_gnn_loss_grad(self.h, u_dev, lam, dlam, uh_dev, lam_h, ut_dev,
                                             C.byref(loss), _p(g)), "fgpu_gnn_loss_grad")
        return loss.value, g

    def predict(self, u_dev, lam, dlam, du_dev, uh_dev=None, lam_h=0.0):
        self._check(lib().fgpu_gnn_predict(self.h, u_dev, lam, dlam, uh_dev, lam_h, du_dev),
                    "fgpu_gnn_predict")

    def train(self, states_dev, lams, nsteps, maxgap=4, lr=1e-3, seed=0, origin_frac=0.1):
        lams = np.ascontiguousarray(lams, dtype=np.float64)
        hist = np.zeros(nsteps, dtype=np.float32)
        self._check(lib().fgpu_gnn_train(self.h, len(lams), states_dev, _p(lams), nsteps, maxgap,
                                         origin_frac, lr, seed, _p(hist)), "fgpu_gnn_train")
        return hist
