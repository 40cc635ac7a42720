/* fgpu.h - C-ABI of the MI355X element-integration / assembly / solve path.
 *
 * The seam is FAST's global element loops and system-matrix interface: a C++
 * analysis driver (linear static, load-control Newton-Raphson) calls these
 * entry points where the reference calls its per-element virtuals and its
 * cSysMatrix.  Plain pointers and sizes only; no exceptions cross the ABI.
 *
 * Conventions (mirroring the reference):
 *   - return 1 = ok, 0 = failure (elc/element.h:813 IntForce, mlib/sysmat.h:90
 *     Solve); fgpu_last_error() describes the failure.  Nothing calls exit().
 *   - equations are the reference's VecDof numbering (node/node.h:319-320):
 *     dofmap[node*3 + dir] = 1-based equation, 0 = fixed (ctrl/ctrl.cpp
 *     2004-2104); vectors passed as "equation vectors" have neq entries.
 *   - nodal vectors are [nnode*3], node-major (cNode::VecDispl, 0..2).
 *   - host pointers unless the name ends in _dev (device pointers on the
 *     context's HIP device, already resident in HBM).
 * One context = one mesh (or one batch of independent meshes) = one HIP
 * stream on the current device, created non-blocking.  Device inputs must be
 * complete when a call starts (synchronise the stream that wrote them); every
 * call has finished its device work, outputs included, when it returns.  A
 * context is not re-entrant; independent contexts may be used from different
 * host threads.
 */
#ifndef FGPU_H
#define FGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fgpu_ctx fgpu_ctx;

/* Element shape: cShapeBrick8 (shp/shpsolid.cpp:173-238), cShapeBspSolid
 * (shp/shpbsp.cpp:873-1067, Bezier extraction per element) or cShapeBrick20
 * (shp/shpsolid.cpp:301-435, 20-node serendipity brick in the node order of
 * its NodeNatCoord). */
enum { FGPU_HEX8 = 0, FGPU_BSP = 1, FGPU_HEX20 = 2 };
/* Formulation: cElmParam (elc/elmparam.cpp:449-597, linear), cElmParamTL
 * (elc/elmparam.cpp:1615-1807, Green-Lagrange + St.Venant-Kirchhoff), or
 * total-Lagrangian compressible neo-Hookean (no reference counterpart). */
enum { FGPU_LINEAR = 0, FGPU_TL_SVK = 1, FGPU_TL_NEOHOOKE = 2 };

/* Mesh description, cf. the reference's model tables read by InpReadData
 * (main/input.cpp:355-401): nodes, one element family, per-element material
 * (mat/material.cpp:846-851 cMatElastIso: E, nu; neo-Hookean: mu, lambda). */
typedef struct {
  int nnode, nelem;
  int shape;            /* FGPU_HEX8 | FGPU_BSP | FGPU_HEX20 */
  int nn;               /* nodes per element: 8, 20, or prod(p_d+1) <= 512 */
  int deg[3];           /* FGPU_BSP: degree per direction, 1..7; (3,3,3) and
                           (3,3,2) on 4x4x4 Gauss run the tuned FP32 MFMA
                           kernels, every other shape FP64 kernels */
  int nip[3];           /* Gauss points per direction (intpoint.cpp:579-617):
                           hex8, hex20 isotropic 1..3; FGPU_BSP 1..16 each, at most
                           512 points per element */
  int form;             /* FGPU_LINEAR | FGPU_TL_SVK | FGPU_TL_NEOHOOKE */
  const double *coords; /* [nnode*3] */
  const int *conn;      /* [nelem*nn] 0-based, reference local node order */
  const int *dofmap;    /* [nnode*3] 1-based equation or 0 */
  const double *ext;    /* FGPU_BSP: per element, per direction d the
                           (p_d+1) x (p_d+1) extraction matrix, row-major,
                           packed: [nelem * sum_d (p_d+1)^2] */
  const double *cpw;    /* FGPU_BSP: [nnode] control-point weights
                           (cNode::GetW); NULL = all weights 1, i.e. a
                           polynomial (non-rational) B-spline solid */
  const double *mat;    /* [nelem*2] */
  /* Batch of independent meshes (0 / NULL for one mesh): mesh k owns nodes
   * [batch_nodeptr[k], batch_nodeptr[k+1]), no element couples two meshes,
   * and its equations are contiguous (node-major numbering). */
  int nbatch;
  const long long *batch_nodeptr;  /* [nbatch+1] */
} fgpu_mesh;

/* Build a context: uploads the mesh, numbers the sparsity pattern and the
 * deterministic assembly maps (replaces cControl::CompProf + CreateMatrix,
 * ctrl/ctrl.cpp:2112-2157, mlib/sysmat.cpp:42-70). */
int fgpu_create(fgpu_ctx **ctx, const fgpu_mesh *mesh);
void fgpu_destroy(fgpu_ctx *ctx);
const char *fgpu_last_error(fgpu_ctx *ctx);
int fgpu_neq(fgpu_ctx *ctx);
long long fgpu_nnz(fgpu_ctx *ctx);
/* Stored 3x3 node blocks of the global matrix (every pair of nodes sharing
 * an element, both orders) and upper node blocks (A <= B) the assembly sums:
 * the sizes its algorithmic byte count is made of. */
int fgpu_block_counts(fgpu_ctx *ctx, long long *nbf, long long *nblk);

/* Global stiffness: element K_e (FP32) + deterministic assembly into the FP64
 * CSR.  Replaces cControl::StiffnessMatrix (ctrl/ctrl.cpp:886-975) and the
 * per-element cElement::StiffMat / AddGlobMat (elc/element.cpp:2975-3002).
 * u_nodal: current nodal displacements for TL forms, NULL = zero. */
int fgpu_assemble_K(fgpu_ctx *ctx, const double *u_nodal);
int fgpu_assemble_K_dev(fgpu_ctx *ctx, const double *u_nodal_dev);

/* Internal force (FP64): replaces cControl::IntForceVector
 * (ctrl/ctrl.cpp:719-813) with cElement::IntForce + AddGlobVec
 * (elc/element.cpp:2953-2971).  g has neq entries.  Returns 0 when a stress
 * update fails (non-positive det C) or an element Jacobian is not positive. */
int fgpu_int_force(fgpu_ctx *ctx, const double *u_nodal, double *g);
int fgpu_int_force_dev(fgpu_ctx *ctx, const double *u_nodal_dev, double *g_dev);

/* The internal force of fgpu_int_force_dev without any allocation: the
 * element vectors go to a fixed device array of 64 MiB that the library
 * holds on every device it runs on until the process ends, instead of the
 * context's buffer; same kernels, same bits.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Limits: nelem * nn * 3 * 8 bytes <= 64
 * MiB (349,525 hex8 elements, 43,690 cubic B-spline elements); B-spline
 * solids only of degree (3,3,3) or (3,3,2) on 4x4x4 Gauss points; larger
 * meshes and other shapes return 0 with an error.  All contexts on one
 * device share the array: calls that overlap in time on one GPU race. */
int fgpu_int_force_dev_noalloc(fgpu_ctx *ctx, const double *u_nodal_dev, double *g_dev);

/* Tangent and internal force of the same state in one call: the element
 * stiffness and the assembly run on the context stream while the internal
 * force (element vectors + gather) runs beside them on a second stream of
 * the context; the results equal those of fgpu_assemble_K followed by
 * fgpu_int_force, bit for bit.  What cCtrlNR::Solver needs of each
 * updated state (ctrl/ctrlnr.cpp:203 IntForceVector, then 157
 * StiffnessMatrix in the next iteration). */
int fgpu_assemble_K_fint(fgpu_ctx *ctx, const double *u_nodal, double *g);
int fgpu_assemble_K_fint_dev(fgpu_ctx *ctx, const double *u_nodal_dev, double *g_dev);

/* Add v[k] to the diagonal of equation eq[k] (1-based), as the prescribed-DOF
 * penalty of ctrl/ctrl.cpp:960-974 does with cSysMatrix::Add. */
int fgpu_add_diag(fgpu_ctx *ctx, int n, const int *eq, const double *v);

/* Jacobi-PCG in FP64 on the assembled CSR: replaces cSysMatrix::Solve with
 * PCGSolver (mlib/sysmat.cpp:1061-1125): relres = |r| / |b| (|b| > 1e11 taken
 * as 1), x0 = 0 unless x0 is given (warm start).  Returns 1 when converged. */
int fgpu_solve(fgpu_ctx *ctx, const double *b, double *x, const double *x0,
               int maxit, double rtol, int *iters, double *relres);
int fgpu_solve_dev(fgpu_ctx *ctx, const double *b_dev, double *x_dev,
                   const double *x0_dev, int maxit, double rtol, int *iters,
                   double *relres);

/* Precision of the matrix values the PCG reads (64, the default, or 32):
 * with 32 every PCG of the context (fgpu_solve, the drivers) multiplies by an
 * FP32 copy of the assembled FP64 matrix, refreshed on the first solve after
 * the matrix changes; arithmetic, vectors and dot products stay FP64.  The
 * matrix is assembled from FP32 element matrices, so the copy adds a
 * rounding of the same order; the drivers' FP64 residuals correct it, and a
 * plain fgpu_solve then solves the rounded matrix.  Halves the bytes of each
 * iteration's matrix-vector product.  No counterpart in the reference. */
int fgpu_set_matrix_precision(fgpu_ctx *ctx, int bits);

/* Preconditioner of every PCG of the context (fgpu_solve, the drivers, the
 * batch solve): 0 (the default) point Jacobi, z_i = r_i / K_ii, the reference's
 * (cSysMatrix PCG, mlib/sysmat.cpp); 1 the inverse of each node's 3x3
 * diagonal block, z_j = K_jj^-1 r_j (the three directions of a node
 * coupled; fixed directions keep their unit rows).  Same stopping rule; the
 * iterates, and so the iteration counts, differ from the reference's with 1.
 * No counterpart in the reference for 1. */
int fgpu_set_preconditioner(fgpu_ctx *ctx, int kind);

/* Direct solve of the assembled global matrix, the reference's skyline
 * solver: cSymSkylMatrix::Solve (mlib/sysmat.cpp:265-277) with CroutSolver
 * (mlib/matvec.cpp:609-664), LDL^T without pivoting in FP64.  Storage is a
 * band over the equation numbering (half bandwidth = the widest row
 * profile, which contains the skyline and all fill); the factors stay on the
 * device until the matrix changes (assembly, fgpu_add_diag,
 * fgpu_set_csr_values).
 *   fgpu_factor        factorize (CroutSolver op 2); returns 0 when a pivot
 *                      is below 1e-12 in magnitude (the reference's TOLC,
 *                      "singular matrix") or the half bandwidth exceeds 16384
 *   fgpu_factor_solve  solve with the current factors (op 3)
 *   fgpu_direct_solve  factorize when needed, then solve (op 1)
 *   fgpu_factor_info   half bandwidth, negative pivots of the last
 *                      factorization (cSymSkylMatrix::GetNgtPvt; -1 if none),
 *                      band bytes
 * b and x are equation vectors and may be the same array. */
int fgpu_factor(fgpu_ctx *ctx);
int fgpu_factor_solve(fgpu_ctx *ctx, const double *b, double *x);
int fgpu_factor_solve_dev(fgpu_ctx *ctx, const double *b_dev, double *x_dev);
int fgpu_direct_solve(fgpu_ctx *ctx, const double *b, double *x);
int fgpu_direct_solve_dev(fgpu_ctx *ctx, const double *b_dev, double *x_dev);
int fgpu_factor_info(fgpu_ctx *ctx, int *half_band, int *neg_pivots, long long *bytes);

/* Batched solve: one Jacobi-PCG per mesh of a batch context, each with its
 * own |b|, scalars and stopping test (the reference solves each model on its
 * own), all meshes advancing in the same launches.  iters / relres (optional)
 * have nbatch entries.  Returns the number of meshes that converged. */
int fgpu_solve_batch(fgpu_ctx *ctx, const double *b, double *x, int maxit, double rtol,
                     int *iters, double *relres);
int fgpu_solve_batch_dev(fgpu_ctx *ctx, const double *b_dev, double *x_dev, int maxit,
                         double rtol, int *iters, double *relres);
int fgpu_nbatch(fgpu_ctx *ctx);

/* y = K x on the assembled CSR (cSymSprsMatrix::MultVect, sysmat.cpp:942). */
int fgpu_spmv_dev(fgpu_ctx *ctx, const double *x_dev, double *y_dev);

/* Overwrite the CSR values (host array of fgpu_nnz entries, the pattern of
 * fgpu_get_csr): the counterpart of filling a cSysMatrix through Add
 * (mlib/sysmat.h:60-99) for callers that assemble on their own. */
int fgpu_set_csr_values(fgpu_ctx *ctx, const double *val);
/* Read back the CSR (full symmetric pattern, 0-based, rows in equation
 * order; within a row the columns follow the node order of the rows'
 * neighbours, which is ascending equation order unless a solver node order
 * numbered the equations).  Any pointer may be NULL to skip it. */
int fgpu_get_csr(fgpu_ctx *ctx, long long *rowptr, int *col, double *val);
/* Element matrices of elements [e0, e0+ne) as dense row-major
 * (3nn x 3nn) FP32, element-local DOF order (node-major). */
int fgpu_get_element_matrices(fgpu_ctx *ctx, int e0, int ne, float *Ke);

/* ---------------------------------------------------------------- drivers
 * Analysis drivers running the whole loop on the device (the caller side of
 * the seam, cf. cLinStat::Solver ctrl/ctrllins.cpp:46-99 and
 * cCtrlNR::Solver ctrl/ctrlnr.cpp:47-267). */

/* Linear static: assemble K, solve K u = f to |f - K u| / |f| <= rtol (u:
 * neq).  K u is evaluated in FP64 element by element (the linear internal
 * force); PCG on the matrix assembled from FP32 element matrices solves for
 * the corrections (iterative refinement), so ill-conditioned models reach
 * FP64 accuracy.  maxit bounds the total PCG iterations (<= 0: 50 neq);
 * iters returns them, relres the final residual. */
int fgpu_linear_static(fgpu_ctx *ctx, const double *f, double *u, int maxit,
                       double rtol, int *iters, double *relres);
/* The same with the device LDL^T (fgpu_factor) in place of PCG: one
 * factorization, then refinements with FP64 residuals until
 * |f - K u| / |f| <= rtol or the residual stops falling (cLinStat::Solver
 * with the reference's default skyline matrix).  refinements returns the
 * number of solves with the factors. */
int fgpu_linear_static_direct(fgpu_ctx *ctx, const double *f, double *u, double rtol,
                              int *refinements, double *relres);

/* Predictor hook for the Newton driver: given the converged equation vector
 * u_dev of the previous step (load factor lam_prev), the converged state
 * before it (u_hist_dev at lam_hist; NULL at the first step) and the new load
 * factor lam, write a predicted displacement increment into du_dev (neq,
 * device).  Return 1 to use it, 0 to start the step from du = 0.  A
 * prediction d is discarded (the step then starts from du = 0) unless it
 * lowers the potential energy by the trapezoidal estimate
 * d . (r(u) + r(u + d)) > 0, r = f - f_int. */
typedef int (*fgpu_predict_fn)(void *user, int step, const double *u_dev, double lam_prev,
                               double lam, const double *u_hist_dev, double lam_hist,
                               double *du_dev);

typedef struct {
  int steps;          /* MaxStep */
  double factor;      /* StepFactor: lambda += factor per step */
  int maxit;          /* MaxIter: Newton iterations per step */
  double tol;         /* Tol: |r| / max(1, |f|) */
  int pcg_maxit;      /* <= 0: neq (as the reference's NR) */
  double pcg_tol;     /* reference NR: 1e-5 (sysmat.h:77-78 default) */
  int adaptive_pcg;   /* 0: fixed pcg_tol, as the reference;
                         1: min(pcg_tol, 0.1 err) after the first iteration;
                         2: the PCG reduction each Newton iteration needs to
                            reach tol: clamp(0.5 tol max(1,|f|) / |r|,
                            1e-12, pcg_tol) */
  fgpu_predict_fn predict;  /* optional */
  void *predict_user;
  const double *lambdas;    /* optional [steps]: load factor of each step
                               (else lambda += factor) */
  const double *u0_dev;     /* optional start state (equation vector, device) */
  double lam0;              /* load factor of the start state */
  double *u_steps_dev;      /* optional [steps * neq] converged states (device) */
  int secant;               /* 1: without a predictor, extrapolate the last
                               converged increment linearly in lambda */
  int galerkin;             /* 1: replace a prediction d by the combination
                               of d and the secant increment that minimises
                               1/2 x.K x - x.r with the tangent K of the
                               previous state (two SpMVs) */
  int direct;               /* 1: solve each correction with the device LDL^T
                               of the tangent (fgpu_factor; the reference's
                               skyline solver) instead of PCG */
  int graph;                /* 1: after the first iteration, replay each
                               Newton iteration from captured hipGraphs (the
                               whole iteration with the direct solver; the
                               tangent + PCG start and the update + f_int +
                               residual around the PCG chunk graphs), one
                               host read per iteration; bit-identical to 0
                               (eager launches) */
  int extrap;               /* p = 1..4, without a predictor callback: start
                               each step from the polynomial in lambda through
                               the last p + 1 converged states, extrapolated
                               to the step's lambda (the start state is the
                               first of them; fewer while fewer exist, none
                               with one); 0: off (then secant applies) */
} fgpu_nr_params;

typedef struct {
  int step, iters, pcg_iters, converged;
  int predicted;      /* 1 prediction used; 0 none; -1 rejected (energy
                         test failed or f_int failed at the prediction) */
  double lam, err;
  double err_start;   /* |r| / max(1, |f|) at the start of the step */
  double ms_stiffness, ms_fint, ms_solve, ms_total;
} fgpu_nr_stat;

/* Load-control Newton-Raphson: f = lambda q; full Newton with the tangent
 * rebuilt every iteration; convergence |r| / max(1, |f|) < tol (ctrl/ctrl.cpp
 * 1366-1380).  q: neq host vector.  u_steps (optional, host): neq * steps,
 * the converged u of each step.  stats (optional): one entry per step.
 * Returns the number of converged steps (the reference stops at the first
 * non-converged step); 0 with fgpu_last_error on a failure. */
int fgpu_newton(fgpu_ctx *ctx, const double *q, const fgpu_nr_params *prm,
                double *u_steps, fgpu_nr_stat *stats);

/* ---------------------------------------------------------------- learned
 * warm start (no counterpart in the reference; for callers that run many
 * related Newton solves on one mesh).  A 4-layer message-passing network
 * (hidden 64, FP16 MFMA with FP32 accumulation) over the mesh graph predicts
 * the displacement increment of a load step from the current state; the
 * unchanged Newton/PCG then drives the residual to tolerance.  Vectors are
 * equation vectors on the device. */
typedef struct fgpu_gnn fgpu_gnn;
/* load_mask: [nnode] bit d set where the load acts in direction d (or NULL) */
int fgpu_gnn_create(fgpu_gnn **net, fgpu_ctx *ctx, const unsigned char *load_mask,
                    unsigned seed);
void fgpu_gnn_destroy(fgpu_gnn *net);
int fgpu_gnn_num_params(fgpu_gnn *net);
/* feature / target normalisation: |u| scale, load-factor scale, output scale */
int fgpu_gnn_set_scales(fgpu_gnn *net, double u_scale, double lam_scale, double out_scale);
/* loss = (sum_free e^2 + beta sum_nodes |e - M e|^2) / neq, e the error of
 * the scaled prediction, M the mean over the neighbours: beta > 0 penalises
 * rough errors, which raise the residual of a predicted state most
 * (default 0) */
int fgpu_gnn_set_loss(fgpu_gnn *net, double beta);
int fgpu_gnn_get_params(fgpu_gnn *net, float *params);
int fgpu_gnn_set_params(fgpu_gnn *net, const float *params);
/* One sample: state u at lam, an earlier state uh at lam_h (NULL: the
 * unloaded state, lam_h ignored), step dlam; equation vectors on the device,
 * u may be NULL (the unloaded state, lam = 0).  The prediction is the secant
 * of the last step (u - uh) / (lam - lam_h) times dlam plus the network's
 * correction. */
/* loss and gradient (host) against the target state ut at lam + dlam */
int fgpu_gnn_loss_grad(fgpu_gnn *net, const double *u_dev, double lam, double dlam,
                       const double *uh_dev, double lam_h, const double *ut_dev, double *loss,
                       float *grad);
/* du = predicted increment of the step lam -> lam + dlam */
int fgpu_gnn_predict(fgpu_gnn *net, const double *u_dev, double lam, double dlam,
                     const double *uh_dev, double lam_h, double *du_dev);
/* Adam training on a converged trajectory: states [nstate * neq] (device) at
 * load factors lams[nstate] (host, increasing).  Each step draws one sample
 * and takes one Adam step on its loss: with probability origin_frac a solve
 * from the unloaded state to a uniformly drawn state, else a triple h < i < j
 * (history, state, target) with log-uniform gaps in [1, maxgap] (index -1 =
 * the unloaded state).  loss_hist (optional, host): the loss of every step. */
int fgpu_gnn_train(fgpu_gnn *net, int nstate, const double *states_dev, const double *lams,
                   int nsteps, int maxgap, double origin_frac, float lr, unsigned seed,
                   float *loss_hist);
/* fgpu_predict_fn adapter: pass with predict_user = the fgpu_gnn handle */
int fgpu_gnn_newton_predictor(void *user, int step, const double *u_dev, double lam_prev,
                              double lam, const double *u_hist_dev, double lam_hist,
                              double *du_dev);

/* Uniform surface loads on element faces, integrated on the device
 * (cFaceUnifLoad, load/load.cpp:1440-1500; cFaceUnifLoadIGA, load/loadiga.cpp
 * 305-354; both through cElemLoad::ExtForce, load/load.cpp:618-691).  Each
 * entry is one face: a tensor-product surface of degree (p1, p2) over
 * (p1+1)(p2+1) nodes, the first surface parameter fastest, with the Bezier
 * extraction rows of its two parameters (N_i = sum_j C_ij B_j, as `ext` of
 * fgpu_mesh) and the context's control-point weights.  At nip x nip Gauss
 * points (first parameter fastest, cIntPoint::CreateQuadPoints,
 * elc/intpoint.cpp:266-300) f_a += w |x_,s x x_,t| R_a load, with
 * load = q in global axes, or (local = 1) q0 u + q1 v + q2 w in the face's
 * frame of cShapeSurf::LocalSys (shp/shpsurf.cpp:121-164): u = x_,s/|x_,s|,
 * w = u x x_,t/|x_,t|, v = w x u.  A brick face is the degree-(1,1) surface
 * over its four corners with identity extraction (ext NULL): the nodes of
 * cShapeBrick8's face n0 n1 n2 n3 (shp/shpsolid.cpp:264-276) given as
 * n0 n1 n3 n2.  An IGA face is the patch's boundary surface over the
 * element's two spans.  shape 1: faces of 20-node bricks, the 8-node
 * serendipity quadrilateral of cShapeQ8 (shp/shpplane.cpp:428-486; deg
 * (2, 2), ext NULL, the eight nodes of cShapeBrick20's face in its order,
 * shp/shpsolid.cpp:459-470). */
typedef struct {
  int nface;           /* faces of this group */
  int deg[2];          /* degree per surface parameter, 1..7 */
  int nip;             /* Gauss points per parameter, 1..16 */
  const int *nodes;    /* [nface * (p1+1)(p2+1)] 0-based nodes (shape 1: 8) */
  const double *ext;   /* [nface * ((p1+1)^2 + (p2+1)^2)] row-major, or NULL
                          (identity: the Bernstein polynomials are the basis) */
  const int *local;    /* [nface] 1 = local axes, or NULL (all global) */
  const double *q;     /* [nface * 3] load per unit area */
  int shape;           /* 0: tensor product (above); 1: 8-node serendipity */
} fgpu_surface;
/* Append a group of surface loads to the context (uploaded once; the
 * entries keep their order across groups); clear all of them. */
int fgpu_add_surface_loads(fgpu_ctx *ctx, const fgpu_surface *loads);
int fgpu_clear_surface_loads(fgpu_ctx *ctx);
/* f = the equation vector of the context's surface loads (zero elsewhere),
 * every node's face contributions summed in entry order (FP64). */
int fgpu_surface_load(fgpu_ctx *ctx, double *f);
int fgpu_surface_load_dev(fgpu_ctx *ctx, double *f_dev);

/* Prescribed equations (cPrescDof) with the penalty of ctrl/ctrl.cpp
 * 958-974 and 794-810: after this call every stiffness assembly adds v[k]
 * to the diagonal of equation eq[k] (1-based) after the element
 * contributions, and every internal force adds v[k] * u at that equation
 * after the element gather (u the current nodal displacement; repeated
 * equations add their values).  n = 0 removes the penalty. */
int fgpu_set_penalty(fgpu_ctx *ctx, int n, const int *eq, const double *v);

/* External load vector of uniform body forces, b [nelem * 3] (force per
 * unit volume of each element): f = sum_e sum_q w detJ N^T b_e on the
 * element's Gauss rule (cElemLoad::ExtForce with cBodyLoad, load/load.cpp
 * 618-705, 969-974).  f: [neq] equation vector. */
int fgpu_body_force(fgpu_ctx *ctx, const double *b, double *f);
int fgpu_body_force_dev(fgpu_ctx *ctx, const double *b_dev, double *f_dev);

/* Stresses of the nodal state u [nnode * 3] (FP64), in Voigt order xx yy zz
 * xy xz yz: the second Piola-Kirchhoff stress of the Green strain for the
 * total-Lagrangian forms, the Cauchy stress of the small strain for the
 * linear form.
 *   sig_ip   [nelem * nq * 6]  at the integration points, in the brick
 *            rule's order (r fastest); nq = fgpu_num_points.  Replaces
 *            cElmParam::IntPntStress / cElmParamTL::IntPntStress
 *            (elc/elmparam.cpp:859-937, 1811-1885) over the element loop of
 *            cControl::PrintIntPntStress (ctrl/ctrl.cpp:2189-2259).
 *   sig_node [nelem * nn * 6]  extrapolated to each element's nodes by the
 *            least-squares linear fit of cElmParam::NodalStress /
 *            CompTRMatrix (elmparam.cpp:1304-1333, 1444-1535; the reference's
 *            default, HEADER.OUTPUT.STRESS.EXTRAPOLATION 1), as printed by
 *            cControl::PrintNodalStress (ctrl.cpp:2264-2330); may be NULL.
 *   sig_avg  [nnode * 6]  element-nodal values averaged per node in element
 *            order (cControl::PrintAvrNodalStress, ctrl.cpp:2397-2470); may be
 *            NULL.
 * Returns 0 on a failed stress update or an invalid Jacobian. */
int fgpu_num_points(fgpu_ctx *ctx);
/* extrapolate 0: sig_node holds the stresses evaluated at each element's
 * nodes instead (HEADER.OUTPUT.STRESS.EXTRAPOLATION 0: NodalStress through
 * cElmParam::PntStress, elmparam.cpp:1206-1290, at the nodes' natural
 * coordinates with the linear strain B u and the linear elastic material);
 * B-spline solids with (E, nu) materials only.  1 (default): extrapolation. */
int fgpu_set_stress_nodes(fgpu_ctx *ctx, int extrapolate);
int fgpu_stresses(fgpu_ctx *ctx, const double *u_nodal, double *sig_ip, double *sig_node,
                  double *sig_avg);
int fgpu_stresses_dev(fgpu_ctx *ctx, const double *u_nodal_dev, double *sig_ip_dev,
                      double *sig_node_dev, double *sig_avg_dev);

/* Launch configuration.  On the first stiffness assembly of a context the
 * autotuner times every valid configuration of the element-stiffness kernel
 * (hex8: elements per workgroup 4..64, or 1 = one element per lane with its
 * point data in registers, linear bricks on 2x2x2 Gauss) and of the assembly (contribution
 * loads in batches of 1, 4, 8 or 16: mode & 3; mode & 4: workgroups ordered
 * so that each of the 8 XCDs sweeps a contiguous range of block rows; mode
 * 8, B-spline solids of at most 64 control points per element: one wave per
 * node row, the row accumulated in LDS element by element) on the actual
 * mesh and state, keeps the fastest and records the timings
 * (environment FGPU_AUTOTUNE=0 skips it).  All configurations give identical
 * results, bit for bit.  fgpu_set_launch fixes a configuration instead
 * (hex8_te 0: the largest; asm_mode -1: the default); 0 if it is not valid
 * for the mesh.  No counterpart in the reference (a CPU loop). */
int fgpu_set_launch(fgpu_ctx *ctx, int hex8_te, int asm_mode);
/* Order of fgpu_assemble_K_fint[_dev]: 0 = the element stiffness and the
 * assembly, then the internal force (element vectors + gather), all on the
 * context stream; 1 = the internal force on the second stream beside the
 * assembly, and for the linear tuned B-spline kernels (degrees (3,3,3) and
 * (3,3,2), 4x4x4 Gauss) finished from the per-point records the stiffness
 * kernel writes; 2 = as 1, but only the element vectors beside the assembly
 * and the gather after it; 3 = the internal force's element vectors first,
 * then its gather on the second stream beside the element stiffness, then
 * the assembly; -1 (default) = the first call times the four and keeps the
 * fastest (FGPU_AUTOTUNE=0: 0).  All orders give the same bits. */
int fgpu_set_overlap(fgpu_ctx *ctx, int mode);
const char *fgpu_tuning_report(fgpu_ctx *ctx);

/* Time (ms, HIP events on the context stream) of the last call of each
 * phase: 0 = element K, 1 = assembly, 2 = element f_int (+ gather),
 * 3 = solve. */
double fgpu_last_time_ms(fgpu_ctx *ctx, int phase);
/* Raw device pointers of internal buffers, for benchmarks and tests:
 * 0 = CSR values (f64), 1 = element matrices (f32), 2 = scratch nodal
 * vector (f64 [nnode*3]). */
void *fgpu_device_buffer(fgpu_ctx *ctx, int which);

#ifdef __cplusplus
}
#endif
#endif
