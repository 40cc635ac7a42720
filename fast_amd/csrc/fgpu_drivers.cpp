// fgpu_drivers.cpp - analysis drivers of libfgpu: linear static and
// load-control Newton-Raphson, with every vector resident on the device.
//
// Linear static follows cLinStat::Solver (ctrl/ctrllins.cpp:46-99):
// assemble, solve; the solve refines with FP64 residuals (below).
// Newton-Raphson follows cCtrlNR::Solver (ctrl/ctrlnr.cpp:47-267):
// g = f_int(0); per step lambda += factor, f = lambda q, r = f - g; per
// iteration K(u), K du = r, u += du, g = f_int(u), r = f - g, converged when
// |r| / max(1, |f|) < Tol (ctrl/ctrl.cpp:1366-1413; |g| replaces |f| when
// |f| <= Tol; the norms leave the prescribed equations out); the loop stops
// at the first step that does not converge.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fgpu_internal.h"

using namespace fgpu;

namespace {

template <class F>
int guarded(fgpu_ctx *ctx, F f) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return 0;
  try {
    return f(*c);
  } catch (const Error &e) {
    c->err = e.msg;
  } catch (const std::exception &e) {
    c->err = e.what();
  }
  return 0;
}

struct DevVec {
  double *p = nullptr;
  explicit DevVec(size_t n) { FGPU_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double))); }
  ~DevVec() {
    if (p) (void)hipFree(p);
  }
  DevVec(const DevVec &) = delete;
  DevVec &operator=(const DevVec &) = delete;
};

// |x| over the free equations: x copied to tmp with the prescribed (penalty)
// entries zeroed (cControl::Convergence, ctrl/ctrl.cpp:1391-1413)
double free_len(Ctx &c, const double *x, double *tmp, int n) {
  if (!c.npen) return std::sqrt(dev_dot(c, x, x, n));
  FGPU_CHECK(hipMemcpyAsync(tmp, x, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
  launch_zero_penalty(c, tmp);
  return std::sqrt(dev_dot(c, tmp, tmp, n));
}

// |g| of the elements' internal force alone (penalty left out) at the nodal
// state u: when the prescribed values are the only load, the reactions it
// holds give the force scale of the residual test
double element_force_len(Ctx &c, const double *u_nodal, double *tmp, int n) {
  struct Restore {
    Ctx &c;
    int npen;
    ~Restore() { c.npen = npen; }
  } keep{c, c.npen};
  c.npen = 0;
  if (!int_force_dev(c, u_nodal, tmp)) throw Error(c.err);
  return std::sqrt(dev_dot(c, tmp, tmp, n));
}

// Normwise backward error |r| / (| |K| |u| | + |f|) over the free equations
// (2-norms) of the equation residual r (prescribed entries zeroed) at the
// nodal state u.  The FP64 residual of an ill-conditioned system stalls at
// a relative size near cond(K) times the rounding of the sums it is formed
// from; there the backward error still shows whether the solution is as
// accurate as FP64 allows.
double backward_error(Ctx &c, const double *r, const double *u_nodal, double f_len, double *tmp,
                      int n) {
  launch_abs_spmv_nodal(c, u_nodal, c.d_w);
  launch_nodal_to_eq(c, c.d_w, tmp);
  launch_zero_penalty(c, tmp);
  const double ku = std::sqrt(dev_dot(c, tmp, tmp, n));
  const double rl = std::sqrt(dev_dot(c, r, r, n));
  return ku + f_len > 0.0 ? rl / (ku + f_len) : 0.0;
}
// a refinement stalled at the FP64 floor with this backward error has
// converged (the solution solves a system within this relative distance)
constexpr double kBackwardTol = 1e-13;

// Graphs of one Newton iteration, captured on first use in a call and
// replayed for the rest of it; destroyed with the call.
struct IterGraphs {
  hipGraphExec_t head = nullptr, tail = nullptr;
  hipEvent_t ev[4] = {};
  ~IterGraphs() {
    if (head) (void)hipGraphExecDestroy(head);
    if (tail) (void)hipGraphExecDestroy(tail);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// captures what enqueue() launches on the context stream into an
// executable graph; a failure inside ends the capture before it propagates
template <class F>
void capture(Ctx &c, hipGraphExec_t &ge, F &&enqueue) {
  hipGraph_t g = nullptr;
  FGPU_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
  try {
    enqueue();
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  FGPU_CHECK(hipStreamEndCapture(c.stream, &g));
  const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  FGPU_CHECK(e);
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  FGPU_CHECK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// y = K x on equation vectors (PCG work vectors as scratch)
void spmv_eq(Ctx &c, const double *x, double *y) {
  launch_eq_to_nodal(c, x, c.d_bn);
  launch_spmv_nodal(c, c.d_bn, c.d_w);
  launch_nodal_to_eq(c, c.d_w, y);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" {

int fgpu_linear_static(fgpu_ctx *ctx, const double *f, double *u, int maxit, double rtol,
                       int *iters, double *relres) {
  return guarded(ctx, [&](Ctx &c) {
    if (!f || !u) throw Error("null argument");
    const int n = c.neq;
    const size_t nb = sizeof(double) * n;
    DevVec df(n), du(n), dr(n), dd(n), dg(n);
    FGPU_CHECK(hipMemcpyAsync(df.p, f, nb, hipMemcpyHostToDevice, c.stream));
    FGPU_CHECK(hipMemsetAsync(du.p, 0, nb, c.stream));
    FGPU_CHECK(hipMemcpyAsync(dr.p, df.p, nb, hipMemcpyDeviceToDevice, c.stream));
    if (!stiffness_dev(c, nullptr)) return 0;
    // Iterative refinement: the PCG matrix is assembled from FP32 element
    // matrices; the residual r = f - K u is formed in FP64 element by element
    // (f_int of the linear model is K u), and each PCG solve only needs to
    // reduce it by the factor the FP32 matrix can deliver.  With prescribed
    // equations the test runs over the free ones: after the first solve the
    // prescribed equations hold their values to the penalty's accuracy, so
    // the corrections see residuals with those entries zeroed.
    double scale = free_len(c, df.p, dd.p, n);
    const int budget = maxit > 0 ? maxit : 50 * n;
    int total = 0, conv = 0;
    double rel = scale > 0.0 || c.npen ? 1.0 : 0.0;
    for (int k = 0; k < 50 && total < budget; k++) {
      if (rel <= rtol) {
        conv = 1;
        break;
      }
      const double ptol = std::min(1e-2, std::max(0.5 * rtol / rel, 1e-13));
      int it = 0;
      double rr = 0.0;
      run_pcg(c, dr.p, dd.p, nullptr, budget - total, ptol, &it, &rr, false);
      total += it;
      launch_axpby(c, n, 1.0, du.p, 1.0, dd.p, du.p);
      launch_eq_to_nodal(c, du.p, c.d_u);
      if (!int_force_dev(c, c.d_u, dg.p)) return 0;
      launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
      launch_zero_penalty(c, dr.p);
      if (!(scale > 0.0)) scale = element_force_len(c, c.d_u, dd.p, n);
      const double prev = rel;
      rel = scale > 0.0 ? std::sqrt(dev_dot(c, dr.p, dr.p, n)) / scale : 0.0;
      // at the floor of the FP64 residual (a first solve with prescribed
      // equations is dominated by their penalty rows and does not count):
      // converged if the backward error is at the rounding level
      if (rel > 0.9 * prev && (k > 0 || !c.npen)) {
        if (rel > rtol &&
            backward_error(c, dr.p, c.d_u, free_len(c, df.p, dd.p, n), dd.p, n) <= kBackwardTol)
          conv = 1;
        break;
      }
    }
    if (!conv && rel <= rtol) conv = 1;
    FGPU_CHECK(hipMemcpyAsync(u, du.p, nb, hipMemcpyDeviceToHost, c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    if (iters) *iters = total;
    if (relres) *relres = rel;
    if (!conv) {
      char msg[160];
      std::snprintf(msg, sizeof msg,
                    "linear static: residual %.3e above the tolerance after %d PCG iterations",
                    rel, total);
      c.err = msg;
    }
    return conv;
  });
}

int fgpu_linear_static_direct(fgpu_ctx *ctx, const double *f, double *u, double rtol,
                              int *refinements, double *relres) {
  return guarded(ctx, [&](Ctx &c) {
    if (!f || !u) throw Error("null argument");
    const int n = c.neq;
    const size_t nb = sizeof(double) * n;
    DevVec df(n), du(n), dr(n), dd(n), dg(n);
    FGPU_CHECK(hipMemcpyAsync(df.p, f, nb, hipMemcpyHostToDevice, c.stream));
    FGPU_CHECK(hipMemsetAsync(du.p, 0, nb, c.stream));
    FGPU_CHECK(hipMemcpyAsync(dr.p, df.p, nb, hipMemcpyDeviceToDevice, c.stream));
    if (!stiffness_dev(c, nullptr)) return 0;
    if (!band_factor(c)) return 0;
    // refinement as in fgpu_linear_static, each correction by the factors of
    // the matrix assembled from FP32 element matrices
    double scale = free_len(c, df.p, dd.p, n);
    double rel = scale > 0.0 || c.npen ? 1.0 : 0.0;
    int k = 0, conv = 0;
    for (; k < 50; k++) {
      if (rel <= rtol) {
        conv = 1;
        break;
      }
      band_solve(c, dr.p, dd.p);
      launch_axpby(c, n, 1.0, du.p, 1.0, dd.p, du.p);
      launch_eq_to_nodal(c, du.p, c.d_u);
      if (!int_force_dev(c, c.d_u, dg.p)) return 0;
      launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
      launch_zero_penalty(c, dr.p);
      if (!(scale > 0.0)) scale = element_force_len(c, c.d_u, dd.p, n);
      const double prev = rel;
      rel = scale > 0.0 ? std::sqrt(dev_dot(c, dr.p, dr.p, n)) / scale : 0.0;
      if (rel > 0.9 * prev && (k > 0 || !c.npen)) {  // at the floor of the FP64 residual
        k++;
        if (rel > rtol &&
            backward_error(c, dr.p, c.d_u, free_len(c, df.p, dd.p, n), dd.p, n) <= kBackwardTol)
          conv = 1;
        break;
      }
    }
    if (!conv && rel <= rtol) conv = 1;
    FGPU_CHECK(hipMemcpyAsync(u, du.p, nb, hipMemcpyDeviceToHost, c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    if (refinements) *refinements = k;
    if (relres) *relres = rel;
    if (!conv) {
      char msg[160];
      std::snprintf(msg, sizeof msg,
                    "linear static (direct): residual %.3e above the tolerance after %d solves",
                    rel, k);
      c.err = msg;
    }
    return conv;
  });
}

int fgpu_newton(fgpu_ctx *ctx, const double *q, const fgpu_nr_params *prm, double *u_steps,
                fgpu_nr_stat *stats) {
  return guarded(ctx, [&](Ctx &c) {
    if (!q || !prm) throw Error("null argument");
    const int n = c.neq;
    const size_t nb = sizeof(double) * n;
    DevVec dq(n), df(n), dg(n), dr(n), du(n), ddu(n), dpred(n), dlast(n), dinc(n), dgsave(n),
        dr0(n), db2(prm->galerkin ? n : 1), dw1(prm->galerkin ? n : 1), dw2(prm->galerkin ? n : 1),
        dfree(c.npen ? n : 1);
    // the prescribed equations are left out of the convergence test
    auto len = [&](const double *x) { return free_len(c, x, dfree.p, n); };
    bool k_valid = false;  // the context holds the tangent of the current state
    FGPU_CHECK(hipMemcpyAsync(dq.p, q, nb, hipMemcpyHostToDevice, c.stream));
    if (prm->u0_dev)
      FGPU_CHECK(hipMemcpyAsync(du.p, prm->u0_dev, nb, hipMemcpyDeviceToDevice, c.stream));
    else
      FGPU_CHECK(hipMemsetAsync(du.p, 0, nb, c.stream));
    double *unod = c.d_u;  // nodal displacement field
    launch_eq_to_nodal(c, du.p, unod);
    if (!int_force_dev(c, unod, dg.p)) return 0;  // initial stresses
    const int pcg_maxit = prm->pcg_maxit > 0 ? prm->pcg_maxit : n;
    // Captured iterations (prm->graph): after one eager iteration has done
    // the lazy set-up (autotuning, scratch buffers), each Newton iteration
    // is replayed from graphs -- with PCG: head = tangent + PCG start + its
    // first CHUNK iterations, then PCG chunk graphs while the device's flag
    // says running, then tail = update + f_int + residual + norms; with the
    // direct solver one graph holds the whole iteration.  One host read per
    // iteration (and one per PCG chunk) replaces the eager path's
    // synchronisations; the kernels and their order are the eager path's.
    const bool use_graph = prm->graph != 0 && c.use_graphs;
    const int abs_exit = prm->adaptive_pcg == 0;
    IterGraphs G;
    bool warm = false;
    double *d_tol = c.d_scal + 10;                      // PCG tolerance of the solve
    double *d_rr = c.d_scal + 11, *d_gg = c.d_scal + 12;  // |r|^2, |g|^2 (free equations)
    auto enqueue_tail = [&](bool pcg_x) {
      if (pcg_x) launch_nodal_to_eq(c, c.d_x, ddu.p);
      launch_axpby(c, n, 1.0, du.p, 1.0, ddu.p, du.p);  // u += du
      launch_eq_to_nodal(c, du.p, unod);
      enqueue_int_force(c, unod, dg.p);
      launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
      for (int k = 0; k < 2; k++) {
        const double *x = k ? dg.p : dr.p;
        if (c.npen) {
          launch_copy(c, x, dfree.p, n);
          launch_zero_penalty(c, dfree.p);
          x = dfree.p;
        }
        enqueue_dot(c, x, x, n, k ? d_gg : d_rr);
      }
    };
    const char *venv = std::getenv("FGPU_NR_VERBOSE");  // per-iteration log on stderr
    const bool verbose = venv && venv[0] == '1';
    double lam = prm->u0_dev ? prm->lam0 : 0.0, lam_last_step = lam;
    bool have_last = false;
    // converged states for the extrapolation predictor: a ring of p + 1
    // vectors, hl[] their load factors oldest first, hs[] their slots
    const int xp = prm->predict ? 0 : std::max(0, std::min(prm->extrap, 4));
    DevVec hist(xp ? (size_t)(xp + 1) * n : 1);
    std::vector<double> hl;
    std::vector<int> hs;
    auto keep_state = [&](double l) {
      if (!xp) return;
      int slot = (int)hs.size();
      if (slot == xp + 1) {  // full: reuse the oldest slot
        slot = hs.front();
        hs.erase(hs.begin());
        hl.erase(hl.begin());
      }
      FGPU_CHECK(hipMemcpyAsync(hist.p + (size_t)slot * n, du.p, nb, hipMemcpyDeviceToDevice,
                                c.stream));
      hs.push_back(slot);
      hl.push_back(l);
    };
    keep_state(lam);
    int done_steps = 0;
    for (int step = 1; step <= prm->steps; step++) {
      const double t_step = now_ms();
      fgpu_nr_stat st{};
      st.step = step;
      const double lam_prev = lam;
      lam = prm->lambdas ? prm->lambdas[step - 1] : lam + prm->factor;
      st.lam = lam;
      launch_axpby(c, n, lam, dq.p, 0.0, nullptr, df.p);  // f = lambda q
      const double flen = len(df.p);
      launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);  // r = f - g(u_prev)
      const double ref0 = flen > prm->tol ? flen : len(dg.p);
      double err = len(dr.p) / std::max(1.0, ref0);
      // predictor (dlast still holds the state the previous step started from)
      bool predicted = false;
      if (prm->predict) {
        predicted = prm->predict(prm->predict_user, step, du.p, lam_prev, lam,
                                 have_last ? dlast.p : nullptr, lam_last_step, dpred.p) != 0;
      } else if (xp && hs.size() >= 2) {
        // Lagrange weights of the kept states at lam (they sum to one)
        const int m = (int)hs.size();
        std::vector<double> w(m, 1.0);
        bool distinct = true;
        for (int j = 0; j < m; j++)
          for (int i = 0; i < m; i++)
            if (i != j) {
              if (hl[j] == hl[i])
                distinct = false;
              else
                w[j] *= (lam - hl[i]) / (hl[j] - hl[i]);
            }
        if (distinct) {
          // d = sum_j w_j U_j - u; the newest kept state is u itself
          launch_axpby(c, n, w[m - 1] - 1.0, du.p, w[0], hist.p + (size_t)hs[0] * n, dpred.p);
          for (int j = 1; j < m - 1; j++)
            launch_axpby(c, n, 1.0, dpred.p, w[j], hist.p + (size_t)hs[j] * n, dpred.p);
          predicted = true;
        }
      } else if (prm->secant && have_last && lam_prev != lam_last_step) {
        launch_axpby(c, n, (lam - lam_prev) / (lam_prev - lam_last_step), dinc.p, 0.0, nullptr,
                     dpred.p);
        predicted = true;
      }
      FGPU_CHECK(hipMemcpyAsync(dlast.p, du.p, nb, hipMemcpyDeviceToDevice, c.stream));
      if (predicted && prm->galerkin) {
        // d = sum_i a_i b_i over b = {prediction, secant of the last step}
        // minimising the quadratic model 1/2 d.K d - d.r of the tangent at
        // the previous state (a Galerkin projection onto the span)
        if (!k_valid) {
          if (!stiffness_dev(c, unod)) return 0;
          st.ms_stiffness += c.last_ms[0] + c.last_ms[1];
          k_valid = true;
        }
        const bool two = have_last && lam_prev != lam_last_step;
        if (two)
          launch_axpby(c, n, (lam - lam_prev) / (lam_prev - lam_last_step), dinc.p, 0.0, nullptr,
                       db2.p);
        spmv_eq(c, dpred.p, dw1.p);
        const double g11 = dev_dot(c, dpred.p, dw1.p, n), r1 = dev_dot(c, dpred.p, dr.p, n);
        double a1 = g11 > 0 ? r1 / g11 : 1.0, a2 = 0.0;
        if (two) {
          spmv_eq(c, db2.p, dw2.p);
          const double g12 = dev_dot(c, dpred.p, dw2.p, n), g22 = dev_dot(c, db2.p, dw2.p, n);
          const double r2 = dev_dot(c, db2.p, dr.p, n);
          const double det = g11 * g22 - g12 * g12;
          if (g11 > 0 && g22 > 0 && det > 1e-12 * g11 * g22) {
            a1 = (r1 * g22 - r2 * g12) / det;
            a2 = (r2 * g11 - r1 * g12) / det;
          }
        }
        launch_axpby(c, n, a1, dpred.p, a2, two ? db2.p : nullptr, dpred.p);
      }
      if (predicted) {
        // accept the prediction d when it lowers the potential energy:
        // Pi(u + d) - Pi(u) = -int_0^1 d . r(u + s d) ds ~ -d . (r0 + r1) / 2
        // (trapezoidal rule on the residuals before and after).  The l2 norm
        // of the residual is not the test: it weights the stiff, fast-to-solve
        // modes most and rejects good extrapolations.
        FGPU_CHECK(hipMemcpyAsync(dgsave.p, dg.p, nb, hipMemcpyDeviceToDevice, c.stream));
        FGPU_CHECK(hipMemcpyAsync(dr0.p, dr.p, nb, hipMemcpyDeviceToDevice, c.stream));
        launch_axpby(c, n, 1.0, du.p, 1.0, dpred.p, du.p);
        launch_eq_to_nodal(c, du.p, unod);
        bool ok = int_force_dev(c, unod, dg.p);
        st.ms_fint += c.last_ms[2];
        double errp = 0, work = 0;
        if (ok) {
          launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
          const double refp = flen > prm->tol ? flen : len(dg.p);
          errp = len(dr.p) / std::max(1.0, refp);
          work = dev_dot(c, dpred.p, dr0.p, n) + dev_dot(c, dpred.p, dr.p, n);
        }
        if (ok && std::isfinite(errp) && work > 0) {
          err = errp;
          st.predicted = 1;
        } else {  // back to the converged state of the previous step
          c.err.clear();
          st.predicted = -1;
          FGPU_CHECK(hipMemcpyAsync(du.p, dlast.p, nb, hipMemcpyDeviceToDevice, c.stream));
          FGPU_CHECK(hipMemcpyAsync(dg.p, dgsave.p, nb, hipMemcpyDeviceToDevice, c.stream));
          launch_eq_to_nodal(c, du.p, unod);
          launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
        }
      }
      st.err_start = err;
      // the reference builds K, solves and updates at least once per step
      // (ctrl/ctrlnr.cpp:153-227); only an accepted prediction that already
      // meets the tolerance ends the step without an iteration
      bool conv = st.predicted == 1 && err < prm->tol;
      int it = 0;
      for (it = 1; it <= prm->maxit && !conv; it++) {
        double ptol = prm->pcg_tol;
        if (prm->adaptive_pcg == 1 && it > 1) ptol = std::min(ptol, std::max(1e-12, 0.1 * err));
        if (prm->adaptive_pcg == 2)
          ptol = std::min(prm->pcg_tol, std::max(1e-12, 0.5 * prm->tol / std::max(err, 1e-300)));
        int pit = 0;
        double prr = 0;
        if (use_graph && warm) {
          if (!G.head) {
            for (hipEvent_t &e : G.ev) FGPU_CHECK(hipEventCreate(&e));
            if (prm->direct) {
              if (!band_setup(c)) return 0;
              band_solve_setup(c);
              capture(c, G.head, [&] {
                enqueue_stiffness(c, unod);
                enqueue_band_factor(c);
                enqueue_band_solve(c, dr.p, ddu.p, c.d_b);
                enqueue_tail(false);
              });
            } else {
              capture(c, G.head, [&] {
                enqueue_stiffness(c, unod);
                enqueue_pcg_start(c, dr.p, d_tol, abs_exit);
                enqueue_pcg_chunk(c);
              });
              capture(c, G.tail, [&] { enqueue_tail(true); });
            }
          }
          k_valid = true;
          if (!prm->direct) launch_fill(c, d_tol, 1, ptol);
          FGPU_CHECK(hipEventRecord(G.ev[0], c.stream));
          FGPU_CHECK(hipGraphLaunch(G.head, c.stream));
          FGPU_CHECK(hipEventRecord(G.ev[1], c.stream));
          if (!prm->direct) {
            pcg_continue(c, pcg_maxit, ptol, abs_exit, &pit, &prr);
            FGPU_CHECK(hipEventRecord(G.ev[2], c.stream));
            FGPU_CHECK(hipGraphLaunch(G.tail, c.stream));
          } else {
            FGPU_CHECK(hipEventRecord(G.ev[2], c.stream));
          }
          FGPU_CHECK(hipEventRecord(G.ev[3], c.stream));
          int info[2] = {0x7fffffff, 0};
          FGPU_CHECK(hipMemcpyAsync(c.h_status, c.d_status, 8 * sizeof(int),
                                    hipMemcpyDeviceToHost, c.stream));
          if (prm->direct)
            FGPU_CHECK(hipMemcpyAsync(info, c.d_finfo, sizeof info, hipMemcpyDeviceToHost,
                                      c.stream));
          FGPU_CHECK(hipMemcpyAsync(c.h_scal + 11, d_rr, 2 * sizeof(double),
                                    hipMemcpyDeviceToHost, c.stream));
          FGPU_CHECK(hipStreamSynchronize(c.stream));
          const float t_head = ev_ms(G.ev[0], G.ev[1]), t_mid = ev_ms(G.ev[1], G.ev[2]),
                      t_tail = ev_ms(G.ev[2], G.ev[3]);
          st.ms_stiffness += t_head;  // the whole iteration with the direct solver
          st.ms_solve += t_mid;
          st.ms_fint += t_tail;
          st.pcg_iters += pit;
          if (!check_status_values(c, "stiffness", c.h_status)) return 0;
          if (prm->direct && !band_factor_result(c, info)) return 0;
          if (verbose)
            std::fprintf(stderr, "fgpu_newton: step %d lambda %.6g iter %d |r|rel %.3e pcg_tol "
                         "%.2e pcg %d (%.2f ms, graph)\n", step, lam, it, err, ptol, pit,
                         t_head + t_mid + t_tail);
          if (!check_status_values(c, "internal force", c.h_status + 4)) {
            st.iters = it;
            if (stats) stats[step - 1] = st;
            return 0;
          }
          const double rlen = std::sqrt(c.h_scal[11]);
          double ref = flen;
          if (!(flen > prm->tol)) ref = std::sqrt(c.h_scal[12]);
          err = rlen / std::max(1.0, ref);
          if (err < prm->tol) conv = true;
          continue;
        }
        if (!stiffness_dev(c, unod)) return 0;
        k_valid = true;
        st.ms_stiffness += c.last_ms[0] + c.last_ms[1];
        FGPU_CHECK(hipEventRecord(c.ev[5], c.stream));
        if (prm->direct) {
          if (!band_factor(c)) return 0;
          band_solve(c, dr.p, ddu.p);
        } else {
          run_pcg(c, dr.p, ddu.p, nullptr, pcg_maxit, ptol, &pit, &prr, prm->adaptive_pcg == 0);
        }
        FGPU_CHECK(hipEventRecord(c.ev[6], c.stream));
        FGPU_CHECK(hipEventSynchronize(c.ev[6]));
        float ms = 0;
        FGPU_CHECK(hipEventElapsedTime(&ms, c.ev[5], c.ev[6]));
        st.ms_solve += ms;
        st.pcg_iters += pit;
        if (verbose)
          std::fprintf(stderr, "fgpu_newton: step %d lambda %.6g iter %d |r|rel %.3e pcg_tol %.2e "
                       "pcg %d (%.2f ms)\n", step, lam, it, err, ptol, pit, ms);
        launch_axpby(c, n, 1.0, du.p, 1.0, ddu.p, du.p);  // u += du
        launch_eq_to_nodal(c, du.p, unod);
        if (!int_force_dev(c, unod, dg.p)) {  // stress update failed: stop
          st.iters = it;
          if (stats) stats[step - 1] = st;
          return 0;
        }
        warm = true;
        st.ms_fint += c.last_ms[2];
        launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
        const double rlen = len(dr.p);
        double ref = flen;
        if (!(flen > prm->tol)) ref = len(dg.p);
        err = rlen / std::max(1.0, ref);
        if (err < prm->tol) conv = true;
      }
      st.iters = conv ? it - 1 : prm->maxit;
      st.err = err;
      st.converged = conv;
      // converged increment of this step (for the secant predictor)
      launch_axpby(c, n, 1.0, du.p, -1.0, dlast.p, dinc.p);
      if (conv) keep_state(lam);
      have_last = true;
      lam_last_step = lam_prev;
      if (prm->u_steps_dev)
        FGPU_CHECK(hipMemcpyAsync(prm->u_steps_dev + (size_t)(step - 1) * n, du.p, nb,
                                  hipMemcpyDeviceToDevice, c.stream));
      if (u_steps)
        FGPU_CHECK(hipMemcpyAsync(u_steps + (size_t)(step - 1) * n, du.p, nb,
                                  hipMemcpyDeviceToHost, c.stream));
      FGPU_CHECK(hipStreamSynchronize(c.stream));
      st.ms_total = now_ms() - t_step;
      if (stats) stats[step - 1] = st;
      if (!conv) {
        c.err = "Newton iterations did not converge at step " + std::to_string(step);
        break;
      }
      done_steps++;
    }
    return done_steps;
  });
}

}  // extern "C"
