// fgpu_drivers.cpp - analysis drivers of libfgpu: linear static and
// load-control Newton-Raphson, with every vector resident on the device.
//
// Linear static follows cLinStat::Solver (ctrl/ctrllins.cpp:46-99):
// assemble, solve.  Newton-Raphson follows cCtrlNR::Solver
// (ctrl/ctrlnr.cpp:47-267): g = f_int(0); per step lambda += factor,
// f = lambda q, r = f - g; per iteration K(u), K du = r, u += du,
// g = f_int(u), r = f - g, converged when |r| / max(1, |f|) < Tol
// (ctrl/ctrl.cpp:1366-1380; |g| replaces |f| when |f| <= Tol); the loop stops
// at the first step that does not converge.
#include <algorithm>
#include <chrono>
#include <cmath>
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
    DevVec df(c.neq), du(c.neq);
    FGPU_CHECK(hipMemcpyAsync(df.p, f, sizeof(double) * c.neq, hipMemcpyHostToDevice, c.stream));
    if (!stiffness_dev(c, nullptr)) return 0;
    int conv = run_pcg(c, df.p, du.p, nullptr, maxit > 0 ? maxit : c.neq, rtol, iters, relres);
    FGPU_CHECK(hipMemcpyAsync(u, du.p, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    if (!conv) c.err = "PCG did not reach the tolerance";
    return conv;
  });
}

int fgpu_newton(fgpu_ctx *ctx, const double *q, const fgpu_nr_params *prm, double *u_steps,
                fgpu_nr_stat *stats) {
  return guarded(ctx, [&](Ctx &c) {
    if (!q || !prm) throw Error("null argument");
    const int n = c.neq;
    const size_t nb = sizeof(double) * n;
    DevVec dq(n), df(n), dg(n), dr(n), du(n), ddu(n), dpred(n);
    FGPU_CHECK(hipMemcpyAsync(dq.p, q, nb, hipMemcpyHostToDevice, c.stream));
    FGPU_CHECK(hipMemsetAsync(du.p, 0, nb, c.stream));
    double *unod = c.d_u;  // nodal displacement field
    launch_eq_to_nodal(c, du.p, unod);
    if (!int_force_dev(c, unod, dg.p)) return 0;  // initial stresses
    const int pcg_maxit = prm->pcg_maxit > 0 ? prm->pcg_maxit : n;
    double lam = 0.0;
    int done_steps = 0;
    for (int step = 1; step <= prm->steps; step++) {
      const double t_step = now_ms();
      fgpu_nr_stat st{};
      st.step = step;
      const double lam_prev = lam;
      lam += prm->factor;
      st.lam = lam;
      launch_axpby(c, n, lam, dq.p, 0.0, nullptr, df.p);  // f = lambda q
      const double flen = std::sqrt(dev_dot(c, df.p, df.p, n));
      if (prm->predict) {
        // predictor: u += du_pred, then refresh g
        if (prm->predict(prm->predict_user, step, du.p, lam_prev, lam, dpred.p)) {
          launch_axpby(c, n, 1.0, du.p, 1.0, dpred.p, du.p);
          launch_eq_to_nodal(c, du.p, unod);
          if (!int_force_dev(c, unod, dg.p)) return 0;
          st.ms_fint += c.last_ms[2];
        }
      }
      launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);  // r = f - g
      double err = 1.0;
      bool conv = false;
      int it;
      for (it = 1; it <= prm->maxit; it++) {
        if (!stiffness_dev(c, unod)) return 0;
        st.ms_stiffness += c.last_ms[0] + c.last_ms[1];
        double ptol = prm->pcg_tol;
        if (prm->adaptive_pcg && it > 1) ptol = std::min(ptol, std::max(1e-12, 0.1 * err));
        int pit = 0;
        double prr = 0;
        FGPU_CHECK(hipEventRecord(c.ev[5], c.stream));
        run_pcg(c, dr.p, ddu.p, nullptr, pcg_maxit, ptol, &pit, &prr);
        FGPU_CHECK(hipEventRecord(c.ev[6], c.stream));
        FGPU_CHECK(hipEventSynchronize(c.ev[6]));
        float ms = 0;
        FGPU_CHECK(hipEventElapsedTime(&ms, c.ev[5], c.ev[6]));
        st.ms_solve += ms;
        st.pcg_iters += pit;
        launch_axpby(c, n, 1.0, du.p, 1.0, ddu.p, du.p);  // u += du
        launch_eq_to_nodal(c, du.p, unod);
        if (!int_force_dev(c, unod, dg.p)) {  // stress update failed: stop
          st.iters = it;
          if (stats) stats[step - 1] = st;
          return 0;
        }
        st.ms_fint += c.last_ms[2];
        launch_axpby(c, n, 1.0, df.p, -1.0, dg.p, dr.p);
        const double rlen = std::sqrt(dev_dot(c, dr.p, dr.p, n));
        double ref = flen;
        if (!(flen > prm->tol)) ref = std::sqrt(dev_dot(c, dg.p, dg.p, n));
        err = rlen / std::max(1.0, ref);
        if (err < prm->tol) {
          conv = true;
          break;
        }
      }
      st.iters = std::min(it, prm->maxit);
      st.err = err;
      st.converged = conv;
      st.ms_total = now_ms() - t_step;
      if (stats) stats[step - 1] = st;
      if (u_steps)
        FGPU_CHECK(hipMemcpyAsync(u_steps + (size_t)(step - 1) * n, du.p, nb,
                                  hipMemcpyDeviceToHost, c.stream));
      FGPU_CHECK(hipStreamSynchronize(c.stream));
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
