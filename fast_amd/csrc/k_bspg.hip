// k_bspg.hip - B-spline solids of any degree (p <= 7 per direction) and any
// Gauss order (<= 16 points per direction, <= 512 per element) on gfx950:
// FP64 vector kernels beside the tuned cubic kernels of k_bsp.hip.
//
// Reference path: the same as k_bsp.hip (cShapeBspSolid::DrvShpRST,
// shp/shpbsp.cpp:985-1067; cShapeBsp::EvalBspBas / EvalBspBasDerv,
// shpbsp.cpp:254-300; nBernsteinBasis::CompPols / CompDervPols,
// geom/bernbasis.cpp:91-124; cElmParam / cElmParamTL StiffMat, IntForce,
// IntPntStress, elc/elmparam.cpp:449-597, 859-937, 1615-1885; cElemLoad::
// ExtForce, load/load.cpp:618-691) for every degree the reference reads.
//
// Per chunk of elements, state-free geometry first (one workgroup per
// element: univariate tables, then per point the rational weight sums,
// Jacobian and inverse, then per (point, control point) R_a and
// G_a = dR_a/dx into global scratch), then, for a state, per point
// L = grad u and the stress terms, then one thread per output entry: a
// stored K block (sums over the points in FP64, kept in FP64: these
// elements serve small, often slender models whose Newton iterations need
// the full tangent), an f_int or body-force vector, a point stress.  The
// K closed forms are those of the hex8 and cubic kernels:
//   linear  K_ab = sum c [lambda g h^T + mu h g^T + mu (g.h) I]
//   SVK     K_ab = sum c [lambda (Fg)(Fh)^T + mu (Fh)(Fg)^T + mu (g.h) F F^T + (g^T S h) I]
//   NH      K_ab = sum c [lambda g'h'^T + (mu - lambda ln J) h'g'^T + mu (g.h) I], g' = F^-T g
#include "fgpu_internal.h"

namespace fgpu {
namespace {

constexpr int GMAXB = 8;   // basis functions per direction
constexpr int GMAXQ = 16;  // Gauss points per direction (the table's size)
// per point: c, L (9), S (6), F F^T (6), mu - lambda ln J, P = F S or sigma (9)
constexpr int NPQ = 32;
enum { PQ_C = 0, PQ_L = 1, PQ_S = 10, PQ_FF = 16, PQ_MU = 22, PQ_P = 23 };

struct GDims {
  int nb[3], nq[3];       // basis functions and points per direction
  int off[3], row[3];     // extraction operator of direction d: row i at off + i row
  int nn, nqt, ext_per;   // control points, points, extraction entries per element
};

// B_j and dB_j / dt (with the 1/2 of t' = (t + 1) / 2), degree p, at u
__device__ inline void bernstein_p(int p, double u, double *B, double *dB) {
  double b[GMAXB];
  b[0] = 1.0;
  for (int k = 1; k < p; k++) {  // degree p - 1 by de Casteljau's recursion
    b[k] = 0.0;
    for (int j = k; j >= 1; j--) b[j] = (1.0 - u) * b[j] + u * b[j - 1];
    b[0] = (1.0 - u) * b[0];
  }
  for (int j = 0; j <= p; j++) {
    const double lo = j >= 1 ? b[j - 1] : 0.0, hi = j < p ? b[j] : 0.0;
    dB[j] = 0.5 * p * (lo - hi);
  }
  if (p == 0) {
    B[0] = 1.0;
    return;
  }
  b[p] = 0.0;
  for (int j = p; j >= 1; j--) b[j] = (1.0 - u) * b[j] + u * b[j - 1];
  b[0] = (1.0 - u) * b[0];
  for (int j = 0; j <= p; j++) B[j] = b[j];
}

__device__ inline double inv3g(const double J[3][3], double I[3][3]) {
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / det;
  I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  I[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) * id;
  I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  I[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) * id;
  I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  I[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) * id;
  I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  I[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) * id;
  I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

struct GLds {
  double *N, *dN, *X, *W, *pt;  // tables [3][GMAXQ][GMAXB], X [nn][3], W [nn], pt [nqt][13]
};

__device__ inline GLds glds(double *smem, const GDims &D) {
  GLds s;
  s.N = smem;
  s.dN = s.N + 3 * GMAXQ * GMAXB;
  s.X = s.dN + 3 * GMAXQ * GMAXB;
  s.W = s.X + 3 * D.nn;
  s.pt = s.W + D.nn;
  return s;
}

// numerators (N w, dN w) of basis a at point q
__device__ inline void gtensor(const GLds &s, const GDims &D, int q, int a, double &n,
                               double d[3]) {
  const int qr = q % D.nq[0], qs = (q / D.nq[0]) % D.nq[1], qt = q / (D.nq[0] * D.nq[1]);
  const int i = a % D.nb[0], j = (a / D.nb[0]) % D.nb[1], k = a / (D.nb[0] * D.nb[1]);
  const double w = s.W[a];
  const double Nr = s.N[(0 * GMAXQ + qr) * GMAXB + i], Ns = s.N[(1 * GMAXQ + qs) * GMAXB + j],
               Nt = s.N[(2 * GMAXQ + qt) * GMAXB + k];
  n = Nr * Ns * Nt * w;
  d[0] = s.dN[(0 * GMAXQ + qr) * GMAXB + i] * Ns * Nt * w;
  d[1] = Nr * s.dN[(1 * GMAXQ + qs) * GMAXB + j] * Nt * w;
  d[2] = Nr * Ns * s.dN[(2 * GMAXQ + qt) * GMAXB + k] * w;
}

// geometry of the chunk's elements: c per point (PQ), R_a and G_a per
// (point, control point)
__global__ __launch_bounds__(256) void gen_geo_kernel(
    int e0, GDims D, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext, const double *__restrict__ gp,
    double *__restrict__ G, double *__restrict__ R, double *__restrict__ PQ,
    int *__restrict__ status) {
  extern __shared__ double smem[];
  const GLds s = glds(smem, D);
  const int le = blockIdx.x, e = e0 + le, nn = D.nn, nqt = D.nqt;
  const int *cn = conn + (size_t)e * nn;
  for (int a = threadIdx.x; a < nn; a += blockDim.x) {
    const int n = cn[a], n0 = cn[0];
    for (int k = 0; k < 3; k++)
      s.X[3 * a + k] = coords[3 * (size_t)n + k] - coords[3 * (size_t)n0 + k];
    s.W[a] = cpw ? cpw[n] : 1.0;
  }
  for (int t = threadIdx.x; t < 3 * GMAXQ * GMAXB; t += blockDim.x) {
    const int d = t / (GMAXQ * GMAXB), q = (t / GMAXB) % GMAXQ, i = t % GMAXB;
    double n = 0.0, dn = 0.0;
    if (q < D.nq[d] && i < D.nb[d]) {
      const int p = D.nb[d] - 1;
      const double u = (gp[(d * 16 + q) * 2] + 1.0) / 2.0;
      double B[GMAXB], dB[GMAXB];
      bernstein_p(p, u, B, dB);
      const double *C = ext + (size_t)e * D.ext_per + D.off[d] + i * D.row[d];
      for (int j = 0; j <= p; j++) {
        n += C[j] * B[j];
        dn += C[j] * dB[j];
      }
    }
    s.N[t] = n;
    s.dN[t] = dn;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nqt; q += blockDim.x) {
    double sw = 0.0, ds[3] = {0, 0, 0};
    for (int a = 0; a < nn; a++) {
      double n, d[3];
      gtensor(s, D, q, a, n, d);
      sw += n;
      for (int k = 0; k < 3; k++) ds[k] += d[k];
    }
    const double is = 1.0 / sw;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < nn; a++) {
      double n, d[3];
      gtensor(s, D, q, a, n, d);
      const double Ra = n * is;
      for (int i = 0; i < 3; i++) {
        const double dR = (d[i] - Ra * ds[i]) * is;
        for (int k = 0; k < 3; k++) J[i][k] += dR * s.X[3 * a + k];
      }
    }
    double Ji[3][3];
    const double det = inv3g(J, Ji);
    count_jacobian(status, det);
    const int qr = q % D.nq[0], qs = (q / D.nq[0]) % D.nq[1], qt = q / (D.nq[0] * D.nq[1]);
    const double w = gp[(0 * 16 + qr) * 2 + 1] * gp[(1 * 16 + qs) * 2 + 1] * gp[(2 * 16 + qt) * 2 + 1];
    PQ[((size_t)le * nqt + q) * NPQ + PQ_C] = w * fabs(det);
    double *P = s.pt + 13 * q;
    P[0] = is;
    for (int k = 0; k < 3; k++) P[1 + k] = ds[k];
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) P[4 + 3 * i + k] = Ji[i][k];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nqt * nn; t += blockDim.x) {
    const int q = t / nn, a = t % nn;
    double n, d[3];
    gtensor(s, D, q, a, n, d);
    const double *P = s.pt + 13 * q;
    const double Ra = n * P[0];
    double dR[3];
    for (int i = 0; i < 3; i++) dR[i] = (d[i] - Ra * P[1 + i]) * P[0];
    const size_t o = ((size_t)le * nqt + q) * nn + a;
    R[o] = Ra;
    for (int i = 0; i < 3; i++)
      G[3 * o + i] = P[4 + 3 * i] * dR[0] + P[4 + 3 * i + 1] * dR[1] + P[4 + 3 * i + 2] * dR[2];
  }
}

// per point of the chunk: L = grad u and the stress terms of FORM
template <int FORM>
__global__ __launch_bounds__(256) void gen_point_kernel(
    int e0, int ne, GDims D, const int *__restrict__ conn, const double *__restrict__ mat,
    const double *__restrict__ unodal, const double *__restrict__ G, double *__restrict__ PQ,
    int *__restrict__ status) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * D.nqt) return;
  const int le = (int)(t / D.nqt), q = (int)(t % D.nqt), e = e0 + le, nn = D.nn;
  const int *cn = conn + (size_t)e * nn;
  double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double *g = G + ((size_t)le * D.nqt + q) * nn * 3;
  for (int a = 0; a < nn; a++) {
    const double *ua = unodal + 3 * (size_t)cn[a];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) L[i][j] += ua[i] * g[3 * a + j];
  }
  double *o = PQ + (size_t)t * NPQ;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[PQ_L + 3 * i + j] = L[i][j];
  const double p0 = mat[2 * (size_t)e], p1 = mat[2 * (size_t)e + 1];
  double F[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1.0 : 0.0) + L[i][j];
  double sv[6];  // Voigt stress that multiplies the gradients
  if (!point_stress<FORM>(L, p0, p1, sv)) atomicAdd(&status[1], 1);
  const double Sm[3][3] = {{sv[0], sv[3], sv[4]}, {sv[3], sv[1], sv[5]}, {sv[4], sv[5], sv[2]}};
  for (int k = 0; k < 6; k++) o[PQ_S + k] = sv[k];
  if (FORM == FGPU_LINEAR) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) o[PQ_P + 3 * i + j] = Sm[i][j];
  } else {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        o[PQ_P + 3 * i + j] = F[i][0] * Sm[0][j] + F[i][1] * Sm[1][j] + F[i][2] * Sm[2][j];
    const int ff[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 6; k++) {
      const int i = ff[k][0], j = ff[k][1];
      o[PQ_FF + k] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
    }
    if (FORM == FGPU_TL_NEOHOOKE) {  // mu - lambda ln J, J^2 = det(F^T F)
      double Fi[3][3];
      const double J = inv3g(F, Fi);
      o[PQ_MU] = p0 - p1 * (J > 0.0 ? log(J) : 0.0);
    }
  }
}

// per (point, control point): F g (St.Venant-Kirchhoff) or F^-T g (neo-Hooke)
template <int FORM>
__global__ __launch_bounds__(256) void gen_h_kernel(int ne, GDims D, const double *__restrict__ G,
                                                   const double *__restrict__ PQ,
                                                   double *__restrict__ H) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * D.nqt * D.nn) return;
  const long long pq = t / D.nn;
  const double *L = PQ + pq * NPQ + PQ_L;
  double F[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1.0 : 0.0) + L[3 * i + j];
  const double *g = G + 3 * t;
  double *h = H + 3 * t;
  if (FORM == FGPU_TL_SVK) {
    for (int i = 0; i < 3; i++) h[i] = F[i][0] * g[0] + F[i][1] * g[1] + F[i][2] * g[2];
  } else {
    double Fi[3][3];
    inv3g(F, Fi);
    for (int i = 0; i < 3; i++) h[i] = Fi[0][i] * g[0] + Fi[1][i] * g[1] + Fi[2][i] * g[2];
  }
}

// one stored block (a <= b) of one element
template <int FORM>
__global__ __launch_bounds__(256) void gen_K_kernel(
    int e0, int ne, GDims D, int nsym, const int2 *__restrict__ sym,
    const double *__restrict__ mat, const double *__restrict__ G, const double *__restrict__ H,
    const double *__restrict__ PQ, double *__restrict__ Ke) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * nsym) return;
  const int le = (int)(t / nsym), sb = (int)(t % nsym), e = e0 + le, nn = D.nn;
  const int a = sym[sb].x, b = sym[sb].y;
  const double p0 = mat[2 * (size_t)e], p1 = mat[2 * (size_t)e + 1];
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, M2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double Bs[6] = {0, 0, 0, 0, 0, 0}, sc = 0.0;
  for (int q = 0; q < D.nqt; q++) {
    const size_t pq = (size_t)le * D.nqt + q;
    const double *P = PQ + pq * NPQ;
    const double c = P[PQ_C];
    const double *ga = G + 3 * (pq * nn + a), *gb = G + 3 * (pq * nn + b);
    if (FORM == FGPU_LINEAR) {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[3 * i + j] += c * (ga[i] * gb[j]);
    } else {
      const double *ha = H + 3 * (pq * nn + a), *hb = H + 3 * (pq * nn + b);
      const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
      if (FORM == FGPU_TL_SVK) {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) M[3 * i + j] += c * (ha[i] * hb[j]);
        for (int k = 0; k < 6; k++) Bs[k] += c * gg * P[PQ_FF + k];
        const double *S = P + PQ_S;
        sc += c * (ga[0] * (S[0] * gb[0] + S[3] * gb[1] + S[4] * gb[2]) +
                   ga[1] * (S[3] * gb[0] + S[1] * gb[1] + S[5] * gb[2]) +
                   ga[2] * (S[4] * gb[0] + S[5] * gb[1] + S[2] * gb[2]));
      } else {
        const double cl = c * p1, cm2 = c * P[PQ_MU];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            const double pr = ha[i] * hb[j];
            M[3 * i + j] += cl * pr;
            M2[3 * i + j] += cm2 * pr;
          }
        sc += c * p0 * gg;
      }
    }
  }
  double K[9];
  if (FORM == FGPU_LINEAR || FORM == FGPU_TL_SVK) {
    const double lam = p0 * p1 / ((1.0 + p1) * (1.0 - 2.0 * p1)), mu = 0.5 * p0 / (1.0 + p1);
    const double Bm[9] = {Bs[0], Bs[3], Bs[4], Bs[3], Bs[1], Bs[5], Bs[4], Bs[5], Bs[2]};
    const double tr = M[0] + M[4] + M[8];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        K[3 * i + j] = FORM == FGPU_LINEAR
                           ? lam * M[3 * i + j] + mu * M[3 * j + i] + (i == j ? mu * tr : 0.0)
                           : lam * M[3 * i + j] + mu * M[3 * j + i] + mu * Bm[3 * i + j] +
                                 (i == j ? sc : 0.0);
  } else {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) K[3 * i + j] = M[3 * i + j] + M2[3 * j + i] + (i == j ? sc : 0.0);
  }
  double *out = Ke + ((size_t)e * nsym + sb) * 9;
  for (int k = 0; k < 9; k++) out[k] = K[k];
}

// g_a = sum_q c P G_a (internal force) or f_a = sum_q c R_a b (body force)
template <bool BODY>
__global__ __launch_bounds__(256) void gen_vec_kernel(int e0, int ne, GDims D,
                                                     const double *__restrict__ G,
                                                     const double *__restrict__ R,
                                                     const double *__restrict__ PQ,
                                                     const double *__restrict__ bvec,
                                                     double *__restrict__ ge) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * D.nn) return;
  const int le = (int)(t / D.nn), a = (int)(t % D.nn), e = e0 + le, nn = D.nn;
  double v[3] = {0, 0, 0}, sR = 0.0;
  for (int q = 0; q < D.nqt; q++) {
    const size_t pq = (size_t)le * D.nqt + q;
    const double *P = PQ + pq * NPQ;
    if (BODY) {
      sR += P[PQ_C] * R[pq * nn + a];
    } else {
      const double *g = G + 3 * (pq * nn + a);
      const double c = P[PQ_C];
      for (int i = 0; i < 3; i++)
        v[i] += c * (P[PQ_P + 3 * i] * g[0] + P[PQ_P + 3 * i + 1] * g[1] + P[PQ_P + 3 * i + 2] * g[2]);
    }
  }
  double *o = ge + ((size_t)e * nn + a) * 3;
  for (int i = 0; i < 3; i++) o[i] = BODY ? sR * bvec[3 * (size_t)e + i] : v[i];
}

// point stresses (Voigt) and their extrapolation to the element's nodes
__global__ __launch_bounds__(256) void gen_sig_kernel(int e0, int ne, GDims D,
                                                     const double *__restrict__ PQ,
                                                     double *__restrict__ sig_ip) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * D.nqt) return;
  const double *P = PQ + (size_t)t * NPQ + PQ_S;
  double *o = sig_ip + ((size_t)e0 * D.nqt + t) * 6;
  for (int k = 0; k < 6; k++) o[k] = P[k];
}

__global__ __launch_bounds__(256) void gen_sig_node_kernel(int e0, int ne, GDims D,
                                                          const double *__restrict__ TR,
                                                          const double *__restrict__ sig_ip,
                                                          double *__restrict__ sig_node) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * D.nn) return;
  const int le = (int)(t / D.nn), a = (int)(t % D.nn), e = e0 + le;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const double *si = sig_ip + (size_t)e * D.nqt * 6;
  for (int q = 0; q < D.nqt; q++) {
    const double w = TR[(size_t)a * D.nqt + q];
    for (int k = 0; k < 6; k++) s[k] += w * si[6 * q + k];
  }
  double *o = sig_node + ((size_t)e * D.nn + a) * 6;
  for (int k = 0; k < 6; k++) o[k] = s[k];
}

GDims gdims(const Ctx &c) {
  GDims D;
  int off = 0;
  const bool packed = bsp_generic(c);  // else the tuned kernels' [e][dir][4][4]
  for (int d = 0; d < 3; d++) {
    D.nb[d] = c.deg[d] + 1;
    D.nq[d] = c.nip[d];
    D.off[d] = packed ? off : 16 * d;
    D.row[d] = packed ? D.nb[d] : 4;
    off += D.nb[d] * D.nb[d];
  }
  D.ext_per = packed ? off : 48;
  D.nn = c.nn;
  D.nqt = c.nip[0] * c.nip[1] * c.nip[2];
  return D;
}

unsigned grid_of(long long n) { return (unsigned)((n + 255) / 256); }

size_t geo_lds(const GDims &D) {
  return sizeof(double) * (2 * 3 * GMAXQ * GMAXB + 4 * (size_t)D.nn + 13 * (size_t)D.nqt);
}

// scratch for a chunk of elements (at most 256 MB of it); c.gen_cap holds the
// (point, control point) entries the buffers were sized for
void gen_scratch(Ctx &c, const GDims &D, bool need_h) {
  const size_t per = sizeof(double) * ((size_t)D.nqt * D.nn * 7 + (size_t)D.nqt * NPQ);
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(c.nelem, (256u << 20) / per));
  const size_t gq = (size_t)chunk * D.nqt * D.nn, pq = (size_t)chunk * D.nqt;
  if (gq > c.gen_cap || pq > c.gen_cap_pq) {
    for (double *p : {c.d_gG, c.d_gR, c.d_gH, c.d_gPQ})
      if (p) (void)hipFree(p);
    c.d_gG = c.d_gR = c.d_gH = c.d_gPQ = nullptr;
    c.gen_cap = gq;
    c.gen_cap_pq = pq;
    FGPU_CHECK(hipMalloc(&c.d_gG, sizeof(double) * 3 * gq));
    FGPU_CHECK(hipMalloc(&c.d_gR, sizeof(double) * gq));
    FGPU_CHECK(hipMalloc(&c.d_gPQ, sizeof(double) * NPQ * pq));
  }
  if (need_h && !c.d_gH) FGPU_CHECK(hipMalloc(&c.d_gH, sizeof(double) * 3 * c.gen_cap));
  c.gen_chunk = chunk;
}

void launch_geo(Ctx &c, const GDims &D, int e0, int ne, const double *gp = nullptr,
                int *status = nullptr) {
  const size_t lds = geo_lds(D);
  FGPU_CHECK(hipFuncSetAttribute((const void *)gen_geo_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(gen_geo_kernel, dim3(ne), dim3(256), lds, c.stream, e0, D, c.d_coords,
                     c.d_conn, c.d_cpw, c.d_ext, gp ? gp : c.d_gauss64, c.d_gG, c.d_gR, c.d_gPQ,
                     status ? status : c.d_status);
}

template <int FORM>
void launch_point(Ctx &c, const GDims &D, int e0, int ne, const double *u) {
  hipLaunchKernelGGL((gen_point_kernel<FORM>), dim3(grid_of((long long)ne * D.nqt)), dim3(256), 0,
                     c.stream, e0, ne, D, c.d_conn, c.d_mat, u, c.d_gG, c.d_gPQ, c.d_status);
}

void launch_point_f(Ctx &c, const GDims &D, int e0, int ne, const double *u) {
  switch (c.form) {
    case FGPU_TL_SVK: launch_point<FGPU_TL_SVK>(c, D, e0, ne, u); break;
    case FGPU_TL_NEOHOOKE: launch_point<FGPU_TL_NEOHOOKE>(c, D, e0, ne, u); break;
    default: launch_point<FGPU_LINEAR>(c, D, e0, ne, u); break;
  }
}

template <int FORM>
void launch_K_chunk(Ctx &c, const GDims &D, int e0, int ne) {
  const int nsym = nsym_blocks(D.nn);
  if (FORM != FGPU_LINEAR)
    hipLaunchKernelGGL((gen_h_kernel<FORM>), dim3(grid_of((long long)ne * D.nqt * D.nn)),
                       dim3(256), 0, c.stream, ne, D, c.d_gG, c.d_gPQ, c.d_gH);
  hipLaunchKernelGGL((gen_K_kernel<FORM>), dim3(grid_of((long long)ne * nsym)), dim3(256), 0,
                     c.stream, e0, ne, D, nsym, (const int2 *)c.d_gsym, c.d_mat, c.d_gG, c.d_gH,
                     c.d_gPQ, c.d_Ke64);
}

}  // namespace

bool bsp_generic(const Ctx &c) {
  return !(c.deg[0] == 3 && c.deg[1] == 3 && (c.deg[2] == 3 || c.deg[2] == 2) && c.nip[0] == 4 &&
           c.nip[1] == 4 && c.nip[2] == 4);
}

void gen_prepare(Ctx &c, const double *h_ext) {
  const GDims D = gdims(c);
  const size_t n = (size_t)c.nelem * D.ext_per;
  FGPU_CHECK(hipMalloc(&c.d_ext, n * sizeof(double)));
  FGPU_CHECK(hipMemcpyAsync(c.d_ext, h_ext, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  const int nsym = nsym_blocks(c.nn);
  std::vector<int2> sym(nsym);
  for (int a = 0, k = 0; a < c.nn; a++)
    for (int b = a; b < c.nn; b++, k++) sym[k] = make_int2(a, b);
  FGPU_CHECK(hipMalloc(&c.d_gsym, sizeof(int2) * nsym));
  FGPU_CHECK(hipMemcpyAsync(c.d_gsym, sym.data(), sizeof(int2) * nsym, hipMemcpyHostToDevice,
                            c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
}

void gen_launch_K(Ctx &c, const double *u) {
  const GDims D = gdims(c);
  gen_scratch(c, D, c.form != FGPU_LINEAR);
  for (int e0 = 0; e0 < c.nelem; e0 += c.gen_chunk) {
    const int ne = std::min(c.gen_chunk, c.nelem - e0);
    launch_geo(c, D, e0, ne);
    if (c.form != FGPU_LINEAR) launch_point_f(c, D, e0, ne, u);
    switch (c.form) {
      case FGPU_TL_SVK: launch_K_chunk<FGPU_TL_SVK>(c, D, e0, ne); break;
      case FGPU_TL_NEOHOOKE: launch_K_chunk<FGPU_TL_NEOHOOKE>(c, D, e0, ne); break;
      default: launch_K_chunk<FGPU_LINEAR>(c, D, e0, ne); break;
    }
  }
  FGPU_CHECK(hipGetLastError());
}

void gen_launch_vec(Ctx &c, const double *u, const double *b) {
  const GDims D = gdims(c);
  gen_scratch(c, D, false);
  for (int e0 = 0; e0 < c.nelem; e0 += c.gen_chunk) {
    const int ne = std::min(c.gen_chunk, c.nelem - e0);
    launch_geo(c, D, e0, ne);
    const unsigned g = grid_of((long long)ne * D.nn);
    if (b) {
      hipLaunchKernelGGL((gen_vec_kernel<true>), dim3(g), dim3(256), 0, c.stream, e0, ne, D,
                         c.d_gG, c.d_gR, c.d_gPQ, b, c.d_ge);
    } else {
      launch_point_f(c, D, e0, ne, u);
      hipLaunchKernelGGL((gen_vec_kernel<false>), dim3(g), dim3(256), 0, c.stream, e0, ne, D,
                         c.d_gG, c.d_gR, c.d_gPQ, nullptr, c.d_ge);
    }
  }
  FGPU_CHECK(hipGetLastError());
}

void gen_launch_stress(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  const GDims D = gdims(c);
  gen_scratch(c, D, false);
  for (int e0 = 0; e0 < c.nelem; e0 += c.gen_chunk) {
    const int ne = std::min(c.gen_chunk, c.nelem - e0);
    launch_geo(c, D, e0, ne);
    launch_point_f(c, D, e0, ne, u);
    hipLaunchKernelGGL(gen_sig_kernel, dim3(grid_of((long long)ne * D.nqt)), dim3(256), 0,
                       c.stream, e0, ne, D, c.d_gPQ, sig_ip);
    if (sig_node)
      hipLaunchKernelGGL(gen_sig_node_kernel, dim3(grid_of((long long)ne * D.nn)), dim3(256), 0,
                         c.stream, e0, ne, D, c.d_TR, sig_ip, sig_node);
  }
  FGPU_CHECK(hipGetLastError());
}

// Stresses at the element's nodes evaluated there (cElmParam::NodalStress
// without extrapolation, elc/elmparam.cpp:1313-1333, through PntStress,
// 1206-1290: e = B u at the nodes' natural coordinates, a uniform grid on
// [-1, 1]^3 for B-spline solids; the linear strain of PntStress also for the
// total-Lagrangian forms).  sig_node [nelem * nn * 6].
void gen_launch_node_stress(Ctx &c, const double *u, double *sig_node) {
  GDims D = gdims(c);
  if (!c.d_gnodes) {
    double t[3 * 16 * 2] = {};
    for (int d = 0; d < 3; d++)
      for (int i = 0; i < D.nb[d]; i++) {
        t[(d * 16 + i) * 2] = -1.0 + 2.0 * i / (D.nb[d] - 1);
        t[(d * 16 + i) * 2 + 1] = 1.0;
      }
    FGPU_CHECK(hipMalloc(&c.d_gnodes, sizeof t));
    FGPU_CHECK(hipMemcpyAsync(c.d_gnodes, t, sizeof t, hipMemcpyHostToDevice, c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
  }
  for (int d = 0; d < 3; d++) D.nq[d] = D.nb[d];
  D.nqt = D.nn;
  gen_scratch(c, D, false);
  for (int e0 = 0; e0 < c.nelem; e0 += c.gen_chunk) {
    const int ne = std::min(c.gen_chunk, c.nelem - e0);
    launch_geo(c, D, e0, ne, c.d_gnodes, c.d_status2);  // (node Jacobians are not checked)
    launch_point<FGPU_LINEAR>(c, D, e0, ne, u);
    hipLaunchKernelGGL(gen_sig_kernel, dim3(grid_of((long long)ne * D.nqt)), dim3(256), 0,
                       c.stream, e0, ne, D, c.d_gPQ, sig_node);
  }
  FGPU_CHECK(hipGetLastError());
}

}  // namespace fgpu
