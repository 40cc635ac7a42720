// k_hex20.hip - 20-node serendipity brick kernels for gfx950, FP64 on the
// vector ALUs: stiffness, internal force, body force and stresses.
//
// Reference path: BRICK20 / BRICK20TL elements (elc/element.cpp:2043-2100)
// through cElmParam::StiffMat / IntForce (elc/elmparam.cpp:449-597) and
// cElmParamTL::StiffMat / IntForce (1615-1807), over cShapeBrick20
// (shp/shpsolid.cpp:301-435) and cShapeSolid::DrvShpXYZ (58-126).
//
// Node order (cShapeBrick20::NodeNatCoord): the t = +1 face counterclockwise
// from (-1, -1), corners and mid-edge nodes alternating (0..7), the four
// mid-height nodes (8..11), then the t = -1 face alike (12..19).  Serendipity
// functions: a corner c has N = (1 + r cr)(1 + s cs)(1 + t ct)(r cr + s cs +
// t ct - 2) / 8; a mid-edge node whose coordinate k is 0 has
// N = (1 - x_k^2) prod_{j != k} (1 + x_j c_j) / 4.
//
// Stiffness: the closed forms of the hex8 kernels for isotropic materials
// (per point only three-vectors per node: sqrt(c) grad N for the linear form;
// F g, g, F F^T, S, c for St.Venant-Kirchhoff; F^-T g, F F^T, c lambda,
// c mu', c mu for neo-Hooke).  Phase 1: one thread per (element, point)
// stages them in LDS for the workgroup's TE elements; phase 2: one thread per
// (element, upper node block, 210 per element) sums over the points and
// stores the 3x3 block.  The element matrices are kept in FP64 (the
// assembly's double path) in the 64-element tile layout of the hex8 kernels
// (KeLayout T = 64).
//
// Internal force, stresses: phase 1 one thread per (element, point) in FP64
// (Jacobian, its inverse, grad u, the stress); phase 2 one thread per
// (element, node) sums over the points in point order.
#include "fgpu_internal.h"

namespace fgpu {
namespace {

constexpr int NN20 = 20;
constexpr int NSYM20 = NN20 * (NN20 + 1) / 2;  // 210 upper node blocks

__constant__ signed char kC20[NN20][3] = {
    {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {-1, 1, 1},
    {-1, 0, 1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, -1, -1},
    {0, -1, -1}, {1, -1, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}, {-1, 1, -1}, {-1, 0, -1}};

// N_a and dN_a / d(r, s, t) at (r, s, t), without indexing by the node's
// zero coordinate (selects only): for a mid-edge node, f_k = 1 makes
// P_j = prod_{i != j} f_i serve both cases.
template <class T>
__device__ inline T h20_shape(T r, T s, T t, int a, T d[3]) {
  const T c0 = T(kC20[a][0]), c1 = T(kC20[a][1]), c2 = T(kC20[a][2]);
  const T f0 = T(1) + r * c0, f1 = T(1) + s * c1, f2 = T(1) + t * c2;
  const T P0 = f1 * f2, P1 = f0 * f2, P2 = f0 * f1;
  const bool corner = c0 != T(0) && c1 != T(0) && c2 != T(0);
  if (corner) {
    const T g = r * c0 + s * c1 + t * c2 - T(2);
    d[0] = T(0.125) * c0 * P0 * (g + f0);
    d[1] = T(0.125) * c1 * P1 * (g + f1);
    d[2] = T(0.125) * c2 * P2 * (g + f2);
    return T(0.125) * f0 * P0 * g;
  }
  const T q = (c0 == T(0) ? T(1) - r * r : T(1)) * (c1 == T(0) ? T(1) - s * s : T(1)) *
              (c2 == T(0) ? T(1) - t * t : T(1));
  d[0] = c0 == T(0) ? T(-0.5) * r * P0 : T(0.25) * q * c0 * P0;
  d[1] = c1 == T(0) ? T(-0.5) * s * P1 : T(0.25) * q * c1 * P1;
  d[2] = c2 == T(0) ? T(-0.5) * t * P2 : T(0.25) * q * c2 * P2;
  // N = q P_k / 4 with P_k the product of the two other factors (f_k = 1)
  return T(0.25) * q * f0 * f1 * f2;
}

template <class T>
__device__ inline T inv3x3(const T J[3][3], T I[3][3]) {
  T det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
          J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  T id = T(1) / det;
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

// (a, b) of upper block blk, packed row-major
__device__ inline void h20_block(int blk, int &a, int &b) {
  int r = blk, i = 0;
  while (r >= NN20 - i) {
    r -= NN20 - i;
    i++;
  }
  a = i;
  b = i + r;
}

// Staged values per element-point:
//   LINEAR:   sqrt(c) g (60)
//   SVK:      F g (60), g (60), F F^T (6), S (6), c       = 133
//   NEOHOOKE: F^-T g (60), F F^T (6), c lambda, c mu', c mu = 69
template <int FORM> struct Stage20;
template <> struct Stage20<FGPU_LINEAR> { static constexpr int NF = 60; };
template <> struct Stage20<FGPU_TL_SVK> { static constexpr int NF = 133; };
template <> struct Stage20<FGPU_TL_NEOHOOKE> { static constexpr int NF = 69; };

constexpr int XP20 = 61;  // per-element node data, padded (60 values)
template <int FORM>
constexpr int h20_elem_floats() { return XP20 * (FORM == FGPU_LINEAR ? 1 : 2) + 2; }

template <class T, int FORM, int NIP, int TE>
__global__ __launch_bounds__(256) void hex20_K_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, T *__restrict__ Ke, KeLayout L, int *__restrict__ status) {
  constexpr int NQ = NIP * NIP * NIP;
  constexpr int NF = Stage20<FORM>::NF;
  extern __shared__ __align__(16) unsigned char smem20[];
  T *sh = reinterpret_cast<T *>(smem20);  // [q][f][TE], then per element X, U, (lambda, mu)
  T *sX = sh + (size_t)NQ * NF * TE;
  T *sU = sX + (size_t)TE * XP20;
  T *sM = sX + (size_t)TE * XP20 * (FORM == FGPU_LINEAR ? 1 : 2);
  const int e0 = blockIdx.x * TE;
  const int tid = threadIdx.x;

  // phase 0: one thread per (element, node)
  for (int task = tid; task < TE * NN20; task += blockDim.x) {
    const int el = task / NN20, a = task % NN20;
    const int e = e0 + el;
    if (e >= nelem) continue;
    const size_t n = conn[(size_t)e * NN20 + a], n0 = conn[(size_t)e * NN20];
    for (int k = 0; k < 3; k++)
      sX[el * XP20 + 3 * a + k] = (T)(coords[3 * n + k] - coords[3 * n0 + k]);
    if constexpr (FORM != FGPU_LINEAR)
      for (int k = 0; k < 3; k++) sU[el * XP20 + 3 * a + k] = (T)unodal[3 * n + k];
    if (a == 0) {
      const double p0 = mat[2 * (size_t)e], p1 = mat[2 * (size_t)e + 1];
      if constexpr (FORM == FGPU_TL_NEOHOOKE) {  // (mu, lambda)
        sM[2 * el] = (T)p1;
        sM[2 * el + 1] = (T)p0;
      } else {
        sM[2 * el] = (T)(p0 * p1 / ((1.0 + p1) * (1.0 - 2.0 * p1)));
        sM[2 * el + 1] = (T)(0.5 * p0 / (1.0 + p1));
      }
    }
  }
  __syncthreads();

  // phase 1: one thread per (element, point)
  for (int task = tid; task < TE * NQ; task += blockDim.x) {
    const int el = task % TE, q = task / TE;
    const int e = e0 + el;
    T *S = sh + (size_t)q * NF * TE + el;
    if (e >= nelem) {
      for (int f = 0; f < NF; f++) S[f * TE] = T(0);
      continue;
    }
    const T *X = sX + el * XP20;
    const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
    const T r = (T)gp[(0 * 16 + ir) * 2], s = (T)gp[(1 * 16 + is) * 2],
            t = (T)gp[(2 * 16 + it) * 2];
    const T w = (T)(gp[(0 * 16 + ir) * 2 + 1] * gp[(1 * 16 + is) * 2 + 1] *
                    gp[(2 * 16 + it) * 2 + 1]);
    T dN[NN20][3];
    T J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < NN20; a++) {
      (void)h20_shape(r, s, t, a, dN[a]);
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) J[i][k] += dN[a][i] * X[3 * a + k];
    }
    T Ji[3][3];
    const T det = inv3x3(J, Ji);
    count_jacobian(status, det);
    const T c = w * fabs(det);
    T g[NN20][3];
#pragma unroll
    for (int a = 0; a < NN20; a++)
      for (int i = 0; i < 3; i++)
        g[a][i] = Ji[i][0] * dN[a][0] + Ji[i][1] * dN[a][1] + Ji[i][2] * dN[a][2];
    if constexpr (FORM == FGPU_LINEAR) {
      const T sc = sqrt(c);
#pragma unroll
      for (int a = 0; a < NN20; a++)
        for (int i = 0; i < 3; i++) S[(3 * a + i) * TE] = sc * g[a][i];
    } else {
      T F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      const T *Ue = sU + el * XP20;
#pragma unroll
      for (int a = 0; a < NN20; a++)
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) F[i][j] += Ue[3 * a + i] * g[a][j];
      T B[6];  // F F^T (xx, yy, zz, xy, xz, yz)
      B[0] = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2];
      B[1] = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2];
      B[2] = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2];
      B[3] = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
      B[4] = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
      B[5] = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
      const T lam = sM[2 * el], mu = sM[2 * el + 1];
      if constexpr (FORM == FGPU_TL_SVK) {
        const T C00 = F[0][0] * F[0][0] + F[1][0] * F[1][0] + F[2][0] * F[2][0];
        const T C11 = F[0][1] * F[0][1] + F[1][1] * F[1][1] + F[2][1] * F[2][1];
        const T C22 = F[0][2] * F[0][2] + F[1][2] * F[1][2] + F[2][2] * F[2][2];
        const T C01 = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
        const T C02 = F[0][0] * F[0][2] + F[1][0] * F[1][2] + F[2][0] * F[2][2];
        const T C12 = F[0][1] * F[0][2] + F[1][1] * F[1][2] + F[2][1] * F[2][2];
        const T E00 = T(0.5) * (C00 - T(1)), E11 = T(0.5) * (C11 - T(1)),
                E22 = T(0.5) * (C22 - T(1));
        const T tr = E00 + E11 + E22;
        const T Sv[6] = {lam * tr + T(2) * mu * E00, lam * tr + T(2) * mu * E11,
                         lam * tr + T(2) * mu * E22, mu * C01, mu * C02, mu * C12};
#pragma unroll
        for (int a = 0; a < NN20; a++)
          for (int i = 0; i < 3; i++) {
            S[(3 * a + i) * TE] = F[i][0] * g[a][0] + F[i][1] * g[a][1] + F[i][2] * g[a][2];
            S[(60 + 3 * a + i) * TE] = g[a][i];
          }
        for (int k = 0; k < 6; k++) {
          S[(120 + k) * TE] = B[k];
          S[(126 + k) * TE] = Sv[k];
        }
        S[132 * TE] = c;
      } else {  // neo-Hooke
        T Fi[3][3];
        T Jd = inv3x3(F, Fi);
        if (!(Jd > T(0))) {
          atomicAdd(&status[1], 1);
          Jd = T(1);
        }
        const T lnJ = log(Jd);
#pragma unroll
        for (int a = 0; a < NN20; a++)
          for (int i = 0; i < 3; i++)  // g' = F^-T g
            S[(3 * a + i) * TE] = Fi[0][i] * g[a][0] + Fi[1][i] * g[a][1] + Fi[2][i] * g[a][2];
        for (int k = 0; k < 6; k++) S[(60 + k) * TE] = B[k];
        S[66 * TE] = c * lam;
        S[67 * TE] = c * (mu - lam * lnJ);
        S[68 * TE] = c * mu;
      }
    }
  }
  __syncthreads();

  // phase 2: one thread per (element, upper node block), consecutive
  // threads = consecutive elements
  for (int task = tid; task < TE * NSYM20; task += blockDim.x) {
    const int el = task % TE, blk = task / TE;
    const int e = e0 + el;
    if (e >= nelem) continue;
    int a, b;
    h20_block(blk, a, b);
    const T lam = sM[2 * el], mu = sM[2 * el + 1];
    T K[9];
    if constexpr (FORM == FGPU_LINEAR) {
      T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = 0; q < NQ; q++) {
        const T *S = sh + (size_t)q * NF * TE + el;
        T va[3], vb[3];
        for (int i = 0; i < 3; i++) {
          va[i] = S[(3 * a + i) * TE];
          vb[i] = S[(3 * b + i) * TE];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) M[3 * i + j] = fma(va[i], vb[j], M[3 * i + j]);
      }
      if (a == b)  // exact symmetry of a diagonal block
        for (int i = 0; i < 3; i++)
          for (int j = i + 1; j < 3; j++) {
            const T m = T(0.5) * (M[3 * i + j] + M[3 * j + i]);
            M[3 * i + j] = m;
            M[3 * j + i] = m;
          }
      const T tr = mu * (M[0] + M[4] + M[8]);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          K[3 * i + j] = lam * M[3 * i + j] + mu * M[3 * j + i] + (i == j ? tr : T(0));
    } else if constexpr (FORM == FGPU_TL_SVK) {
      T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      T Bs[6] = {0, 0, 0, 0, 0, 0};
      T sg = T(0);
      for (int q = 0; q < NQ; q++) {
        const T *S = sh + (size_t)q * NF * TE + el;
        const T c = S[132 * TE];
        T fa[3], fb[3], ga[3], gb[3];
        for (int i = 0; i < 3; i++) {
          fa[i] = S[(3 * a + i) * TE];
          fb[i] = S[(3 * b + i) * TE];
          ga[i] = c * S[(60 + 3 * a + i) * TE];
          gb[i] = S[(60 + 3 * b + i) * TE];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) M[3 * i + j] += c * (fa[i] * fb[j]);
        const T gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        for (int k = 0; k < 6; k++) Bs[k] += gg * S[(120 + k) * TE];
        const T s0 = S[126 * TE], s1 = S[127 * TE], s2 = S[128 * TE], s3 = S[129 * TE],
                s4 = S[130 * TE], s5 = S[131 * TE];
        sg += ga[0] * (s0 * gb[0] + s3 * gb[1] + s4 * gb[2]) +
              ga[1] * (s3 * gb[0] + s1 * gb[1] + s5 * gb[2]) +
              ga[2] * (s4 * gb[0] + s5 * gb[1] + s2 * gb[2]);
      }
      const T Bm[9] = {Bs[0], Bs[3], Bs[4], Bs[3], Bs[1], Bs[5], Bs[4], Bs[5], Bs[2]};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          K[3 * i + j] = lam * M[3 * i + j] + mu * M[3 * j + i] + mu * Bm[3 * i + j] +
                         (i == j ? sg : T(0));
    } else {
      T M1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, M2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      T sg = T(0);
      for (int q = 0; q < NQ; q++) {
        const T *S = sh + (size_t)q * NF * TE + el;
        T ga[3], gb[3];
        for (int i = 0; i < 3; i++) {
          ga[i] = S[(3 * a + i) * TE];
          gb[i] = S[(3 * b + i) * TE];
        }
        const T cl = S[66 * TE], cm2 = S[67 * TE], cm = S[68 * TE];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            const T p = ga[i] * gb[j];
            M1[3 * i + j] += cl * p;
            M2[3 * i + j] += cm2 * p;
          }
        // g.h = g'^T (F F^T) h'
        const T b0 = S[60 * TE], b1 = S[61 * TE], b2 = S[62 * TE], b3 = S[63 * TE],
                b4 = S[64 * TE], b5 = S[65 * TE];
        sg += cm * (ga[0] * (b0 * gb[0] + b3 * gb[1] + b4 * gb[2]) +
                    ga[1] * (b3 * gb[0] + b1 * gb[1] + b5 * gb[2]) +
                    ga[2] * (b4 * gb[0] + b5 * gb[1] + b2 * gb[2]));
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          K[3 * i + j] = M1[3 * i + j] + M2[3 * j + i] + (i == j ? sg : T(0));
    }
    T *out = Ke + L.offset(e, blk);
    for (int k = 0; k < 9; k++) out[k] = K[k];
  }
}

// ---------------------------------------------------------------- FP64
// The stress acting on material gradients at a point with L = grad u:
// g_a += c P grad N_a.  Linear (p0 = E, p1 = nu): P = sigma; St.Venant-
// Kirchhoff: P = F S with S = lambda tr(E) I + 2 mu E; neo-Hooke (p0 = mu,
// p1 = lambda): P = F S with S = mu (I - C^-1) + lambda ln J C^-1.  Returns
// false when det C <= 0 (S formed with det C = 1).
template <int FORM>
__device__ inline bool point_P(const double Lg[3][3], double p0, double p1, double P[3][3]) {
  if constexpr (FORM == FGPU_LINEAR) {
    const double lam = p0 * p1 / ((1.0 + p1) * (1.0 - 2.0 * p1)), mu = 0.5 * p0 / (1.0 + p1);
    const double tr = Lg[0][0] + Lg[1][1] + Lg[2][2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) P[i][j] = mu * (Lg[i][j] + Lg[j][i]) + (i == j ? lam * tr : 0.0);
    return true;
  } else {
    double F[3][3], C[3][3], Sm[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) F[i][j] = Lg[i][j] + (i == j ? 1.0 : 0.0);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    bool ok = true;
    if constexpr (FORM == FGPU_TL_SVK) {
      const double lam = p0 * p1 / ((1.0 + p1) * (1.0 - 2.0 * p1)), mu = 0.5 * p0 / (1.0 + p1);
      const double tr = 0.5 * (C[0][0] + C[1][1] + C[2][2] - 3.0);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          Sm[i][j] = mu * (C[i][j] - (i == j ? 1.0 : 0.0)) + (i == j ? lam * tr : 0.0);
    } else {
      double Ci[3][3];
      double detC = inv3x3(C, Ci);
      if (!(detC > 0.0)) {
        ok = false;
        detC = 1.0;
      }
      const double lnJ = 0.5 * log(detC);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          Sm[i][j] = p0 * ((i == j ? 1.0 : 0.0) - Ci[i][j]) + p1 * lnJ * Ci[i][j];
    }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        P[i][j] = F[i][0] * Sm[0][j] + F[i][1] * Sm[1][j] + F[i][2] * Sm[2][j];
    return ok;
  }
}

// Phase 1 of the FP64 kernels: the Jacobian, its inverse, c = w |det J| and
// (WITH_U) grad u of one element point from the staged coordinates and
// displacements (relative to node 0).
template <int NIP, bool WITH_U>
__device__ inline void h20_point(const double *X, const double *U, const double *gp, int q,
                                 double Ji[3][3], double &c, double Lg[3][3], double &det) {
  const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
  const double r = gp[(0 * 16 + ir) * 2], s = gp[(1 * 16 + is) * 2], t = gp[(2 * 16 + it) * 2];
  const double w = gp[(0 * 16 + ir) * 2 + 1] * gp[(1 * 16 + is) * 2 + 1] * gp[(2 * 16 + it) * 2 + 1];
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < NN20; a++) {
    double d[3];
    (void)h20_shape(r, s, t, a, d);
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) J[i][k] += d[i] * X[3 * a + k];
  }
  det = inv3x3(J, Ji);
  c = w * fabs(det);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Lg[i][j] = 0.0;
  if constexpr (!WITH_U) return;
  for (int a = 0; a < NN20; a++) {
    double d[3], G[3];
    (void)h20_shape(r, s, t, a, d);
    for (int i = 0; i < 3; i++) G[i] = Ji[i][0] * d[0] + Ji[i][1] * d[1] + Ji[i][2] * d[2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Lg[i][j] += U[3 * a + i] * G[j];
  }
}

constexpr int TF20 = 8;  // elements per workgroup of the FP64 kernels

// Internal force: g_a = sum_q c P grad N_a, P = sigma (linear) or F S (TL).
template <int FORM, int NIP>
__global__ __launch_bounds__(256) void hex20_fint_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, double *__restrict__ ge, int *__restrict__ status) {
  constexpr int NQ = NIP * NIP * NIP;
  __shared__ double sX[TF20][60], sU[TF20][60];
  __shared__ double sQ[TF20][NQ][18];  // J^-1, then c P, per point
  const int e0 = blockIdx.x * TF20, tid = threadIdx.x;
  for (int task = tid; task < TF20 * NN20; task += blockDim.x) {
    const int el = task / NN20, a = task % NN20, e = e0 + el;
    if (e >= nelem) continue;
    const size_t n = conn[(size_t)e * NN20 + a], n0 = conn[(size_t)e * NN20];
    for (int k = 0; k < 3; k++) {
      sX[el][3 * a + k] = coords[3 * n + k] - coords[3 * n0 + k];
      sU[el][3 * a + k] = unodal[3 * n + k];
    }
  }
  __syncthreads();
  for (int task = tid; task < TF20 * NQ; task += blockDim.x) {
    const int el = task % TF20, q = task / TF20, e = e0 + el;
    if (e >= nelem) continue;
    double Ji[3][3], c, Lg[3][3], det;
    h20_point<NIP, true>(sX[el], sU[el], gp, q, Ji, c, Lg, det);
    count_jacobian(status, det);
    double P[3][3];
    if (!point_P<FORM>(Lg, mat[2 * (size_t)e], mat[2 * (size_t)e + 1], P))
      atomicAdd(&status[1], 1);
    double *o = sQ[el][q];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        o[3 * i + j] = Ji[i][j];
        o[9 + 3 * i + j] = c * P[i][j];
      }
  }
  __syncthreads();
  for (int task = tid; task < TF20 * NN20; task += blockDim.x) {
    const int el = task % TF20, a = task / TF20, e = e0 + el;
    if (e >= nelem) continue;
    double g[3] = {0, 0, 0};
    for (int q = 0; q < NQ; q++) {
      const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
      double d[3];
      (void)h20_shape(gp[(0 * 16 + ir) * 2], gp[(1 * 16 + is) * 2], gp[(2 * 16 + it) * 2], a, d);
      const double *o = sQ[el][q];
      double G[3];
      for (int i = 0; i < 3; i++) G[i] = o[3 * i] * d[0] + o[3 * i + 1] * d[1] + o[3 * i + 2] * d[2];
      for (int i = 0; i < 3; i++)
        g[i] += o[9 + 3 * i] * G[0] + o[9 + 3 * i + 1] * G[1] + o[9 + 3 * i + 2] * G[2];
    }
    double *out = ge + (size_t)e * 60 + 3 * a;
    for (int i = 0; i < 3; i++) out[i] = g[i];
  }
}

// Body-force element vectors (cElemLoad::ExtForce with cBodyLoad, load/
// load.cpp:618-691): f_a = sum_q w |det J| N_a b.  b: [nelem * 3].
template <int NIP>
__global__ __launch_bounds__(256) void hex20_body_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ bvec, const double *__restrict__ gp, double *__restrict__ ge,
    int *__restrict__ status) {
  constexpr int NQ = NIP * NIP * NIP;
  __shared__ double sX[TF20][60];
  __shared__ double sC[TF20][NQ];
  const int e0 = blockIdx.x * TF20, tid = threadIdx.x;
  for (int task = tid; task < TF20 * NN20; task += blockDim.x) {
    const int el = task / NN20, a = task % NN20, e = e0 + el;
    if (e >= nelem) continue;
    const size_t n = conn[(size_t)e * NN20 + a], n0 = conn[(size_t)e * NN20];
    for (int k = 0; k < 3; k++) sX[el][3 * a + k] = coords[3 * n + k] - coords[3 * n0 + k];
  }
  __syncthreads();
  for (int task = tid; task < TF20 * NQ; task += blockDim.x) {
    const int el = task % TF20, q = task / TF20, e = e0 + el;
    if (e >= nelem) continue;
    double Ji[3][3], c, Lg[3][3], det;
    h20_point<NIP, false>(sX[el], nullptr, gp, q, Ji, c, Lg, det);
    count_jacobian(status, det);
    sC[el][q] = c;
  }
  __syncthreads();
  for (int task = tid; task < TF20 * NN20; task += blockDim.x) {
    const int el = task % TF20, a = task / TF20, e = e0 + el;
    if (e >= nelem) continue;
    double sN = 0.0;
    for (int q = 0; q < NQ; q++) {
      const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
      double d[3];
      const double N =
          h20_shape(gp[(0 * 16 + ir) * 2], gp[(1 * 16 + is) * 2], gp[(2 * 16 + it) * 2], a, d);
      sN += sC[el][q] * N;
    }
    const double *b = bvec + 3 * (size_t)e;
    double *out = ge + (size_t)e * 60 + 3 * a;
    for (int i = 0; i < 3; i++) out[i] = sN * b[i];
  }
}

// Stresses at the points (cElmParam::IntPntStress, elc/elmparam.cpp:859-937;
// cElmParamTL::IntPntStress, 1811-1885) and, when sig_node is not null,
// extrapolated to the nodes by TR [20][NQ] (cElmParam::NodalStress,
// 1304-1333, with CompTRMatrix, 1444-1535).
template <int FORM, int NIP>
__global__ __launch_bounds__(256) void hex20_stress_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, const double *__restrict__ TR, double *__restrict__ sig_ip,
    double *__restrict__ sig_node, int *__restrict__ status) {
  constexpr int NQ = NIP * NIP * NIP;
  __shared__ double sX[TF20][60], sU[TF20][60];
  __shared__ double sS[TF20][NQ][6];
  const int e0 = blockIdx.x * TF20, tid = threadIdx.x;
  for (int task = tid; task < TF20 * NN20; task += blockDim.x) {
    const int el = task / NN20, a = task % NN20, e = e0 + el;
    if (e >= nelem) continue;
    const size_t n = conn[(size_t)e * NN20 + a], n0 = conn[(size_t)e * NN20];
    for (int k = 0; k < 3; k++) {
      sX[el][3 * a + k] = coords[3 * n + k] - coords[3 * n0 + k];
      sU[el][3 * a + k] = unodal[3 * n + k];
    }
  }
  __syncthreads();
  for (int task = tid; task < TF20 * NQ; task += blockDim.x) {
    const int el = task % TF20, q = task / TF20, e = e0 + el;
    if (e >= nelem) continue;
    double Ji[3][3], c, Lg[3][3], det;
    h20_point<NIP, true>(sX[el], sU[el], gp, q, Ji, c, Lg, det);
    count_jacobian(status, det);
    double sg[6];
    if (!point_stress<FORM>(Lg, mat[2 * (size_t)e], mat[2 * (size_t)e + 1], sg))
      atomicAdd(&status[1], 1);
    double *o = sig_ip + ((size_t)e * NQ + q) * 6;
    for (int k = 0; k < 6; k++) {
      o[k] = sg[k];
      sS[el][q][k] = sg[k];
    }
  }
  if (!sig_node) return;  // uniform across the grid
  __syncthreads();
  for (int task = tid; task < TF20 * NN20; task += blockDim.x) {
    const int el = task % TF20, a = task / TF20, e = e0 + el;
    if (e >= nelem) continue;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < NQ; q++) {
      const double w = TR[a * NQ + q];
      for (int k = 0; k < 6; k++) acc[k] += w * sS[el][q][k];
    }
    double *o = sig_node + ((size_t)e * NN20 + a) * 6;
    for (int k = 0; k < 6; k++) o[k] = acc[k];
  }
}

// largest elements per workgroup (a divisor of the 64-element storage tile)
// whose staged point data (values of b bytes) fit in 64 KB of LDS
constexpr int h20_te(int nq, int nf, int ne, int b) {
  return (65536 / ((nq * nf + ne) * b)) >= 32   ? 32
         : (65536 / ((nq * nf + ne) * b)) >= 16 ? 16
         : (65536 / ((nq * nf + ne) * b)) >= 8  ? 8
         : (65536 / ((nq * nf + ne) * b)) >= 4  ? 4
         : (65536 / ((nq * nf + ne) * b)) >= 2  ? 2
                                                : 1;
}

// The element matrices of 20-node bricks are formed and kept in FP64
// (d_Ke64): bending-dominated meshes of these elements reach condition
// numbers (4e9 on the reference's BeamB20 deck) at which FP32 element
// matrices leave the refined solve above the reference's accuracy.
template <int FORM, int NIP>
void launch_K20(Ctx &c, const double *u) {
  using T = double;
  constexpr int NQ = NIP * NIP * NIP, NF = Stage20<FORM>::NF, NE = h20_elem_floats<FORM>();
  constexpr int TE = h20_te(NQ, NF, NE, (int)sizeof(T));
  static_assert((size_t)(NQ * NF + NE) * TE * sizeof(T) <= 65536, "hex20 K: LDS");
  const size_t lds = (size_t)(NQ * NF + NE) * TE * sizeof(T);
  const int grid = (c.nelem + TE - 1) / TE;
  if (!c.d_Ke64) throw Error("hex20: FP64 element matrices not allocated");
  hipLaunchKernelGGL((hex20_K_kernel<T, FORM, NIP, TE>), dim3(grid), dim3(256), lds, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_mat, u, c.d_gauss64, c.d_Ke64, c.kel,
                     c.d_status);
  FGPU_CHECK(hipGetLastError());
}

template <int FORM>
void launch_K20_f(Ctx &c, const double *u) {
  switch (c.nip[0]) {
    case 1: launch_K20<FORM, 1>(c, u); break;
    case 2: launch_K20<FORM, 2>(c, u); break;
    default: launch_K20<FORM, 3>(c, u); break;
  }
}

template <int FORM, int NIP>
void launch_f20(Ctx &c, const double *u) {
  const int grid = (c.nelem + TF20 - 1) / TF20;
  hipLaunchKernelGGL((hex20_fint_kernel<FORM, NIP>), dim3(grid), dim3(256), 0, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_mat, u, c.d_gauss64, c.d_ge, c.d_status);
  FGPU_CHECK(hipGetLastError());
  c.ge_sorted = false;
}

template <int FORM>
void launch_f20_f(Ctx &c, const double *u) {
  switch (c.nip[0]) {
    case 1: launch_f20<FORM, 1>(c, u); break;
    case 2: launch_f20<FORM, 2>(c, u); break;
    default: launch_f20<FORM, 3>(c, u); break;
  }
}

template <int NIP>
void launch_b20(Ctx &c, const double *b) {
  const int grid = (c.nelem + TF20 - 1) / TF20;
  hipLaunchKernelGGL((hex20_body_kernel<NIP>), dim3(grid), dim3(256), 0, c.stream, c.nelem,
                     c.d_coords, c.d_conn, b, c.d_gauss64, c.d_ge, c.d_status);
  FGPU_CHECK(hipGetLastError());
  c.ge_sorted = false;
}

template <int FORM, int NIP>
void launch_s20(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  const int grid = (c.nelem + TF20 - 1) / TF20;
  hipLaunchKernelGGL((hex20_stress_kernel<FORM, NIP>), dim3(grid), dim3(256), 0, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_mat, u, c.d_gauss64, c.d_TR, sig_ip,
                     sig_node, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

template <int FORM>
void launch_s20_f(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  switch (c.nip[0]) {
    case 1: launch_s20<FORM, 1>(c, u, sig_ip, sig_node); break;
    case 2: launch_s20<FORM, 2>(c, u, sig_ip, sig_node); break;
    default: launch_s20<FORM, 3>(c, u, sig_ip, sig_node); break;
  }
}

}  // namespace

void launch_hex20_K(Ctx &c, const double *u) {
  if (c.form != FGPU_LINEAR && !u) {  // zero displacement: a zero nodal field
    launch_fill(c, c.d_u, 3 * (long long)c.nnode, 0.0);
    u = c.d_u;
  }
  switch (c.form) {
    case FGPU_LINEAR: launch_K20_f<FGPU_LINEAR>(c, u); break;
    case FGPU_TL_SVK: launch_K20_f<FGPU_TL_SVK>(c, u); break;
    default: launch_K20_f<FGPU_TL_NEOHOOKE>(c, u); break;
  }
}

void launch_hex20_fint(Ctx &c, const double *u) {
  switch (c.form) {
    case FGPU_LINEAR: launch_f20_f<FGPU_LINEAR>(c, u); break;
    case FGPU_TL_SVK: launch_f20_f<FGPU_TL_SVK>(c, u); break;
    default: launch_f20_f<FGPU_TL_NEOHOOKE>(c, u); break;
  }
}

void launch_hex20_body(Ctx &c, const double *b) {
  switch (c.nip[0]) {
    case 1: launch_b20<1>(c, b); break;
    case 2: launch_b20<2>(c, b); break;
    default: launch_b20<3>(c, b); break;
  }
}

void launch_hex20_stress(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  switch (c.form) {
    case FGPU_LINEAR: launch_s20_f<FGPU_LINEAR>(c, u, sig_ip, sig_node); break;
    case FGPU_TL_SVK: launch_s20_f<FGPU_TL_SVK>(c, u, sig_ip, sig_node); break;
    default: launch_s20_f<FGPU_TL_NEOHOOKE>(c, u, sig_ip, sig_node); break;
  }
}

}  // namespace fgpu
