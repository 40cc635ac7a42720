// k_hex8.hip - 8-node brick element kernels (stiffness in FP32, internal
// force in FP64) for gfx950.
//
// Reference path: cElmParam::StiffMat / IntForce (elc/elmparam.cpp:449-597)
// and cElmParamTL::StiffMat / IntForce (elc/elmparam.cpp:1615-1807) over
// cShapeBrick8 (shp/shpsolid.cpp:173-238), cShapeSolid::DrvShpXYZ
// (shp/shpsolid.cpp:58-126) and cSolid B / Bl / Bnl / S matrices
// (elc/anmodel.cpp:1418-1568) with the isotropic Q of anmodel.cpp:1286-1298.
//
// The dense contraction B^T C B is evaluated in closed form for isotropic
// materials.  With g = grad N_a, h = grad N_b (per integration point):
//   linear:  K_ab = lambda g h^T + mu h g^T + mu (g.h) I
//   TL SVK:  K_ab = lambda (Fg)(Fh)^T + mu (Fh)(Fg)^T + mu (g.h) F F^T
//                   + (g^T S h) I            [= (B+Bl)^T C (B+Bl) + Bnl^T S Bnl]
//   TL neo-Hooke: with spatial gradients g' = F^-T g, lambda' = lambda,
//                 mu' = mu - lambda ln J:
//            K_ab = lambda g' h'^T + mu' h' g'^T + mu (g.h) I
// so only three-vectors per node and point are needed.  A workgroup stages
// those vectors for TE elements x all points in LDS (phase 1: one thread per
// element-point), then each thread forms whole 3x3 node blocks (phase 2:
// one thread per element-block, consecutive threads = consecutive elements,
// so the block stores of a wave are contiguous).
#include "fgpu_internal.h"

namespace fgpu {
namespace {

__constant__ signed char kNodeR[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
__constant__ signed char kNodeS[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
__constant__ signed char kNodeT[8] = {1, 1, 1, 1, -1, -1, -1, -1};
// upper node blocks (a <= b) of a hex8, packed row-major
__constant__ unsigned char kBlkA[36] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2,
                                        2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 6, 6, 7};
__constant__ unsigned char kBlkB[36] = {0, 1, 2, 3, 4, 5, 6, 7, 1, 2, 3, 4, 5, 6, 7, 2, 3, 4,
                                        5, 6, 7, 3, 4, 5, 6, 7, 4, 5, 6, 7, 5, 6, 7, 6, 7, 7};

template <class T>
__device__ inline void hex8_dN(T r, T s, T t, int a, T &dr, T &ds, T &dt) {
  const T q = T(0.125);
  T fr = T(1) + T(kNodeR[a]) * r, fs = T(1) + T(kNodeS[a]) * s, ft = T(1) + T(kNodeT[a]) * t;
  dr = (kNodeR[a] > 0 ? q : -q) * fs * ft;
  ds = (kNodeS[a] > 0 ? q : -q) * fr * ft;
  dt = (kNodeT[a] > 0 ? q : -q) * fr * fs;
}

template <class T>
__device__ inline T inv3(const T J[3][3], T I[3][3]) {
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

// FP32 point geometry of the stiffness kernels, every product and sum spelt
// out (fmaf) so that each kernel using it rounds alike, whether its shape
// derivatives are run-time values or compile-time constants:
//   J += dN_a (x) X_a over the nodes; J^-1 by cofactors; g = J^-1 dN
__device__ inline void hex8_jac_add(float J[3][3], const float d[3], const float *Xa) {
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) J[i][k] = fmaf(d[i], Xa[k], J[i][k]);
}

__device__ inline float inv3_fma(const float J[3][3], float I[3][3]) {
  const float c00 = fmaf(J[1][1], J[2][2], -(J[1][2] * J[2][1]));
  const float c01 = fmaf(J[1][0], J[2][2], -(J[1][2] * J[2][0]));
  const float c02 = fmaf(J[1][0], J[2][1], -(J[1][1] * J[2][0]));
  const float det = fmaf(J[0][0], c00, fmaf(-J[0][1], c01, J[0][2] * c02));
  const float id = 1.f / det;
  I[0][0] = c00 * id;
  I[0][1] = -fmaf(J[0][1], J[2][2], -(J[0][2] * J[2][1])) * id;
  I[0][2] = fmaf(J[0][1], J[1][2], -(J[0][2] * J[1][1])) * id;
  I[1][0] = -c01 * id;
  I[1][1] = fmaf(J[0][0], J[2][2], -(J[0][2] * J[2][0])) * id;
  I[1][2] = -fmaf(J[0][0], J[1][2], -(J[0][2] * J[1][0])) * id;
  I[2][0] = c02 * id;
  I[2][1] = -fmaf(J[0][0], J[2][1], -(J[0][1] * J[2][0])) * id;
  I[2][2] = fmaf(J[0][0], J[1][1], -(J[0][1] * J[1][0])) * id;
  return det;
}

__device__ inline float hex8_grad_i(const float Ji[3][3], int i, const float d[3]) {
  return fmaf(Ji[i][2], d[2], fmaf(Ji[i][1], d[1], Ji[i][0] * d[0]));
}

// Gradients g_a = J^-T dN_a of the 8 shape functions at (r, s, t) in FP32,
// from the node coordinates X relative to node 0; returns det J
__device__ inline float hex8_point_grad(const float *X, float r, float s, float t,
                                        float g[8][3]) {
  float dN[8][3];
  float J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < 8; a++) {
    hex8_dN(r, s, t, a, dN[a][0], dN[a][1], dN[a][2]);
    hex8_jac_add(J, dN[a], X + 3 * a);
  }
  float Ji[3][3];
  const float det = inv3_fma(J, Ji);
  for (int a = 0; a < 8; a++)
    for (int i = 0; i < 3; i++) g[a][i] = hex8_grad_i(Ji, i, dN[a]);
  return det;
}

// Linear isotropic block from M = sum_q s_a s_b^T:
// K = lambda M + mu M^T + mu tr(M) I, every product and sum spelt out (fmaf)
// so that each kernel forming blocks rounds them alike
__device__ inline void lin_block(const float M[9], float lam, float mu, float K[9]) {
  const float trs = (M[0] + M[4]) + M[8];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const float v = fmaf(lam, M[3 * i + j], mu * M[3 * j + i]);
      K[3 * i + j] = i == j ? fmaf(mu, trs, v) : v;
    }
}

// Per element-point vectors staged in LDS (floats per point):
//   LINEAR:   sqrt(c) g (24), c = w detJ: M_ab = sum_q (sqrt(c) g_a)(sqrt(c) g_b)^T
//             costs 9 FMA per point and block, and a diagonal block comes out
//             exactly symmetric (each product and sum is formed alike for ij, ji)
//   SVK:      Fg (24), g (24), FF^T (6), S (6), c
//   NEOHOOKE: g' (24), FF^T (6), c*lambda, c*mu', c*mu
template <int FORM> struct Stage;
template <> struct Stage<FGPU_LINEAR> { static constexpr int NF = 24; };
template <> struct Stage<FGPU_TL_SVK> { static constexpr int NF = 61; };
template <> struct Stage<FGPU_TL_NEOHOOKE> { static constexpr int NF = 33; };

__device__ inline void lame(double E, double nu, float &lam, float &mu) {
  lam = (float)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
  mu = (float)(0.5 * E / (1.0 + nu));
}

// Per element in LDS beside the staged point data: node coordinates
// relative to node 0 (FP64 subtraction, then FP32) and, for the
// total-Lagrangian forms, nodal displacements, 8 x 3 floats each, padded to
// 25 per element so that consecutive elements fall in different banks; the
// two material constants.
constexpr int XPAD = 25;
template <int FORM>
constexpr int elem_lds_floats() { return XPAD * (FORM == FGPU_LINEAR ? 1 : 2) + 2; }

// Linear blocks of rows p and 7 - p (p = 0..3: 8 - p and p + 1 blocks, 9 in
// all) from the staged vectors s_a = sqrt(c) grad N_a of each point (S: the
// element's column of the [q][f][TE] stage): M_ab = sum_q s_a s_b^T,
// K_ab = lambda M + mu M^T + mu tr(M) I, stored as upper node blocks.
template <int P>
__device__ inline void hex8_lin_rows(const float *S, int TE, int NQ, float lam, float mu,
                                     float *__restrict__ Ke, KeLayout L, int e) {
  constexpr int p = P, a1 = P, a2 = 7 - P;
  float M[9][9];
  for (int k = 0; k < 9; k++)
    for (int m = 0; m < 9; m++) M[k][m] = 0.f;
#pragma unroll 1
  for (int q = 0; q < NQ; q++) {
    const float *Sq = S + (size_t)q * 24 * TE;
    float v[8][3];
#pragma unroll
    for (int b = 0; b < 8; b++)
      if (b >= p)
        for (int i = 0; i < 3; i++) v[b][i] = Sq[(3 * b + i) * TE];
#pragma unroll
    for (int b = 0; b < 8; b++) {
      // block slot: row a1 holds b = a1..7 in slots 0..7-a1, row a2 the rest
      if (b >= a1) {
        float *Mk = M[b - a1];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Mk[3 * i + j] = fmaf(v[a1][i], v[b][j], Mk[3 * i + j]);
      }
      if (b >= a2) {
        float *Mk = M[8 - a1 + b - a2];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Mk[3 * i + j] = fmaf(v[a2][i], v[b][j], Mk[3 * i + j]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int a = k < 8 - a1 ? a1 : a2;
    const int b = k < 8 - a1 ? a1 + k : a2 + (k - (8 - a1));
    float K[9];
    lin_block(M[k], lam, mu, K);
    float *out = Ke + L.offset(e, sym_index(a, b, 8));
    for (int m = 0; m < 9; m++) out[m] = K[m];
  }
}

template <int FORM, int NIP, int TE>
__global__ __launch_bounds__(256) void hex8_K_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const float *__restrict__ gp, float *__restrict__ Ke, KeLayout L,
    int *__restrict__ status) {
  constexpr int NQ = NIP * NIP * NIP;
  constexpr int NF = Stage<FORM>::NF;
  extern __shared__ float sh[];  // [q][f][TE], then the per-element data
  float *sX = sh + (size_t)NQ * NF * TE;                       // [TE][XPAD]
  float *sU = sX + (size_t)TE * XPAD;                          // [TE][XPAD] (TL)
  float *sM = sX + (size_t)TE * XPAD * (FORM == FGPU_LINEAR ? 1 : 2);  // [TE][2]
  const int e0 = blockIdx.x * TE;
  const int tid = threadIdx.x;

  // ---------------- phase 0: one thread per (element, node): each node's
  // coordinates (and displacement) are read once per element; consecutive
  // threads read consecutive connectivity entries
  for (int task = tid; task < TE * 8; task += blockDim.x) {
    const int el = task / 8, a = task % 8;
    const int e = e0 + el;
    if (e >= nelem) continue;
    const size_t n = conn[(size_t)e * 8 + a], n0 = conn[(size_t)e * 8];
    for (int k = 0; k < 3; k++) sX[el * XPAD + 3 * a + k] = (float)(coords[3 * n + k] - coords[3 * n0 + k]);
    if constexpr (FORM != FGPU_LINEAR)
      for (int k = 0; k < 3; k++) sU[el * XPAD + 3 * a + k] = (float)unodal[3 * n + k];
    if (a == 0) {
      float lam, mu;
      if constexpr (FORM == FGPU_TL_NEOHOOKE) {  // mat = (mu, lambda)
        mu = (float)mat[2 * (size_t)e];
        lam = (float)mat[2 * (size_t)e + 1];
      } else {
        lame(mat[2 * (size_t)e], mat[2 * (size_t)e + 1], lam, mu);
      }
      sM[2 * el] = lam;
      sM[2 * el + 1] = mu;
    }
  }
  __syncthreads();

  // ---------------- phase 1: one thread per (element, point)
  for (int task = tid; task < TE * NQ; task += blockDim.x) {
    const int el = task % TE, q = task / TE;
    const int e = e0 + el;
    float *S = sh + (size_t)q * NF * TE + el;
    if (e >= nelem) {
      for (int f = 0; f < NF; f++) S[f * TE] = 0.f;
      continue;
    }
    const float *X = sX + el * XPAD;
    const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
    const float r = gp[(0 * 16 + ir) * 2], s = gp[(1 * 16 + is) * 2], t = gp[(2 * 16 + it) * 2];
    const float w = gp[(0 * 16 + ir) * 2 + 1] * gp[(1 * 16 + is) * 2 + 1] * gp[(2 * 16 + it) * 2 + 1];
    float g[8][3];
    const float det = hex8_point_grad(X, r, s, t, g);
    count_jacobian(status, det);
    const float c = w * fabsf(det);  // |det J|: a left-handed mesh is accepted whole
    if constexpr (FORM == FGPU_LINEAR) {
      const float sc = sqrtf(c);
      for (int a = 0; a < 8; a++)
        for (int i = 0; i < 3; i++) S[(3 * a + i) * TE] = sc * g[a][i];
    } else {
      float F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      const float *Ue = sU + el * XPAD;
      for (int a = 0; a < 8; a++) {
        const float u0 = Ue[3 * a], u1 = Ue[3 * a + 1], u2 = Ue[3 * a + 2];
        for (int j = 0; j < 3; j++) {
          F[0][j] += u0 * g[a][j];
          F[1][j] += u1 * g[a][j];
          F[2][j] += u2 * g[a][j];
        }
      }
      float B[6];  // F F^T (xx, yy, zz, xy, xz, yz)
      B[0] = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2];
      B[1] = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2];
      B[2] = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2];
      B[3] = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
      B[4] = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
      B[5] = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
      const float lam = sM[2 * el], mu = sM[2 * el + 1];
      if constexpr (FORM == FGPU_TL_SVK) {
        // Green-Lagrange E = (F^T F - I) / 2 ; S = lambda tr(E) I + 2 mu E
        float C00 = F[0][0] * F[0][0] + F[1][0] * F[1][0] + F[2][0] * F[2][0];
        float C11 = F[0][1] * F[0][1] + F[1][1] * F[1][1] + F[2][1] * F[2][1];
        float C22 = F[0][2] * F[0][2] + F[1][2] * F[1][2] + F[2][2] * F[2][2];
        float C01 = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
        float C02 = F[0][0] * F[0][2] + F[1][0] * F[1][2] + F[2][0] * F[2][2];
        float C12 = F[0][1] * F[0][2] + F[1][1] * F[1][2] + F[2][1] * F[2][2];
        float E00 = 0.5f * (C00 - 1.f), E11 = 0.5f * (C11 - 1.f), E22 = 0.5f * (C22 - 1.f);
        float tr = E00 + E11 + E22;
        float Sv[6] = {lam * tr + 2.f * mu * E00, lam * tr + 2.f * mu * E11,
                       lam * tr + 2.f * mu * E22, mu * C01, mu * C02, mu * C12};
        for (int a = 0; a < 8; a++)
          for (int i = 0; i < 3; i++) {
            S[(3 * a + i) * TE] = F[i][0] * g[a][0] + F[i][1] * g[a][1] + F[i][2] * g[a][2];
            S[(24 + 3 * a + i) * TE] = g[a][i];
          }
        for (int k = 0; k < 6; k++) {
          S[(48 + k) * TE] = B[k];
          S[(54 + k) * TE] = Sv[k];
        }
        S[60 * TE] = c;
      } else {  // neo-Hooke
        float Fi[3][3];
        float Jd = inv3(F, Fi);
        if (!(Jd > 0.f)) {
          atomicAdd(&status[1], 1);
          Jd = 1.f;
        }
        float lnJ = logf(Jd);
        for (int a = 0; a < 8; a++)
          for (int i = 0; i < 3; i++)  // g' = F^-T g
            S[(3 * a + i) * TE] = Fi[0][i] * g[a][0] + Fi[1][i] * g[a][1] + Fi[2][i] * g[a][2];
        for (int k = 0; k < 6; k++) S[(24 + k) * TE] = B[k];
        S[30 * TE] = c * lam;
        S[31 * TE] = c * (mu - lam * lnJ);
        S[32 * TE] = c * mu;
      }
    }
  }
  __syncthreads();

  if constexpr (FORM == FGPU_LINEAR) {
    // ---------------- phase 2 (linear): one thread per (element, row pair):
    // rows p and 7 - p of the element's upper blocks, 9 blocks together,
    // share their loads of the staged vectors; consecutive threads are
    // consecutive elements, so each block store of a wave is contiguous
    for (int task = tid; task < TE * 4; task += blockDim.x) {
      const int el = task % TE, p = task / TE;
      const int e = e0 + el;
      if (e >= nelem) continue;
      const float lam = sM[2 * el], mu = sM[2 * el + 1];
      switch (p) {  // uniform across a wave when TE >= 64
        case 0: hex8_lin_rows<0>(sh + el, TE, NQ, lam, mu, Ke, L, e); break;
        case 1: hex8_lin_rows<1>(sh + el, TE, NQ, lam, mu, Ke, L, e); break;
        case 2: hex8_lin_rows<2>(sh + el, TE, NQ, lam, mu, Ke, L, e); break;
        default: hex8_lin_rows<3>(sh + el, TE, NQ, lam, mu, Ke, L, e); break;
      }
    }
    return;
  }
  // ---------------- phase 2: one thread per (element, upper node block)
  for (int task = tid; task < TE * 36; task += blockDim.x) {
    const int el = task % TE, blk = task / TE;
    const int e = e0 + el;
    if (e >= nelem) continue;
    const int a = kBlkA[blk], b = kBlkB[blk];
    float K[9];
    if constexpr (FORM == FGPU_TL_SVK) {
      float M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      float Bs[6] = {0, 0, 0, 0, 0, 0};
      float s = 0.f;
      for (int q = 0; q < NQ; q++) {
        const float *S = sh + (size_t)q * NF * TE + el;
        float c = S[60 * TE];
        float fa[3], fb[3], ga[3], gb[3];
        for (int i = 0; i < 3; i++) {
          fa[i] = S[(3 * a + i) * TE];
          fb[i] = S[(3 * b + i) * TE];
          ga[i] = c * S[(24 + 3 * a + i) * TE];
          gb[i] = S[(24 + 3 * b + i) * TE];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) M[3 * i + j] += c * (fa[i] * fb[j]);
        float gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        for (int k = 0; k < 6; k++) Bs[k] += gg * S[(48 + k) * TE];
        float s0 = S[54 * TE], s1 = S[55 * TE], s2 = S[56 * TE], s3 = S[57 * TE],
              s4 = S[58 * TE], s5 = S[59 * TE];
        s += ga[0] * (s0 * gb[0] + s3 * gb[1] + s4 * gb[2]) +
             ga[1] * (s3 * gb[0] + s1 * gb[1] + s5 * gb[2]) +
             ga[2] * (s4 * gb[0] + s5 * gb[1] + s2 * gb[2]);
      }
      const float lam = sM[2 * el], mu = sM[2 * el + 1];
      const float Bm[9] = {Bs[0], Bs[3], Bs[4], Bs[3], Bs[1], Bs[5], Bs[4], Bs[5], Bs[2]};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          K[3 * i + j] = lam * M[3 * i + j] + mu * M[3 * j + i] + mu * Bm[3 * i + j] +
                         (i == j ? s : 0.f);
    } else {
      float M1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, M2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      float s = 0.f;
      for (int q = 0; q < NQ; q++) {
        const float *S = sh + (size_t)q * NF * TE + el;
        float ga[3], gb[3];
        for (int i = 0; i < 3; i++) {
          ga[i] = S[(3 * a + i) * TE];
          gb[i] = S[(3 * b + i) * TE];
        }
        float cl = S[30 * TE], cm2 = S[31 * TE], cm = S[32 * TE];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            float p = ga[i] * gb[j];
            M1[3 * i + j] += cl * p;
            M2[3 * i + j] += cm2 * p;
          }
        // g.h = g'^T (F F^T) h'
        float b0 = S[24 * TE], b1 = S[25 * TE], b2 = S[26 * TE], b3 = S[27 * TE],
              b4 = S[28 * TE], b5 = S[29 * TE];
        s += cm * (ga[0] * (b0 * gb[0] + b3 * gb[1] + b4 * gb[2]) +
                   ga[1] * (b3 * gb[0] + b1 * gb[1] + b5 * gb[2]) +
                   ga[2] * (b4 * gb[0] + b5 * gb[1] + b2 * gb[2]));
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          K[3 * i + j] = M1[3 * i + j] + M2[3 * j + i] + (i == j ? s : 0.f);
    }
    float *out = Ke + L.offset(e, blk);
    for (int k = 0; k < 9; k++) out[k] = K[k];
  }
}

// ---------------------------------------------------------------- one element per lane
// Linear bricks on 2x2x2 Gauss: a lane computes one element, its 8 x 24
// scaled gradients held in registers, then the 36 upper blocks one at a
// time, M = sum_q s_a s_b^T over the points in order and lin_block --
// hex8_lin_rows' arithmetic, so the same blocks bit for bit.  The wave's 64
// elements are one KeLayout tile: each block of the tile is one run of 576
// floats, which the wave stages in LDS and writes as 9 contiguous rows of 64
// floats (written straight from the registers, lane by lane 36 bytes apart,
// the same kernel took 4.08 ms on 10.3 M elements instead of 2.86 ms).  Two
// waves per SIMD (register budget 256).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void hex8_K_lane_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, float *__restrict__ Ke, KeLayout L,
    int *__restrict__ status) {
  __shared__ float sK[2][64 * 9];  // a block of the wave's 64 elements, [lane][9]
  const int lane = threadIdx.x;
  const int e0 = blockIdx.x * 64;
  const int e = min(e0 + lane, nelem - 1);  // lanes past the end repeat the last element
  const bool live = e0 + lane < nelem;
  float X[24];
  {
    const int *cn = conn + (size_t)e * 8;
    const size_t n0 = cn[0];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const size_t n = cn[a];
      for (int k = 0; k < 3; k++) X[3 * a + k] = (float)(coords[3 * n + k] - coords[3 * n0 + k]);
    }
  }
  float lam, mu;
  lame(mat[2 * (size_t)e], mat[2 * (size_t)e + 1], lam, mu);
  // per point J^-1 and sqrt(c) first (80 floats; X dies), then the scaled
  // gradients point by point as each J^-1 dies -- the values and operations
  // of hex8_point_grad, in an order that keeps fewer registers live
  // the 2-point rule's points and unit weights as constants (the values the
  // context's FP32 table holds: -/+0.577350269189626 rounded to FP32, 1), so
  // the shape-function derivatives fold at compile time
  constexpr float G2 = 0.577350269189626f;
  float Ji[8][3][3], sq[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int ir = q % 2, is = (q / 2) % 2, it = q / 4;
    const float r = ir ? G2 : -G2, s = is ? G2 : -G2, t = it ? G2 : -G2;
    const float w = 1.f;
    float J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; a++) {
      float d[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      hex8_jac_add(J, d, X + 3 * a);
    }
    const float det = inv3_fma(J, Ji[q]);
    if (live) count_jacobian(status, det);
    sq[q] = sqrtf(w * fabsf(det));
  }
  float S[8][24];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int ir = q % 2, is = (q / 2) % 2, it = q / 4;
    const float r = ir ? G2 : -G2, s = is ? G2 : -G2, t = it ? G2 : -G2;
    for (int a = 0; a < 8; a++) {
      float d[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++) S[q][3 * a + i] = sq[q] * hex8_grad_i(Ji[q], i, d);
    }
  }
  // upper blocks in packed row-major order: (0,0) (0,1) .. (0,7) (1,1) ..
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = a; b < 8; b++) {
      const int blk = a * 8 - a * (a - 1) / 2 + (b - a);
      float M[9];
#pragma unroll
      for (int m = 0; m < 9; m++) M[m] = 0.f;
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++)
            M[3 * i + j] = fmaf(S[q][3 * a + i], S[q][3 * b + j], M[3 * i + j]);
      float K[9];
      lin_block(M, lam, mu, K);
      // the wave's 64 blocks are one run of 576 floats in the tile: staged
      // in LDS ([lane][9], stride 9: no bank conflicts) and stored as 9
      // contiguous rows of 64 floats
      float *sk = sK[blk & 1];
#pragma unroll
      for (int m = 0; m < 9; m++) sk[9 * lane + m] = K[m];
      __syncthreads();
      float *out = Ke + L.offset(e0, blk);
      if (e0 + 64 <= nelem) {
#pragma unroll
        for (int k = 0; k < 9; k++) out[64 * k + lane] = sk[64 * k + lane];
      } else {
        for (int k = 0; k < 9; k++) {
          const int v = 64 * k + lane;
          if (e0 + v / 9 < nelem) out[v] = sk[v];
        }
      }
    }
}

// ---------------------------------------------------------------- f_int
// One thread per element, FP64 throughout.  Linear: g = sum c B^T (C B u);
// TL: g_a = sum c P grad N_a with P = F S (the B-bar^T S of the reference).
template <int FORM, int NIP>
__global__ __launch_bounds__(128) void hex8_fint_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, double *__restrict__ ge, int *__restrict__ status,
    const int *__restrict__ epos) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int *cn = conn + (size_t)e * 8;
  double X[8][3], U[8][3], g[8][3];
  const int n0 = cn[0];
  for (int a = 0; a < 8; a++) {
    int n = cn[a];
    for (int k = 0; k < 3; k++) {
      X[a][k] = coords[3 * (size_t)n + k] - coords[3 * (size_t)n0 + k];
      U[a][k] = unodal[3 * (size_t)n + k];
      g[a][k] = 0.0;
    }
  }
  double p0 = mat[2 * (size_t)e], p1 = mat[2 * (size_t)e + 1];
  int neg = 0, zero = 0, fail = 0;
  for (int q = 0; q < NIP * NIP * NIP; q++) {
    const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
    const double r = gp[(0 * 16 + ir) * 2], s = gp[(1 * 16 + is) * 2], t = gp[(2 * 16 + it) * 2];
    const double w = gp[(0 * 16 + ir) * 2 + 1] * gp[(1 * 16 + is) * 2 + 1] * gp[(2 * 16 + it) * 2 + 1];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; a++) {
      double d[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) J[i][k] += d[i] * X[a][k];
    }
    double Ji[3][3];
    double det = inv3(J, Ji);
    neg += det < 0.0;
    zero += !(det > 0.0) && !(det < 0.0);
    const double c = w * fabs(det);
    // L = grad u (L_ij = d u_i / d X_j)
    double Lg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; a++) {
      double d[3], G[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++) G[i] = Ji[i][0] * d[0] + Ji[i][1] * d[1] + Ji[i][2] * d[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Lg[i][j] += U[a][i] * G[j];
    }
    double P[3][3];  // stress acting on material gradients: g_a += c P grad N_a
    if constexpr (FORM == FGPU_LINEAR) {
      double nu = p1, E = p0;
      double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 0.5 * E / (1.0 + nu);
      double tr = Lg[0][0] + Lg[1][1] + Lg[2][2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          P[i][j] = mu * (Lg[i][j] + Lg[j][i]) + (i == j ? lam * tr : 0.0);
    } else {
      double F[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) F[i][j] = Lg[i][j] + (i == j ? 1.0 : 0.0);
      double C[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
      double Sm[3][3];
      if constexpr (FORM == FGPU_TL_SVK) {
        double nu = p1, E = p0;
        double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 0.5 * E / (1.0 + nu);
        double tr = 0.5 * (C[0][0] + C[1][1] + C[2][2] - 3.0);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            Sm[i][j] = mu * (C[i][j] - (i == j ? 1.0 : 0.0)) + (i == j ? lam * tr : 0.0);
      } else {
        double mu = p0, lam = p1;
        double Ci[3][3];
        double detC = inv3(C, Ci);
        if (!(detC > 0.0)) {
          fail++;
          detC = 1.0;
        }
        double lnJ = 0.5 * log(detC);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            Sm[i][j] = mu * ((i == j ? 1.0 : 0.0) - Ci[i][j]) + lam * lnJ * Ci[i][j];
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          P[i][j] = F[i][0] * Sm[0][j] + F[i][1] * Sm[1][j] + F[i][2] * Sm[2][j];
    }
    for (int a = 0; a < 8; a++) {
      double d[3], G[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++) G[i] = Ji[i][0] * d[0] + Ji[i][1] * d[1] + Ji[i][2] * d[2];
      for (int i = 0; i < 3; i++) g[a][i] += c * (P[i][0] * G[0] + P[i][1] * G[1] + P[i][2] * G[2]);
    }
  }
  if (neg) atomicAdd(&status[0], neg);
  if (zero) atomicAdd(&status[3], zero);
  if (fail) atomicAdd(&status[1], fail);
  if (epos) {  // node-sorted: entry (e, a) at its place in the node's list
    for (int a = 0; a < 8; a++) {
      double *out = ge + 3 * (size_t)epos[(size_t)e * 8 + a];
      for (int i = 0; i < 3; i++) out[i] = g[a][i];
    }
  } else {
    double *out = ge + (size_t)e * 24;
    for (int a = 0; a < 8; a++)
      for (int i = 0; i < 3; i++) out[3 * a + i] = g[a][i];
  }
}

// Body-force element vectors (cElemLoad::ExtForce with cBodyLoad, load/
// load.cpp:618-691, 969-974): f_a = sum_q w detJ N_a b, one thread per
// element, FP64.  b: [nelem * 3].
template <int NIP>
__global__ __launch_bounds__(128) void hex8_body_kernel(int nelem, const double *__restrict__ coords,
                                                        const int *__restrict__ conn,
                                                        const double *__restrict__ bvec,
                                                        const double *__restrict__ gp,
                                                        double *__restrict__ ge,
                                                        int *__restrict__ status) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int *cn = conn + (size_t)e * 8;
  double X[8][3];
  const int n0 = cn[0];
  for (int a = 0; a < 8; a++)
    for (int k = 0; k < 3; k++) X[a][k] = coords[3 * (size_t)cn[a] + k] - coords[3 * (size_t)n0 + k];
  double sN[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // sum_q w detJ N_a
  int neg = 0, zero = 0;
  for (int q = 0; q < NIP * NIP * NIP; q++) {
    const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
    const double r = gp[(0 * 16 + ir) * 2], s = gp[(1 * 16 + is) * 2], t = gp[(2 * 16 + it) * 2];
    const double w = gp[(0 * 16 + ir) * 2 + 1] * gp[(1 * 16 + is) * 2 + 1] * gp[(2 * 16 + it) * 2 + 1];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; a++) {
      double d[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) J[i][k] += d[i] * X[a][k];
    }
    double Ji[3][3];
    const double det = inv3(J, Ji);
    neg += det < 0.0;
    zero += !(det > 0.0) && !(det < 0.0);
    const double c = w * fabs(det);
    for (int a = 0; a < 8; a++) {
      const double N = 0.125 * (1.0 + kNodeR[a] * r) * (1.0 + kNodeS[a] * s) * (1.0 + kNodeT[a] * t);
      sN[a] += c * N;
    }
  }
  if (neg) atomicAdd(&status[0], neg);
  if (zero) atomicAdd(&status[3], zero);
  const double *b = bvec + 3 * (size_t)e;
  double *out = ge + (size_t)e * 24;
  for (int a = 0; a < 8; a++)
    for (int i = 0; i < 3; i++) out[3 * a + i] = sN[a] * b[i];
}

// Stress output: cElmParam::IntPntStress (elc/elmparam.cpp:859-937) and
// cElmParamTL::IntPntStress (1811-1885) at every point, and, when sig_node is
// not null, cElmParam::NodalStress (1304-1333) by the extrapolation matrix TR
// [8][NIP^3] (CompTRMatrix, 1444-1535).  One thread per element, FP64.
template <int FORM, int NIP>
__global__ __launch_bounds__(128) void hex8_stress_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, const double *__restrict__ TR, double *__restrict__ sig_ip,
    double *__restrict__ sig_node, int *__restrict__ status) {
  constexpr int NQ = NIP * NIP * NIP;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const int *cn = conn + (size_t)e * 8;
  double X[8][3], U[8][3];
  const int n0 = cn[0];
  for (int a = 0; a < 8; a++) {
    const int n = cn[a];
    for (int k = 0; k < 3; k++) {
      X[a][k] = coords[3 * (size_t)n + k] - coords[3 * (size_t)n0 + k];
      U[a][k] = unodal[3 * (size_t)n + k];
    }
  }
  const double p0 = mat[2 * (size_t)e], p1 = mat[2 * (size_t)e + 1];
  double acc[8][6];
  for (int a = 0; a < 8; a++)
    for (int k = 0; k < 6; k++) acc[a][k] = 0.0;
  int neg = 0, zero = 0, fail = 0;
  for (int q = 0; q < NQ; q++) {
    const int ir = q % NIP, is = (q / NIP) % NIP, it = q / (NIP * NIP);
    const double r = gp[(0 * 16 + ir) * 2], s = gp[(1 * 16 + is) * 2], t = gp[(2 * 16 + it) * 2];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; a++) {
      double d[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) J[i][k] += d[i] * X[a][k];
    }
    double Ji[3][3];
    const double det = inv3(J, Ji);
    neg += det < 0.0;
    zero += !(det > 0.0) && !(det < 0.0);
    double Lg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 8; a++) {
      double d[3], G[3];
      hex8_dN(r, s, t, a, d[0], d[1], d[2]);
      for (int i = 0; i < 3; i++) G[i] = Ji[i][0] * d[0] + Ji[i][1] * d[1] + Ji[i][2] * d[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Lg[i][j] += U[a][i] * G[j];
    }
    double sg[6];
    fail += !point_stress<FORM>(Lg, p0, p1, sg);
    double *o = sig_ip + ((size_t)e * NQ + q) * 6;
    for (int k = 0; k < 6; k++) o[k] = sg[k];
    if (sig_node)
      for (int a = 0; a < 8; a++) {
        const double w = TR[a * NQ + q];
        for (int k = 0; k < 6; k++) acc[a][k] += w * sg[k];
      }
  }
  if (sig_node) {
    double *o = sig_node + (size_t)e * 48;
    for (int a = 0; a < 8; a++)
      for (int k = 0; k < 6; k++) o[6 * a + k] = acc[a][k];
  }
  if (neg) atomicAdd(&status[0], neg);
  if (zero) atomicAdd(&status[3], zero);
  if (fail) atomicAdd(&status[1], fail);
}

template <int FORM, int NIP, int TE>
void launch_K_te(Ctx &c, const double *u) {
  constexpr int NQ = NIP * NIP * NIP;
  constexpr int NF = Stage<FORM>::NF;
  size_t lds = ((size_t)NQ * NF + elem_lds_floats<FORM>()) * TE * sizeof(float);
  int grid = (c.nelem + TE - 1) / TE;
  hipLaunchKernelGGL((hex8_K_kernel<FORM, NIP, TE>), dim3(grid), dim3(256), lds, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_mat, u, c.d_gauss, c.d_Ke, c.kel,
                     c.d_status);
  FGPU_CHECK(hipGetLastError());
}

// largest number of elements per workgroup (a divisor of the 64-element
// storage tile) whose staged point data and per-element data fit in 64 KB of
// LDS (ne floats per element beside the nq * nf of the points)
constexpr int te_max(int nq, int nf, int ne) {
  return (65536 / ((nq * nf + ne) * 4)) >= 64   ? 64
         : (65536 / ((nq * nf + ne) * 4)) >= 32 ? 32
         : (65536 / ((nq * nf + ne) * 4)) >= 16 ? 16
         : (65536 / ((nq * nf + ne) * 4)) >= 8  ? 8
                                                : 4;
}

template <int FORM, int NIP>
void launch_K_t(Ctx &c, const double *u) {
  constexpr int TMAX = te_max(NIP * NIP * NIP, Stage<FORM>::NF, elem_lds_floats<FORM>());
  if constexpr (FORM == FGPU_LINEAR && NIP == 2) {
    if (c.hex8_te == 1) {  // one element per lane, point data in registers
      hipLaunchKernelGGL(hex8_K_lane_kernel, dim3((c.nelem + 63) / 64), dim3(64), 0, c.stream,
                         c.nelem, c.d_coords, c.d_conn, c.d_mat, c.d_Ke, c.kel, c.d_status);
      FGPU_CHECK(hipGetLastError());
      return;
    }
  }
  const int te = c.hex8_te > 0 && c.hex8_te <= TMAX ? c.hex8_te : TMAX;
  switch (te) {
    case 4: launch_K_te<FORM, NIP, 4>(c, u); break;
    case 8: if constexpr (TMAX >= 8) launch_K_te<FORM, NIP, 8>(c, u); break;
    case 16: if constexpr (TMAX >= 16) launch_K_te<FORM, NIP, 16>(c, u); break;
    case 32: if constexpr (TMAX >= 32) launch_K_te<FORM, NIP, 32>(c, u); break;
    default: if constexpr (TMAX >= 64) launch_K_te<FORM, NIP, 64>(c, u); break;
  }
}

template <int FORM>
int te_limit(int nip) {
  constexpr int ne = elem_lds_floats<FORM>();
  return nip == 1 ? te_max(1, Stage<FORM>::NF, ne)
                  : nip == 2 ? te_max(8, Stage<FORM>::NF, ne) : te_max(27, Stage<FORM>::NF, ne);
}

template <int FORM>
void launch_K_f(Ctx &c, const double *u) {
  switch (c.nip[0]) {
    case 1: launch_K_t<FORM, 1>(c, u); break;
    case 2: launch_K_t<FORM, 2>(c, u); break;
    default: launch_K_t<FORM, 3>(c, u); break;
  }
}

template <int FORM, int NIP>
void launch_f_t(Ctx &c, const double *u) {
  int grid = (c.nelem + 127) / 128;
  hipLaunchKernelGGL((hex8_fint_kernel<FORM, NIP>), dim3(grid), dim3(128), 0, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_mat, u, c.d_gauss64, c.d_ge,
                     c.d_status, c.d_epos);
  FGPU_CHECK(hipGetLastError());
  c.ge_sorted = c.d_epos != nullptr;
}

template <int FORM>
void launch_f_f(Ctx &c, const double *u) {
  switch (c.nip[0]) {
    case 1: launch_f_t<FORM, 1>(c, u); break;
    case 2: launch_f_t<FORM, 2>(c, u); break;
    default: launch_f_t<FORM, 3>(c, u); break;
  }
}

}  // namespace

void launch_hex8_K(Ctx &c, const double *u) {
  if (c.form != FGPU_LINEAR && !u) {
    // zero displacement: stage a zero nodal field
    FGPU_CHECK(hipMemsetAsync(c.d_u, 0, sizeof(double) * 3 * c.nnode, c.stream));
    u = c.d_u;
  }
  switch (c.form) {
    case FGPU_LINEAR: launch_K_f<FGPU_LINEAR>(c, u); break;
    case FGPU_TL_SVK: launch_K_f<FGPU_TL_SVK>(c, u); break;
    default: launch_K_f<FGPU_TL_NEOHOOKE>(c, u); break;
  }
}

template <int NIP>
static void launch_body_t(Ctx &c, const double *b) {
  hipLaunchKernelGGL((hex8_body_kernel<NIP>), dim3((c.nelem + 127) / 128), dim3(128), 0, c.stream,
                     c.nelem, c.d_coords, c.d_conn, b, c.d_gauss64, c.d_ge, c.d_status);
  FGPU_CHECK(hipGetLastError());
  c.ge_sorted = false;
}

template <int FORM, int NIP>
static void launch_stress_t(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  hipLaunchKernelGGL((hex8_stress_kernel<FORM, NIP>), dim3((c.nelem + 127) / 128), dim3(128), 0,
                     c.stream, c.nelem, c.d_coords, c.d_conn, c.d_mat, u, c.d_gauss64, c.d_TR,
                     sig_ip, sig_node, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

template <int FORM>
static void launch_stress_f(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  switch (c.nip[0]) {
    case 1: launch_stress_t<FORM, 1>(c, u, sig_ip, sig_node); break;
    case 2: launch_stress_t<FORM, 2>(c, u, sig_ip, sig_node); break;
    default: launch_stress_t<FORM, 3>(c, u, sig_ip, sig_node); break;
  }
}

void launch_hex8_stress(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  switch (c.form) {
    case FGPU_LINEAR: launch_stress_f<FGPU_LINEAR>(c, u, sig_ip, sig_node); break;
    case FGPU_TL_SVK: launch_stress_f<FGPU_TL_SVK>(c, u, sig_ip, sig_node); break;
    default: launch_stress_f<FGPU_TL_NEOHOOKE>(c, u, sig_ip, sig_node); break;
  }
}

void launch_hex8_body(Ctx &c, const double *b) {
  switch (c.nip[0]) {
    case 1: launch_body_t<1>(c, b); break;
    case 2: launch_body_t<2>(c, b); break;
    default: launch_body_t<3>(c, b); break;
  }
}

std::vector<int> hex8_te_options(const Ctx &c) {
  int tmax = c.form == FGPU_LINEAR   ? te_limit<FGPU_LINEAR>(c.nip[0])
             : c.form == FGPU_TL_SVK ? te_limit<FGPU_TL_SVK>(c.nip[0])
                                     : te_limit<FGPU_TL_NEOHOOKE>(c.nip[0]);
  std::vector<int> out;
  // 1: one element per lane with its point data in registers (linear, 2x2x2)
  if (c.form == FGPU_LINEAR && c.nip[0] == 2 && c.nip[1] == 2 && c.nip[2] == 2) out.push_back(1);
  for (int te = 4; te <= tmax; te *= 2) out.push_back(te);
  return out;
}

void launch_hex8_fint(Ctx &c, const double *u) {
  switch (c.form) {
    case FGPU_LINEAR: launch_f_f<FGPU_LINEAR>(c, u); break;
    case FGPU_TL_SVK: launch_f_f<FGPU_TL_SVK>(c, u); break;
    default: launch_f_f<FGPU_TL_NEOHOOKE>(c, u); break;
  }
}

}  // namespace fgpu
