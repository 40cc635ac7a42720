// k_bsp.hip - cubic B-spline solid element kernels (Bezier extraction) for
// gfx950: stiffness on FP32 MFMA, internal force in FP64.
//
// Reference path: cShapeBspSolid::DrvShpRST (shp/shpbsp.cpp:985-1067, NURBS
// quotient rule, control points k(t) slowest / i(r) fastest),
// cShapeBsp::EvalBspBas / EvalBspBasDerv (shp/shpbsp.cpp:254-300:
// t' = (t+1)/2, N_i = sum_j C_ij B_j(t'), derivatives halved),
// nBernsteinBasis::CompPols / CompDervPols (geom/bernbasis.cpp:91-124),
// cShapeSolid::DrvShpXYZ (shp/shpsolid.cpp:58-126) and
// cElmParam::StiffMat / IntForce (elc/elmparam.cpp:449-597).
//
// Stiffness of an isotropic element (64 control points, 64 Gauss points):
//   P_ij = G_i^T diag(w detJ) G_j      (G_i[q][a] = dN_a/dx_i at point q)
//   K_ab[i][j] = lambda P_ij[a][b] + mu P_ji[a][b] + mu delta_ij sum_k P_kk[a][b]
// One workgroup (4 waves) per element.  Phase 1 evaluates the geometry in
// FP64 and stages G (FP32, 48 KB) in LDS.  Phase 2: the 64 x 64 node-block
// matrix is cut into 16 x 16 tiles; each wave takes upper tile pairs
// (ta <= tb) and, for its pair, runs all nine products P_ij (i, j = x, y, z)
// on v_mfma_f32_16x16x4f32 (K = 64 points), sharing the three A and three B
// operands of every k-step between the nine chains.  All nine accumulators
// of a tile sit in the same lanes, so K_ab is formed in registers and each
// lane stores whole 3x3 blocks; no P matrices go through LDS, which keeps the
// workgroup small enough for three of them to share a CU.
//
// Total-Lagrangian tangents (cElmParamTL::StiffMat, elc/elmparam.cpp
// 1711-1807, with the same closed forms as the hex8 kernel: St.Venant-
// Kirchhoff K_ab = sum c [lambda (Fg)(Fh)^T + mu (Fh)(Fg)^T + mu (g.h) F F^T
// + (g^T S h) I]; neo-Hooke K_ab = sum c [lambda g'h'^T + mu' h'g'^T +
// mu (g.h) I], g' = F^-T g, mu' = mu - lambda ln J) carry per-point matrices
// (F F^T, S) between the two gradients, which do not factor into a few
// products over the points; bsp_Kt_kernel forms them on the vector ALUs
// instead: phase 1 stages the mapped gradients per point in LDS, phase 2
// runs one thread per stored block over the 64 points.
#include "fgpu_internal.h"

namespace fgpu {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Degree 3 along r and s; along t degree 3 (NB2 = 4 basis functions) or 2
// (NB2 = 3).  Points: 4 x 4 x 4 Gauss.
constexpr int NB = 4, NQ1 = 4;   // basis (at most) and points per direction
constexpr int NQP = 64;
template <int NB2>
struct Bsp {
  static constexpr int NBAS = 16 * NB2;                  // 64 or 48 control points
  static constexpr int NSYM = NBAS * (NBAS + 1) / 2;     // 2080 or 1176 stored blocks
  static constexpr int NPART = NBAS / 4;                 // control points per quarter
  static constexpr int NT = NBAS / 16;                   // 16-wide tiles
};

// degree 2: B_i and d B_i / dt (with the 1/2 of t' = (t+1)/2)
__device__ inline void bernstein2(double u, double *B, double *dB) {
  double v = 1.0 - u;
  B[0] = v * v;
  B[1] = 2.0 * u * v;
  B[2] = u * u;
  B[3] = 0.0;
  dB[0] = 0.5 * 2.0 * (0.0 - v);
  dB[1] = 0.5 * 2.0 * (v - u);
  dB[2] = 0.5 * 2.0 * (u - 0.0);
  dB[3] = 0.0;
}

__device__ inline void bernstein3(double u, double *B, double *dB) {
  double v = 1.0 - u;
  B[0] = v * v * v;
  B[1] = 3.0 * u * v * v;
  B[2] = 3.0 * u * u * v;
  B[3] = u * u * u;
  // n (B_{i-1,n-1} - B_{i,n-1}) with n = 3, then the 1/2 of t' = (t+1)/2
  double b0 = v * v, b1 = 2.0 * u * v, b2 = u * u;
  dB[0] = 0.5 * 3.0 * (0.0 - b0);
  dB[1] = 0.5 * 3.0 * (b0 - b1);
  dB[2] = 0.5 * 3.0 * (b1 - b2);
  dB[3] = 0.5 * 3.0 * (b2 - 0.0);
}

__device__ inline double inv3d(const double J[3][3], double I[3][3]) {
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double id = 1.0 / det;
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

// Element geometry staged in LDS: univariate tables at the Gauss points,
// control points (relative to the first) and weights.
template <int NB2>
struct Geo {
  double N[3][NQ1][NB], dN[3][NQ1][NB];  // [dir][point][basis]
  double X[Bsp<NB2>::NBAS][3];
  double W[Bsp<NB2>::NBAS];
};

// extraction rows on the device: [e][dir][4][4], zero-padded for degree 2
template <int NB2>
__device__ inline void geo_load(Geo<NB2> &g, int e, const double *__restrict__ coords,
                                const int *__restrict__ conn, const double *__restrict__ cpw,
                                const double *__restrict__ ext, const double *__restrict__ gp) {
  constexpr int NBAS = Bsp<NB2>::NBAS;
  const int tid = threadIdx.x;
  const int *cn = conn + (size_t)e * NBAS;
  if (tid < NBAS) {
    int n = cn[tid], n0 = cn[0];
    for (int k = 0; k < 3; k++) g.X[tid][k] = coords[3 * (size_t)n + k] - coords[3 * (size_t)n0 + k];
    g.W[tid] = cpw ? cpw[n] : 1.0;
  } else if (tid < NBAS + 3 * NQ1 * NB) {
    int t = tid - NBAS;
    int d = t / (NQ1 * NB), q = (t / NB) % NQ1, i = t % NB;
    double u = (gp[(d * 16 + q) * 2] + 1.0) / 2.0;
    double B[4], dB[4];
    if (d == 2 && NB2 == 3)
      bernstein2(u, B, dB);
    else
      bernstein3(u, B, dB);
    const double *C = ext + ((size_t)e * 3 + d) * 16 + i * 4;
    double n = 0.0, dn = 0.0;
    for (int j = 0; j < 4; j++) {
      n += C[j] * B[j];
      dn += C[j] * dB[j];
    }
    g.N[d][q][i] = n;
    g.dN[d][q][i] = dn;
  }
}

// The element's nodal displacements into LDS by threads 128.. (idle in
// geo_load), so that their gathers overlap the geometry's loads.
template <int NB2>
__device__ inline void stage_u(double (*U)[3], const int *__restrict__ cn,
                               const double *__restrict__ unodal) {
  const int t = threadIdx.x - 128;
  if (t >= 0 && t < Bsp<NB2>::NBAS) {
    const int n = cn[t];
    for (int k = 0; k < 3; k++) U[t][k] = unodal[3 * (size_t)n + k];
  }
}

// Numerators (N w, dN w) of the tensor-product basis a at point q;
// a = i + 4 j + 16 k with i along r, q likewise.
template <int NB2>
__device__ inline void tensor(const Geo<NB2> &g, int q, int a, double &n, double d[3]) {
  const int qr = q % 4, qs = (q / 4) % 4, qt = q / 16;
  const int i = a % 4, j = (a / 4) % 4, k = a / 16;
  const double w = g.W[a];
  const double Nr = g.N[0][qr][i], Ns = g.N[1][qs][j], Nt = g.N[2][qt][k];
  n = Nr * Ns * Nt * w;
  d[0] = g.dN[0][qr][i] * Ns * Nt * w;
  d[1] = Nr * g.dN[1][qs][j] * Nt * w;
  d[2] = Nr * Ns * g.dN[2][qt][k] * w;
}

// Rational basis derivatives of the 16 control points [16 part, 16 part + 16)
// at point q (quotient rule of shpbsp.cpp:1049-1066), their Jacobian and its
// inverse.  The four lanes of a point (consecutive lanes) reduce their
// partial sums with shuffles, so all of them hold the same J^-1.
struct PointGeo {
  double Ji[3][3];
  double detJ, c, inv_s, dW[3];
};

// d R_a / d(r,s,t) at point q (rational quotient with the point's sums)
template <int NB2>
__device__ inline void rational_dR(const Geo<NB2> &g, int q, int a, const PointGeo &P,
                                   double dR[3]) {
  double n, d[3];
  tensor(g, q, a, n, d);
  const double R = n * P.inv_s;
  for (int i = 0; i < 3; i++) dR[i] = (d[i] - R * P.dW[i]) * P.inv_s;
}

template <int NB2>
__device__ inline void point_geometry(const Geo<NB2> &g, int q, int part, const double *gp,
                                      PointGeo &P) {
  constexpr int NPART = Bsp<NB2>::NPART;
  double s = 0.0, ds[3] = {0, 0, 0};
  #pragma unroll 2
  for (int k = 0; k < NPART; k++) {
    double n, d[3];
    tensor(g, q, part * NPART + k, n, d);
    s += n;
    ds[0] += d[0];
    ds[1] += d[1];
    ds[2] += d[2];
  }
  for (int m = 1; m < 4; m <<= 1) {
    s += __shfl_xor(s, m);
    for (int k = 0; k < 3; k++) ds[k] += __shfl_xor(ds[k], m);
  }
  P.inv_s = 1.0 / s;
  for (int k = 0; k < 3; k++) P.dW[k] = ds[k];
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  #pragma unroll 2
  for (int k = 0; k < NPART; k++) {
    const int a = part * NPART + k;
    double dR[3];
    rational_dR(g, q, a, P, dR);
    for (int i = 0; i < 3; i++)
      for (int l = 0; l < 3; l++) J[i][l] += dR[i] * g.X[a][l];
  }
  for (int m = 1; m < 4; m <<= 1)
    for (int i = 0; i < 3; i++)
      for (int l = 0; l < 3; l++) J[i][l] += __shfl_xor(J[i][l], m);
  P.detJ = inv3d(J, P.Ji);
  const int qr = q % 4, qs = (q / 4) % 4, qt = q / 16;
  // |det J|: a mesh parametrized left-handed throughout is accepted
  P.c = gp[(0 * 16 + qr) * 2 + 1] * gp[(1 * 16 + qs) * 2 + 1] * gp[(2 * 16 + qt) * 2 + 1] *
        fabs(P.detJ);
}

// c sigma (Voigt order xx yy zz xy xz yz) of the linear form at a point with
// displacement gradient L
__device__ inline void linear_csig(const double *__restrict__ mat, int e, double c,
                                   const double L[3][3], double *sg) {
  const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 0.5 * E / (1.0 + nu);
  const double tr = L[0][0] + L[1][1] + L[2][2];
  sg[0] = c * (lam * tr + 2.0 * mu * L[0][0]);
  sg[1] = c * (lam * tr + 2.0 * mu * L[1][1]);
  sg[2] = c * (lam * tr + 2.0 * mu * L[2][2]);
  sg[3] = c * mu * (L[0][1] + L[1][0]);
  sg[4] = c * mu * (L[0][2] + L[2][0]);
  sg[5] = c * mu * (L[1][2] + L[2][1]);
}

template <int NB2, bool FP>
struct KShared {
  Geo<NB2> g;
  float G[3][NQP][Bsp<NB2>::NBAS];  // dN_a/dx_i at point q
  float c[NQP];                      // w detJ
  double U[FP ? Bsp<NB2>::NBAS : 1][3];  // FP: the element's nodal displacements
};

// Per-point record of the internal force, [element][point][PREC]: J^-1 (9),
// 1 / W and dW/d(r,s,t) (4), c sigma (Voigt, 6) -- what the internal-force
// kernel's second phase reads (FShared's Ji, rs, sig)
constexpr int PREC = 19;

// FP (linear form): the geometry phase also forms grad u = sum_a u_a (x) G_a
// from the same FP64 gradients, in the order of bsp_fint_kernel's first
// phase, and writes each point's record to pts; bsp_fint_pts_kernel then
// finishes the internal force without recomputing the points' geometry.
template <int NB2, bool FP>
__global__ __launch_bounds__(256, 3) void bsp_K_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ gp, float *__restrict__ Ke,
    int *__restrict__ status, const double *__restrict__ unodal, double *__restrict__ pts) {
  constexpr int NBAS = Bsp<NB2>::NBAS, NSYM = Bsp<NB2>::NSYM, NPART = Bsp<NB2>::NPART;
  constexpr int NT = Bsp<NB2>::NT, NPAIR = NT * (NT + 1) / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  KShared<NB2, FP> &S = *reinterpret_cast<KShared<NB2, FP> *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  if constexpr (FP) stage_u<NB2>(S.U, conn + (size_t)e * NBAS, unodal);
  __syncthreads();
  {  // phase 1: G and c
    const int q = tid / 4, part = tid % 4;
    PointGeo P;
    point_geometry(S.g, q, part, gp, P);
    if (part == 0) {
      S.c[q] = (float)P.c;
      count_jacobian(status, P.detJ);
    }
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      double dR[3], G[3];
      rational_dR(S.g, q, part * NPART + k, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++) S.G[i][q][part * NPART + k] = (float)G[i];
      if constexpr (FP) {
        const double *ua = S.U[part * NPART + k];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) L[i][j] += ua[i] * G[j];
      }
    }
    if constexpr (FP) {
      for (int m = 1; m < 4; m <<= 1)
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) L[i][j] += __shfl_xor(L[i][j], m);
      if (part == 0) {
        double *r = pts + ((size_t)e * NQP + q) * PREC;
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) r[3 * i + j] = P.Ji[i][j];
        r[9] = P.inv_s;
        for (int k = 0; k < 3; k++) r[10 + k] = P.dW[k];
        linear_csig(mat, e, P.c, L, r + 13);
      }
    }
  }
  __syncthreads();
  // phase 2: nine 16x16x64 products per tile pair, combined in registers
  const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
  const float lam = (float)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
  const float mu = (float)(0.5 * E / (1.0 + nu));
  const int wave = tid / 64, lane = tid % 64;
  const int ln = lane % 16, lk = lane / 16;
  float *out = Ke + (size_t)e * NSYM * 9;
  // waves take the tile pairs round-robin from a start that rotates with the
  // element, so that the waves holding one pair more than the others are not
  // always on the same SIMDs of a CU
  for (int tp = (wave + e) % 4; tp < NPAIR; tp += 4) {
    // upper 16x16 tile pair (ta <= tb) number tp, row-major
    int ta = 0, r = tp;
    while (r >= NT - ta) {
      r -= NT - ta;
      ta++;
    }
    const int a0 = 16 * ta, b0 = 16 * (ta + r);
    floatx4 acc[9];
    for (int t = 0; t < 9; t++)
      for (int r = 0; r < 4; r++) acc[t][r] = 0.f;
#pragma unroll 4
    for (int k0 = 0; k0 < NQP; k0 += 4) {
      const int q = k0 + lk;
      const float cq = S.c[q];
      float A[3], B[3];
      for (int i = 0; i < 3; i++) {
        A[i] = cq * S.G[i][q][a0 + ln];
        B[i] = S.G[i][q][b0 + ln];
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          acc[3 * i + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[i], B[j], acc[3 * i + j], 0, 0, 0);
    }
    // lane holds D[m][n] with n = lane % 16, m = 4 (lane / 16) + r
    for (int r = 0; r < 4; r++) {
      const int a = a0 + 4 * lk + r, b = b0 + ln;
      if (a > b) continue;
      float P[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) P[i][j] = acc[3 * i + j][r];
      if (a == b) {  // exact symmetry of diagonal blocks
        for (int i = 0; i < 3; i++)
          for (int j = i + 1; j < 3; j++) {
            float s = 0.5f * (P[i][j] + P[j][i]);
            P[i][j] = s;
            P[j][i] = s;
          }
      }
      const float tr = mu * (P[0][0] + P[1][1] + P[2][2]);
      float *o = out + (size_t)sym_index(a, b, NBAS) * 9;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          o[3 * i + j] = lam * P[i][j] + mu * P[j][i] + (i == j ? tr : 0.f);
    }
  }
}

// ---------------------------------------------------------------- TL K
// (a, b) of the stored blocks in packed upper order: a * 64 + b
template <int NB2>
struct SymTable {
  unsigned short ab[Bsp<NB2>::NSYM];
};
__constant__ SymTable<4> kSym4;
__constant__ SymTable<3> kSym3;

template <int NB2>
__device__ inline unsigned short sym_ab(int blk) {
  if constexpr (NB2 == 4)
    return kSym4.ab[blk];
  else
    return kSym3.ab[blk];
}

template <int FORM, int NB2>
struct KtShared {
  Geo<NB2> g;
  double U[Bsp<NB2>::NBAS][3];
  float A[3][NQP][Bsp<NB2>::NBAS];                            // F g (SVK) / F^-T g (neo-Hooke)
  float G[FORM == FGPU_TL_SVK ? 3 : 1][NQP][Bsp<NB2>::NBAS];  // g (SVK only)
  float pp[NQP][16];  // SVK: c, F F^T (6), S (6);  neo-Hooke: F F^T (6), c lam, c mu', c mu
};

template <int FORM, int NB2>
__global__ __launch_bounds__(256) void bsp_Kt_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, float *__restrict__ Ke, int *__restrict__ status) {
  constexpr int NBAS = Bsp<NB2>::NBAS, NSYM = Bsp<NB2>::NSYM, NPART = Bsp<NB2>::NPART;
  extern __shared__ __align__(16) unsigned char smem[];
  KtShared<FORM, NB2> &S = *reinterpret_cast<KtShared<FORM, NB2> *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  const int *cn = conn + (size_t)e * NBAS;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  stage_u<NB2>(S.U, cn, unodal);
  __syncthreads();
  double lam, mu;
  if constexpr (FORM == FGPU_TL_SVK) {
    const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    mu = 0.5 * E / (1.0 + nu);
  } else {  // neo-Hooke: mat = (mu, lambda)
    mu = mat[2 * (size_t)e];
    lam = mat[2 * (size_t)e + 1];
  }
  {  // phase 1, thread (q, part): F at q, mapped gradients of 16 nodes
    const int q = tid / 4, part = tid % 4;
    PointGeo P;
    point_geometry(S.g, q, part, gp, P);
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      const int a = part * NPART + k;
      const double *ua = S.U[a];
      double dR[3], G[3];
      rational_dR(S.g, q, a, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += ua[i] * G[j];
    }
    for (int m = 1; m < 4; m <<= 1)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += __shfl_xor(L[i][j], m);
    double F[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1.0 : 0.0) + L[i][j];
    double B[6];  // F F^T (xx, yy, zz, xy, xz, yz)
    B[0] = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2];
    B[1] = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2];
    B[2] = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2];
    B[3] = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
    B[4] = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
    B[5] = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
    double M[3][3];  // the map of the gradients: F (SVK) or F^-T (neo-Hooke)
    if constexpr (FORM == FGPU_TL_SVK) {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = F[i][j];
      if (part == 0) {
        // Green-Lagrange E = (F^T F - I) / 2 ; S = lambda tr(E) I + 2 mu E
        double C[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
        const double tr = 0.5 * (C[0][0] - 1.0) + 0.5 * (C[1][1] - 1.0) + 0.5 * (C[2][2] - 1.0);
        float *pp = S.pp[q];
        pp[0] = (float)P.c;
        for (int k = 0; k < 6; k++) pp[1 + k] = (float)B[k];
        pp[7] = (float)(lam * tr + mu * (C[0][0] - 1.0));
        pp[8] = (float)(lam * tr + mu * (C[1][1] - 1.0));
        pp[9] = (float)(lam * tr + mu * (C[2][2] - 1.0));
        pp[10] = (float)(mu * C[0][1]);
        pp[11] = (float)(mu * C[0][2]);
        pp[12] = (float)(mu * C[1][2]);
      }
    } else {
      double Fi[3][3];
      double J = inv3d(F, Fi);
      if (!(J > 0.0)) {
        if (part == 0) atomicAdd(&status[1], 1);
        J = 1.0;
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = Fi[j][i];
      if (part == 0) {
        float *pp = S.pp[q];
        for (int k = 0; k < 6; k++) pp[k] = (float)B[k];
        pp[6] = (float)(P.c * lam);
        pp[7] = (float)(P.c * (mu - lam * log(J)));
        pp[8] = (float)(P.c * mu);
      }
    }
    if (part == 0) count_jacobian(status, P.detJ);
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      const int a = part * NPART + k;
      double dR[3], G[3];
      rational_dR(S.g, q, a, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++) {
        S.A[i][q][a] = (float)(M[i][0] * G[0] + M[i][1] * G[1] + M[i][2] * G[2]);
        if constexpr (FORM == FGPU_TL_SVK) S.G[i][q][a] = (float)G[i];
      }
    }
  }
  __syncthreads();
  // phase 2: one thread per stored block, all 64 points
  const float flam = (float)lam, fmu = (float)mu;
  float *out = Ke + (size_t)e * NSYM * 9;
  for (int blk = tid; blk < NSYM; blk += 256) {
    const unsigned short ab = sym_ab<NB2>(blk);
    const int a = ab >> 6, b = ab & 63;
    float K[9];
    if constexpr (FORM == FGPU_TL_SVK) {
      float Mm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Bs[6] = {0, 0, 0, 0, 0, 0}, sg = 0.f;
      for (int q = 0; q < NQP; q++) {
        const float *pp = S.pp[q];
        const float c = pp[0];
        float fa[3], fb[3], ga[3], gb[3];
        for (int i = 0; i < 3; i++) {
          fa[i] = S.A[i][q][a];
          fb[i] = S.A[i][q][b];
          ga[i] = c * S.G[i][q][a];
          gb[i] = S.G[i][q][b];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Mm[3 * i + j] += c * (fa[i] * fb[j]);
        const float gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        for (int k = 0; k < 6; k++) Bs[k] += gg * pp[1 + k];
        sg += ga[0] * (pp[7] * gb[0] + pp[10] * gb[1] + pp[11] * gb[2]) +
              ga[1] * (pp[10] * gb[0] + pp[8] * gb[1] + pp[12] * gb[2]) +
              ga[2] * (pp[11] * gb[0] + pp[12] * gb[1] + pp[9] * gb[2]);
      }
      const float Bm[9] = {Bs[0], Bs[3], Bs[4], Bs[3], Bs[1], Bs[5], Bs[4], Bs[5], Bs[2]};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          K[3 * i + j] = flam * Mm[3 * i + j] + fmu * Mm[3 * j + i] + fmu * Bm[3 * i + j] +
                         (i == j ? sg : 0.f);
    } else {
      float M1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, M2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sg = 0.f;
      for (int q = 0; q < NQP; q++) {
        const float *pp = S.pp[q];
        float ga[3], gb[3];
        for (int i = 0; i < 3; i++) {
          ga[i] = S.A[i][q][a];
          gb[i] = S.A[i][q][b];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            const float p = ga[i] * gb[j];
            M1[3 * i + j] += pp[6] * p;
            M2[3 * i + j] += pp[7] * p;
          }
        // g.h = g'^T (F F^T) h'
        sg += pp[8] * (ga[0] * (pp[0] * gb[0] + pp[3] * gb[1] + pp[4] * gb[2]) +
                       ga[1] * (pp[3] * gb[0] + pp[1] * gb[1] + pp[5] * gb[2]) +
                       ga[2] * (pp[4] * gb[0] + pp[5] * gb[1] + pp[2] * gb[2]));
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) K[3 * i + j] = M1[3 * i + j] + M2[3 * j + i] + (i == j ? sg : 0.f);
    }
    float *o = out + (size_t)blk * 9;
    for (int k = 0; k < 9; k++) o[k] = K[k];
  }
}

// ------------------------------------------------------- TL K on MFMA
// The total-Lagrangian tangents as sums of weighted products over the 64
// points, each a 16x16x64 MFMA chain over a tile pair like the linear
// kernel's nine:
//   St.Venant-Kirchhoff: with f = F g,
//     M_ij  = sum_q c f_ai f_bj                      (9 chains)
//     BB_ij = sum_q sum_k (c (F F^T)_ij g_ak) g_bk    (6 x 3 chains)
//     sg    = sum_q sum_l (sum_k c S_kl g_ak) g_bl    (3 chains)
//     K_ab  = lambda M + mu M^T + mu BB + sg I
//   (the initial-stress term sg has its own chains: folded into the
//   diagonal BB chains, its operands would be rounded against the much
//   larger mu B terms, and thin shells lose their geometric stiffness)
//   neo-Hooke: with g' = F^-T g,
//     M1_ij = sum_q (c lambda g'_ai) g'_bj,  M2_ij = sum_q (c mu' g'_ai) g'_bj
//     sg    = sum_q sum_l (sum_k c mu (F F^T)_kl g'_ak) g'_bl
//     K_ab  = M1 + M2^T + sg I
// Phase 1 stages the gradients (g for SVK, g' for neo-Hooke) in FP32 and the
// point's factors (SVK: sqrt(c) F, c F F^T, c S; neo-Hooke: c lambda, c mu',
// c mu F F^T) in the LDS the geometry used; phase 2 forms the operands on the
// vector ALUs per k-step (f = sqrt(c) F g for both sides of M) and runs the
// chains (30 per k-step for SVK, 21 for neo-Hooke).  Upper tile pairs,
// diagonal blocks made exactly symmetric, blocks stored as the linear
// kernel stores them.
template <int FORM>
struct KtmPoint {
  static constexpr int NPP = FORM == FGPU_TL_SVK ? 21 : 8;  // floats per point
};
template <int FORM, int NB2>
struct KtmShared {
  union {
    struct {
      Geo<NB2> g;
      double U[Bsp<NB2>::NBAS][3];
    } in;                                      // phase 1
    float pp[NQP][KtmPoint<FORM>::NPP];        // phase 2
  } u;
  float G[3][NQP][Bsp<NB2>::NBAS];  // g (SVK) or F^-T g (neo-Hooke)
};

template <int FORM, int NB2>
__global__ __launch_bounds__(256, 3) void bsp_Ktm_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, float *__restrict__ Ke, int *__restrict__ status) {
  constexpr int NBAS = Bsp<NB2>::NBAS, NSYM = Bsp<NB2>::NSYM, NPART = Bsp<NB2>::NPART;
  constexpr int NT = Bsp<NB2>::NT, NPAIR = NT * (NT + 1) / 2;
  constexpr int NPP = KtmPoint<FORM>::NPP;
  constexpr bool SVK = FORM == FGPU_TL_SVK;
  extern __shared__ __align__(16) unsigned char smem[];
  KtmShared<FORM, NB2> &S = *reinterpret_cast<KtmShared<FORM, NB2> *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  const int *cn = conn + (size_t)e * NBAS;
  geo_load(S.u.in.g, e, coords, conn, cpw, ext, gp);
  stage_u<NB2>(S.u.in.U, cn, unodal);
  __syncthreads();
  double lam, mu;
  if constexpr (SVK) {
    const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    mu = 0.5 * E / (1.0 + nu);
  } else {  // neo-Hooke: mat = (mu, lambda)
    mu = mat[2 * (size_t)e];
    lam = mat[2 * (size_t)e + 1];
  }
  const int q1 = tid / 4, part = tid % 4;
  float ppv[NPP];  // the point's factors (lanes with part == 0)
  {  // phase 1, thread (q1, part): F at q1, mapped gradients of 16 nodes
    const Geo<NB2> &g = S.u.in.g;
    PointGeo P;
    point_geometry(g, q1, part, gp, P);
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      const int a = part * NPART + k;
      const double *ua = S.u.in.U[a];
      double dR[3], G[3];
      rational_dR(g, q1, a, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += ua[i] * G[j];
    }
    for (int m = 1; m < 4; m <<= 1)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += __shfl_xor(L[i][j], m);
    double F[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1.0 : 0.0) + L[i][j];
    double B[6];  // F F^T (xx, yy, zz, xy, xz, yz)
    B[0] = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2];
    B[1] = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2];
    B[2] = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2];
    B[3] = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
    B[4] = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
    B[5] = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
    double M[3][3];  // the map of the staged gradients: I (SVK) or F^-T (neo-Hooke)
    if constexpr (SVK) {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = i == j ? 1.0 : 0.0;
      // Green-Lagrange E = (F^T F - I) / 2 ; S = lambda tr(E) I + 2 mu E
      double C[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
      const double tr = 0.5 * (C[0][0] - 1.0) + 0.5 * (C[1][1] - 1.0) + 0.5 * (C[2][2] - 1.0);
      const double rc = sqrt(P.c);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ppv[3 * i + j] = (float)(rc * F[i][j]);
      for (int k = 0; k < 6; k++) ppv[9 + k] = (float)(P.c * B[k]);
      ppv[15] = (float)(P.c * (lam * tr + mu * (C[0][0] - 1.0)));
      ppv[16] = (float)(P.c * (lam * tr + mu * (C[1][1] - 1.0)));
      ppv[17] = (float)(P.c * (lam * tr + mu * (C[2][2] - 1.0)));
      ppv[18] = (float)(P.c * (mu * C[0][1]));
      ppv[19] = (float)(P.c * (mu * C[0][2]));
      ppv[20] = (float)(P.c * (mu * C[1][2]));
    } else {
      double Fi[3][3];
      double J = inv3d(F, Fi);
      if (!(J > 0.0)) {
        if (part == 0) atomicAdd(&status[1], 1);
        J = 1.0;
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = Fi[j][i];
      ppv[0] = (float)(P.c * lam);
      ppv[1] = (float)(P.c * (mu - lam * log(J)));
      for (int k = 0; k < 6; k++) ppv[2 + k] = (float)(P.c * mu * B[k]);
    }
    if (part == 0) count_jacobian(status, P.detJ);
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      const int a = part * NPART + k;
      double dR[3], G[3];
      rational_dR(g, q1, a, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++)
        S.G[i][q1][a] = (float)(M[i][0] * G[0] + M[i][1] * G[1] + M[i][2] * G[2]);
    }
  }
  __syncthreads();  // the geometry is dead: its LDS takes the points' factors
  if (part == 0)
    for (int k = 0; k < NPP; k++) S.u.pp[q1][k] = ppv[k];
  __syncthreads();
  const float flam = (float)lam, fmu = (float)mu;
  const int wave = tid / 64, lane = tid % 64;
  const int ln = lane % 16, lk = lane / 16;
  float *out = Ke + (size_t)e * NSYM * 9;
  for (int tp = (wave + e) % 4; tp < NPAIR; tp += 4) {
    int ta = 0, r = tp;
    while (r >= NT - ta) {
      r -= NT - ta;
      ta++;
    }
    const int a0 = 16 * ta, b0 = 16 * (ta + r);
    constexpr int NACC = SVK ? 16 : 19;
    floatx4 acc[NACC];
    for (int t = 0; t < NACC; t++)
      for (int m = 0; m < 4; m++) acc[t][m] = 0.f;
#pragma unroll 2
    for (int k0 = 0; k0 < NQP; k0 += 4) {
      const int q = k0 + lk;
      const float *pp = S.u.pp[q];
      float ga[3], gb[3];
      for (int i = 0; i < 3; i++) {
        ga[i] = S.G[i][q][a0 + ln];
        gb[i] = S.G[i][q][b0 + ln];
      }
      if constexpr (SVK) {
        float fa[3], fb[3];
        for (int i = 0; i < 3; i++) {
          fa[i] = pp[3 * i] * ga[0] + pp[3 * i + 1] * ga[1] + pp[3 * i + 2] * ga[2];
          fb[i] = pp[3 * i] * gb[0] + pp[3 * i + 1] * gb[1] + pp[3 * i + 2] * gb[2];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            acc[3 * i + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[i], fb[j], acc[3 * i + j], 0, 0, 0);
        for (int ij = 0; ij < 6; ij++)
          for (int k = 0; k < 3; k++)
            acc[9 + ij] = __builtin_amdgcn_mfma_f32_16x16x4f32(pp[9 + ij] * ga[k], gb[k], acc[9 + ij], 0, 0, 0);
        // c S g_a (S symmetric: xx yy zz xy xz yz at pp[15..20])
        const float s0 = pp[15] * ga[0] + pp[18] * ga[1] + pp[19] * ga[2];
        const float s1 = pp[18] * ga[0] + pp[16] * ga[1] + pp[20] * ga[2];
        const float s2 = pp[19] * ga[0] + pp[20] * ga[1] + pp[17] * ga[2];
        acc[15] = __builtin_amdgcn_mfma_f32_16x16x4f32(s0, gb[0], acc[15], 0, 0, 0);
        acc[15] = __builtin_amdgcn_mfma_f32_16x16x4f32(s1, gb[1], acc[15], 0, 0, 0);
        acc[15] = __builtin_amdgcn_mfma_f32_16x16x4f32(s2, gb[2], acc[15], 0, 0, 0);
      } else {
        for (int i = 0; i < 3; i++) {
          const float a1 = pp[0] * ga[i], a2 = pp[1] * ga[i];
          for (int j = 0; j < 3; j++) {
            acc[3 * i + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, gb[j], acc[3 * i + j], 0, 0, 0);
            acc[9 + 3 * i + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, gb[j], acc[9 + 3 * i + j], 0, 0, 0);
          }
        }
        // c mu (F F^T) g'_a (xx yy zz xy xz yz at pp[2..7])
        const float s0 = pp[2] * ga[0] + pp[5] * ga[1] + pp[6] * ga[2];
        const float s1 = pp[5] * ga[0] + pp[3] * ga[1] + pp[7] * ga[2];
        const float s2 = pp[6] * ga[0] + pp[7] * ga[1] + pp[4] * ga[2];
        acc[18] = __builtin_amdgcn_mfma_f32_16x16x4f32(s0, gb[0], acc[18], 0, 0, 0);
        acc[18] = __builtin_amdgcn_mfma_f32_16x16x4f32(s1, gb[1], acc[18], 0, 0, 0);
        acc[18] = __builtin_amdgcn_mfma_f32_16x16x4f32(s2, gb[2], acc[18], 0, 0, 0);
      }
    }
    // lane holds D[m][n] with n = lane % 16, m = 4 (lane / 16) + r
    for (int r = 0; r < 4; r++) {
      const int a = a0 + 4 * lk + r, b = b0 + ln;
      if (a > b) continue;
      float K[9];
      if constexpr (SVK) {
        float P[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) P[i][j] = acc[3 * i + j][r];
        if (a == b)
          for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++) {
              const float s = 0.5f * (P[i][j] + P[j][i]);
              P[i][j] = s;
              P[j][i] = s;
            }
        const float Bm[9] = {acc[9][r],  acc[12][r], acc[13][r], acc[12][r], acc[10][r],
                             acc[14][r], acc[13][r], acc[14][r], acc[11][r]};
        const float sg = acc[15][r];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            K[3 * i + j] = flam * P[i][j] + fmu * P[j][i] + fmu * Bm[3 * i + j] + (i == j ? sg : 0.f);
      } else {
        float P1[3][3], P2[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            P1[i][j] = acc[3 * i + j][r];
            P2[i][j] = acc[9 + 3 * i + j][r];
          }
        if (a == b)
          for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++) {
              const float s1 = 0.5f * (P1[i][j] + P1[j][i]), s2 = 0.5f * (P2[i][j] + P2[j][i]);
              P1[i][j] = P1[j][i] = s1;
              P2[i][j] = P2[j][i] = s2;
            }
        const float sg = acc[18][r];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) K[3 * i + j] = P1[i][j] + P2[j][i] + (i == j ? sg : 0.f);
      }
      float *o = out + (size_t)sym_index(a, b, NBAS) * 9;
      for (int k = 0; k < 9; k++) o[k] = K[k];
    }
  }
}

// ---------------------------------------------------------------- f_int
// Phase 1, thread (q, part): geometry, then grad u at q from its 16 control
// points (shuffle-reduced), then c times the stress that multiplies the
// gradients: sigma (linear) or P = F S (total Lagrangian; the reference's
// (B + Bl)^T S, elmparam.cpp:1615-1707).
//
// Phase 2 by sum factorization.  With T = (c sigma or c P) J^-1 per point,
// 1/W and dW the rational sums and d_a, n_a the tensor-product numerators
// (w_a times dN_r N_s N_t, N_r dN_s N_t, N_r N_s dN_t and N_r N_s N_t):
//   g_a[i] = sum_q sum_k T_ik dR_a,k = sum_q (sum_{k<3} Z_ik d_a,k + Z_i3 n_a),
//   Z_ik = T_ik / W (k < 3),  Z_i3 = -(sum_k T_ik dW_k) / W^2,
// and every term factors over the three point directions: the sums over
// q_t, then q_s, then q_r are three small contractions with the
// univariate tables (6.9 k multiply-adds per cubic element instead of
// 64 x 64 gradients).
template <int NB2, int FORM>
struct FShared {
  Geo<NB2> g;
  double U[Bsp<NB2>::NBAS][3];  // the element's nodal displacements
  double Ji[NQP][9];
  double rs[NQP][4];   // 1 / W, dW/dr, dW/ds, dW/dt
  // c * sigma (Voigt, linear) or c * P (row-major, total-Lagrangian)
  double sig[NQP][FORM == FGPU_LINEAR ? 6 : 9];
  double Z[3][4][NQP];                 // [i][k][q]
  double Y[3][3][NQ1 * NQ1 * NB];      // after the q_t sums: [i][term][(q_r, q_s, k_t)]
  double V[3][2][NQ1 * NB * NB];       // after the q_s sums: [i][term][(q_r, j_s, k_t)]
};

template <int NB2, int FORM>
__device__ inline void fint_points(FShared<NB2, FORM> &S, int e, double *__restrict__ ge) {
  constexpr int NBAS = Bsp<NB2>::NBAS;
  const int tid = threadIdx.x;
  // Z per point and component: thread (q, i), 192 of them
  if (tid < 3 * NQP) {
    const int q = tid / 3, i = tid - 3 * q;
    const double *s = S.sig[q];
    double si[3];  // row i of c sigma / c P
    if constexpr (FORM == FGPU_LINEAR) {
      const int r0[3] = {0, 3, 4}, r1[3] = {3, 1, 5}, r2[3] = {4, 5, 2};
      si[0] = s[r0[i]];
      si[1] = s[r1[i]];
      si[2] = s[r2[i]];
    } else {
      si[0] = s[3 * i];
      si[1] = s[3 * i + 1];
      si[2] = s[3 * i + 2];
    }
    const double is = S.rs[q][0];
    double t[3];
    for (int k = 0; k < 3; k++)
      t[k] = si[0] * S.Ji[q][k] + si[1] * S.Ji[q][3 + k] + si[2] * S.Ji[q][6 + k];
    for (int k = 0; k < 3; k++) S.Z[i][k][q] = t[k] * is;
    S.Z[i][3][q] = -(t[0] * S.rs[q][1] + t[1] * S.rs[q][2] + t[2] * S.rs[q][3]) * is * is;
  }
  __syncthreads();
  // sums over q_t: thread (i, q_r, q_s, k_t)
  constexpr int NY = NQ1 * NQ1 * NB2;
  if (tid < 3 * NY) {
    const int i = tid / NY, r = tid - NY * i;
    const int kt = r / (NQ1 * NQ1), qrs = r - kt * NQ1 * NQ1;
    double y0 = 0.0, y1 = 0.0, y23 = 0.0;
    for (int qt = 0; qt < NQ1; qt++) {
      const int q = qrs + NQ1 * NQ1 * qt;
      const double nt = S.g.N[2][qt][kt], dnt = S.g.dN[2][qt][kt];
      y0 += S.Z[i][0][q] * nt;
      y1 += S.Z[i][1][q] * nt;
      y23 += S.Z[i][2][q] * dnt + S.Z[i][3][q] * nt;
    }
    const int o = qrs + NQ1 * NQ1 * kt;
    S.Y[i][0][o] = y0;
    S.Y[i][1][o] = y1;
    S.Y[i][2][o] = y23;
  }
  __syncthreads();
  // sums over q_s: thread (i, q_r, j_s, k_t)
  constexpr int NV = NQ1 * NB * NB2;
  if (tid < 3 * NV) {
    const int i = tid / NV, r = tid - NV * i;
    const int kt = r / (NQ1 * NB), rj = r - kt * NQ1 * NB;
    const int js = rj / NQ1, qr = rj - NQ1 * js;
    double v0 = 0.0, v1 = 0.0;
    for (int qs = 0; qs < NQ1; qs++) {
      const int o = qr + NQ1 * qs + NQ1 * NQ1 * kt;
      const double ns = S.g.N[1][qs][js], dns = S.g.dN[1][qs][js];
      v0 += S.Y[i][0][o] * ns;
      v1 += S.Y[i][1][o] * dns + S.Y[i][2][o] * ns;
    }
    const int o = qr + NQ1 * js + NQ1 * NB * kt;
    S.V[i][0][o] = v0;
    S.V[i][1][o] = v1;
  }
  __syncthreads();
  // sums over q_r: thread (a, i), consecutive threads writing consecutive entries
  if (tid < 3 * NBAS) {
    const int a = tid / 3, i = tid - 3 * a;
    const int ir = a % 4, js = (a / 4) % 4, kt = a / 16;
    double g = 0.0;
    for (int qr = 0; qr < NQ1; qr++) {
      const int o = qr + NQ1 * js + NQ1 * NB * kt;
      g += S.V[i][0][o] * S.g.dN[0][qr][ir] + S.V[i][1][o] * S.g.N[0][qr][ir];
    }
    ge[((size_t)e * NBAS + a) * 3 + i] = S.g.W[a] * g;
  }
}

template <int FORM, int NB2>
__global__ __launch_bounds__(256) void bsp_fint_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, double *__restrict__ ge, int *__restrict__ status) {
  constexpr int NBAS = Bsp<NB2>::NBAS, NPART = Bsp<NB2>::NPART;
  extern __shared__ __align__(16) unsigned char smem[];
  FShared<NB2, FORM> &S = *reinterpret_cast<FShared<NB2, FORM> *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  const int *cn = conn + (size_t)e * NBAS;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  stage_u<NB2>(S.U, cn, unodal);
  __syncthreads();
  {
    const int q = tid / 4, part = tid % 4;
    PointGeo P;
    point_geometry(S.g, q, part, gp, P);
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      const double *ua = S.U[part * NPART + k];
      double dR[3], G[3];
      rational_dR(S.g, q, part * NPART + k, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += ua[i] * G[j];
    }
    for (int m = 1; m < 4; m <<= 1)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += __shfl_xor(L[i][j], m);
    if (part == 0) {
      count_jacobian(status, P.detJ);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S.Ji[q][3 * i + j] = P.Ji[i][j];
      S.rs[q][0] = P.inv_s;
      for (int k = 0; k < 3; k++) S.rs[q][1 + k] = P.dW[k];
      const double c = P.c;
      double *sg = S.sig[q];
      if constexpr (FORM == FGPU_LINEAR) {
        linear_csig(mat, e, c, L, sg);
      } else {
        double F[3][3], C[3][3], Sp[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1.0 : 0.0) + L[i][j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
        if constexpr (FORM == FGPU_TL_SVK) {
          const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
          const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)),
                       mu = 0.5 * E / (1.0 + nu);
          const double tr = 0.5 * (C[0][0] - 1.0) + 0.5 * (C[1][1] - 1.0) + 0.5 * (C[2][2] - 1.0);
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
              Sp[i][j] = (i == j ? lam * tr : 0.0) + mu * (C[i][j] - (i == j ? 1.0 : 0.0));
        } else {  // neo-Hooke, mat = (mu, lambda): S = mu (I - C^-1) + lambda ln J C^-1
          const double mu = mat[2 * (size_t)e], lam = mat[2 * (size_t)e + 1];
          double Ci[3][3];
          const double detC = inv3d(C, Ci);
          if (!(detC > 0.0)) atomicAdd(&status[1], 1);
          const double lnJ = detC > 0.0 ? 0.5 * log(detC) : 0.0;
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
              Sp[i][j] = mu * ((i == j ? 1.0 : 0.0) - Ci[i][j]) + lam * lnJ * Ci[i][j];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            sg[3 * i + j] = c * (F[i][0] * Sp[0][j] + F[i][1] * Sp[1][j] + F[i][2] * Sp[2][j]);
      }
    }
  }
  __syncthreads();
  fint_points<NB2, FORM>(S, e, ge);
}

// The internal force of the linear form from the per-point records the
// stiffness kernel wrote (bsp_K_kernel<NB2, true>): the second phase of
// bsp_fint_kernel on the same data, so the same result.
template <int NB2>
__global__ __launch_bounds__(256) void bsp_fint_pts_kernel(
    const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ gp, const double *__restrict__ pts, double *__restrict__ ge) {
  extern __shared__ __align__(16) unsigned char smem[];
  FShared<NB2, FGPU_LINEAR> &S = *reinterpret_cast<FShared<NB2, FGPU_LINEAR> *>(smem);
  const int e = blockIdx.x;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  const double *r0 = pts + (size_t)e * NQP * PREC;
  for (int i = threadIdx.x; i < NQP * PREC; i += blockDim.x) {
    const int q = i / PREC, f = i - PREC * q;
    const double v = r0[i];
    if (f < 9)
      S.Ji[q][f] = v;
    else if (f < 13)
      S.rs[q][f - 9] = v;
    else
      S.sig[q][f - 13] = v;
  }
  __syncthreads();
  fint_points<NB2, FGPU_LINEAR>(S, e, ge);
}

// dynamic LDS above 64 KB must be allowed per kernel (cheap; set per launch)
// ---------------------------------------------------------------- body force
// cElemLoad::ExtForce with cBodyLoad (load/load.cpp:618-691, 969-974):
// f_a = sum_q w detJ R_a b.  Phase 1, thread (q, part): the point's w detJ
// and rational denominator; phase 2, thread (a, quarter): sum over every fourth point.
template <int NB2>
struct BShared {
  Geo<NB2> g;
  double c[NQP], inv_s[NQP];
};

template <int NB2>
__global__ __launch_bounds__(256) void bsp_body_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ bvec, const double *__restrict__ gp, double *__restrict__ ge,
    int *__restrict__ status) {
  constexpr int NBAS = Bsp<NB2>::NBAS;
  extern __shared__ __align__(16) unsigned char smem[];
  BShared<NB2> &S = *reinterpret_cast<BShared<NB2> *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  __syncthreads();
  {
    const int q = tid / 4, part = tid % 4;
    PointGeo P;
    point_geometry(S.g, q, part, gp, P);
    if (part == 0) {
      S.c[q] = P.c;
      S.inv_s[q] = P.inv_s;
      count_jacobian(status, P.detJ);
    }
  }
  __syncthreads();
  if (tid < 4 * NBAS) {  // whole waves
    const int a = tid / 4, quarter = tid % 4;
    double sN = 0.0;
    #pragma unroll
    for (int q = quarter; q < NQP; q += 4) {  // the quarters' points side by side in LDS
      double n, d[3];
      tensor(S.g, q, a, n, d);
      sN += S.c[q] * (n * S.inv_s[q]);
    }
    for (int m = 1; m < 4; m <<= 1) sN += __shfl_xor(sN, m);
    if (quarter == 0) {
      const double *b = bvec + 3 * (size_t)e;
      double *o = ge + ((size_t)e * NBAS + a) * 3;
      for (int i = 0; i < 3; i++) o[i] = sN * b[i];
    }
  }
}

// ---------------------------------------------------------------- stresses
// cElmParam::IntPntStress (elc/elmparam.cpp:859-937) / cElmParamTL::
// IntPntStress (1811-1885) at the 64 points and, when sig_node is not null,
// cElmParam::NodalStress (1304-1333): s_a = sum_q TR[a][q] s_q with the
// extrapolation matrix of CompTRMatrix (1444-1535).  Phase 1, thread
// (q, part): grad u at q as in the f_int kernel, then the point's stress;
// phase 2, thread (a, quarter): every fourth point of the extrapolation sum.
template <int NB2>
struct SShared {
  Geo<NB2> g;
  double U[Bsp<NB2>::NBAS][3];
  double sig[NQP][6];
};

template <int FORM, int NB2>
__global__ __launch_bounds__(256) void bsp_stress_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, const double *__restrict__ TR, double *__restrict__ sig_ip,
    double *__restrict__ sig_node, int *__restrict__ status) {
  constexpr int NBAS = Bsp<NB2>::NBAS, NPART = Bsp<NB2>::NPART;
  extern __shared__ __align__(16) unsigned char smem[];
  SShared<NB2> &S = *reinterpret_cast<SShared<NB2> *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  const int *cn = conn + (size_t)e * NBAS;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  stage_u<NB2>(S.U, cn, unodal);
  __syncthreads();
  {
    const int q = tid / 4, part = tid % 4;
    PointGeo P;
    point_geometry(S.g, q, part, gp, P);
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    #pragma unroll 2
    for (int k = 0; k < NPART; k++) {
      const double *ua = S.U[part * NPART + k];
      double dR[3], G[3];
      rational_dR(S.g, q, part * NPART + k, P, dR);
      for (int i = 0; i < 3; i++)
        G[i] = P.Ji[i][0] * dR[0] + P.Ji[i][1] * dR[1] + P.Ji[i][2] * dR[2];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += ua[i] * G[j];
    }
    for (int m = 1; m < 4; m <<= 1)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += __shfl_xor(L[i][j], m);
    if (part == 0) {
      count_jacobian(status, P.detJ);
      double sg[6];
      if (!point_stress<FORM>(L, mat[2 * (size_t)e], mat[2 * (size_t)e + 1], sg))
        atomicAdd(&status[1], 1);
      double *o = sig_ip + ((size_t)e * NQP + q) * 6;
      for (int k = 0; k < 6; k++) {
        o[k] = sg[k];
        S.sig[q][k] = sg[k];
      }
    }
  }
  if (!sig_node) return;
  __syncthreads();
  if (tid < 4 * NBAS) {  // whole waves
    const int a = tid / 4, quarter = tid % 4;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    #pragma unroll
    for (int q = quarter; q < NQP; q += 4) {  // the quarters' points side by side in LDS
      const double w = TR[a * NQP + q];
      for (int k = 0; k < 6; k++) acc[k] += w * S.sig[q][k];
    }
    for (int m = 1; m < 4; m <<= 1)
      for (int k = 0; k < 6; k++) acc[k] += __shfl_xor(acc[k], m);
    if (quarter == 0) {
      double *o = sig_node + ((size_t)e * NBAS + a) * 6;
      for (int k = 0; k < 6; k++) o[k] = acc[k];
    }
  }
}

template <class K>
void set_lds(K kernel, size_t lds) {
  FGPU_CHECK(hipFuncSetAttribute((const void *)kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

}  // namespace

void bsp_prepare_tables(Ctx &c, const double *h_ext, const double *h_cpw) {
  if (h_cpw) {
    FGPU_CHECK(hipMalloc(&c.d_cpw, c.nnode * sizeof(double)));
    FGPU_CHECK(hipMemcpyAsync(c.d_cpw, h_cpw, c.nnode * sizeof(double), hipMemcpyHostToDevice,
                              c.stream));
  }
  if (bsp_generic(c)) {
    gen_prepare(c, h_ext);  // the caller's packed layout as it is
    return;
  }
  // extraction rows stay FP64 on the device, [e][dir][4][4]: the caller's
  // packed (p_d + 1)^2 per direction, zero-padded for degree 2
  const size_t n = (size_t)c.nelem * 3 * 16;
  std::vector<double> pad(n, 0.0);
  size_t per = 0;
  for (int d = 0; d < 3; d++) per += (size_t)(c.deg[d] + 1) * (c.deg[d] + 1);
  for (size_t e = 0; e < (size_t)c.nelem; e++) {
    const double *src = h_ext + e * per;
    for (int d = 0; d < 3; d++) {
      const int q = c.deg[d] + 1;
      for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) pad[(e * 3 + d) * 16 + i * 4 + j] = src[i * q + j];
      src += q * q;
    }
  }
  FGPU_CHECK(hipMalloc(&c.d_ext, n * sizeof(double)));
  FGPU_CHECK(hipMemcpyAsync(c.d_ext, pad.data(), n * sizeof(double), hipMemcpyHostToDevice,
                            c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
}

namespace {

template <int NB2>
void upload_sym_table() {
  SymTable<NB2> t;
  constexpr int NBAS = Bsp<NB2>::NBAS;
  for (int a = 0; a < NBAS; a++)
    for (int b = a; b < NBAS; b++) t.ab[sym_index(a, b, NBAS)] = (unsigned short)(a * 64 + b);
  if constexpr (NB2 == 4)
    FGPU_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(kSym4), &t, sizeof t));
  else
    FGPU_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(kSym3), &t, sizeof t));
}

template <int FORM, int NB2>
void launch_Kt(Ctx &c, const double *u) {
  static int table_dev = -1;  // the packed (a, b) table, once per device
  if (table_dev != c.device) {
    upload_sym_table<NB2>();
    table_dev = c.device;
  }
  if (c.bsp_tl_valu) {  // the vector-ALU kernel, kept for comparison
    const size_t lds = sizeof(KtShared<FORM, NB2>);
    set_lds(bsp_Kt_kernel<FORM, NB2>, lds);
    hipLaunchKernelGGL((bsp_Kt_kernel<FORM, NB2>), dim3(c.nelem), dim3(256), lds, c.stream,
                       c.nelem, c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_mat, u, c.d_gauss64,
                       c.d_Ke, c.d_status);
  } else {
    const size_t lds = sizeof(KtmShared<FORM, NB2>);
    set_lds(bsp_Ktm_kernel<FORM, NB2>, lds);
    hipLaunchKernelGGL((bsp_Ktm_kernel<FORM, NB2>), dim3(c.nelem), dim3(256), lds, c.stream,
                       c.nelem, c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_mat, u, c.d_gauss64,
                       c.d_Ke, c.d_status);
  }
  FGPU_CHECK(hipGetLastError());
}

template <int NB2>
void launch_K_lin(Ctx &c) {
  const size_t lds = sizeof(KShared<NB2, false>);
  set_lds(bsp_K_kernel<NB2, false>, lds);
  hipLaunchKernelGGL((bsp_K_kernel<NB2, false>), dim3(c.nelem), dim3(256), lds, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_mat, c.d_gauss64,
                     c.d_Ke, c.d_status, nullptr, nullptr);
  FGPU_CHECK(hipGetLastError());
}

template <int NB2>
void launch_K_lin_pts(Ctx &c, const double *u) {
  const size_t lds = sizeof(KShared<NB2, true>);
  set_lds(bsp_K_kernel<NB2, true>, lds);
  hipLaunchKernelGGL((bsp_K_kernel<NB2, true>), dim3(c.nelem), dim3(256), lds, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_mat, c.d_gauss64,
                     c.d_Ke, c.d_status, u, c.d_pts);
  FGPU_CHECK(hipGetLastError());
}

template <int NB2>
void launch_fint_pts(Ctx &c) {
  const size_t lds = sizeof(FShared<NB2, FGPU_LINEAR>);
  hipLaunchKernelGGL(bsp_fint_pts_kernel<NB2>, dim3(c.nelem), dim3(256), lds, c.stream,
                     c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_gauss64, c.d_pts, c.d_ge);
  FGPU_CHECK(hipGetLastError());
}

template <int FORM, int NB2>
void launch_fint(Ctx &c, const double *u) {
  const size_t lds = sizeof(FShared<NB2, FORM>);
  hipLaunchKernelGGL((bsp_fint_kernel<FORM, NB2>), dim3(c.nelem), dim3(256), lds, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_mat, u, c.d_gauss64,
                     c.d_ge, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

template <int NB2>
void launch_K_nb(Ctx &c, const double *u) {
  if (c.form == FGPU_TL_SVK)
    launch_Kt<FGPU_TL_SVK, NB2>(c, u);
  else if (c.form == FGPU_TL_NEOHOOKE)
    launch_Kt<FGPU_TL_NEOHOOKE, NB2>(c, u);
  else
    launch_K_lin<NB2>(c);
}

template <int NB2>
void launch_fint_nb(Ctx &c, const double *u) {
  switch (c.form) {
    case FGPU_TL_SVK: launch_fint<FGPU_TL_SVK, NB2>(c, u); break;
    case FGPU_TL_NEOHOOKE: launch_fint<FGPU_TL_NEOHOOKE, NB2>(c, u); break;
    default: launch_fint<FGPU_LINEAR, NB2>(c, u); break;
  }
}

template <int NB2>
void launch_body_nb(Ctx &c, const double *b) {
  const size_t lds = sizeof(BShared<NB2>);
  hipLaunchKernelGGL(bsp_body_kernel<NB2>, dim3(c.nelem), dim3(256), lds, c.stream, c.nelem,
                     c.d_coords, c.d_conn, c.d_cpw, c.d_ext, b, c.d_gauss64, c.d_ge, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

template <int FORM, int NB2>
void launch_stress(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  const size_t lds = sizeof(SShared<NB2>);
  hipLaunchKernelGGL((bsp_stress_kernel<FORM, NB2>), dim3(c.nelem), dim3(256), lds, c.stream,
                     c.nelem, c.d_coords, c.d_conn, c.d_cpw, c.d_ext, c.d_mat, u, c.d_gauss64,
                     c.d_TR, sig_ip, sig_node, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

template <int NB2>
void launch_stress_nb(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  switch (c.form) {
    case FGPU_TL_SVK: launch_stress<FGPU_TL_SVK, NB2>(c, u, sig_ip, sig_node); break;
    case FGPU_TL_NEOHOOKE: launch_stress<FGPU_TL_NEOHOOKE, NB2>(c, u, sig_ip, sig_node); break;
    default: launch_stress<FGPU_LINEAR, NB2>(c, u, sig_ip, sig_node); break;
  }
}

}  // namespace

void launch_bsp_stress(Ctx &c, const double *u, double *sig_ip, double *sig_node) {
  if (bsp_generic(c))
    gen_launch_stress(c, u, sig_ip, sig_node);
  else if (c.deg[2] == 2)
    launch_stress_nb<3>(c, u, sig_ip, sig_node);
  else
    launch_stress_nb<4>(c, u, sig_ip, sig_node);
}

void launch_bsp_body(Ctx &c, const double *b) {
  if (bsp_generic(c))
    gen_launch_vec(c, nullptr, b);
  else if (c.deg[2] == 2)
    launch_body_nb<3>(c, b);
  else
    launch_body_nb<4>(c, b);
}

void launch_bsp_K(Ctx &c, const double *u) {
  if (c.form != FGPU_LINEAR && !u) {
    // zero displacement: stage a zero nodal field
    FGPU_CHECK(hipMemsetAsync(c.d_u, 0, sizeof(double) * 3 * c.nnode, c.stream));
    u = c.d_u;
  }
  if (bsp_generic(c))
    gen_launch_K(c, u);
  else if (c.deg[2] == 2)
    launch_K_nb<3>(c, u);
  else
    launch_K_nb<4>(c, u);
}

// the point records are allocated with the context (bsp_alloc_pts); without
// them (over the budget, or the allocation failed) f_int takes the unfused
// kernels, which need no extra memory
bool bsp_fused_fint(const Ctx &c) {
  return c.shape == FGPU_BSP && !bsp_generic(c) && c.form == FGPU_LINEAR && c.d_pts;
}

void bsp_alloc_pts(Ctx &c) {
  if (c.shape != FGPU_BSP || bsp_generic(c) || c.form != FGPU_LINEAR || c.d_pts) return;
  const size_t bytes = sizeof(double) * (size_t)c.nelem * NQP * PREC;  // 9.7 KB per element
  if ((double)bytes > 4.0 * (1 << 30)) return;
  if (hipMalloc(&c.d_pts, bytes) != hipSuccess) {
    c.d_pts = nullptr;
    (void)hipGetLastError();  // clear the failed allocation's error
  }
}

void launch_bsp_K_pts(Ctx &c, const double *u) {
  if (!bsp_fused_fint(c)) throw Error("fused internal force: linear tuned B-spline kernels only");
  if (c.deg[2] == 2)
    launch_K_lin_pts<3>(c, u);
  else
    launch_K_lin_pts<4>(c, u);
}

void launch_bsp_fint_pts(Ctx &c) {
  if (c.deg[2] == 2)
    launch_fint_pts<3>(c);
  else
    launch_fint_pts<4>(c);
}

void launch_bsp_fint(Ctx &c, const double *u) {
  if (bsp_generic(c))
    gen_launch_vec(c, u, nullptr);
  else if (c.deg[2] == 2)
    launch_fint_nb<3>(c, u);
  else
    launch_fint_nb<4>(c, u);
}

}  // namespace fgpu
