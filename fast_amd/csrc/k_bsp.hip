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
// The six 64x64x64 products P_ij (i <= j; the diagonal ones only on their
// upper 32x32 tiles) run on v_mfma_f32_32x32x2f32 with FP32 accumulation.
// One workgroup (4 waves) per element: phase 1 stages G (FP32) in LDS from an
// FP64 geometry pass, phase 2 runs the 21 output tiles on MFMA into LDS,
// phase 3 forms the 2080 upper 3x3 node blocks and stores them contiguously.
#include "fgpu_internal.h"

namespace fgpu {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int P3 = 3, NB = 4, NQ1 = 4;   // degree, basis and points per direction
constexpr int NBAS = 64, NQP = 64;
constexpr int NSYM = NBAS * (NBAS + 1) / 2;  // 2080
constexpr int PLD = 65;                      // padded row of a P tile in LDS

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

// Shared geometry pass for one element: univariate tables, then for each
// point q the rational basis derivatives, J, J^-1 and w detJ.  Threads are
// (q, part) with part owning 16 control points; J partial sums are reduced
// across the 4 lanes of a point with shuffles.
struct Geo {
  double N[3][NQ1][NB], dN[3][NQ1][NB];  // [dir][point][basis]
  double X[NBAS][3];
  double W[NBAS];
};

__device__ inline void geo_load(Geo &g, int e, const double *__restrict__ coords,
                                const int *__restrict__ conn, const double *__restrict__ cpw,
                                const double *__restrict__ ext, const double *__restrict__ gp) {
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

// For point q and control point a: numerator values (N w, dN w) of the
// tensor-product basis; a = i + 4 j + 16 k with i along r.
__device__ inline void tensor(const Geo &g, int q, int a, double &n, double d[3]) {
  const int qr = q % 4, qs = (q / 4) % 4, qt = q / 16;
  const int i = a % 4, j = (a / 4) % 4, k = a / 16;
  const double w = g.W[a];
  const double Nr = g.N[0][qr][i], Ns = g.N[1][qs][j], Nt = g.N[2][qt][k];
  n = Nr * Ns * Nt * w;
  d[0] = g.dN[0][qr][i] * Ns * Nt * w;
  d[1] = Nr * g.dN[1][qs][j] * Nt * w;
  d[2] = Nr * Ns * g.dN[2][qt][k] * w;
}

// Computes, for the calling thread's (q, part): J^-1 (all lanes of the point
// agree) and w detJ.  Returns the rational sums through Wsum, dW.
__device__ inline void point_geometry(const Geo &g, int q, int part, const double *gp,
                                      double Ji[3][3], double &c, double &Wsum, double dW[3],
                                      double &detJ) {
  double s = 0.0, ds[3] = {0, 0, 0};
  for (int a = part * 16; a < part * 16 + 16; a++) {
    double n, d[3];
    tensor(g, q, a, n, d);
    s += n;
    ds[0] += d[0];
    ds[1] += d[1];
    ds[2] += d[2];
  }
  for (int m = 1; m < 4; m <<= 1) {
    s += __shfl_xor(s, m);
    for (int k = 0; k < 3; k++) ds[k] += __shfl_xor(ds[k], m);
  }
  Wsum = s;
  for (int k = 0; k < 3; k++) dW[k] = ds[k];
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = part * 16; a < part * 16 + 16; a++) {
    double n, d[3];
    tensor(g, q, a, n, d);
    double R = n / s;
    for (int i = 0; i < 3; i++) {
      double dR = (d[i] - R * dW[i]) / s;
      for (int k = 0; k < 3; k++) J[i][k] += dR * g.X[a][k];
    }
  }
  for (int m = 1; m < 4; m <<= 1)
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) J[i][k] += __shfl_xor(J[i][k], m);
  detJ = inv3d(J, Ji);
  const int qr = q % 4, qs = (q / 4) % 4, qt = q / 16;
  c = gp[(0 * 16 + qr) * 2 + 1] * gp[(1 * 16 + qs) * 2 + 1] * gp[(2 * 16 + qt) * 2 + 1] * detJ;
}

struct KShared {
  Geo g;
  float G[3][NQP][NBAS];   // dN_a/dx_i at point q
  float c[NQP];            // w detJ
  float P[6][NBAS][PLD];   // xx yy zz xy xz yz
};

__constant__ unsigned char kTileI[21] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
__constant__ unsigned char kTileJ[21] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2};
__constant__ unsigned char kTileP[21] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5};
__constant__ unsigned char kTileA[21] = {0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1};
__constant__ unsigned char kTileB[21] = {0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1};

__global__ __launch_bounds__(256) void bsp_K_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ gp,
    const unsigned short *__restrict__ blkab, float *__restrict__ Ke, int *__restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  KShared &S = *reinterpret_cast<KShared *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  __syncthreads();
  {  // phase 1: G and c
    const int q = tid / 4, part = tid % 4;
    double Ji[3][3], c, Ws, dW[3], det;
    point_geometry(S.g, q, part, gp, Ji, c, Ws, dW, det);
    if (part == 0) {
      S.c[q] = (float)c;
      if (!(det > 0.0)) atomicAdd(&status[0], 1);
    }
    for (int a = part * 16; a < part * 16 + 16; a++) {
      double n, d[3];
      tensor(S.g, q, a, n, d);
      double R = n / Ws, dR[3];
      for (int i = 0; i < 3; i++) dR[i] = (d[i] - R * dW[i]) / Ws;
      for (int i = 0; i < 3; i++)
        S.G[i][q][a] = (float)(Ji[i][0] * dR[0] + Ji[i][1] * dR[1] + Ji[i][2] * dR[2]);
    }
  }
  __syncthreads();
  {  // phase 2: 21 tiles of 32x32 (K = 64 points) on MFMA
    const int wave = tid / 64, lane = tid % 64;
    const int lm = lane % 32, lk = lane / 32;
    for (int t = wave; t < 21; t += 4) {
      const int i = kTileI[t], j = kTileJ[t], p = kTileP[t];
      const int a0 = 32 * kTileA[t], b0 = 32 * kTileB[t];
      floatx16 acc;
      for (int r = 0; r < 16; r++) acc[r] = 0.f;
#pragma unroll 8
      for (int k = 0; k < NQP; k += 2) {
        const int q = k + lk;
        float av = S.c[q] * S.G[i][q][a0 + lm];
        float bv = S.G[j][q][b0 + lm];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
      }
      // D[m][n]: n = lane % 32, m = 8 (r / 4) + 4 (lane / 32) + r % 4
      for (int r = 0; r < 16; r++) {
        const int m = 8 * (r / 4) + 4 * lk + (r % 4);
        S.P[p][a0 + m][b0 + lm] = acc[r];
      }
    }
  }
  __syncthreads();
  {  // phase 3: K_ab for a <= b
    const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
    const float lam = (float)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    const float mu = (float)(0.5 * E / (1.0 + nu));
    float *out = Ke + (size_t)e * NSYM * 9;
    for (int blk = tid; blk < NSYM; blk += blockDim.x) {
      const unsigned short ab = blkab[blk];
      const int a = ab >> 8, b = ab & 255;
      float Pab[3][3];
      Pab[0][0] = S.P[0][a][b];
      Pab[1][1] = S.P[1][a][b];
      Pab[2][2] = S.P[2][a][b];
      Pab[0][1] = S.P[3][a][b];
      Pab[1][0] = S.P[3][b][a];
      Pab[0][2] = S.P[4][a][b];
      Pab[2][0] = S.P[4][b][a];
      Pab[1][2] = S.P[5][a][b];
      Pab[2][1] = S.P[5][b][a];
      const float tr = mu * (Pab[0][0] + Pab[1][1] + Pab[2][2]);
      float *o = out + (size_t)blk * 9;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          o[3 * i + j] = lam * Pab[i][j] + mu * Pab[j][i] + (i == j ? tr : 0.f);
    }
  }
}

// ---------------------------------------------------------------- f_int
struct FShared {
  Geo g;
  double G[3][NQP][NBAS];
  double sig[NQP][6];  // c * sigma (Voigt xx yy zz xy xz yz)
};

__global__ __launch_bounds__(256) void bsp_fint_kernel(
    int nelem, const double *__restrict__ coords, const int *__restrict__ conn,
    const double *__restrict__ cpw, const double *__restrict__ ext,
    const double *__restrict__ mat, const double *__restrict__ unodal,
    const double *__restrict__ gp, double *__restrict__ ge, int *__restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  FShared &S = *reinterpret_cast<FShared *>(smem);
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  const int *cn = conn + (size_t)e * NBAS;
  geo_load(S.g, e, coords, conn, cpw, ext, gp);
  __syncthreads();
  {
    const int q = tid / 4, part = tid % 4;
    double Ji[3][3], c, Ws, dW[3], det;
    point_geometry(S.g, q, part, gp, Ji, c, Ws, dW, det);
    if (part == 0 && !(det > 0.0)) atomicAdd(&status[0], 1);
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = part * 16; a < part * 16 + 16; a++) {
      double n, d[3];
      tensor(S.g, q, a, n, d);
      double R = n / Ws, dR[3], G[3];
      for (int i = 0; i < 3; i++) dR[i] = (d[i] - R * dW[i]) / Ws;
      for (int i = 0; i < 3; i++) {
        G[i] = Ji[i][0] * dR[0] + Ji[i][1] * dR[1] + Ji[i][2] * dR[2];
        S.G[i][q][a] = G[i];
      }
      const double *ua = unodal + 3 * (size_t)cn[a];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += ua[i] * G[j];
    }
    for (int m = 1; m < 4; m <<= 1)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) L[i][j] += __shfl_xor(L[i][j], m);
    if (part == 0) {
      const double E = mat[2 * (size_t)e], nu = mat[2 * (size_t)e + 1];
      const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 0.5 * E / (1.0 + nu);
      double tr = L[0][0] + L[1][1] + L[2][2];
      S.sig[q][0] = c * (lam * tr + 2.0 * mu * L[0][0]);
      S.sig[q][1] = c * (lam * tr + 2.0 * mu * L[1][1]);
      S.sig[q][2] = c * (lam * tr + 2.0 * mu * L[2][2]);
      S.sig[q][3] = c * mu * (L[0][1] + L[1][0]);
      S.sig[q][4] = c * mu * (L[0][2] + L[2][0]);
      S.sig[q][5] = c * mu * (L[1][2] + L[2][1]);
    }
  }
  __syncthreads();
  if (tid < NBAS * 3) {
    const int a = tid / 3, i = tid % 3;
    // sigma row i in Voigt positions
    const int v0 = i == 0 ? 0 : (i == 1 ? 3 : 4);
    const int v1 = i == 0 ? 3 : (i == 1 ? 1 : 5);
    const int v2 = i == 0 ? 4 : (i == 1 ? 5 : 2);
    double s = 0.0;
    for (int q = 0; q < NQP; q++)
      s += S.sig[q][v0] * S.G[0][q][a] + S.sig[q][v1] * S.G[1][q][a] + S.sig[q][v2] * S.G[2][q][a];
    ge[((size_t)e * NBAS + a) * 3 + i] = s;
  }
}

unsigned short *g_blkab = nullptr;  // shared (a << 8 | b) table, per process
int g_blkab_dev = -1;

unsigned short *blkab_table(Ctx &c) {
  if (g_blkab && g_blkab_dev == c.device) return g_blkab;
  std::vector<unsigned short> t;
  for (int a = 0; a < NBAS; a++)
    for (int b = a; b < NBAS; b++) t.push_back((unsigned short)(a << 8 | b));
  FGPU_CHECK(hipMalloc(&g_blkab, t.size() * sizeof(unsigned short)));
  FGPU_CHECK(hipMemcpy(g_blkab, t.data(), t.size() * sizeof(unsigned short),
                       hipMemcpyHostToDevice));
  g_blkab_dev = c.device;
  return g_blkab;
}

}  // namespace

void bsp_prepare_tables(Ctx &c, const double *h_ext, const double *h_cpw) {
  // extraction rows and weights stay FP64 on the device
  size_t n = (size_t)c.nelem * 3 * 16;
  double *d_ext = nullptr, *d_w = nullptr;
  FGPU_CHECK(hipMalloc(&d_ext, n * sizeof(double)));
  FGPU_CHECK(hipMemcpyAsync(d_ext, h_ext, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (h_cpw) {
    FGPU_CHECK(hipMalloc(&d_w, c.nnode * sizeof(double)));
    FGPU_CHECK(hipMemcpyAsync(d_w, h_cpw, c.nnode * sizeof(double), hipMemcpyHostToDevice,
                              c.stream));
  }
  c.d_ext = d_ext;
  c.d_cpw = d_w;
  blkab_table(c);
}

void launch_bsp_K(Ctx &c, const double *u) {
  (void)u;
  size_t lds = sizeof(KShared);
  static bool attr = false;
  if (!attr) {
    FGPU_CHECK(hipFuncSetAttribute((const void *)bsp_K_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  hipLaunchKernelGGL(bsp_K_kernel, dim3(c.nelem), dim3(256), lds, c.stream, c.nelem,
                     c.d_coords, c.d_conn, c.d_cpw,
                     c.d_ext, c.d_mat, c.d_gauss64,
                     blkab_table(c), c.d_Ke, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

void launch_bsp_fint(Ctx &c, const double *u) {
  size_t lds = sizeof(FShared);
  static bool attr = false;
  if (!attr) {
    FGPU_CHECK(hipFuncSetAttribute((const void *)bsp_fint_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  hipLaunchKernelGGL(bsp_fint_kernel, dim3(c.nelem), dim3(256), lds, c.stream, c.nelem,
                     c.d_coords, c.d_conn, c.d_cpw,
                     c.d_ext, c.d_mat, u, c.d_gauss64,
                     c.d_ge, c.d_status);
  FGPU_CHECK(hipGetLastError());
}

}  // namespace fgpu
