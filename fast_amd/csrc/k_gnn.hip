// k_gnn.hip - mesh graph network for warm-starting the Newton/PCG solves
// (no counterpart in the reference): forward, backward and Adam, all
// hand-written for gfx950.  Dense layers run on v_mfma_f32_16x16x16f16
// (FP16 operands converted on load, FP32 accumulation, FP32 master weights).
//
// Graph: the mesh nodes; neighbours = nodes sharing an element (the node-block
// rows of the global matrix, self included).  Per node features (16, padded):
//   x, y, z (box-normalised), u_state / u_scale (3), lambda / lam_scale,
//   dlambda / lam_scale, free-direction flags (3), load flags (3), 1.
// Network (H = 64, L = 4):
//   h0 = relu(X W_in + b_in)
//   h_{l+1} = h_l + relu(h_l W_s,l + mean_N(h_l) W_n,l + b_l)
//   d = relu(h_L W_d1 + b_d1);  y = d W_d2 + b_d2   (first 3 columns used)
//   du = dlambda * out_scale * y   (zero on fixed directions)
// Loss: mean squared error of y against (u_target - u_state) / (dlambda *
// out_scale) over free directions.  Every reduction (weight gradients, bias
// gradients, loss) is a block-partial sum added in a fixed order.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "fgpu_dev.h"

namespace fgpu {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GH = 64;   // hidden width
constexpr int GF = 16;   // padded feature / output width
constexpr int GL = 4;    // message-passing layers

// parameter offsets in the flat FP32 vector
struct GnnLayout {
  int w_in, b_in, ws[GL], wn[GL], bl[GL], w_d1, b_d1, w_d2, b_d2, total;
  GnnLayout() {
    int o = 0;
    w_in = o; o += GF * GH;
    b_in = o; o += GH;
    for (int l = 0; l < GL; l++) {
      ws[l] = o; o += GH * GH;
      wn[l] = o; o += GH * GH;
      bl[l] = o; o += GH;
    }
    w_d1 = o; o += GH * GH;
    b_d1 = o; o += GH;
    w_d2 = o; o += GH * GF;
    b_d2 = o; o += GF;
    total = o;
  }
};

struct Gnn {
  Ctx *c = nullptr;
  GnnLayout lay;
  int N = 0;              // nodes
  float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;  // params, grads, Adam moments
  float *X = nullptr;     // [N x 16] features
  float *H[GL + 1] = {};  // [N x 64] h_l
  float *A[GL] = {};      // [N x 64] relu outputs of the layers
  float *AG[GL] = {};     // [N x 64] mean aggregations
  float *H0 = nullptr;    // relu(X W_in + b)  (== H[0])
  float *D1 = nullptr;    // [N x 64]
  float *Y = nullptr;     // [N x 16]
  float *T = nullptr;     // [N x 16] target (scaled)
  float *dY = nullptr, *dD1 = nullptr, *dH = nullptr, *dZ = nullptr, *tmp = nullptr,
        *tmp2 = nullptr;  // [N x 64] work
  float *part = nullptr;  // partial sums
  float *box = nullptr;   // [6] bounding box (min, 1/extent)
  double *dbuf = nullptr; // loss scalar
  double u_scale = 1.0, lam_scale = 1.0, out_scale = 1.0;
  int nblk_rows = 0;
  int step = 0;
  unsigned seed = 0;
};

namespace {

constexpr int ROWS_PER_BLOCK = 64;   // rowgemm: 4 waves x 16 rows
constexpr int WG_ROWS = 2048;        // rows per weight-gradient partial

struct Bop {  // B[k][n] = p[k * sk + n * sm]
  const float *p;
  int sk, sm;
};

__device__ inline half4 load_a(const float *__restrict__ X, int ld, long row, int k, long N) {
  half4 h;
  if (row < N) {
    const float *x = X + row * (long)ld + k;
    h[0] = (_Float16)x[0];
    h[1] = (_Float16)x[1];
    h[2] = (_Float16)x[2];
    h[3] = (_Float16)x[3];
  } else {
    h[0] = h[1] = h[2] = h[3] = (_Float16)0.f;
  }
  return h;
}

__device__ inline half4 load_b(Bop B, int k, int n) {
  half4 h;
  for (int q = 0; q < 4; q++) h[q] = (_Float16)B.p[(k + q) * B.sk + n * B.sm];
  return h;
}

// Y = act(X1 B1 + X2 B2 + bias); Yres = R + Y when R is given.
// One wave: 16 rows x M columns; K1, K2 multiples of 16.
template <int M, bool RELU>
__global__ __launch_bounds__(256) void rowgemm_kernel(long N, const float *__restrict__ X1, int K1,
                                                      Bop B1, const float *__restrict__ X2, int K2,
                                                      Bop B2, const float *__restrict__ bias,
                                                      const float *__restrict__ R,
                                                      float *__restrict__ Y,
                                                      float *__restrict__ Yres) {
  constexpr int NT = M / 16;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const long r0 = (long)blockIdx.x * ROWS_PER_BLOCK + wave * 16;
  const int lm = lane % 16, lk = 4 * (lane / 16);
  floatx4 acc[NT];
  for (int t = 0; t < NT; t++)
    for (int r = 0; r < 4; r++) acc[t][r] = 0.f;
  for (int kb = 0; kb < K1; kb += 16) {
    half4 a = load_a(X1, K1, r0 + lm, kb + lk, N);
    for (int t = 0; t < NT; t++) {
      half4 b = load_b(B1, kb + lk, 16 * t + lm);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc[t], 0, 0, 0);
    }
  }
  if (X2)
    for (int kb = 0; kb < K2; kb += 16) {
      half4 a = load_a(X2, K2, r0 + lm, kb + lk, N);
      for (int t = 0; t < NT; t++) {
        half4 b = load_b(B2, kb + lk, 16 * t + lm);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc[t], 0, 0, 0);
      }
    }
  for (int t = 0; t < NT; t++)
    for (int r = 0; r < 4; r++) {
      const long row = r0 + lk + r;
      const int col = 16 * t + lm;
      if (row >= N) continue;
      float y = acc[t][r] + (bias ? bias[col] : 0.f);
      if (RELU) y = y > 0.f ? y : 0.f;
      Y[row * M + col] = y;
      if (Yres) Yres[row * M + col] = R[row * M + col] + y;
    }
}

// partial[blk][i][o] = sum over the block's rows of X[row][i] dY[row][o]
template <int KI, int MO>
__global__ __launch_bounds__(256) void wgrad_kernel(long N, const float *__restrict__ X,
                                                    const float *__restrict__ dY,
                                                    float *__restrict__ part) {
  constexpr int TI = KI / 16, TO = MO / 16;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int lm = lane % 16, lk = 4 * (lane / 16);
  const long rb = (long)blockIdx.x * WG_ROWS;
  float *out = part + (size_t)blockIdx.x * KI * MO;
  for (int t = wave; t < TI * TO; t += 4) {
    const int ti = t / TO, to = t % TO;
    floatx4 acc;
    for (int r = 0; r < 4; r++) acc[r] = 0.f;
    for (int kb = 0; kb < WG_ROWS; kb += 16) {
      half4 a, b;
      for (int q = 0; q < 4; q++) {
        const long row = rb + kb + lk + q;
        const bool in = row < N;
        a[q] = (_Float16)(in ? X[row * KI + 16 * ti + lm] : 0.f);
        b[q] = (_Float16)(in ? dY[row * MO + 16 * to + lm] : 0.f);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc, 0, 0, 0);
    }
    for (int r = 0; r < 4; r++) out[(16 * ti + lk + r) * MO + 16 * to + lm] = acc[r];
  }
}

// out[e] (+)= sum_b part[b][e], fixed order in b
__global__ void reduce_parts_kernel(int nb, int n, const float *__restrict__ part,
                                    float *__restrict__ out, int accumulate) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int b = 0; b < nb; b++) s += part[(size_t)b * n + e];
  out[e] = accumulate ? out[e] + s : s;
}

// column sums: part[blk][col] over the block's rows
template <int M>
__global__ __launch_bounds__(256) void colsum_kernel(long N, const float *__restrict__ dY,
                                                     float *__restrict__ part) {
  __shared__ float sh[256];
  const int col = threadIdx.x % M, rsub = threadIdx.x / M, rstep = 256 / M;
  const long rb = (long)blockIdx.x * WG_ROWS;
  float s = 0.f;
  for (long row = rb + rsub; row < rb + WG_ROWS && row < N; row += rstep) s += dY[row * M + col];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < M) {
    float t = 0.f;
    for (int k = 0; k < rstep; k++) t += sh[k * M + threadIdx.x];
    part[(size_t)blockIdx.x * M + threadIdx.x] = t;
  }
}

// agg[i] = (1/deg_i) sum_{j in N(i)} h[j]   (16 lanes per node, 4 features each)
__global__ __launch_bounds__(256) void agg_fwd_kernel(int N, const long long *__restrict__ bptr,
                                                      const int *__restrict__ bcol,
                                                      const float *__restrict__ h,
                                                      float *__restrict__ agg) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t / 16), f = 4 * (int)(t % 16);
  if (i >= N) return;
  const long long k0 = bptr[i], k1 = bptr[i + 1];
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (long long k = k0; k < k1; k++) {
    const float *x = h + (long)bcol[k] * GH + f;
    s0 += x[0];
    s1 += x[1];
    s2 += x[2];
    s3 += x[3];
  }
  const float inv = 1.f / (float)(k1 - k0);
  float *o = agg + (long)i * GH + f;
  o[0] = s0 * inv;
  o[1] = s1 * inv;
  o[2] = s2 * inv;
  o[3] = s3 * inv;
}

// dh[j] += sum_{i in N(j)} dagg[i] / deg_i  (the graph is symmetric)
__global__ __launch_bounds__(256) void agg_bwd_kernel(int N, const long long *__restrict__ bptr,
                                                      const int *__restrict__ bcol,
                                                      const float *__restrict__ dagg,
                                                      float *__restrict__ dh) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (int)(t / 16), f = 4 * (int)(t % 16);
  if (j >= N) return;
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (long long k = bptr[j]; k < bptr[j + 1]; k++) {
    const int i = bcol[k];
    const float inv = 1.f / (float)(bptr[i + 1] - bptr[i]);
    const float *x = dagg + (long)i * GH + f;
    s0 += x[0] * inv;
    s1 += x[1] * inv;
    s2 += x[2] * inv;
    s3 += x[3] * inv;
  }
  float *o = dh + (long)j * GH + f;
  o[0] += s0;
  o[1] += s1;
  o[2] += s2;
  o[3] += s3;
}

// dz = dy * (a > 0)
__global__ void relu_bwd_kernel(long n, const float *__restrict__ dy, const float *__restrict__ a,
                                float *__restrict__ dz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = a[i] > 0.f ? dy[i] : 0.f;
}

__global__ void add_kernel(long n, float *__restrict__ y, const float *__restrict__ x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

// features of node i for state u (nodal FP64), lam, dlam
__global__ void features_kernel(int N, const double *__restrict__ coords,
                                const double *__restrict__ u, const unsigned char *__restrict__ fmask,
                                const unsigned char *__restrict__ lmask,
                                const float *__restrict__ box, double u_scale, double lam_scale,
                                double lam, double dlam, float *__restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float *x = X + (long)i * GF;
  for (int d = 0; d < 3; d++) x[d] = (float)((coords[3 * i + d] - box[d]) * box[3 + d]);
  for (int d = 0; d < 3; d++) x[3 + d] = (float)(u[3 * i + d] / u_scale);
  x[6] = (float)(lam / lam_scale);
  x[7] = (float)(dlam / lam_scale);
  const unsigned fm = fmask[i], lm = lmask[i];
  for (int d = 0; d < 3; d++) x[8 + d] = (fm >> d) & 1u ? 1.f : 0.f;
  for (int d = 0; d < 3; d++) x[11 + d] = (lm >> d) & 1u ? 1.f : 0.f;
  x[14] = 1.f;
  x[15] = 0.f;
}

// target T = (u_tgt - u) / (dlam out_scale) on free directions, 0 elsewhere
__global__ void target_kernel(int N, const double *__restrict__ u, const double *__restrict__ ut,
                              const unsigned char *__restrict__ fmask, double scale,
                              float *__restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < GF; c++) {
    float v = 0.f;
    if (c < 3 && ((fmask[i] >> c) & 1u)) v = (float)((ut[3 * i + c] - u[3 * i + c]) * scale);
    T[(long)i * GF + c] = v;
  }
}

// dY = 2 (Y - T) / count on free directions; partial loss sums per block
__global__ __launch_bounds__(256) void loss_kernel(int N, const float *__restrict__ Y,
                                                   const float *__restrict__ T,
                                                   const unsigned char *__restrict__ fmask,
                                                   double inv_count, float *__restrict__ dY,
                                                   double *__restrict__ part) {
  __shared__ double sh[16];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < GF; c++) {
      float g = 0.f;
      if (c < 3 && ((fmask[i] >> c) & 1u)) {
        const double e = (double)Y[i * GF + c] - (double)T[i * GF + c];
        s += e * e;
        // d(loss)/dy times count: keeps FP16 backward operands O(1); the
        // count is divided out when gradients are reported (Adam is
        // invariant to the scale)
        g = (float)(2.0 * e);
      }
      dY[i * GF + c] = g;
    }
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void loss_finish_kernel(const double *__restrict__ part, double inv_count,
                                   double *__restrict__ out) {
  __shared__ double sh[16];
  double s = reduce_partials(part, RED, sh);
  if (threadIdx.x == 0) *out = s * inv_count;
}

// du = dlam out_scale y (nodal FP64), zero on fixed directions
__global__ void output_kernel(int N, const float *__restrict__ Y,
                              const unsigned char *__restrict__ fmask, double scale,
                              double *__restrict__ du) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < 3; c++)
    du[3 * i + c] = ((fmask[i] >> c) & 1u) ? scale * (double)Y[(long)i * GF + c] : 0.0;
}

__global__ void adam_kernel(int n, float *__restrict__ P, const float *__restrict__ G,
                            float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------- host side
inline unsigned grid_of(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// copies and fills on the context stream (created non-blocking: the legacy
// default stream does not order against it), completed before returning
void copy_sync(Ctx &c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  FGPU_CHECK(hipMemcpyAsync(dst, src, bytes, kind, c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
}

void fill_sync(Ctx &c, void *dst, int v, size_t bytes) {
  FGPU_CHECK(hipMemsetAsync(dst, v, bytes, c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
}

template <int M, bool RELU>
void rowgemm(Gnn &g, const float *X1, int K1, Bop B1, const float *X2, int K2, Bop B2,
             const float *bias, const float *R, float *Y, float *Yres) {
  hipLaunchKernelGGL((rowgemm_kernel<M, RELU>), dim3(grid_of(g.N, ROWS_PER_BLOCK)), dim3(256), 0,
                     g.c->stream, (long)g.N, X1, K1, B1, X2, K2, B2, bias, R, Y, Yres);
}

Bop fwdW(const float *W, int out) { return Bop{W, out, 1}; }  // B[k][n] = W[k][n]
Bop bwdW(const float *W, int out) { return Bop{W, 1, out}; }  // B[k][n] = W[n][k]

template <int KI, int MO>
void wgrad(Gnn &g, const float *X, const float *dY, float *dW) {
  hipLaunchKernelGGL((wgrad_kernel<KI, MO>), dim3(g.nblk_rows), dim3(256), 0, g.c->stream,
                     (long)g.N, X, dY, g.part);
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(grid_of(KI * MO)), dim3(256), 0, g.c->stream,
                     g.nblk_rows, KI * MO, g.part, dW, 0);
}

template <int M>
void bgrad(Gnn &g, const float *dY, float *db) {
  hipLaunchKernelGGL((colsum_kernel<M>), dim3(g.nblk_rows), dim3(256), 0, g.c->stream, (long)g.N,
                     dY, g.part);
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(64), 0, g.c->stream, g.nblk_rows, M,
                     g.part, db, 0);
}

void agg_fwd(Gnn &g, const float *h, float *agg) {
  hipLaunchKernelGGL(agg_fwd_kernel, dim3(grid_of((long)g.N * 16)), dim3(256), 0, g.c->stream,
                     g.N, g.c->d_bptr, g.c->d_bcol, h, agg);
}

void agg_bwd(Gnn &g, const float *dagg, float *dh) {
  hipLaunchKernelGGL(agg_bwd_kernel, dim3(grid_of((long)g.N * 16)), dim3(256), 0, g.c->stream,
                     g.N, g.c->d_bptr, g.c->d_bcol, dagg, dh);
}

}  // namespace

void gnn_forward(Gnn &g) {
  const GnnLayout &L = g.lay;
  const float *P = g.P;
  rowgemm<GH, true>(g, g.X, GF, fwdW(P + L.w_in, GH), nullptr, 0, Bop{}, P + L.b_in, nullptr,
                    g.H[0], nullptr);
  for (int l = 0; l < GL; l++) {
    agg_fwd(g, g.H[l], g.AG[l]);
    rowgemm<GH, true>(g, g.H[l], GH, fwdW(P + L.ws[l], GH), g.AG[l], GH, fwdW(P + L.wn[l], GH),
                      P + L.bl[l], g.H[l], g.A[l], g.H[l + 1]);
  }
  rowgemm<GH, true>(g, g.H[GL], GH, fwdW(P + L.w_d1, GH), nullptr, 0, Bop{}, P + L.b_d1, nullptr,
                    g.D1, nullptr);
  rowgemm<GF, false>(g, g.D1, GH, fwdW(P + L.w_d2, GF), nullptr, 0, Bop{}, P + L.b_d2, nullptr,
                     g.Y, nullptr);
  FGPU_CHECK(hipGetLastError());
}

// gradient of the loss with respect to every parameter, from g.dY
void gnn_backward(Gnn &g) {
  const GnnLayout &L = g.lay;
  const float *P = g.P;
  float *G = g.G;
  const long NH = (long)g.N * GH;
  hipStream_t s = g.c->stream;
  wgrad<GH, GF>(g, g.D1, g.dY, G + L.w_d2);
  bgrad<GF>(g, g.dY, G + L.b_d2);
  // dD1 = (dY W_d2^T) * (D1 > 0)
  rowgemm<GH, false>(g, g.dY, GF, bwdW(P + L.w_d2, GF), nullptr, 0, Bop{}, nullptr, nullptr,
                     g.tmp, nullptr);
  hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_of(NH)), dim3(256), 0, s, NH, g.tmp, g.D1, g.dD1);
  wgrad<GH, GH>(g, g.H[GL], g.dD1, G + L.w_d1);
  bgrad<GH>(g, g.dD1, G + L.b_d1);
  rowgemm<GH, false>(g, g.dD1, GH, bwdW(P + L.w_d1, GH), nullptr, 0, Bop{}, nullptr, nullptr,
                     g.dH, nullptr);
  for (int l = GL - 1; l >= 0; l--) {
    // h_{l+1} = h_l + a_l ; a_l = relu(z_l) ; z_l = h_l Ws + agg_l Wn + b
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_of(NH)), dim3(256), 0, s, NH, g.dH, g.A[l], g.dZ);
    wgrad<GH, GH>(g, g.H[l], g.dZ, G + L.ws[l]);
    wgrad<GH, GH>(g, g.AG[l], g.dZ, G + L.wn[l]);
    bgrad<GH>(g, g.dZ, G + L.bl[l]);
    // dh_l = dh_{l+1} + dZ Ws^T + agg^T(dZ Wn^T)
    rowgemm<GH, false>(g, g.dZ, GH, bwdW(P + L.wn[l], GH), nullptr, 0, Bop{}, nullptr, nullptr,
                       g.tmp, nullptr);
    rowgemm<GH, false>(g, g.dZ, GH, bwdW(P + L.ws[l], GH), nullptr, 0, Bop{}, nullptr, g.dH,
                       g.tmp2, g.dH);
    agg_bwd(g, g.tmp, g.dH);
  }
  // h0 = relu(X W_in + b_in)
  hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_of(NH)), dim3(256), 0, s, NH, g.dH, g.H[0], g.dZ);
  wgrad<GF, GH>(g, g.X, g.dZ, G + L.w_in);
  bgrad<GH>(g, g.dZ, G + L.b_in);
  FGPU_CHECK(hipGetLastError());
}

double gnn_loss(Gnn &g) {
  const double inv = 1.0 / (double)std::max(1, g.c->neq);  // free directions
  hipLaunchKernelGGL(loss_kernel, dim3(RED), dim3(256), 0, g.c->stream, g.N, g.Y, g.T,
                     g.c->d_fmask, inv, g.dY, g.c->d_part);
  hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(256), 0, g.c->stream, g.c->d_part, inv,
                     g.dbuf);
  double h = 0;
  FGPU_CHECK(hipMemcpyAsync(&h, g.dbuf, sizeof(double), hipMemcpyDeviceToHost, g.c->stream));
  FGPU_CHECK(hipStreamSynchronize(g.c->stream));
  return h;
}

void gnn_features(Gnn &g, const double *u, double lam, double dlam, const unsigned char *lmask) {
  hipLaunchKernelGGL(features_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N,
                     g.c->d_coords, u, g.c->d_fmask, lmask, g.box, g.u_scale, g.lam_scale, lam,
                     dlam, g.X);
}

void gnn_target(Gnn &g, const double *u, const double *ut, double dlam) {
  hipLaunchKernelGGL(target_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N, u, ut,
                     g.c->d_fmask, 1.0 / (dlam * g.out_scale), g.T);
}

void gnn_output(Gnn &g, double dlam, double *du) {
  hipLaunchKernelGGL(output_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N, g.Y,
                     g.c->d_fmask, dlam * g.out_scale, du);
}

void gnn_adam(Gnn &g, float lr) {
  g.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)g.step), c2 = 1.f - powf(b2, (float)g.step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_of(g.lay.total)), dim3(256), 0, g.c->stream,
                     g.lay.total, g.P, g.G, g.Mo, g.Vo, lr, b1, b2, 1e-8f, c1, c2);
}

void gnn_alloc(Gnn &g, Ctx &c, unsigned seed) {
  g.c = &c;
  g.N = c.nnode;
  g.seed = seed;
  g.nblk_rows = (int)((g.N + WG_ROWS - 1) / WG_ROWS);
  const size_t NH = (size_t)g.N * GH, NF = (size_t)g.N * GF;
  auto alloc = [](size_t n) {
    float *p = nullptr;
    FGPU_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(float)));
    return p;
  };
  g.P = alloc(g.lay.total);
  g.G = alloc(g.lay.total);
  g.Mo = alloc(g.lay.total);
  g.Vo = alloc(g.lay.total);
  g.X = alloc(NF);
  for (int l = 0; l <= GL; l++) g.H[l] = alloc(NH);
  for (int l = 0; l < GL; l++) {
    g.A[l] = alloc(NH);
    g.AG[l] = alloc(NH);
  }
  g.D1 = alloc(NH);
  g.Y = alloc(NF);
  g.T = alloc(NF);
  g.dY = alloc(NF);
  g.dD1 = alloc(NH);
  g.dH = alloc(NH);
  g.dZ = alloc(NH);
  g.tmp = alloc(NH);
  g.tmp2 = alloc(NH);
  g.part = alloc((size_t)g.nblk_rows * GH * GH);
  g.box = alloc(8);
  FGPU_CHECK(hipMalloc(&g.dbuf, 8 * sizeof(double)));
  // Glorot-uniform initialisation (host RNG, seeded), zero biases
  std::vector<float> P(g.lay.total, 0.f);
  std::mt19937 rng(seed);
  auto init = [&](int off, int fin, int fout) {
    const float a = std::sqrt(6.f / (float)(fin + fout));
    std::uniform_real_distribution<float> U(-a, a);
    for (int k = 0; k < fin * fout; k++) P[off + k] = U(rng);
  };
  const GnnLayout &L = g.lay;
  init(L.w_in, GF, GH);
  for (int l = 0; l < GL; l++) {
    init(L.ws[l], GH, GH);
    init(L.wn[l], GH, GH);
    // residual branches start small so the stack starts near identity
    for (int k = 0; k < GH * GH; k++) {
      P[L.ws[l] + k] *= 0.5f;
      P[L.wn[l] + k] *= 0.5f;
    }
  }
  init(L.w_d1, GH, GH);
  init(L.w_d2, GH, GF);
  for (int r = 0; r < GH; r++)  // only the first 3 output columns are used
    for (int k = 3; k < GF; k++) P[L.w_d2 + r * GF + k] = 0.f;
  copy_sync(c, g.P, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice);
  fill_sync(c, g.Mo, 0, P.size() * sizeof(float));
  fill_sync(c, g.Vo, 0, P.size() * sizeof(float));
  fill_sync(c, g.G, 0, P.size() * sizeof(float));
  // bounding box of the coordinates
  std::vector<double> X(3 * (size_t)g.N);
  copy_sync(c, X.data(), c.d_coords, X.size() * sizeof(double), hipMemcpyDeviceToHost);
  float box[8] = {0, 0, 0, 1, 1, 1, 0, 0};
  for (int d = 0; d < 3; d++) {
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < g.N; i++) {
      lo = std::min(lo, X[3 * i + d]);
      hi = std::max(hi, X[3 * i + d]);
    }
    box[d] = (float)lo;
    box[3 + d] = (float)(hi > lo ? 1.0 / (hi - lo) : 1.0);
  }
  copy_sync(c, g.box, box, sizeof(box), hipMemcpyHostToDevice);
}

void gnn_free(Gnn &g) {
  float *bufs[] = {g.P, g.G, g.Mo, g.Vo, g.X, g.D1, g.Y, g.T, g.dY, g.dD1, g.dH, g.dZ, g.tmp,
                   g.tmp2, g.part, g.box};
  for (float *p : bufs)
    if (p) (void)hipFree(p);
  for (int l = 0; l <= GL; l++)
    if (g.H[l]) (void)hipFree(g.H[l]);
  for (int l = 0; l < GL; l++) {
    if (g.A[l]) (void)hipFree(g.A[l]);
    if (g.AG[l]) (void)hipFree(g.AG[l]);
  }
  if (g.dbuf) (void)hipFree(g.dbuf);
}

// ---------------------------------------------------------------- C-ABI
namespace {

struct GnnHandle {
  Gnn g;
  unsigned char *lmask = nullptr;  // load flags per node
  double *un = nullptr, *utn = nullptr, *dun = nullptr;  // nodal scratch
};

template <class F>
int gguard(fgpu_gnn *h, F f) {
  GnnHandle *H = reinterpret_cast<GnnHandle *>(h);
  if (!H) return 0;
  try {
    return f(*H);
  } catch (const Error &e) {
    H->g.c->err = e.msg;
  }
  return 0;
}

void to_nodal(GnnHandle &H, const double *u_eq, double *u_nodal) {
  if (u_eq)
    launch_eq_to_nodal(*H.g.c, u_eq, u_nodal);
  else
    FGPU_CHECK(hipMemsetAsync(u_nodal, 0, sizeof(double) * 3 * H.g.N, H.g.c->stream));
}

double sample_loss_grad(GnnHandle &H, const double *u_eq, double lam, double dlam,
                        const double *ut_eq, bool grad) {
  Gnn &g = H.g;
  to_nodal(H, u_eq, H.un);
  to_nodal(H, ut_eq, H.utn);
  gnn_features(g, H.un, lam, dlam, H.lmask);
  gnn_target(g, H.un, H.utn, dlam);
  gnn_forward(g);
  double loss = gnn_loss(g);
  if (grad) gnn_backward(g);
  return loss;
}

}  // namespace
}  // namespace fgpu

using namespace fgpu;

extern "C" {

int fgpu_gnn_create(fgpu_gnn **out, fgpu_ctx *ctx, const unsigned char *load_mask,
                    unsigned seed) {
  if (!out || !ctx) return 0;
  *out = nullptr;
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  GnnHandle *H = new GnnHandle();
  try {
    gnn_alloc(H->g, c, seed);
    FGPU_CHECK(hipMalloc(&H->lmask, std::max(1, c.nnode)));
    if (load_mask)
      copy_sync(c, H->lmask, load_mask, c.nnode, hipMemcpyHostToDevice);
    else
      fill_sync(c, H->lmask, 0, c.nnode);
    for (double **p : {&H->un, &H->utn, &H->dun})
      FGPU_CHECK(hipMalloc(p, sizeof(double) * 3 * std::max(1, c.nnode)));
  } catch (const Error &e) {
    c.err = e.msg;
    gnn_free(H->g);
    delete H;
    return 0;
  }
  *out = reinterpret_cast<fgpu_gnn *>(H);
  return 1;
}

void fgpu_gnn_destroy(fgpu_gnn *h) {
  GnnHandle *H = reinterpret_cast<GnnHandle *>(h);
  if (!H) return;
  gnn_free(H->g);
  for (void *p : {(void *)H->lmask, (void *)H->un, (void *)H->utn, (void *)H->dun})
    if (p) (void)hipFree(p);
  delete H;
}

int fgpu_gnn_num_params(fgpu_gnn *h) {
  return h ? reinterpret_cast<GnnHandle *>(h)->g.lay.total : 0;
}

int fgpu_gnn_set_scales(fgpu_gnn *h, double u_scale, double lam_scale, double out_scale) {
  return gguard(h, [&](GnnHandle &H) {
    if (!(u_scale > 0 && lam_scale > 0 && out_scale > 0)) throw Error("scales must be > 0");
    H.g.u_scale = u_scale;
    H.g.lam_scale = lam_scale;
    H.g.out_scale = out_scale;
    return 1;
  });
}

int fgpu_gnn_get_params(fgpu_gnn *h, float *p) {
  return gguard(h, [&](GnnHandle &H) {
    copy_sync(*H.g.c, p, H.g.P, sizeof(float) * H.g.lay.total, hipMemcpyDeviceToHost);
    return 1;
  });
}

int fgpu_gnn_set_params(fgpu_gnn *h, const float *p) {
  return gguard(h, [&](GnnHandle &H) {
    copy_sync(*H.g.c, H.g.P, p, sizeof(float) * H.g.lay.total, hipMemcpyHostToDevice);
    return 1;
  });
}

int fgpu_gnn_loss_grad(fgpu_gnn *h, const double *u_dev, double lam, double dlam,
                       const double *ut_dev, double *loss, float *grad) {
  return gguard(h, [&](GnnHandle &H) {
    double l = sample_loss_grad(H, u_dev, lam, dlam, ut_dev, grad != nullptr);
    if (loss) *loss = l;
    if (grad) {
      copy_sync(*H.g.c, grad, H.g.G, sizeof(float) * H.g.lay.total, hipMemcpyDeviceToHost);
      const float inv = 1.f / (float)std::max(1, H.g.c->neq);
      for (int k = 0; k < H.g.lay.total; k++) grad[k] *= inv;
    }
    return 1;
  });
}

int fgpu_gnn_predict(fgpu_gnn *h, const double *u_dev, double lam, double dlam, double *du_dev) {
  return gguard(h, [&](GnnHandle &H) {
    Gnn &g = H.g;
    to_nodal(H, u_dev, H.un);
    gnn_features(g, H.un, lam, dlam, H.lmask);
    gnn_forward(g);
    gnn_output(g, dlam, H.dun);
    launch_nodal_to_eq(*g.c, H.dun, du_dev);
    FGPU_CHECK(hipStreamSynchronize(g.c->stream));
    return 1;
  });
}

int fgpu_gnn_train(fgpu_gnn *h, int nstate, const double *states_dev, const double *lams,
                   int nsteps, int maxgap, float lr, unsigned seed, float *loss_hist) {
  return gguard(h, [&](GnnHandle &H) {
    if (nstate < 2) throw Error("need at least two states");
    Gnn &g = H.g;
    const size_t n = (size_t)g.c->neq;
    std::mt19937 rng(seed);
    maxgap = std::max(1, std::min(maxgap, nstate - 1));
    for (int s = 0; s < nsteps; s++) {
      // a pair i < j with a log-uniform gap j - i in [1, maxgap]; state
      // i = -1 is the unloaded state
      std::uniform_int_distribution<int> Ui(-1, nstate - 2);
      std::uniform_real_distribution<double> Ug(0.0, std::log((double)maxgap + 1.0));
      const int i = Ui(rng);
      const int gap = std::max(1, std::min(maxgap, (int)std::exp(Ug(rng))));
      const int j = std::min(nstate - 1, i + gap);
      const double li = i < 0 ? 0.0 : lams[i], lj = lams[j];
      const double *ui = i < 0 ? nullptr : states_dev + (size_t)i * n;
      double l = sample_loss_grad(H, ui, li, lj - li, states_dev + (size_t)j * n, true);
      // cosine-decayed learning rate
      const float f = 0.5f * (1.f + cosf(3.14159265f * (float)s / (float)nsteps));
      gnn_adam(g, lr * (0.05f + 0.95f * f));
      if (loss_hist) loss_hist[s] = (float)l;
    }
    FGPU_CHECK(hipStreamSynchronize(g.c->stream));
    return 1;
  });
}

// fgpu_predict_fn for fgpu_newton: user = the fgpu_gnn handle
int fgpu_gnn_newton_predictor(void *user, int step, const double *u_dev, double lam_prev,
                              double lam, double *du_dev) {
  (void)step;
  return fgpu_gnn_predict(reinterpret_cast<fgpu_gnn *>(user), u_dev, lam_prev, lam - lam_prev,
                          du_dev);
}

}  // extern "C"
