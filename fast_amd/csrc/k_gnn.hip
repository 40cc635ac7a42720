// k_gnn.hip - mesh graph network for warm-starting the Newton/PCG solves
// (no counterpart in the reference): forward, backward and Adam, all
// hand-written for gfx950.  Dense layers run on v_mfma_f32_16x16x16f16
// (FP16 operands converted on load, FP32 accumulation, FP32 master weights).
//
// Graph: the mesh nodes; neighbours = nodes sharing an element (the node-block
// rows of the global matrix, self included).  A sample is a converged state u
// at load factor lambda, an earlier converged state u_h at lambda_h (the
// unloaded state when there is none) and the step dlambda.  Per node features
// (32, the last 14 zero):
//   0-2   x, y, z (box-normalised)       3-5  u / u_scale
//   6-8   s = u / (lambda out_scale), the slope from the unloaded state
//   9-11  v = (u - u_h) / ((lambda - lambda_h) out_scale), the slope of the
//         last step (the secant predictor; s when there is no history)
//   12    lambda / lam_scale   13  dlambda / lam_scale
//   14    (lambda - lambda_h) / lam_scale
//   15    free-direction bit mask / 7    16  load bit mask / 7    17  1
// (s, v are 0 at lambda = 0.)  Network (H = 64, L = 4):
//   h0 = relu(X W_in + b_in)
//   h_{l+1} = h_l + relu(h_l W_s,l + mean_N(h_l) W_n,l + b_l)
//   d = relu(h_L W_d1 + b_d1);  y = M M (d W_d2 + b_d2)   (first 3 columns,
//   output layer in FP32; M = mean over the neighbours: the correction is
//   smoothed twice after the last rounding to FP16)
//   du = dlambda * out_scale * (v + y)   (zero on fixed directions)
// so the network learns the correction of the secant predictor.  Loss: the
// squared error e of v + y against (u_target - u) / (dlambda out_scale) over
// free directions, plus beta times the squared roughness |e - M e|^2 over all
// nodes, divided by neq.  Every reduction (weight gradients, bias gradients,
// loss) is a block-partial sum added in a fixed order.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "fgpu_dev.h"

namespace fgpu {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GH = 64;   // hidden width
constexpr int GFI = 32;  // input features per node
constexpr int GF = 16;   // padded output width
constexpr int GL = 4;    // message-passing layers
constexpr int GS = 2;    // smoothing passes (mean aggregation) of the decoder

// parameter offsets in the flat FP32 vector
struct GnnLayout {
  int w_in, b_in, ws[GL], wn[GL], bl[GL], w_d1, b_d1, w_d2, b_d2, total;
  GnnLayout() {
    int o = 0;
    w_in = o; o += GFI * GH;
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
  float *X = nullptr;     // [N x 32] features
  float *H[GL + 1] = {};  // [N x 64] h_l
  float *A[GL] = {};      // [N x 64] relu outputs of the layers
  float *AG[GL] = {};     // [N x 64] mean aggregations
  float *H0 = nullptr;    // relu(X W_in + b)  (== H[0])
  float *D1 = nullptr;    // [N x 64] decoder hidden layer
  float *YS[2] = {};      // [N x 16] output before / between smoothing passes
  float *Y = nullptr;     // [N x 16]
  float *T = nullptr;     // [N x 16] target (scaled)
  float *dY = nullptr, *dD1 = nullptr, *dH = nullptr, *dZ = nullptr, *tmp = nullptr,
        *tmp2 = nullptr;  // [N x 64] work
  float *part = nullptr;  // partial sums
  float *box = nullptr;   // [6] bounding box (min, 1/extent)
  double *dbuf = nullptr; // loss scalar
  double u_scale = 1.0, lam_scale = 1.0, out_scale = 1.0;
  double beta = 0.0;      // loss weight of the roughness of the error
  float *E = nullptr, *ME = nullptr;  // [N x 16] error, its aggregation
  int nblk_rows = 0;
  int step = 0;
  unsigned seed = 0;
};

namespace {

constexpr int ROWS_PER_BLOCK = 64;   // rowgemm: 4 waves x 16 rows
constexpr int WG_ROWS = 256;         // rows per weight-gradient partial

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

// agg[i] = (1/deg_i) sum_{j in N(i)} h[j]   (W/4 lanes per node, 4 features each)
template <int W>
__global__ __launch_bounds__(256) void agg_fwd_kernel(int N, const long long *__restrict__ bptr,
                                                      const int *__restrict__ bcol,
                                                      const float *__restrict__ h,
                                                      float *__restrict__ agg) {
  constexpr int LPN = W / 4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t / LPN), f = 4 * (int)(t % LPN);
  if (i >= N) return;
  const long long k0 = bptr[i], k1 = bptr[i + 1];
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (long long k = k0; k < k1; k++) {
    const float *x = h + (long)bcol[k] * W + f;
    s0 += x[0];
    s1 += x[1];
    s2 += x[2];
    s3 += x[3];
  }
  const float inv = 1.f / (float)(k1 - k0);
  float *o = agg + (long)i * W + f;
  o[0] = s0 * inv;
  o[1] = s1 * inv;
  o[2] = s2 * inv;
  o[3] = s3 * inv;
}

// dh[j] (+)= sum_{i in N(j)} dagg[i] / deg_i  (the graph is symmetric)
template <int W>
__global__ __launch_bounds__(256) void agg_bwd_kernel(int N, const long long *__restrict__ bptr,
                                                      const int *__restrict__ bcol,
                                                      const float *__restrict__ dagg,
                                                      float *__restrict__ dh, int accumulate) {
  constexpr int LPN = W / 4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (int)(t / LPN), f = 4 * (int)(t % LPN);
  if (j >= N) return;
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (long long k = bptr[j]; k < bptr[j + 1]; k++) {
    const int i = bcol[k];
    const float inv = 1.f / (float)(bptr[i + 1] - bptr[i]);
    const float *x = dagg + (long)i * W + f;
    s0 += x[0] * inv;
    s1 += x[1] * inv;
    s2 += x[2] * inv;
    s3 += x[3] * inv;
  }
  float *o = dh + (long)j * W + f;
  if (accumulate) {
    o[0] += s0;
    o[1] += s1;
    o[2] += s2;
    o[3] += s3;
  } else {
    o[0] = s0;
    o[1] = s1;
    o[2] = s2;
    o[3] = s3;
  }
}

// output layer in FP32: y[i][c] = sum_k d[i][k] W[k][c] + b[c] for the three
// used columns (0 elsewhere).  It runs on the vector ALUs in FP32 so that no
// FP16 rounding enters the prediction after its last smoothing input.
__global__ __launch_bounds__(256) void head_kernel(int N, const float *__restrict__ D,
                                                   const float *__restrict__ W,
                                                   const float *__restrict__ b,
                                                   float *__restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float *d = D + (long)i * GH;
  float y0 = b[0], y1 = b[1], y2 = b[2];
  for (int k = 0; k < GH; k++) {
    const float dk = d[k];
    y0 = fmaf(dk, W[k * GF + 0], y0);
    y1 = fmaf(dk, W[k * GF + 1], y1);
    y2 = fmaf(dk, W[k * GF + 2], y2);
  }
  float *o = Y + (long)i * GF;
  o[0] = y0;
  o[1] = y1;
  o[2] = y2;
  for (int c = 3; c < GF; c++) o[c] = 0.f;
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

constexpr int SKIP = 9;  // feature offset of the skip term v

// features of node i (nodal FP64 states; uh == nullptr: the unloaded state)
__global__ void features_kernel(int N, const double *__restrict__ coords,
                                const double *__restrict__ u, const double *__restrict__ uh,
                                const unsigned char *__restrict__ fmask,
                                const unsigned char *__restrict__ lmask,
                                const float *__restrict__ box, double u_scale, double lam_scale,
                                double out_scale, double lam, double lamh, double dlam,
                                float *__restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float *x = X + (long)i * GFI;
  const double ss = lam != 0.0 ? 1.0 / (lam * out_scale) : 0.0;
  const double sv = lam > lamh ? 1.0 / ((lam - lamh) * out_scale) : 0.0;
  for (int d = 0; d < 3; d++) {
    const double ui = u[3 * i + d], uhi = uh ? uh[3 * i + d] : 0.0;
    x[d] = (float)((coords[3 * i + d] - box[d]) * box[3 + d]);
    x[3 + d] = (float)(ui / u_scale);
    x[6 + d] = (float)(ui * ss);
    x[SKIP + d] = (float)((ui - uhi) * sv);
  }
  x[12] = (float)(lam / lam_scale);
  x[13] = (float)(dlam / lam_scale);
  x[14] = (float)((lam - lamh) / lam_scale);
  x[15] = (float)fmask[i] / 7.f;
  x[16] = (float)lmask[i] / 7.f;
  x[17] = 1.f;
  for (int k = 18; k < GFI; k++) x[k] = 0.f;
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

// e = v + y - T on free directions (0 elsewhere, also stored in E);
// dY = 2 e (d loss / dy times the count); partial loss sums per block
__global__ __launch_bounds__(256) void loss_kernel(int N, const float *__restrict__ Y,
                                                   const float *__restrict__ X,
                                                   const float *__restrict__ T,
                                                   const unsigned char *__restrict__ fmask,
                                                   float *__restrict__ dY, float *__restrict__ E,
                                                   double *__restrict__ part) {
  __shared__ double sh[16];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < GF; c++) {
      float g = 0.f;
      if (c < 3 && ((fmask[i] >> c) & 1u)) {
        const double e =
            (double)Y[i * GF + c] + (double)X[i * GFI + SKIP + c] - (double)T[i * GF + c];
        s += e * e;
        // the count is divided out when gradients are reported: keeps the
        // FP16 backward operands O(1) (Adam is invariant to the scale)
        g = (float)(2.0 * e);
      }
      dY[i * GF + c] = g;
      E[i * GF + c] = 0.5f * g;
    }
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// roughness of the error: D = E - M E (in place over ME); partial sums of D^2
__global__ __launch_bounds__(256) void rough_kernel(int N, const float *__restrict__ E,
                                                    float *__restrict__ ME,
                                                    double *__restrict__ part) {
  __shared__ double sh[16];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; c++) {
      const float d = E[i * GF + c] - ME[i * GF + c];
      ME[i * GF + c] = d;
      s += (double)d * (double)d;
    }
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// dY += 2 beta (D - M^T D) on free directions
__global__ void rough_grad_kernel(int N, const float *__restrict__ D, const float *__restrict__ MTD,
                                  const unsigned char *__restrict__ fmask, float beta,
                                  float *__restrict__ dY) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < 3; c++)
    if ((fmask[i] >> c) & 1u)
      dY[(long)i * GF + c] += 2.f * beta * (D[(long)i * GF + c] - MTD[(long)i * GF + c]);
}

// out = (sum part[0:RED] + beta sum part[RED:2 RED]) * inv_count
__global__ void loss_finish2_kernel(const double *__restrict__ part, double beta,
                                    double inv_count, double *__restrict__ out) {
  __shared__ double sh[16];
  double a = reduce_partials(part, RED, sh);
  double b = reduce_partials(part + RED, RED, sh);
  if (threadIdx.x == 0) *out = (a + beta * b) * inv_count;
}

__global__ void loss_finish_kernel(const double *__restrict__ part, double inv_count,
                                   double *__restrict__ out) {
  __shared__ double sh[16];
  double s = reduce_partials(part, RED, sh);
  if (threadIdx.x == 0) *out = s * inv_count;
}

// du = dlam out_scale (v + y) (nodal FP64), zero on fixed directions
__global__ void output_kernel(int N, const float *__restrict__ Y, const float *__restrict__ X,
                              const unsigned char *__restrict__ fmask, double scale,
                              double *__restrict__ du) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < 3; c++)
    du[3 * i + c] = ((fmask[i] >> c) & 1u)
                        ? scale * ((double)Y[(long)i * GF + c] + (double)X[(long)i * GFI + SKIP + c])
                        : 0.0;
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

template <int W = GH>
void agg_fwd(Gnn &g, const float *h, float *agg) {
  hipLaunchKernelGGL(agg_fwd_kernel<W>, dim3(grid_of((long)g.N * (W / 4))), dim3(256), 0,
                     g.c->stream, g.N, g.c->d_bptr, g.c->d_bcol, h, agg);
}

template <int W = GH>
void agg_bwd(Gnn &g, const float *dagg, float *dh, int accumulate = 1) {
  hipLaunchKernelGGL(agg_bwd_kernel<W>, dim3(grid_of((long)g.N * (W / 4))), dim3(256), 0,
                     g.c->stream, g.N, g.c->d_bptr, g.c->d_bcol, dagg, dh, accumulate);
}

}  // namespace

void gnn_forward(Gnn &g) {
  const GnnLayout &L = g.lay;
  const float *P = g.P;
  rowgemm<GH, true>(g, g.X, GFI, fwdW(P + L.w_in, GH), nullptr, 0, Bop{}, P + L.b_in, nullptr,
                    g.H[0], nullptr);
  for (int l = 0; l < GL; l++) {
    agg_fwd(g, g.H[l], g.AG[l]);
    rowgemm<GH, true>(g, g.H[l], GH, fwdW(P + L.ws[l], GH), g.AG[l], GH, fwdW(P + L.wn[l], GH),
                      P + L.bl[l], g.H[l], g.A[l], g.H[l + 1]);
  }
  rowgemm<GH, true>(g, g.H[GL], GH, fwdW(P + L.w_d1, GH), nullptr, 0, Bop{}, P + L.b_d1, nullptr,
                    g.D1, nullptr);
  // smoothed FP32 output: y = M^GS (d1 W_d2 + b_d2) (M: mean over the neighbours)
  float *out[2] = {g.YS[0], g.YS[1]};
  hipLaunchKernelGGL(head_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N, g.D1,
                     P + L.w_d2, P + L.b_d2, GS ? out[0] : g.Y);
  for (int k = 0; k < GS; k++)
    agg_fwd<GF>(g, out[k % 2], k == GS - 1 ? g.Y : out[(k + 1) % 2]);
  FGPU_CHECK(hipGetLastError());
}

// gradient of the loss with respect to every parameter, from g.dY
void gnn_backward(Gnn &g) {
  const GnnLayout &L = g.lay;
  const float *P = g.P;
  float *G = g.G;
  const long NH = (long)g.N * GH;
  hipStream_t s = g.c->stream;
  // dY0 = (M^T)^GS dY, the gradient at the unsmoothed output
  const float *dy0 = g.dY;
  for (int k = 0; k < GS; k++) {
    agg_bwd<GF>(g, dy0, g.YS[k % 2], 0);
    dy0 = g.YS[k % 2];
  }
  wgrad<GH, GF>(g, g.D1, dy0, G + L.w_d2);
  bgrad<GF>(g, dy0, G + L.b_d2);
  // dD1 = (dY0 W_d2^T) * (D1 > 0)
  rowgemm<GH, false>(g, dy0, GF, bwdW(P + L.w_d2, GF), nullptr, 0, Bop{}, nullptr, nullptr,
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
  wgrad<GFI, GH>(g, g.X, g.dZ, G + L.w_in);
  bgrad<GH>(g, g.dZ, G + L.b_in);
  FGPU_CHECK(hipGetLastError());
}

// loss = (sum_free e^2 + beta sum_nodes |e - M e|^2) / neq: the second term
// weights the local roughness of the error, which drives the residual of a
// predicted state far more than its size
double gnn_loss(Gnn &g) {
  const double inv = 1.0 / (double)std::max(1, g.c->neq);  // free directions
  hipStream_t s = g.c->stream;
  hipLaunchKernelGGL(loss_kernel, dim3(RED), dim3(256), 0, s, g.N, g.Y, g.X, g.T, g.c->d_fmask,
                     g.dY, g.E, g.c->d_part);
  if (g.beta > 0) {
    agg_fwd<GF>(g, g.E, g.ME);
    hipLaunchKernelGGL(rough_kernel, dim3(RED), dim3(256), 0, s, g.N, g.E, g.ME,
                       g.c->d_part + RED);
    agg_bwd<GF>(g, g.ME, g.YS[0], 0);
    hipLaunchKernelGGL(rough_grad_kernel, dim3(grid_of(g.N)), dim3(256), 0, s, g.N, g.ME,
                       g.YS[0], g.c->d_fmask, (float)g.beta, g.dY);
    hipLaunchKernelGGL(loss_finish2_kernel, dim3(1), dim3(256), 0, s, g.c->d_part, g.beta, inv,
                       g.dbuf);
  } else {
    hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(256), 0, s, g.c->d_part, inv, g.dbuf);
  }
  double h = 0;
  FGPU_CHECK(hipMemcpyAsync(&h, g.dbuf, sizeof(double), hipMemcpyDeviceToHost, g.c->stream));
  FGPU_CHECK(hipStreamSynchronize(g.c->stream));
  return h;
}

void gnn_features(Gnn &g, const double *u, double lam, const double *uh, double lamh, double dlam,
                  const unsigned char *lmask) {
  hipLaunchKernelGGL(features_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N,
                     g.c->d_coords, u, uh, g.c->d_fmask, lmask, g.box, g.u_scale, g.lam_scale,
                     g.out_scale, lam, lamh, dlam, g.X);
}

void gnn_target(Gnn &g, const double *u, const double *ut, double dlam) {
  hipLaunchKernelGGL(target_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N, u, ut,
                     g.c->d_fmask, 1.0 / (dlam * g.out_scale), g.T);
}

void gnn_output(Gnn &g, double dlam, double *du) {
  hipLaunchKernelGGL(output_kernel, dim3(grid_of(g.N)), dim3(256), 0, g.c->stream, g.N, g.Y, g.X,
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
  const size_t NH = (size_t)g.N * GH, NF = (size_t)g.N * GF, NX = (size_t)g.N * GFI;
  auto alloc = [](size_t n) {
    float *p = nullptr;
    FGPU_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(float)));
    return p;
  };
  g.P = alloc(g.lay.total);
  g.G = alloc(g.lay.total);
  g.Mo = alloc(g.lay.total);
  g.Vo = alloc(g.lay.total);
  g.X = alloc(NX);
  for (int l = 0; l <= GL; l++) g.H[l] = alloc(NH);
  for (int l = 0; l < GL; l++) {
    g.A[l] = alloc(NH);
    g.AG[l] = alloc(NH);
  }
  g.D1 = alloc(NH);
  for (int k = 0; k < 2; k++) g.YS[k] = alloc(NF);
  g.E = alloc(NF);
  g.ME = alloc(NF);
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
  init(L.w_in, GFI, GH);
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
  float *bufs[] = {g.P, g.G, g.Mo, g.Vo, g.X, g.D1, g.Y, g.T, g.E, g.ME, g.dY, g.dD1, g.dH, g.dZ, g.tmp,
                   g.tmp2, g.part, g.box};
  for (float *p : bufs)
    if (p) (void)hipFree(p);
  for (int l = 0; l <= GL; l++)
    if (g.H[l]) (void)hipFree(g.H[l]);
  for (int k = 0; k < 2; k++)
    if (g.YS[k]) (void)hipFree(g.YS[k]);
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
  double *un = nullptr, *uhn = nullptr, *utn = nullptr, *dun = nullptr;  // nodal scratch
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

// one sample: state u at lam, history uh at lamh (nullptr: the unloaded
// state), target ut at lam + dlam (equation vectors; u == nullptr: zero)
double sample_loss_grad(GnnHandle &H, const double *u_eq, double lam, const double *uh_eq,
                        double lamh, double dlam, const double *ut_eq, bool grad) {
  Gnn &g = H.g;
  to_nodal(H, u_eq, H.un);
  to_nodal(H, ut_eq, H.utn);
  if (uh_eq) to_nodal(H, uh_eq, H.uhn);
  gnn_features(g, H.un, lam, uh_eq ? H.uhn : nullptr, uh_eq ? lamh : 0.0, dlam, H.lmask);
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
    for (double **p : {&H->un, &H->uhn, &H->utn, &H->dun})
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
  for (void *p : {(void *)H->lmask, (void *)H->un, (void *)H->uhn, (void *)H->utn, (void *)H->dun})
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

int fgpu_gnn_set_loss(fgpu_gnn *h, double beta) {
  return gguard(h, [&](GnnHandle &H) {
    if (!(beta >= 0)) throw Error("beta must be >= 0");
    H.g.beta = beta;
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
                       const double *uh_dev, double lam_h, const double *ut_dev, double *loss,
                       float *grad) {
  return gguard(h, [&](GnnHandle &H) {
    double l = sample_loss_grad(H, u_dev, lam, uh_dev, lam_h, dlam, ut_dev, grad != nullptr);
    if (loss) *loss = l;
    if (grad) {
      copy_sync(*H.g.c, grad, H.g.G, sizeof(float) * H.g.lay.total, hipMemcpyDeviceToHost);
      const float inv = 1.f / (float)std::max(1, H.g.c->neq);
      for (int k = 0; k < H.g.lay.total; k++) grad[k] *= inv;
    }
    return 1;
  });
}

int fgpu_gnn_predict(fgpu_gnn *h, const double *u_dev, double lam, double dlam,
                     const double *uh_dev, double lam_h, double *du_dev) {
  return gguard(h, [&](GnnHandle &H) {
    Gnn &g = H.g;
    to_nodal(H, u_dev, H.un);
    if (uh_dev) to_nodal(H, uh_dev, H.uhn);
    gnn_features(g, H.un, lam, uh_dev ? H.uhn : nullptr, uh_dev ? lam_h : 0.0, dlam, H.lmask);
    gnn_forward(g);
    gnn_output(g, dlam, H.dun);
    launch_nodal_to_eq(*g.c, H.dun, du_dev);
    FGPU_CHECK(hipStreamSynchronize(g.c->stream));
    return 1;
  });
}

int fgpu_gnn_train(fgpu_gnn *h, int nstate, const double *states_dev, const double *lams,
                   int nsteps, int maxgap, double origin_frac, float lr, unsigned seed,
                   float *loss_hist) {
  return gguard(h, [&](GnnHandle &H) {
    if (nstate < 2) throw Error("need at least two states");
    Gnn &g = H.g;
    const size_t n = (size_t)g.c->neq;
    std::mt19937 rng(seed);
    maxgap = std::max(1, std::min(maxgap, nstate - 1));
    const char *venv = std::getenv("FGPU_GNN_VERBOSE");  // progress log on stderr
    const bool verbose = venv && venv[0] == '1';
    for (int s = 0; s < nsteps; s++) {
      // a triple h < i < j with log-uniform gaps i - h, j - i in [1, maxgap];
      // index -1 is the unloaded state.  With probability origin_frac the
      // sample is a solve from the unloaded state to a uniformly drawn state
      std::uniform_int_distribution<int> Ui(0, nstate - 2), Uj(0, nstate - 1);
      std::uniform_real_distribution<double> Ug(0.0, std::log((double)maxgap + 1.0));
      std::uniform_real_distribution<double> U01(0.0, 1.0);
      const bool from_origin = U01(rng) < origin_frac;
      const int i = from_origin ? -1 : Ui(rng);
      const int gap = from_origin ? 1 + Uj(rng)
                                  : std::max(1, std::min(maxgap, (int)std::exp(Ug(rng))));
      const int gaph = std::max(1, std::min(maxgap, (int)std::exp(Ug(rng))));
      const int j = std::min(nstate - 1, i + gap);
      const int hh = std::max(-1, i - gaph);
      const double li = i < 0 ? 0.0 : lams[i], lj = lams[j];
      const double *ui = i < 0 ? nullptr : states_dev + (size_t)i * n;
      const double *uh = hh < 0 ? nullptr : states_dev + (size_t)hh * n;
      const double lh = hh < 0 ? 0.0 : lams[hh];
      double l = sample_loss_grad(H, ui, li, i < 0 ? nullptr : uh, lh, lj - li,
                                  states_dev + (size_t)j * n, true);
      // cosine-decayed learning rate
      const float f = 0.5f * (1.f + cosf(3.14159265f * (float)s / (float)nsteps));
      gnn_adam(g, lr * (0.05f + 0.95f * f));
      if (loss_hist) loss_hist[s] = (float)l;
      if (verbose && (s + 1) % 250 == 0)
        std::fprintf(stderr, "fgpu_gnn_train: step %d/%d loss %.3e\n", s + 1, nsteps, l);
    }
    FGPU_CHECK(hipStreamSynchronize(g.c->stream));
    return 1;
  });
}

// fgpu_predict_fn for fgpu_newton: user = the fgpu_gnn handle
int fgpu_gnn_newton_predictor(void *user, int step, const double *u_dev, double lam_prev,
                              double lam, const double *u_hist_dev, double lam_hist,
                              double *du_dev) {
  (void)step;
  return fgpu_gnn_predict(reinterpret_cast<fgpu_gnn *>(user), u_dev, lam_prev, lam - lam_prev,
                          u_hist_dev, lam_hist, du_dev);
}

}  // extern "C"
