// k_direct.hip - direct solve of the assembled equation system on gfx950:
// LDL^T without pivoting, FP64, in band storage.
//
// Reference path: cSymSkylMatrix (mlib/sysmat.cpp:225-277: skyline storage,
// Solve = CroutSolver op 1 on the first call, op 3 afterwards) and
// CroutSolver (mlib/matvec.cpp:609-664: column-by-column LDL^T reduction,
// "singular" when a pivot is below TOLC = 1e-12 in magnitude, then forward
// reduction, diagonal scaling and back substitution).
//
// Storage: the lower triangle inside the half bandwidth bw (the widest row
// profile of the equation matrix) as columns of bw + 1 entries,
// A(i, j) at band[j (bw + 1) + i - j]; the envelope of the reference's
// skyline lies inside it, and so does every fill entry of the factors.
//
// Factorization, right-looking in panels of NB = 32 columns (per panel three
// launches): the diagonal block factorizes in LDS (one workgroup), every row
// of the panel below it solves x L11^T D = a (one thread per row), and the
// trailing band takes the rank-32 update A22 -= L21 D L21^T in 32 x 32 tiles
// (one workgroup per tile, only tiles that meet the band).
//
// Solves, narrow bands (bw < WIDE_BW): one workgroup sweeps the panels; the
// bw + NB live entries of the right-hand side sit in an LDS ring, each
// panel's triangle is solved by one wave through shuffles and the coupling
// to the bw rows below (forward) or above (backward) is a 32-wide product
// over all threads.  Wide bands: one launch per panel over as many
// workgroups as the band has rows of 256 below the panel; every workgroup
// solves the panel's triangle itself (the same operations, so the same
// values; workgroup 0 stores them) and then its own 256 rows: forward, the
// update y_i -= sum_j L_ij y_j; backward, its partial sums of L_ij x_i for
// the next panel up, which the next launch adds in workgroup order.
#include "fgpu_dev.h"

namespace fgpu {
namespace {

constexpr int NB = 32;
constexpr int MAX_BW_SOLVE = 16384;  // LDS ring of the solve: (bw + NB) doubles
constexpr int WIDE_BW = 256;         // from this half bandwidth on, the solves go wide
constexpr int WROWS = 256;           // rows per workgroup of the wide solves

struct Band {
  double *a;
  int n, bw;
  long long ld;  // bw + 1
  __device__ double &at(int i, int j) const { return a[(long long)j * ld + (i - j)]; }
  __device__ bool in(int i, int j) const { return i >= j && i - j <= bw && i < n; }
};

// A panel's unit triangle, loaded up front so that the loads are in flight
// together: lane t < kb gets row k0 + t, Lr[j] = L(k0 + t, k0 + j) for j < t
// (0 elsewhere), or column k0 + t, Lc[j] = L(k0 + j, k0 + t) for t < j < kb.
__device__ inline void panel_row(const Band &B, int k0, int kb, int t, double *Lr) {
#pragma unroll
  for (int j = 0; j < NB; j++)
    Lr[j] = (t < kb && j < t && t - j <= B.bw) ? B.at(k0 + t, k0 + j) : 0.0;
}

__device__ inline void panel_col(const Band &B, int k0, int kb, int t, double *Lc) {
#pragma unroll
  for (int j = 0; j < NB; j++)
    Lc[j] = (t < kb && j > t && j < kb && j - t <= B.bw) ? B.at(k0 + j, k0 + t) : 0.0;
}

// lower triangle of the node-block matrix -> band (fixed directions dropped)
__global__ __launch_bounds__(256) void band_fill_kernel(int nnode, const long long *__restrict__ bptr,
                                                        const int *__restrict__ bcol,
                                                        const double *__restrict__ bval,
                                                        const int *__restrict__ eqmap, Band B) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nnode) return;
  int ea[3];
  for (int i = 0; i < 3; i++) ea[i] = eqmap[3 * A + i];
  for (long long k = bptr[A]; k < bptr[A + 1]; k++) {
    const int C = bcol[k];
    const double *v = bval + 9 * k;
    for (int j = 0; j < 3; j++) {
      const int ej = eqmap[3 * C + j];
      if (ej < 0) continue;
      for (int i = 0; i < 3; i++)
        if (ea[i] >= ej && ea[i] - ej <= B.bw) B.at(ea[i], ej) = v[3 * i + j];
    }
  }
}

// diagonal block k0..k0+kb of the panel: unblocked right-looking LDL^T in LDS
__global__ __launch_bounds__(1024) void band_diag_kernel(Band B, int k0, int kb, double tolc,
                                                         int *__restrict__ info) {
  __shared__ double T[NB][NB + 1];
  const int r = threadIdx.x / NB, c = threadIdx.x % NB;
  const bool live = r < kb && c <= r && r - c <= B.bw;
  T[r][c] = live ? B.at(k0 + r, k0 + c) : 0.0;
  for (int j = 0; j < kb; j++) {
    __syncthreads();
    const double d = T[j][j];
    if (threadIdx.x == 0) {
      if (!(fabs(d) >= tolc)) atomicMin(&info[0], k0 + j + 1);
      if (d < 0.0) atomicAdd(&info[1], 1);
    }
    if (r > j && c > j && c <= r && r < kb) T[r][c] -= T[r][j] * T[c][j] / d;
    __syncthreads();
    if (c == j && r > j && r < kb) T[r][j] /= d;
  }
  __syncthreads();
  if (live) B.at(k0 + r, k0 + c) = T[r][c];
}

// rows k0+kb .. k0+kb+m of the panel's columns: x L11^T D = a
__global__ __launch_bounds__(256) void band_panel_kernel(Band B, int k0, int kb, int m) {
  __shared__ double L[NB][NB + 1], d[NB];
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    L[r][c] = (r < kb && c < r && r - c <= B.bw) ? B.at(k0 + r, k0 + c) : 0.0;
    if (r == c) d[r] = r < kb ? B.at(k0 + r, k0 + r) : 1.0;
  }
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int i = k0 + kb + t;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    double a = (j < kb && B.in(i, k0 + j)) ? B.at(i, k0 + j) : 0.0;
#pragma unroll
    for (int k = 0; k < j; k++) a -= x[k] * d[k] * L[j][k];
    x[j] = j < kb ? a / d[j] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < NB; j++)
    if (j < kb && B.in(i, k0 + j)) B.at(i, k0 + j) = x[j];
}

// trailing update of the rows/columns k0+kb .. k0+kb+m: A -= L21 D L21^T,
// one 32 x 32 tile (ti >= tj) per workgroup
__global__ __launch_bounds__(1024) void band_update_kernel(Band B, int k0, int kb, int m) {
  // tile pair from the linear index: ti (ti + 1) / 2 + tj
  const int q = blockIdx.x;
  int ti = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= q) ti++;
  while (ti * (ti + 1) / 2 > q) ti--;
  const int tj = q - ti * (ti + 1) / 2;
  const int r0 = k0 + kb + ti * NB, c0 = k0 + kb + tj * NB;
  if (r0 - (c0 + NB - 1) > B.bw) return;  // the tile lies outside the band
  __shared__ double Li[NB][NB + 1], Lj[NB][NB + 1];
  const int r = threadIdx.x / NB, c = threadIdx.x % NB;
  {
    const int k = c;  // column of the panel
    const int ir = r0 + r, jr = c0 + r;
    const double dk = k < kb ? B.at(k0 + k, k0 + k) : 0.0;
    Li[r][k] = (k < kb && ir < k0 + kb + m && B.in(ir, k0 + k)) ? B.at(ir, k0 + k) * dk : 0.0;
    Lj[r][k] = (k < kb && jr < k0 + kb + m && B.in(jr, k0 + k)) ? B.at(jr, k0 + k) : 0.0;
  }
  __syncthreads();
  const int i = r0 + r, j = c0 + c;
  if (i >= k0 + kb + m || j > i || i - j > B.bw) return;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NB; k++) s += Li[r][k] * Lj[c][k];
  B.at(i, j) -= s;
}

// forward reduction and diagonal scaling, z = D^-1 L^-1 b (one workgroup)
__global__ __launch_bounds__(1024) void band_forward_kernel(Band B, const double *__restrict__ b,
                                                            double *__restrict__ z) {
  extern __shared__ double ring[];  // bw + NB entries, entry e at e % W
  const int W = B.bw + NB, n = B.n, tid = threadIdx.x;
  for (int e = tid; e < W && e < n; e += blockDim.x) ring[e % W] = b[e];
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = min(NB, n - k0);
    __syncthreads();
    if (tid < 64) {  // the panel's triangle: one wave
      double Lr[NB];
      panel_row(B, k0, kb, tid, Lr);
      double v = tid < kb ? ring[(k0 + tid) % W] : 0.0;
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const double vj = __shfl(v, j);
        v -= Lr[j] * vj;
      }
      if (tid < kb) ring[(k0 + tid) % W] = v;
    }
    __syncthreads();
    const int m = min(B.bw, n - k0 - kb);
    for (int t = tid; t < m; t += blockDim.x) {
      const int i = k0 + kb + t;
      double s = 0.0;
      for (int j = 0; j < kb; j++)
        if (i - (k0 + j) <= B.bw) s += B.at(i, k0 + j) * ring[(k0 + j) % W];
      ring[i % W] -= s;
    }
    __syncthreads();
    if (tid < NB) {
      if (tid < kb) z[k0 + tid] = ring[(k0 + tid) % W] / B.at(k0 + tid, k0 + tid);
      const int e = k0 + W + tid;  // takes the slot just written out
      if (e < n) ring[e % W] = b[e];
    }
  }
}

// back substitution, x = L^-T z (one workgroup)
__global__ __launch_bounds__(1024) void band_backward_kernel(Band B, const double *__restrict__ z,
                                                             double *__restrict__ x) {
  extern __shared__ double ring[];
  __shared__ double red[NB][NB + 1];
  const int W = B.bw + NB, n = B.n, tid = threadIdx.x;
  const int klast = (n - 1) / NB * NB;
  for (int e = klast + tid; e < n; e += blockDim.x) ring[e % W] = z[e];
  for (int k0 = klast; k0 >= 0; k0 -= NB) {
    const int kb = min(NB, n - k0);
    const int m = min(B.bw, n - k0 - kb);
    __syncthreads();
    {  // coupling to the solved rows below: s_j = sum_r L(r, k0 + j) x_r
      const int j = tid / NB, ch = tid % NB;
      double s = 0.0;
      if (j < kb)
        for (int t = ch; t < m; t += NB) {
          const int i = k0 + kb + t;
          if (i - (k0 + j) <= B.bw) s += B.at(i, k0 + j) * ring[i % W];
        }
      red[j][ch] = s;
    }
    __syncthreads();
    if (tid < 64) {
      double v = 0.0;
      if (tid < kb) {
        double s = 0.0;
        for (int ch = 0; ch < NB; ch++) s += red[tid][ch];
        v = ring[(k0 + tid) % W] - s;
      }
      double Lc[NB];
      panel_col(B, k0, kb, tid, Lc);
#pragma unroll
      for (int j = NB - 1; j > 0; j--) {
        const double xj = __shfl(v, j);
        v -= Lc[j] * xj;
      }
      if (tid < kb) ring[(k0 + tid) % W] = v;
    }
    __syncthreads();
    if (tid < kb) x[k0 + tid] = ring[(k0 + tid) % W];
    __syncthreads();  // (with bw < NB the next entries share slots with these)
    if (tid < NB) {
      const int e = k0 - NB + tid;  // the next panel's entries
      if (e >= 0) ring[e % W] = z[e];
    }
  }
}

// Wide forward step for the panel at k0: every workgroup solves the panel's
// triangle from y (final for these rows), workgroup 0 stores z = y / d for
// them, and each workgroup updates its WROWS rows below the panel.  y of the
// panel itself is left as it was (other workgroups read it).
__global__ __launch_bounds__(WROWS) void band_fwd_wide_kernel(Band B, int k0,
                                                              double *__restrict__ y,
                                                              double *__restrict__ z) {
  __shared__ double yp[NB];
  const int n = B.n, tid = threadIdx.x;
  const int kb = min(NB, n - k0);
  if (tid < 64) {
    double Lr[NB];
    panel_row(B, k0, kb, tid, Lr);
    double v = tid < kb ? y[k0 + tid] : 0.0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const double vj = __shfl(v, j);
      v -= Lr[j] * vj;
    }
    if (tid < kb) yp[tid] = v;
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < kb) z[k0 + tid] = yp[tid] / B.at(k0 + tid, k0 + tid);
  const int m = min(B.bw, n - k0 - kb);
  const int t = blockIdx.x * WROWS + tid;
  if (t < m) {
    const int i = k0 + kb + t;
    double s = 0.0;
    for (int j = 0; j < kb; j++)
      if (i - (k0 + j) <= B.bw) s += B.at(i, k0 + j) * yp[j];
    y[i] -= s;
  }
}

// Wide backward step for the panel at k0 (panels from the last one up):
// every workgroup adds the nin partial sums of the panel (written by the
// previous step, in workgroup order), solves its triangle (L^T, unit
// diagonal) from z, workgroup 0 stores x for it, and each workgroup writes
// its partial sums of L(i, k0 - NB + j) x_i over its WROWS rows i >= k0 for
// the panel above (pout[blockIdx.x][j]).
__global__ __launch_bounds__(WROWS) void band_bwd_wide_kernel(Band B, int k0,
                                                              const double *__restrict__ z,
                                                              double *__restrict__ x,
                                                              const double *__restrict__ pin,
                                                              int nin, double *__restrict__ pout) {
  __shared__ double xp[NB];
  __shared__ double red[NB][9];
  const int n = B.n, tid = threadIdx.x;
  const int kb = min(NB, n - k0);
  if (tid < 64) {
    double v = 0.0;
    if (tid < kb) {
      double s = 0.0;
      for (int w = 0; w < nin; w++) s += pin[w * NB + tid];
      v = z[k0 + tid] - s;
    }
    double Lc[NB];
    panel_col(B, k0, kb, tid, Lc);
#pragma unroll
    for (int j = NB - 1; j > 0; j--) {
      const double xj = __shfl(v, j);
      v -= Lc[j] * xj;
    }
    if (tid < kb) xp[tid] = v;
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < kb) x[k0 + tid] = xp[tid];
  if (k0 == 0) return;  // no panel above (uniform over the grid)
  // partial sums for the panel above (columns k0 - NB + j, a full panel):
  // thread (j, g) takes the workgroup's rows g, g + 8, ...
  const int ka = k0 - NB;
  const int m = min(B.bw, n - k0);  // rows below that panel
  const int j = tid / 8, g = tid % 8;
  const int r0 = blockIdx.x * WROWS;
  double s = 0.0;
  for (int r = g; r < WROWS; r += 8) {
    const int t = r0 + r;
    if (t >= m) break;
    const int i = k0 + t;
    if (i - (ka + j) > B.bw) continue;
    const double xi = i < k0 + kb ? xp[i - k0] : x[i];
    s += B.at(i, ka + j) * xi;
  }
  red[j][g] = s;
  __syncthreads();
  if (tid < NB) {
    double t = 0.0;
    for (int q = 0; q < 8; q++) t += red[tid][q];
    pout[blockIdx.x * NB + tid] = t;
  }
}

}  // namespace

int band_width(Ctx &c) {
  if (c.band_bw >= 0) return c.band_bw;
  std::vector<long long> bptr(c.nnode + 1);
  std::vector<int> bcol(c.nbf);
  FGPU_CHECK(hipMemcpyAsync(bptr.data(), c.d_bptr, sizeof(long long) * bptr.size(),
                            hipMemcpyDeviceToHost, c.stream));
  FGPU_CHECK(hipMemcpyAsync(bcol.data(), c.d_bcol, sizeof(int) * bcol.size(),
                            hipMemcpyDeviceToHost, c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  // lowest equation of each node, then per row node the widest reach
  std::vector<int> lo(c.nnode, -1);
  for (int A = 0; A < c.nnode; A++)
    for (int d = 0; d < 3; d++) {
      const int e = c.eqmap[3 * A + d];
      if (e >= 0 && (lo[A] < 0 || e < lo[A])) lo[A] = e;
    }
  int bw = 0;
  for (int A = 0; A < c.nnode; A++) {
    int hi = -1;
    for (int d = 0; d < 3; d++) hi = std::max(hi, c.eqmap[3 * A + d]);
    if (hi < 0) continue;
    for (long long k = bptr[A]; k < bptr[A + 1]; k++) {
      const int C = bcol[k];
      if (lo[C] >= 0 && lo[C] <= hi) bw = std::max(bw, hi - lo[C]);
    }
  }
  c.band_bw = bw;
  return bw;
}

// Width check and buffers of the band factorization (host work; before any
// graph capture).  0 with c.err when the band is too wide.
int band_setup(Ctx &c) {
  const int n = c.neq, bw = band_width(c);
  if (n == 0) return 1;
  if (bw > MAX_BW_SOLVE) {
    c.err = "direct solve: half bandwidth " + std::to_string(bw) + " above the limit " +
            std::to_string(MAX_BW_SOLVE) + " (renumber the equations or use PCG)";
    return 0;
  }
  if (!c.d_band) {
    const size_t entries = (size_t)n * (bw + 1);
    FGPU_CHECK(hipMalloc(&c.d_band, entries * sizeof(double)));
    FGPU_CHECK(hipMalloc(&c.d_finfo, 2 * sizeof(int)));
    if (bw >= WIDE_BW)
      FGPU_CHECK(hipMalloc(&c.d_bpart, sizeof(double) * 2 * NB * ((bw + WROWS - 1) / WROWS)));
  }
  return 1;
}

// The factorization's launches only (graph-safe: fills by kernels): the
// band is cleared and filled from the node blocks, then panels of NB
// columns; d_finfo[0] = first tiny pivot (or INT_MAX), d_finfo[1] = the
// number of negative pivots.
void enqueue_band_factor(Ctx &c) {
  const int n = c.neq, bw = c.band_bw;
  if (n == 0) return;
  const size_t entries = (size_t)n * (bw + 1);
  Band B{c.d_band, n, bw, (long long)bw + 1};
  launch_fill(c, c.d_band, (long long)entries, 0.0);
  launch_fill(c, c.d_finfo, 1, 0x7fffffff);
  launch_fill(c, c.d_finfo + 1, 1, 0);
  hipLaunchKernelGGL(band_fill_kernel, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_bptr, c.d_bcol, c.d_bval, c.d_eqmap, B);
  FGPU_CHECK(hipGetLastError());
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    hipLaunchKernelGGL(band_diag_kernel, dim3(1), dim3(NB * NB), 0, c.stream, B, k0, kb, 1e-12,
                       c.d_finfo);
    const int m = std::min(bw, n - k0 - kb);
    if (m > 0) {
      hipLaunchKernelGGL(band_panel_kernel, dim3((m + 255) / 256), dim3(256), 0, c.stream, B, k0,
                         kb, m);
      const int mt = (m + NB - 1) / NB;
      hipLaunchKernelGGL(band_update_kernel, dim3(mt * (mt + 1) / 2), dim3(NB * NB), 0, c.stream,
                         B, k0, kb, m);
    }
  }
  FGPU_CHECK(hipGetLastError());
}

int band_factor_result(Ctx &c, const int *info) {
  c.neg_pivots = info[1];
  if (info[0] != 0x7fffffff) {
    c.factored = false;
    c.err = "direct solve: singular matrix (pivot below 1e-12) at equation " +
            std::to_string(info[0]);
    return 0;
  }
  c.factored = true;
  return 1;
}

int band_factor(Ctx &c) {
  if (c.neq == 0) {
    c.factored = true;
    return 1;
  }
  if (!band_setup(c)) return 0;
  enqueue_band_factor(c);
  int info[2];
  FGPU_CHECK(hipMemcpyAsync(info, c.d_finfo, sizeof info, hipMemcpyDeviceToHost, c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  return band_factor_result(c, info);
}

// the solves' LDS size for the current band (host work, before a capture)
void band_solve_setup(Ctx &c) {
  const size_t lds = sizeof(double) * (size_t)(c.band_bw + NB);
  FGPU_CHECK(hipFuncSetAttribute((const void *)band_forward_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  FGPU_CHECK(hipFuncSetAttribute((const void *)band_backward_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// forward and backward substitution, x = (L D L^T)^-1 b through z (n
// doubles, distinct from b and x); launches only
void enqueue_band_solve(Ctx &c, const double *b, double *x, double *z) {
  const int n = c.neq, bw = c.band_bw;
  if (n == 0) return;
  Band B{c.d_band, n, bw, (long long)bw + 1};
  if (bw >= WIDE_BW) {
    // forward in place on x (its entries are rewritten by the backward pass
    // from the bottom up), z = D^-1 y; then backward from z into x
    if (!c.d_bpart) throw Error("direct solve: wide-solve buffers not allocated");
    launch_copy(c, b, x, n);
    for (int k0 = 0; k0 < n; k0 += NB) {
      const int m = std::min(bw, n - k0 - std::min(NB, n - k0));
      const int grid = std::max(1, (m + WROWS - 1) / WROWS);
      hipLaunchKernelGGL(band_fwd_wide_kernel, dim3(grid), dim3(WROWS), 0, c.stream, B, k0, x, z);
    }
    const int klast = (n - 1) / NB * NB;
    const int half = NB * ((bw + WROWS - 1) / WROWS);
    int nin = 0, par = 0;
    for (int k0 = klast; k0 >= 0; k0 -= NB) {
      const int m = k0 > 0 ? std::min(bw, n - k0) : 0;
      const int grid = std::max(1, (m + WROWS - 1) / WROWS);
      double *pin = c.d_bpart + (size_t)par * half, *pout = c.d_bpart + (size_t)(1 - par) * half;
      hipLaunchKernelGGL(band_bwd_wide_kernel, dim3(grid), dim3(WROWS), 0, c.stream, B, k0, z, x,
                         pin, nin, pout);
      nin = grid;
      par = 1 - par;
    }
    FGPU_CHECK(hipGetLastError());
    return;
  }
  const size_t lds = sizeof(double) * (size_t)(bw + NB);
  hipLaunchKernelGGL(band_forward_kernel, dim3(1), dim3(1024), lds, c.stream, B, b, z);
  hipLaunchKernelGGL(band_backward_kernel, dim3(1), dim3(1024), lds, c.stream, B, z, x);
  FGPU_CHECK(hipGetLastError());
}

void band_solve(Ctx &c, const double *b, double *x) {
  const int n = c.neq;
  if (n == 0) return;
  double *z = c.d_b == b || c.d_b == x ? nullptr : c.d_b;
  double *tmp = nullptr;
  if (!z) {
    FGPU_CHECK(hipMalloc(&tmp, sizeof(double) * n));
    z = tmp;
  }
  band_solve_setup(c);
  enqueue_band_solve(c, b, x, z);
  if (tmp) {
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(tmp);
  }
}

}  // namespace fgpu
