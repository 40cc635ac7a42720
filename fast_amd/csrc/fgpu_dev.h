// fgpu_dev.h - device helpers shared by the solver translation units:
// deterministic block reductions and the node-block SpMV row loop.
#pragma once

#include "fgpu_internal.h"

namespace fgpu {

constexpr int RED = 1024;  // partial-sum slots (== blocks of reduction kernels)
constexpr int TPB = 256;

// Fixed-shape tree over a block (wave shuffles, then waves in order):
// deterministic for a given block size.  Result valid in thread 0.
__device__ inline double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x / 64); k++) r += sh[k];
  return r;
}

// Sum of n partials, identical in every block that calls it; result in all
// threads.
__device__ inline double reduce_partials(const double *__restrict__ part, int n, double *sh) {
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
  double r = block_sum(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) sh[8] = r;
  __syncthreads();
  return sh[8];
}

// Rows [r0, r1) of y = K x, G lanes per row; returns this thread's share of
// dot(x, y) (nonzero only in lane 0 of a row group).  T: the stored value
// type (the FP64 matrix, or its FP32 copy); the arithmetic is FP64.
template <int G, class T = double>
__device__ inline double bspmv_rows(long long r0, long long r1, long long stride,
                                    const long long *__restrict__ bptr,
                                    const int *__restrict__ bcol,
                                    const T *__restrict__ bval,
                                    const double *__restrict__ x, double *__restrict__ y) {
  const int sub = threadIdx.x % G;
  double dot = 0.0;
  for (long long r = r0 + threadIdx.x / G; r < r1; r += stride) {
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    const long long k1 = bptr[r + 1];
    for (long long k = bptr[r] + sub; k < k1; k += G) {
      const T *vp = bval + 9 * k;
      double v[9];
#pragma unroll
      for (int m = 0; m < 9; m++) v[m] = (double)vp[m];
      const int B = bcol[k];
      const double x0 = x[3 * B], x1 = x[3 * B + 1], x2 = x[3 * B + 2];
      y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
      y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
      y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    for (int o = G / 2; o > 0; o >>= 1) {
      y0 += __shfl_down(y0, o, G);
      y1 += __shfl_down(y1, o, G);
      y2 += __shfl_down(y2, o, G);
    }
    if (sub == 0) {
      y[3 * r] = y0;
      y[3 * r + 1] = y1;
      y[3 * r + 2] = y2;
      dot += y0 * x[3 * r] + y1 * x[3 * r + 1] + y2 * x[3 * r + 2];
    }
  }
  return dot;
}

// lanes per row from the mean number of blocks per row
inline int spmv_group(const Ctx &c) {
  double avg = c.nnode ? (double)c.nbf / c.nnode : 1.0;
  if (avg > 96) return 64;
  if (avg > 40) return 32;
  if (avg > 20) return 16;
  if (avg > 6) return 8;
  return 4;
}

}  // namespace fgpu
