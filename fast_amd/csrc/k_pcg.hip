// k_pcg.hip - FP64 Jacobi-PCG of libfgpu on the node-block matrix: one
// system, or a batch of independent meshes with per-mesh convergence.
//
// Restates PCGSolver (mlib/sysmat.cpp:1061-1125): Jacobi preconditioner,
// x0 = 0 (or a caller warm start), stop when |r| / |b| < tol with |b| > 1e11
// replaced by 1, immediate return when |b| < tol.  Dot products are block
// partial sums reduced in one fixed order by every consumer block, so results
// do not depend on scheduling.  All scalars (alpha, beta, z.r, the iteration
// counter, the tolerance) live on the device; every kernel is a no-op once
// its system has converged, so the iterate returned is the first one meeting
// the tolerance.  A chunk of iterations is captured once into a hipGraph and
// replayed; the host polls one flag per chunk.
//
// Batch mode: the meshes of a batch own contiguous node ranges.  The node
// space is cut into chunks of at most CHN nodes that never cross a mesh; each
// chunk is one workgroup, and the per-mesh sums are the chunk partials of
// that mesh added in chunk order by every chunk of the mesh.
#include <cmath>

#include "fgpu_dev.h"

namespace fgpu {
namespace {

constexpr int CHUNK_IT = 16;  // iterations per captured graph

// ---------------------------------------------------------------- single
// scal: [0] lenf  [1],[2] z.r by iteration parity  [3] |r|/lenf  [4] |b|
//       [6] tol ; done[0] = iteration at convergence, done[1] = breakdown,
//       done[2] = current iteration.  An iteration is three launches: the
//       product (whose block 0 opens the iteration, done[2] += 1; no other
//       block of it reads done[2]), the update of x, r, z and the direction
//       (whose blocks all find convergence alike, block 0 recording it).
template <int G, class T>
__global__ __launch_bounds__(256) void pcg_spmv_kernel(
    int nnode, const long long *__restrict__ bptr, const int *__restrict__ bcol,
    const T *__restrict__ bval, const double *__restrict__ p, double *__restrict__ w,
    double *__restrict__ part, int *__restrict__ done) {
  __shared__ double sh[16];
  if (done[0]) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) done[2] = done[2] + 1;
  const int rows_per_block = blockDim.x / G;
  double dot = bspmv_rows<G, T>((long long)blockIdx.x * rows_per_block, nnode,
                             (long long)gridDim.x * rows_per_block, bptr, bcol, bval, p, w);
  double t = block_sum(dot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void pcg_init_kernel(
    int n, const double *__restrict__ b, double *__restrict__ r, const double *__restrict__ dinv,
    double *__restrict__ z, double *__restrict__ p, double *__restrict__ part) {
  __shared__ double sh[16];
  double zr = 0.0, bb = 0.0, rr = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double ri = r[i], zi = ri * dinv[i];  // r holds b - K x0 (or b)
    z[i] = zi;
    p[i] = zi;
    zr += zi * ri;
    bb += b[i] * b[i];
    rr += ri * ri;
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(bb, sh);
  double t2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = t0;
    part[RED + blockIdx.x] = t1;
    part[2 * RED + blockIdx.x] = t2;
  }
}

// Node-block Jacobi (fgpu_set_preconditioner 1): z_j = B_j^-1 r_j with B_j
// node j's 3x3 diagonal block, binv = [nnode][9].  One thread per node (its
// three entries; no thread reads an entry another writes), otherwise as
// pcg_init_kernel and pcg_update_kernel.
__device__ inline void bj_apply(const double *__restrict__ binv, long long j, const double r3[3],
                                double z3[3]) {
  const double *B = binv + 9 * j;
  for (int d = 0; d < 3; d++) z3[d] = B[3 * d] * r3[0] + B[3 * d + 1] * r3[1] + B[3 * d + 2] * r3[2];
}

__global__ __launch_bounds__(256) void pcg_init_bj_kernel(
    int nnode, const double *__restrict__ b, double *__restrict__ r,
    const double *__restrict__ binv, double *__restrict__ z, double *__restrict__ p,
    double *__restrict__ part) {
  __shared__ double sh[16];
  double zr = 0.0, bb = 0.0, rr = 0.0;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nnode;
       j += (long long)gridDim.x * blockDim.x) {
    double r3[3], z3[3];
    for (int d = 0; d < 3; d++) r3[d] = r[3 * j + d];
    bj_apply(binv, j, r3, z3);
    for (int d = 0; d < 3; d++) {
      z[3 * j + d] = z3[d];
      p[3 * j + d] = z3[d];
      zr += z3[d] * r3[d];
      bb += b[3 * j + d] * b[3 * j + d];
      rr += r3[d] * r3[d];
    }
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(bb, sh);
  double t2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = t0;
    part[RED + blockIdx.x] = t1;
    part[2 * RED + blockIdx.x] = t2;
  }
}

__global__ __launch_bounds__(256) void pcg_update_bj_kernel(
    int nnode, double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
    const double *__restrict__ p, const double *__restrict__ w, const double *__restrict__ binv,
    double *__restrict__ part, const double *__restrict__ scal, const int *__restrict__ done) {
  __shared__ double sh[16];
  if (done[0]) return;
  const int it = done[2];
  const double pw = reduce_partials(part, RED, sh);
  const double alpha = scal[1 + ((it - 1) & 1)] / pw;
  double zr = 0.0, rr = 0.0;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nnode;
       j += (long long)gridDim.x * blockDim.x) {
    double r3[3], z3[3];
    for (int d = 0; d < 3; d++) {
      const long long i = 3 * j + d;
      x[i] = x[i] + alpha * p[i];
      r3[d] = r[i] + (-alpha) * w[i];
      r[i] = r3[d];
    }
    bj_apply(binv, j, r3, z3);
    for (int d = 0; d < 3; d++) {
      z[3 * j + d] = z3[d];
      zr += z3[d] * r3[d];
      rr += r3[d] * r3[d];
    }
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(rr, sh);
  __syncthreads();
  if (threadIdx.x == 0) {
    part[RED + blockIdx.x] = t0;
    part[2 * RED + blockIdx.x] = t1;
  }
}

// inverse of each node's 3x3 diagonal block, FP64 cofactors
__global__ void block_inv_kernel(int nnode, const long long *__restrict__ bdiag,
                                 const double *__restrict__ bval, double *__restrict__ binv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnode) return;
  const double *A = bval + 9 * bdiag[j];
  const double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8],
               c02 = A[3] * A[7] - A[4] * A[6];
  const double id = 1.0 / (A[0] * c00 + A[1] * c01 + A[2] * c02);
  double *I = binv + 9 * (size_t)j;
  I[0] = c00 * id;
  I[1] = (A[2] * A[7] - A[1] * A[8]) * id;
  I[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  I[3] = c01 * id;
  I[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  I[5] = (A[2] * A[3] - A[0] * A[5]) * id;
  I[6] = c02 * id;
  I[7] = (A[1] * A[6] - A[0] * A[7]) * id;
  I[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

__global__ void pcg_start_kernel(const double *__restrict__ part, double *__restrict__ scal,
                                 int *__restrict__ done, double tol) {
  __shared__ double sh[16];
  double zr = reduce_partials(part, RED, sh);
  double bb = reduce_partials(part + RED, RED, sh);
  double rr = reduce_partials(part + 2 * RED, RED, sh);
  if (threadIdx.x == 0) {
    double lenf = sqrt(bb);
    scal[4] = lenf;
    if (lenf > 10e+10) lenf = 1.0;
    scal[0] = lenf;
    scal[1] = zr;
    scal[3] = sqrt(rr) / lenf;
    scal[6] = tol;
    done[0] = 0;
    done[1] = 0;
    done[2] = 0;
  }
}

// The start of a solve whose tolerance lives on the device (a replayed
// graph): as pcg_start_kernel, and the immediate exits of run_pcg (|b| < tol
// under abs_exit, |b| = 0, or the start already converged) mark done[0] = -1,
// which turns every iteration kernel into a no-op.
__global__ void pcg_start_dev_kernel(const double *__restrict__ part, double *__restrict__ scal,
                                     int *__restrict__ done, const double *__restrict__ tolp,
                                     int abs_exit) {
  __shared__ double sh[16];
  double zr = reduce_partials(part, RED, sh);
  double bb = reduce_partials(part + RED, RED, sh);
  double rr = reduce_partials(part + 2 * RED, RED, sh);
  if (threadIdx.x == 0) {
    const double tol = *tolp;
    double lenf = sqrt(bb);
    scal[4] = lenf;
    if (lenf > 10e+10) lenf = 1.0;
    scal[0] = lenf;
    scal[1] = zr;
    scal[3] = sqrt(rr) / lenf;
    scal[6] = tol;
    const bool early = (abs_exit && scal[4] < tol) || scal[4] == 0.0 || scal[3] < tol;
    done[0] = early ? -1 : 0;
    done[1] = 0;
    done[2] = 0;
  }
}

__global__ __launch_bounds__(256) void pcg_update_kernel(
    int n, double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
    const double *__restrict__ p, const double *__restrict__ w, const double *__restrict__ dinv,
    double *__restrict__ part, const double *__restrict__ scal, const int *__restrict__ done) {
  __shared__ double sh[16];
  if (done[0]) return;
  const int it = done[2];
  const double pw = reduce_partials(part, RED, sh);
  const double alpha = scal[1 + ((it - 1) & 1)] / pw;
  double zr = 0.0, rr = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    x[i] = x[i] + alpha * p[i];
    double ri = r[i] + (-alpha) * w[i];
    r[i] = ri;
    double zi = ri * dinv[i];
    z[i] = zi;
    zr += zi * ri;
    rr += ri * ri;
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(rr, sh);
  __syncthreads();
  if (threadIdx.x == 0) {
    part[RED + blockIdx.x] = t0;
    part[2 * RED + blockIdx.x] = t1;
  }
}

__global__ __launch_bounds__(256) void pcg_dir_kernel(
    int n, double *__restrict__ p, const double *__restrict__ z, const double *__restrict__ part,
    double *__restrict__ scal, int *__restrict__ done) {
  __shared__ double sh[16];
  if (done[0]) return;
  const int it = done[2];
  const double zr = reduce_partials(part + RED, RED, sh);
  const double rr = reduce_partials(part + 2 * RED, RED, sh);
  const double err = sqrt(rr) / scal[0];
  if (err < scal[6] || !(err == err)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      scal[3] = err;
      done[1] = !(err == err);
      done[0] = it;  // the other blocks reach the same test and return alike
    }
    return;
  }
  const double beta = zr / scal[1 + ((it - 1) & 1)];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[1 + (it & 1)] = zr;
    scal[3] = err;
  }
}

__global__ void diag_inv_kernel(int nnode, const long long *__restrict__ bdiag,
                                const double *__restrict__ bval, double *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * nnode) dinv[i] = 1.0 / bval[9 * bdiag[i / 3] + 4 * (i % 3)];
}

__global__ void residual_kernel(int n, const double *__restrict__ b, const double *__restrict__ Ax,
                                double *__restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - Ax[i];
}

template <int G>
void pcg_spmv_launch(Ctx &c) {
  if (c.mat_bits == 32)
    hipLaunchKernelGGL((pcg_spmv_kernel<G, float>), dim3(RED), dim3(TPB), 0, c.stream, c.nnode,
                       c.d_bptr, c.d_bcol, c.d_bval32, c.d_p, c.d_w, c.d_part, c.d_done);
  else
    hipLaunchKernelGGL((pcg_spmv_kernel<G, double>), dim3(RED), dim3(TPB), 0, c.stream, c.nnode,
                       c.d_bptr, c.d_bcol, c.d_bval, c.d_p, c.d_w, c.d_part, c.d_done);
}

void enqueue_iteration(Ctx &c) {
  const int n = 3 * c.nnode;
  switch (spmv_group(c)) {
    case 64: pcg_spmv_launch<64>(c); break;
    case 32: pcg_spmv_launch<32>(c); break;
    case 16: pcg_spmv_launch<16>(c); break;
    case 8: pcg_spmv_launch<8>(c); break;
    default: pcg_spmv_launch<4>(c); break;
  }
  if (c.precond == 1)
    hipLaunchKernelGGL(pcg_update_bj_kernel, dim3(RED), dim3(TPB), 0, c.stream, c.nnode, c.d_x,
                       c.d_r, c.d_z, c.d_p, c.d_w, c.d_binv, c.d_part, c.d_scal, c.d_done);
  else
    hipLaunchKernelGGL(pcg_update_kernel, dim3(RED), dim3(TPB), 0, c.stream, n, c.d_x, c.d_r,
                       c.d_z, c.d_p, c.d_w, c.d_dinv, c.d_part, c.d_scal, c.d_done);
  hipLaunchKernelGGL(pcg_dir_kernel, dim3(RED), dim3(TPB), 0, c.stream, n, c.d_p, c.d_z,
                     c.d_part, c.d_scal, c.d_done);
}

// ---------------------------------------------------------------- batch
// per-mesh scalars mscal[8 m + k]: [0] lenf [1],[2] z.r by parity [3] err
// [4] |b| [6] tol; mdone[4 m + k]: [0] converged at, [1] breakdown,
// [2] current iteration (per mesh), [3] started
__device__ inline double mesh_sum(const double *__restrict__ part, int c0, int c1) {
  double s = 0.0;
  for (int c = c0; c < c1; c++) s += part[c];
  return s;
}

template <int G, class T>
__global__ __launch_bounds__(256) void bpcg_spmv_kernel(
    const long long *__restrict__ chn0, const long long *__restrict__ chn1,
    const int *__restrict__ chmesh, const long long *__restrict__ bptr,
    const int *__restrict__ bcol, const T *__restrict__ bval,
    const double *__restrict__ p, double *__restrict__ w, double *__restrict__ part,
    const int *__restrict__ mdone) {
  __shared__ double sh[16];
  const int ch = blockIdx.x;
  if (mdone[4 * chmesh[ch]]) return;
  const int rows_per_block = blockDim.x / G;
  double dot = bspmv_rows<G, T>(chn0[ch], chn1[ch], rows_per_block, bptr, bcol, bval, p, w);
  double t = block_sum(dot, sh);
  if (threadIdx.x == 0) part[ch] = t;
}

__global__ __launch_bounds__(256) void bpcg_init_kernel(
    const long long *__restrict__ chn0, const long long *__restrict__ chn1,
    const double *__restrict__ b, double *__restrict__ r, const double *__restrict__ dinv,
    double *__restrict__ z, double *__restrict__ p, double *__restrict__ part, int nch) {
  __shared__ double sh[16];
  const int ch = blockIdx.x;
  double zr = 0.0, bb = 0.0, rr = 0.0;
  for (long long i = 3 * chn0[ch] + threadIdx.x; i < 3 * chn1[ch]; i += blockDim.x) {
    double ri = r[i], zi = ri * dinv[i];
    z[i] = zi;
    p[i] = zi;
    zr += zi * ri;
    bb += b[i] * b[i];
    rr += ri * ri;
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(bb, sh);
  double t2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[ch] = t0;
    part[nch + ch] = t1;
    part[2 * nch + ch] = t2;
  }
}

__global__ void bpcg_start_kernel(int nmesh, const int *__restrict__ meshch,
                                  const double *__restrict__ part, int nch,
                                  double *__restrict__ mscal, int *__restrict__ mdone,
                                  int *__restrict__ ndone, double tol) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0) *ndone = 0;
  if (m >= nmesh) return;
  const int c0 = meshch[m], c1 = meshch[m + 1];
  const double zr = mesh_sum(part, c0, c1), bb = mesh_sum(part + nch, c0, c1),
               rr = mesh_sum(part + 2 * nch, c0, c1);
  double lenf = sqrt(bb);
  double *s = mscal + 8 * m;
  s[4] = lenf;
  if (lenf > 10e+10) lenf = 1.0;
  s[0] = lenf;
  s[1] = zr;
  s[3] = sqrt(rr) / lenf;
  s[6] = tol;
  int *d = mdone + 4 * m;
  d[1] = 0;
  d[2] = 1;
  // |b| < tol or a warm start already converged: finished before iterating
  d[0] = (s[4] < tol || s[3] < tol) ? -1 : 0;
}

__global__ void bpcg_count_kernel(int nmesh, const int *__restrict__ mdone, int *__restrict__ ndone) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int k = 0;
  for (int m = 0; m < nmesh; m++) k += mdone[4 * m] != 0;
  *ndone = k;
}

// Node-block Jacobi for the batch (fgpu_set_preconditioner 1): one thread
// per node of the chunk, otherwise as bpcg_init_kernel / bpcg_update_kernel.
__global__ __launch_bounds__(256) void bpcg_init_bj_kernel(
    const long long *__restrict__ chn0, const long long *__restrict__ chn1,
    const double *__restrict__ b, double *__restrict__ r, const double *__restrict__ binv,
    double *__restrict__ z, double *__restrict__ p, double *__restrict__ part, int nch) {
  __shared__ double sh[16];
  const int ch = blockIdx.x;
  double zr = 0.0, bb = 0.0, rr = 0.0;
  for (long long j = chn0[ch] + threadIdx.x; j < chn1[ch]; j += blockDim.x) {
    double r3[3], z3[3];
    for (int d = 0; d < 3; d++) r3[d] = r[3 * j + d];
    bj_apply(binv, j, r3, z3);
    for (int d = 0; d < 3; d++) {
      z[3 * j + d] = z3[d];
      p[3 * j + d] = z3[d];
      zr += z3[d] * r3[d];
      bb += b[3 * j + d] * b[3 * j + d];
      rr += r3[d] * r3[d];
    }
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(bb, sh);
  double t2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[ch] = t0;
    part[nch + ch] = t1;
    part[2 * nch + ch] = t2;
  }
}

__global__ __launch_bounds__(256) void bpcg_update_bj_kernel(
    const long long *__restrict__ chn0, const long long *__restrict__ chn1,
    const int *__restrict__ chmesh, const int *__restrict__ meshch, double *__restrict__ x,
    double *__restrict__ r, double *__restrict__ z, const double *__restrict__ p,
    const double *__restrict__ w, const double *__restrict__ binv, double *__restrict__ part,
    int nch, const double *__restrict__ mscal, const int *__restrict__ mdone) {
  __shared__ double sh[16];
  const int ch = blockIdx.x, m = chmesh[ch];
  const int *dn = mdone + 4 * m;
  if (dn[0]) return;
  const int it = dn[2];
  const double pw = mesh_sum(part, meshch[m], meshch[m + 1]);
  const double alpha = mscal[8 * m + 1 + ((it - 1) & 1)] / pw;
  double zr = 0.0, rr = 0.0;
  for (long long j = chn0[ch] + threadIdx.x; j < chn1[ch]; j += blockDim.x) {
    double r3[3], z3[3];
    for (int d = 0; d < 3; d++) {
      const long long i = 3 * j + d;
      x[i] = x[i] + alpha * p[i];
      r3[d] = r[i] + (-alpha) * w[i];
      r[i] = r3[d];
    }
    bj_apply(binv, j, r3, z3);
    for (int d = 0; d < 3; d++) {
      z[3 * j + d] = z3[d];
      zr += z3[d] * r3[d];
      rr += r3[d] * r3[d];
    }
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[nch + ch] = t0;
    part[2 * nch + ch] = t1;
  }
}

__global__ __launch_bounds__(256) void bpcg_update_kernel(
    const long long *__restrict__ chn0, const long long *__restrict__ chn1,
    const int *__restrict__ chmesh, const int *__restrict__ meshch, double *__restrict__ x,
    double *__restrict__ r, double *__restrict__ z, const double *__restrict__ p,
    const double *__restrict__ w, const double *__restrict__ dinv, double *__restrict__ part,
    int nch, const double *__restrict__ mscal, const int *__restrict__ mdone) {
  __shared__ double sh[16];
  const int ch = blockIdx.x, m = chmesh[ch];
  const int *d = mdone + 4 * m;
  if (d[0]) return;
  const int it = d[2];
  const double pw = mesh_sum(part, meshch[m], meshch[m + 1]);
  const double alpha = mscal[8 * m + 1 + ((it - 1) & 1)] / pw;
  double zr = 0.0, rr = 0.0;
  for (long long i = 3 * chn0[ch] + threadIdx.x; i < 3 * chn1[ch]; i += blockDim.x) {
    x[i] = x[i] + alpha * p[i];
    double ri = r[i] + (-alpha) * w[i];
    r[i] = ri;
    double zi = ri * dinv[i];
    z[i] = zi;
    zr += zi * ri;
    rr += ri * ri;
  }
  double t0 = block_sum(zr, sh);
  double t1 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[nch + ch] = t0;
    part[2 * nch + ch] = t1;
  }
}

__global__ __launch_bounds__(256) void bpcg_dir_kernel(
    const long long *__restrict__ chn0, const long long *__restrict__ chn1,
    const int *__restrict__ chmesh, const int *__restrict__ meshch, double *__restrict__ p,
    const double *__restrict__ z, const double *__restrict__ part, int nch,
    const double *__restrict__ mscal, const int *__restrict__ mdone) {
  const int ch = blockIdx.x, m = chmesh[ch];
  const int *d = mdone + 4 * m;
  if (d[0]) return;
  const int it = d[2];
  const double *s = mscal + 8 * m;
  const double zr = mesh_sum(part + nch, meshch[m], meshch[m + 1]);
  const double rr = mesh_sum(part + 2 * nch, meshch[m], meshch[m + 1]);
  const double err = sqrt(rr) / s[0];
  if (err < s[6] || !(err == err)) return;
  const double beta = zr / s[1 + ((it - 1) & 1)];
  for (long long i = 3 * chn0[ch] + threadIdx.x; i < 3 * chn1[ch]; i += blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// per mesh: close the iteration (store z.r, err; converged -> done)
__global__ void bpcg_step_kernel(int nmesh, const int *__restrict__ meshch,
                                 const double *__restrict__ part, int nch,
                                 double *__restrict__ mscal, int *__restrict__ mdone,
                                 int *__restrict__ ndone) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmesh) return;
  int *d = mdone + 4 * m;
  if (d[0]) return;
  double *s = mscal + 8 * m;
  const int it = d[2];
  const double zr = mesh_sum(part + nch, meshch[m], meshch[m + 1]);
  const double rr = mesh_sum(part + 2 * nch, meshch[m], meshch[m + 1]);
  const double err = sqrt(rr) / s[0];
  s[3] = err;
  if (err < s[6] || !(err == err)) {
    d[1] = !(err == err);
    d[0] = it;
    atomicAdd(ndone, 1);  // integer count of finished meshes
  } else {
    s[1 + (it & 1)] = zr;
    d[2] = it + 1;
  }
}

template <int G>
void bpcg_spmv_launch(Ctx &c) {
  if (c.mat_bits == 32)
    hipLaunchKernelGGL((bpcg_spmv_kernel<G, float>), dim3(c.nchunk), dim3(TPB), 0, c.stream,
                       c.d_chn0, c.d_chn1, c.d_chmesh, c.d_bptr, c.d_bcol, c.d_bval32, c.d_p,
                       c.d_w, c.d_part, c.d_mdone);
  else
    hipLaunchKernelGGL((bpcg_spmv_kernel<G, double>), dim3(c.nchunk), dim3(TPB), 0, c.stream,
                       c.d_chn0, c.d_chn1, c.d_chmesh, c.d_bptr, c.d_bcol, c.d_bval, c.d_p,
                       c.d_w, c.d_part, c.d_mdone);
}

void enqueue_batch_iteration(Ctx &c) {
  switch (spmv_group(c)) {
    case 64: bpcg_spmv_launch<64>(c); break;
    case 32: bpcg_spmv_launch<32>(c); break;
    case 16: bpcg_spmv_launch<16>(c); break;
    case 8: bpcg_spmv_launch<8>(c); break;
    default: bpcg_spmv_launch<4>(c); break;
  }
  if (c.precond == 1)
    hipLaunchKernelGGL(bpcg_update_bj_kernel, dim3(c.nchunk), dim3(TPB), 0, c.stream, c.d_chn0,
                       c.d_chn1, c.d_chmesh, c.d_meshch, c.d_x, c.d_r, c.d_z, c.d_p, c.d_w,
                       c.d_binv, c.d_part, c.nchunk, c.d_mscal, c.d_mdone);
  else
    hipLaunchKernelGGL(bpcg_update_kernel, dim3(c.nchunk), dim3(TPB), 0, c.stream, c.d_chn0,
                       c.d_chn1, c.d_chmesh, c.d_meshch, c.d_x, c.d_r, c.d_z, c.d_p, c.d_w,
                       c.d_dinv, c.d_part, c.nchunk, c.d_mscal, c.d_mdone);
  hipLaunchKernelGGL(bpcg_dir_kernel, dim3(c.nchunk), dim3(TPB), 0, c.stream, c.d_chn0,
                     c.d_chn1, c.d_chmesh, c.d_meshch, c.d_p, c.d_z, c.d_part, c.nchunk,
                     c.d_mscal, c.d_mdone);
  hipLaunchKernelGGL(bpcg_step_kernel, dim3((c.nbatch + 255) / 256), dim3(256), 0, c.stream,
                     c.nbatch, c.d_meshch, c.d_part, c.nchunk, c.d_mscal, c.d_mdone, c.d_ndone);
}

// Capture CHUNK_IT iterations once per context into a graph; replay it.
void run_chunk(Ctx &c, bool batch) {
  hipGraphExec_t &ge = batch ? c.g_batch : c.g_single;
  if (!c.use_graphs) {
    for (int k = 0; k < CHUNK_IT; k++) batch ? enqueue_batch_iteration(c) : enqueue_iteration(c);
    FGPU_CHECK(hipGetLastError());
    return;
  }
  int &bits = batch ? c.g_batch_bits : c.g_single_bits;
  const int key = c.mat_bits + 1000 * c.precond;
  if (ge && bits != key) {  // captured for the other matrix precision or preconditioner
    FGPU_CHECK(hipGraphExecDestroy(ge));
    ge = nullptr;
  }
  if (!ge) {
    bits = key;
    hipGraph_t g;
    FGPU_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < CHUNK_IT; k++) batch ? enqueue_batch_iteration(c) : enqueue_iteration(c);
    FGPU_CHECK(hipStreamEndCapture(c.stream, &g));
    FGPU_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    FGPU_CHECK(hipGraphDestroy(g));
  }
  FGPU_CHECK(hipGraphLaunch(ge, c.stream));
}

void prepare(Ctx &c, const double *b, const double *x0) {
  const int n = 3 * c.nnode;
  const int eb = (n + 255) / 256;
  hipStream_t s = c.stream;
  if (c.mat_bits == 32) ensure_bval32(c);
  hipLaunchKernelGGL(diag_inv_kernel, dim3(eb), dim3(256), 0, s, c.nnode, c.d_bdiag, c.d_bval,
                     c.d_dinv);
  if (c.precond == 1)
    hipLaunchKernelGGL(block_inv_kernel, dim3((c.nnode + 255) / 256), dim3(256), 0, s, c.nnode,
                       c.d_bdiag, c.d_bval, c.d_binv);
  launch_eq_to_nodal(c, b, c.d_bn);
  if (x0) {
    launch_eq_to_nodal(c, x0, c.d_x);
    launch_spmv_nodal(c, c.d_x, c.d_w);
    hipLaunchKernelGGL(residual_kernel, dim3(eb), dim3(256), 0, s, n, c.d_bn, c.d_w, c.d_r);
  } else {
    FGPU_CHECK(hipMemsetAsync(c.d_x, 0, sizeof(double) * n, s));
    FGPU_CHECK(hipMemcpyAsync(c.d_r, c.d_bn, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  }
  FGPU_CHECK(hipGetLastError());
}

void launch_pcg_init(Ctx &c) {
  if (c.precond == 1)
    hipLaunchKernelGGL(pcg_init_bj_kernel, dim3(RED), dim3(TPB), 0, c.stream, c.nnode, c.d_bn,
                       c.d_r, c.d_binv, c.d_z, c.d_p, c.d_part);
  else
    hipLaunchKernelGGL(pcg_init_kernel, dim3(RED), dim3(TPB), 0, c.stream, 3 * c.nnode, c.d_bn,
                       c.d_r, c.d_dinv, c.d_z, c.d_p, c.d_part);
}

}  // namespace

// Graph-safe start of a solve of K x = b from x0 = 0 (the Newton
// corrections): no memset or copy nodes, the tolerance read from d_tol.
void enqueue_pcg_start(Ctx &c, const double *b, const double *d_tol, int abs_exit) {
  const int n = 3 * c.nnode;
  const int eb = (n + 255) / 256;
  hipStream_t s = c.stream;
  if (c.mat_bits == 32) ensure_bval32(c);
  hipLaunchKernelGGL(diag_inv_kernel, dim3(eb), dim3(256), 0, s, c.nnode, c.d_bdiag, c.d_bval,
                     c.d_dinv);
  if (c.precond == 1)
    hipLaunchKernelGGL(block_inv_kernel, dim3((c.nnode + 255) / 256), dim3(256), 0, s, c.nnode,
                       c.d_bdiag, c.d_bval, c.d_binv);
  launch_eq_to_nodal(c, b, c.d_bn);
  launch_fill(c, c.d_x, n, 0.0);
  launch_copy(c, c.d_bn, c.d_r, n);
  launch_pcg_init(c);
  hipLaunchKernelGGL(pcg_start_dev_kernel, dim3(1), dim3(TPB), 0, s, c.d_part, c.d_scal,
                     c.d_done, d_tol, abs_exit);
  FGPU_CHECK(hipGetLastError());
}

// CHUNK_IT iterations, launches only (inside a graph being captured)
void enqueue_pcg_chunk(Ctx &c) {
  for (int k = 0; k < CHUNK_IT; k++) enqueue_iteration(c);
  FGPU_CHECK(hipGetLastError());
}

// After enqueue_pcg_start and one chunk: further chunks until the device
// flags convergence or maxit iterations ran, with run_pcg's results (the
// solution stays in the node vector d_x).
int pcg_continue(Ctx &c, int maxit, double rtol, int abs_exit, int *iters, double *relres) {
  hipStream_t s = c.stream;
  int launched = CHUNK_IT;
  while (true) {
    FGPU_CHECK(hipMemcpyAsync(c.h_done, c.d_done, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    FGPU_CHECK(hipMemcpyAsync(c.h_scal, c.d_scal, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    FGPU_CHECK(hipStreamSynchronize(s));
    if (c.h_done[0] || c.h_done[2] >= maxit || launched >= maxit) break;
    run_chunk(c, false);
    launched += CHUNK_IT;
  }
  if (c.h_done[0] < 0) {  // the immediate exits of run_pcg
    if (iters) *iters = 0;
    if (relres)
      *relres = (abs_exit && c.h_scal[4] < rtol) || c.h_scal[4] == 0.0 ? 0.0 : c.h_scal[3];
    return 1;
  }
  const int conv_it = c.h_done[0], broke = c.h_done[1];
  if (iters) *iters = conv_it ? conv_it : c.h_done[2];
  if (relres) *relres = c.h_scal[3];
  return conv_it > 0 && conv_it <= maxit && !broke;
}

// b, x, x0: equation vectors on the device.  The iteration runs on node
// vectors (3 nnode entries, fixed directions identically zero).
int run_pcg(Ctx &c, const double *b, double *x, const double *x0, int maxit, double rtol,
            int *iters, double *relres, bool abs_exit) {
  hipStream_t s = c.stream;
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (c.neq == 0) return 1;
  prepare(c, b, x0);
  launch_pcg_init(c);
  hipLaunchKernelGGL(pcg_start_kernel, dim3(1), dim3(TPB), 0, s, c.d_part, c.d_scal, c.d_done,
                     rtol);
  FGPU_CHECK(hipGetLastError());
  FGPU_CHECK(hipMemcpyAsync(c.h_scal, c.d_scal, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
  FGPU_CHECK(hipStreamSynchronize(s));
  // reference: |b| < tol returns at once (x stays at its start value)
  const bool early = (abs_exit && c.h_scal[4] < rtol) || c.h_scal[4] == 0.0 || c.h_scal[3] < rtol;
  int conv_it = 0, broke = 0, launched = 0;
  while (!early && launched < maxit) {
    run_chunk(c, false);
    launched += CHUNK_IT;
    FGPU_CHECK(hipMemcpyAsync(c.h_done, c.d_done, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    FGPU_CHECK(hipMemcpyAsync(c.h_scal, c.d_scal, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    FGPU_CHECK(hipStreamSynchronize(s));
    if (c.h_done[0]) {
      conv_it = c.h_done[0];
      broke = c.h_done[1];
      break;
    }
    if (c.h_done[2] >= maxit) break;
  }
  // a chunk may run past maxit: the iteration count is what the device did
  launch_nodal_to_eq(c, c.d_x, x);
  FGPU_CHECK(hipStreamSynchronize(s));
  if (early) {
    if (relres) *relres = (abs_exit && c.h_scal[4] < rtol) || c.h_scal[4] == 0.0 ? 0.0 : c.h_scal[3];
    return 1;
  }
  if (iters) *iters = conv_it ? conv_it : c.h_done[2];
  if (relres) *relres = c.h_scal[3];
  return conv_it > 0 && conv_it <= maxit && !broke;
}

int run_pcg_batch(Ctx &c, const double *b, double *x, int maxit, double rtol, int *iters,
                  double *relres) {
  hipStream_t s = c.stream;
  if (c.nbatch <= 0) throw Error("not a batch context");
  prepare(c, b, nullptr);
  if (c.precond == 1)
    hipLaunchKernelGGL(bpcg_init_bj_kernel, dim3(c.nchunk), dim3(TPB), 0, s, c.d_chn0, c.d_chn1,
                       c.d_bn, c.d_r, c.d_binv, c.d_z, c.d_p, c.d_part, c.nchunk);
  else
    hipLaunchKernelGGL(bpcg_init_kernel, dim3(c.nchunk), dim3(TPB), 0, s, c.d_chn0, c.d_chn1,
                       c.d_bn, c.d_r, c.d_dinv, c.d_z, c.d_p, c.d_part, c.nchunk);
  hipLaunchKernelGGL(bpcg_start_kernel, dim3((c.nbatch + 255) / 256), dim3(256), 0, s, c.nbatch,
                     c.d_meshch, c.d_part, c.nchunk, c.d_mscal, c.d_mdone, c.d_ndone, rtol);
  hipLaunchKernelGGL(bpcg_count_kernel, dim3(1), dim3(64), 0, s, c.nbatch, c.d_mdone, c.d_ndone);
  FGPU_CHECK(hipGetLastError());
  int launched = 0;
  while (true) {
    FGPU_CHECK(hipMemcpyAsync(c.h_done, c.d_ndone, sizeof(int), hipMemcpyDeviceToHost, s));
    FGPU_CHECK(hipStreamSynchronize(s));
    if (c.h_done[0] >= c.nbatch || launched >= maxit) break;
    run_chunk(c, true);
    launched += CHUNK_IT;
  }
  launch_nodal_to_eq(c, c.d_x, x);
  std::vector<int> md(4 * (size_t)c.nbatch);
  std::vector<double> ms(8 * (size_t)c.nbatch);
  FGPU_CHECK(hipMemcpyAsync(md.data(), c.d_mdone, md.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  FGPU_CHECK(hipMemcpyAsync(ms.data(), c.d_mscal, ms.size() * sizeof(double),
                            hipMemcpyDeviceToHost, s));
  FGPU_CHECK(hipStreamSynchronize(s));
  int nconv = 0;
  for (int m = 0; m < c.nbatch; m++) {
    const int d0 = md[4 * m], it = d0 > 0 ? d0 : (d0 < 0 ? 0 : md[4 * m + 2] - 1);
    const bool ok = d0 < 0 || (d0 > 0 && d0 <= maxit && !md[4 * m + 1]);
    nconv += ok;
    if (iters) iters[m] = it;
    if (relres) relres[m] = ms[8 * m + 4] < rtol ? 0.0 : ms[8 * m + 3];
  }
  return nconv;
}

}  // namespace fgpu
