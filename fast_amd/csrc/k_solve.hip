// k_solve.hip - deterministic assembly, internal-force gather, block SpMV,
// vector helpers and format conversions of libfgpu (gfx950).
//
// Global matrix storage: node-block rows of 3x3 FP64 blocks ("block CSR"),
// one row per node, blocks in increasing neighbour order.  A fixed direction
// keeps a unit diagonal and no coupling, so the node-space system restricted
// to the free directions is exactly the reference's equation system; the
// equation CSR (rows = equations) is produced on export.  Per stored block
// the SpMV reads 72 B of values and one 4 B column index (12 B per nonzero
// in scalar CSR).
//
// Assembly replaces the element loop + cElement::AddGlobMat of
// ctrl/ctrl.cpp:886-975 / elc/element.cpp:2975-3002: one thread per upper
// node block (A <= B) sums, in FP64 and in element order, the FP32 element
// blocks listed for it (a segmented reduction over contributions pre-sorted
// by row block), then stores the 3x3 result at (A, B) and its transpose at
// (B, A), each as 9 contiguous doubles.  Every value is written exactly once;
// no atomics.
//
// The PCG itself lives in k_pcg.hip.
#include <cmath>

#include <algorithm>

#include "fgpu_dev.h"

namespace fgpu {
namespace {

// Fixed directions, in both assembly kernels: entry (i, j) of block (A, B)
// is kept where direction i of A and direction j of B are free (fmask bits),
// otherwise it is 1 on the diagonal of a diagonal block and 0 elsewhere.

// One thread per upper block (A, B): sum its contributions (sorted by
// element), NB at a time: the NB codes are read, then all NB blocks are
// loaded, then added in order (the loads of a batch are in flight together).
// XCD: consecutive workgroups are dispatched round-robin to the 8 XCDs, each
// with its own L2; with XCD set, workgroup w takes the range
// (w % 8) * (grid / 8) + w / 8 (grid a multiple of 8), so each XCD sweeps one
// contiguous eighth of the block rows and the element blocks its rows share
// stay in that XCD's L2.
// T: the element matrices' storage type (FP32; FP64 where the element
// kernels compute and keep FP64)
template <int NB, bool XCD, class T = float>
__global__ __launch_bounds__(256) void assemble_kernel(
    long long nblk, const unsigned char *__restrict__ bmask, const int *__restrict__ posAB,
    const int *__restrict__ posBA, const int *__restrict__ cptr,
    const int *__restrict__ contrib, const T *__restrict__ Ke, KeLayout L,
    double *__restrict__ bval) {
  __shared__ double sv[256 * 9];  // the workgroup's summed blocks, masked
  __shared__ int sAB[256], sBA[256];
  const long long w = XCD ? (long long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3)
                          : (long long)blockIdx.x;
  const int tid = threadIdx.x;
  const long long k0 = w * blockDim.x;
  const long long k = k0 + tid;
  const int nk = (int)min((long long)blockDim.x, nblk - k0);  // blocks of this workgroup
  if (nk <= 0) return;  // the whole workgroup
  double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int c0 = k < nblk ? cptr[k] : 0, c1 = k < nblk ? cptr[k + 1] : 0;
  for (int tb = c0; tb < c1; tb += NB) {
    int code[NB];
#pragma unroll
    for (int u = 0; u < NB; u++) code[u] = tb + u < c1 ? contrib[tb + u] : -1;
    T v[NB][9];
#pragma unroll
    for (int u = 0; u < NB; u++) {
      if (code[u] < 0) continue;
      const int x = code[u] >> 1;
      const long e = x / L.nsym;
      const int sb = x - (int)e * L.nsym;
      const T *p = Ke + L.offset(e, sb);
#pragma unroll
      for (int m = 0; m < 9; m++) v[u][m] = p[m];
    }
#pragma unroll
    for (int u = 0; u < NB; u++) {
      if (code[u] < 0) continue;
      if (code[u] & 1) {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) s[3 * i + j] += (double)v[u][3 * j + i];
      } else {
        for (int m = 0; m < 9; m++) s[m] += (double)v[u][m];
      }
    }
  }
  if (k < nblk) {
    const unsigned mk = bmask[k], mA = mk & 7u, mB = (mk >> 3) & 7u;
    const bool diag = (mk >> 6) & 1u;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        const bool live = ((mA >> i) & 1u) && ((mB >> j) & 1u);
        sv[9 * tid + 3 * i + j] = live ? s[3 * i + j] : ((diag && i == j) ? 1.0 : 0.0);
      }
    sAB[tid] = posAB[k];
    sBA[tid] = diag ? -1 : posBA[k];
  }
  __syncthreads();
  // write-out value by value (consecutive lanes store consecutive doubles):
  // value i is component m = i % 9 of the workgroup's block t = i / 9, at
  // (A, B) as summed and at (B, A) transposed
  for (int i = tid; i < 9 * nk; i += blockDim.x) {
    const int t = i / 9, m = i - 9 * t;
    bval[9 * (size_t)sAB[t] + m] = sv[i];
    if (sBA[t] >= 0) bval[9 * (size_t)sBA[t] + m] = sv[9 * t + 3 * (m % 3) + m / 3];
  }
}

// Row assembly (elements of at most 64 nodes): one wave per node row A, its
// upper blocks accumulated in LDS, 9 doubles per block, one column per block
// in the row's order.  The wave walks the elements of A in element order, U
// at a time; for element e with A at local node a, lane b reads the stored
// block coupling a and b (row a of the element's blocks when b >= a, read as
// it lies; column a, transposed, when b < a) and adds it to the column
// rslot[(t, b)] names, skipping nodes b that precede A (0xFFFF).  Lanes of
// one element add to distinct columns; elements follow one another in
// order, so every column receives its contributions in element order, as the
// per-block lists of assemble_kernel hold them: the same sums, bit for bit.
// A lane's reads of consecutive b are consecutive blocks of one element.
template <class T, int U>
__global__ __launch_bounds__(64) void assemble_rows_kernel(
    int nn, const int *__restrict__ nptr, const int *__restrict__ nadj,
    const unsigned short *__restrict__ rslot, const long long *__restrict__ kptr,
    const int *__restrict__ blkB, const int *__restrict__ posAB, const int *__restrict__ posBA,
    const T *__restrict__ Ke, KeLayout L, const unsigned char *__restrict__ fmask,
    double *__restrict__ bval, int maxs) {
  extern __shared__ double acc[];  // [9][maxs], then [maxs] positions of (B, A), [maxs] masks
  int *sBA = reinterpret_cast<int *>(acc + 9 * maxs);
  unsigned char *sMB = reinterpret_cast<unsigned char *>(sBA + maxs);
  const int A = blockIdx.x;
  const int b = threadIdx.x;
  const long long k0 = kptr[A];
  const int ns = (int)(kptr[A + 1] - k0);
  for (int j = b; j < ns; j += 64) {
    for (int m = 0; m < 9; m++) acc[m * maxs + j] = 0.0;
    sBA[j] = posBA[k0 + j];
    sMB[j] = fmask[blkB[k0 + j]];
  }
  __syncthreads();
  const int t0 = nptr[A], t1 = nptr[A + 1];
  for (int t = t0; t < t1; t += U) {
    int col[U], tr[U];
    T v[U][9];
#pragma unroll
    for (int u = 0; u < U; u++) {
      col[u] = 0xFFFF;
      tr[u] = 0;
      if (t + u < t1 && b < nn) {
        const int x = nadj[t + u];
        const int e = x / nn, a = x - e * nn;
        col[u] = rslot[(size_t)(t + u) * nn + b];
        // block (a, b) lies in row a of the element's blocks; it is read
        // without waiting on the column table (its column exists unless the
        // numbering runs against the element's local order)
        if (b >= a) {
          const T *p = Ke + L.offset(e, sym_index(a, b, nn));
#pragma unroll
          for (int m = 0; m < 9; m++) v[u][m] = p[m];
        } else if (col[u] != 0xFFFF) {  // block (b, a), read transposed
          tr[u] = 1;
          const T *p = Ke + L.offset(e, sym_index(b, a, nn));
#pragma unroll
          for (int m = 0; m < 9; m++) v[u][m] = p[m];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (col[u] == 0xFFFF) continue;
      double *c = acc + col[u];
      if (tr[u]) {
        for (int q = 0; q < 3; q++)
          for (int r = 0; r < 3; r++) c[(3 * q + r) * maxs] += (double)v[u][3 * r + q];
      } else {
        for (int m = 0; m < 9; m++) c[m * maxs] += (double)v[u][m];
      }
    }
  }
  __syncthreads();
  // write-out value by value, so that consecutive lanes store consecutive
  // doubles: value i is component m = i % 9 of block j = i / 9.  The row's
  // upper blocks are consecutive in the node-block rows, the first being the
  // diagonal (A, A); the transposed copy (B, A) of block j goes to sBA[j].
  // Fixed directions as above.
  const unsigned mA = fmask[A];
  const size_t pAB0 = ns > 0 ? 9 * (size_t)posAB[k0] : 0;
  for (int i = b; i < 9 * ns; i += 64) {
    const int j = i / 9, m = i - 9 * j;
    const int q = m / 3, r = m - 3 * q;
    const unsigned mB = sMB[j];
    const bool live = (mA >> q) & (mB >> r) & 1u;
    bval[pAB0 + i] = live ? acc[m * maxs + j] : ((j == 0 && q == r) ? 1.0 : 0.0);
    if (j > 0) {  // (B, A)[q][r] = (A, B)[r][q]
      const bool liveT = (mA >> r) & (mB >> q) & 1u;
      bval[9 * (size_t)sBA[j] + m] = liveT ? acc[(3 * r + q) * maxs + j] : 0.0;
    }
  }
}

// One thread per node direction (3 nnode threads; the three of a node read
// one 24-byte element vector entry together), its element contributions
// added in adjacency order, eight in flight at a time.
__global__ __launch_bounds__(256) void gather_fint_kernel(
    int nnode, int nn, const int *__restrict__ nptr, const int *__restrict__ nadj,
    const double *__restrict__ ge, const int *__restrict__ eqmap, double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nnode) return;
  const int A = i / 3, d = i - 3 * A;
  double s = 0;
  const int t0 = nptr[A], t1 = nptr[A + 1];
  for (int tb = t0; tb < t1; tb += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (tb + u < t1) v[u] = ge[(size_t)nadj[tb + u] * 3 + d];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (tb + u < t1) s += v[u];
  }
  const int q = eqmap[i];
  if (q >= 0) g[q] = s;
}

// The same sums from node-sorted element vectors (entry t of node A's
// adjacency at ge[3 t]): each node reads one contiguous run, in the same
// order, so the results are the unsorted gather's bits.
__global__ __launch_bounds__(256) void gather_sorted_kernel(
    int nnode, const int *__restrict__ nptr, const double *__restrict__ ge,
    const int *__restrict__ eqmap, double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nnode) return;
  const int A = i / 3, d = i - 3 * A;
  double s = 0;
  const int t0 = nptr[A], t1 = nptr[A + 1];
  for (int tb = t0; tb < t1; tb += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (tb + u < t1) v[u] = ge[(size_t)(tb + u) * 3 + d];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (tb + u < t1) s += v[u];
  }
  const int q = eqmap[i];
  if (q >= 0) g[q] = s;
}

__global__ __launch_bounds__(256) void to_f32_kernel(long long n, const double *__restrict__ a,
                                                     float *__restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// cControl::PrintAvrNodalStress (ctrl/ctrl.cpp:2397-2470): the element-nodal
// stresses of each node summed in element order, divided by their count.
__global__ __launch_bounds__(256) void avg_stress_kernel(int nnode, const int *__restrict__ nptr,
                                                         const int *__restrict__ nadj,
                                                         const double *__restrict__ sig_node,
                                                         double *__restrict__ out) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nnode) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const int t0 = nptr[A], t1 = nptr[A + 1];
  for (int t = t0; t < t1; t++) {
    const double *p = sig_node + (size_t)nadj[t] * 6;
    for (int k = 0; k < 6; k++) s[k] += p[k];
  }
  const int n = t1 - t0;
  for (int k = 0; k < 6; k++) out[6 * (size_t)A + k] = n > 0 ? s[k] / (double)n : 0.0;
}

__global__ void add_diag_kernel(int n, const int *__restrict__ eq, const double *__restrict__ v,
                                const int *__restrict__ eqnode,
                                const long long *__restrict__ bdiag, double *__restrict__ bval) {
  // sequential in k so repeated equations add in order
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int k = 0; k < n; k++) {
      const int nd = eqnode[eq[k] - 1];
      bval[9 * bdiag[nd / 3] + 4 * (nd % 3)] += v[k];
    }
}

template <class T>
__global__ void expand_Ke_kernel(int e0, int ne, int nn, KeLayout L, const T *__restrict__ Ke,
                                 float *__restrict__ out) {
  const long long nd = 3LL * nn;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)ne * nd * nd) return;
  const int el = (int)(idx / (nd * nd));
  const int rc = (int)(idx % (nd * nd));
  const int r = rc / (int)nd, cc = rc % (int)nd;
  const int a = r / 3, i = r % 3, b = cc / 3, j = cc % 3;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const T *p = Ke + L.offset(e0 + el, sym_index(lo, hi, nn));
  out[idx] = (float)(a <= b ? p[3 * i + j] : p[3 * j + i]);
}

__global__ void eq_to_nodal_kernel(int nnode, const int *__restrict__ eqmap,
                                   const double *__restrict__ x, double *__restrict__ u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nnode) return;
  const int q = eqmap[i];
  u[i] = q >= 0 ? x[q] : 0.0;
}

__global__ void nodal_to_eq_kernel(int nnode, const int *__restrict__ eqmap,
                                   const double *__restrict__ u, double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nnode) return;
  const int q = eqmap[i];
  if (q >= 0) x[q] = u[i];
}

// equation CSR values <-> node-block values (export / import)
__global__ void bsr_to_csr_kernel(int nnode, const long long *__restrict__ bptr,
                                  const int *__restrict__ bcol, const int *__restrict__ bcoloff,
                                  const double *__restrict__ bval, const int *__restrict__ eqmap,
                                  const long long *__restrict__ rowptr, double *__restrict__ val,
                                  int to_csr) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nnode) return;
  for (long long k = bptr[A]; k < bptr[A + 1]; k++) {
    const int B = bcol[k];
    for (int i = 0; i < 3; i++) {
      const int qi = eqmap[3 * A + i];
      if (qi < 0) continue;
      long long p = rowptr[qi] + bcoloff[k];
      for (int j = 0; j < 3; j++) {
        if (eqmap[3 * B + j] < 0) continue;
        if (to_csr)
          val[p] = bval[9 * k + 3 * i + j];
        else
          const_cast<double *>(bval)[9 * k + 3 * i + j] = val[p];
        p++;
      }
    }
    if (!to_csr)  // fixed directions: unit diagonal, no coupling
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          if (eqmap[3 * A + i] < 0 || eqmap[3 * B + j] < 0)
            const_cast<double *>(bval)[9 * k + 3 * i + j] = (A == B && i == j) ? 1.0 : 0.0;
  }
}

// equation CSR columns: rows of node A list the free directions of its
// neighbours in node order (the same list for each free direction of A)
__global__ void fill_eq_cols_kernel(int nnode, const long long *__restrict__ bptr,
                                    const int *__restrict__ bcol, const int *__restrict__ eqmap,
                                    const long long *__restrict__ rowptr, int *__restrict__ col) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nnode) return;
  for (int i = 0; i < 3; i++) {
    const int qi = eqmap[3 * A + i];
    if (qi < 0) continue;
    long long p = rowptr[qi];
    for (long long k = bptr[A]; k < bptr[A + 1]; k++) {
      const int B = bcol[k];
      for (int j = 0; j < 3; j++) {
        const int qj = eqmap[3 * B + j];
        if (qj >= 0) col[p++] = qj;
      }
    }
  }
}

__global__ __launch_bounds__(256) void axpby_kernel(int n, double a, const double *__restrict__ x,
                                                    double b, const double *__restrict__ y,
                                                    double *__restrict__ z) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    z[i] = a * x[i] + b * (y ? y[i] : 0.0);
}

__global__ __launch_bounds__(256) void dot_partial_kernel(int n, const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          double *__restrict__ part) {
  __shared__ double sh[16];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void dot_finish_kernel(const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[16];
  double s = reduce_partials(part, RED, sh);
  if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(256) void copy_kernel(long long n, const double *__restrict__ x,
                                                   double *__restrict__ y) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i];
}

// fills for code that may be captured into a graph (no memset nodes there)
template <class T>
__global__ __launch_bounds__(256) void fill_kernel(long long n, T v, T *__restrict__ p) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}


template <int G>
__global__ __launch_bounds__(256) void bspmv_kernel(int nnode, const long long *__restrict__ bptr,
                                                    const int *__restrict__ bcol,
                                                    const double *__restrict__ bval,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ y) {
  const int rows_per_block = blockDim.x / G;
  bspmv_rows<G>((long long)blockIdx.x * rows_per_block, nnode,
                (long long)gridDim.x * rows_per_block, bptr, bcol, bval, x, y);
}

// y = |K| |x| on node vectors (one thread per node direction): the scale of
// the rounding in a residual f - K x, for backward-error tests
__global__ __launch_bounds__(256) void abs_spmv_kernel(int nnode, const long long *__restrict__ bptr,
                                                       const int *__restrict__ bcol,
                                                       const double *__restrict__ bval,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nnode) return;
  const int A = i / 3, d = i - 3 * A;
  double s = 0.0;
  for (long long k = bptr[A]; k < bptr[A + 1]; k++) {
    const double *b = bval + 9 * k + 3 * d;
    const double *xb = x + 3 * (size_t)bcol[k];
    s += fabs(b[0]) * fabs(xb[0]) + fabs(b[1]) * fabs(xb[1]) + fabs(b[2]) * fabs(xb[2]);
  }
  y[i] = s;
}

template <int G>
void bspmv_launch(Ctx &c, const double *x, double *y) {
  const int rows_per_block = TPB / G;
  long long want = (c.nnode + rows_per_block - 1) / rows_per_block;
  int grid = (int)std::min<long long>(want, 4 * RED);
  hipLaunchKernelGGL(bspmv_kernel<G>, dim3(grid), dim3(TPB), 0, c.stream, c.nnode, c.d_bptr,
                     c.d_bcol, c.d_bval, x, y);
}

}  // namespace

// assembly modes 0..7: contribution loads in batches of 1, 4, 8, 16
// (mode & 3), XCD-contiguous workgroup order when mode & 4; mode 8: one wave
// per node row (meshes with the row tables)
std::vector<int> assembly_options(const Ctx &c) {
  std::vector<int> m = {0, 1, 2, 3, 4, 5, 6, 7};
  if (c.d_rslot) m.push_back(8);
  return m;
}

// U elements per batch of loads (FP32 element matrices; FP64 ones take half)
template <int U>
void launch_rows(Ctx &c) {
  if (!c.d_rslot) throw Error("row assembly: no row tables for this mesh");
  const size_t lds = (sizeof(double) * 9 + sizeof(int) + 1) * (size_t)c.row_maxs;
  if (c.d_Ke64)
    hipLaunchKernelGGL((assemble_rows_kernel<double, U / 2>), dim3(c.nnode), dim3(64), lds,
                       c.stream, c.nn, c.d_nptr, c.d_nadj, c.d_rslot, c.d_kptr, c.d_blkB,
                       c.d_offAB, c.d_offBA, c.d_Ke64, c.kel, c.d_fmask, c.d_bval, c.row_maxs);
  else
    hipLaunchKernelGGL((assemble_rows_kernel<float, U>), dim3(c.nnode), dim3(64), lds, c.stream,
                       c.nn, c.d_nptr, c.d_nadj, c.d_rslot, c.d_kptr, c.d_blkB, c.d_offAB,
                       c.d_offBA, c.d_Ke, c.kel, c.d_fmask, c.d_bval, c.row_maxs);
}

template <int NB, bool XCD>
void launch_contrib(Ctx &c) {
  long long grid = (c.nblk + 255) / 256;
  if (XCD) grid = (grid + 7) / 8 * 8;  // the remap is a bijection on multiples of 8
  if (c.d_Ke64)
    hipLaunchKernelGGL((assemble_kernel<NB, XCD, double>), dim3((unsigned)grid), dim3(256), 0,
                       c.stream, c.nblk, c.d_bmask, c.d_offAB, c.d_offBA, c.d_cptr, c.d_contrib,
                       c.d_Ke64, c.kel, c.d_bval);
  else
    hipLaunchKernelGGL((assemble_kernel<NB, XCD, float>), dim3((unsigned)grid), dim3(256), 0,
                       c.stream, c.nblk, c.d_bmask, c.d_offAB, c.d_offBA, c.d_cptr, c.d_contrib,
                       c.d_Ke, c.kel, c.d_bval);
}

void launch_assemble_K(Ctx &c) {
  switch (c.asm_mode < 0 ? 2 : c.asm_mode) {
    case 0: launch_contrib<1, false>(c); break;
    case 1: launch_contrib<4, false>(c); break;
    case 3: launch_contrib<16, false>(c); break;
    case 4: launch_contrib<1, true>(c); break;
    case 5: launch_contrib<4, true>(c); break;
    case 6: launch_contrib<8, true>(c); break;
    case 7: launch_contrib<16, true>(c); break;
    case 8: launch_rows<8>(c); break;
    default: launch_contrib<8, false>(c); break;
  }
  FGPU_CHECK(hipGetLastError());
}

void launch_gather_fint(Ctx &c, double *g) {
  // 64-thread workgroups: a B-spline mesh has few nodes (13,475 on the
  // 32x32x8 cubic patch), and its 632 workgroups spread over the CUs
  if (c.ge_sorted)
    hipLaunchKernelGGL(gather_sorted_kernel, dim3((3 * c.nnode + 63) / 64), dim3(64), 0,
                       c.stream, c.nnode, c.d_nptr, c.d_ge, c.d_eqmap, g);
  else
    hipLaunchKernelGGL(gather_fint_kernel, dim3((3 * c.nnode + 63) / 64), dim3(64), 0,
                       c.stream, c.nnode, c.nn, c.d_nptr, c.d_nadj, c.d_ge, c.d_eqmap, g);
  FGPU_CHECK(hipGetLastError());
}

void ensure_bval32(Ctx &c) {
  if (c.bval32_valid) return;
  const long long n = c.nbf * 9;
  if (!c.d_bval32) FGPU_CHECK(hipMalloc(&c.d_bval32, sizeof(float) * n));
  hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n,
                     c.d_bval, c.d_bval32);
  FGPU_CHECK(hipGetLastError());
  c.bval32_valid = true;
}

void launch_avg_stress(Ctx &c, const double *sig_node, double *sig_avg) {
  hipLaunchKernelGGL(avg_stress_kernel, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_nptr, c.d_nadj, sig_node, sig_avg);
  FGPU_CHECK(hipGetLastError());
}

void launch_add_diag(Ctx &c, int n, const int *d_eq, const double *d_v) {
  hipLaunchKernelGGL(add_diag_kernel, dim3(1), dim3(64), 0, c.stream, n, d_eq, d_v, c.d_eqnode,
                     c.d_bdiag, c.d_bval);
  FGPU_CHECK(hipGetLastError());
}

void launch_expand_Ke(Ctx &c, int e0, int ne, float *out) {
  long long tot = (long long)ne * 9 * c.nn * c.nn;
  if (c.d_Ke64)
    hipLaunchKernelGGL(expand_Ke_kernel<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                       0, c.stream, e0, ne, c.nn, c.kel, c.d_Ke64, out);
  else
    hipLaunchKernelGGL(expand_Ke_kernel<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       c.stream, e0, ne, c.nn, c.kel, c.d_Ke, out);
  FGPU_CHECK(hipGetLastError());
}

void ensure_eq_csr(Ctx &c) {
  if (c.d_col) return;
  FGPU_CHECK(hipMalloc(&c.d_col, std::max<long long>(c.nnz, 1) * sizeof(int)));
  hipLaunchKernelGGL(fill_eq_cols_kernel, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_bptr, c.d_bcol, c.d_eqmap, c.d_rowptr, c.d_col);
  FGPU_CHECK(hipGetLastError());
}

void launch_csr_convert(Ctx &c, double *d_csrval, int to_csr) {
  hipLaunchKernelGGL(bsr_to_csr_kernel, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_bptr, c.d_bcol, c.d_bcoloff, c.d_bval, c.d_eqmap, c.d_rowptr,
                     d_csrval, to_csr);
  FGPU_CHECK(hipGetLastError());
}

void launch_eq_to_nodal(Ctx &c, const double *x, double *u) {
  hipLaunchKernelGGL(eq_to_nodal_kernel, dim3((3 * c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_eqmap, x, u);
  FGPU_CHECK(hipGetLastError());
}

void launch_nodal_to_eq(Ctx &c, const double *u, double *x) {
  hipLaunchKernelGGL(nodal_to_eq_kernel, dim3((3 * c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_eqmap, u, x);
  FGPU_CHECK(hipGetLastError());
}

void launch_spmv_nodal(Ctx &c, const double *x, double *y) {
  switch (spmv_group(c)) {
    case 64: bspmv_launch<64>(c, x, y); break;
    case 32: bspmv_launch<32>(c, x, y); break;
    case 16: bspmv_launch<16>(c, x, y); break;
    case 8: bspmv_launch<8>(c, x, y); break;
    default: bspmv_launch<4>(c, x, y); break;
  }
  FGPU_CHECK(hipGetLastError());
}

void launch_abs_spmv_nodal(Ctx &c, const double *x, double *y) {
  hipLaunchKernelGGL(abs_spmv_kernel, dim3((3 * c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, c.d_bptr, c.d_bcol, c.d_bval, x, y);
  FGPU_CHECK(hipGetLastError());
}

void launch_axpby(Ctx &c, int n, double a, const double *x, double b, const double *y,
                  double *z) {
  int grid = std::min(RED, (n + 255) / 256);
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid), dim3(256), 0, c.stream, n, a, x, b, y, z);
  FGPU_CHECK(hipGetLastError());
}

template <class T>
static void fill_launch(Ctx &c, T *p, long long n, T v) {
  if (n <= 0) return;
  const long long blocks = std::min<long long>(4 * RED, (n + 255) / 256);
  hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, c.stream, n, v, p);
  FGPU_CHECK(hipGetLastError());
}
void launch_fill(Ctx &c, double *p, long long n, double v) { fill_launch(c, p, n, v); }
void launch_fill(Ctx &c, int *p, long long n, int v) { fill_launch(c, p, n, v); }

void launch_copy(Ctx &c, const double *x, double *y, long long n) {
  if (n <= 0) return;
  const long long blocks = std::min<long long>(4 * RED, (n + 255) / 256);
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, c.stream, n, x, y);
  FGPU_CHECK(hipGetLastError());
}

void enqueue_dot(Ctx &c, const double *x, const double *y, int n, double *out) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(RED), dim3(256), 0, c.stream, n, x, y,
                     c.d_part + 3 * RED);
  hipLaunchKernelGGL(dot_finish_kernel, dim3(1), dim3(256), 0, c.stream, c.d_part + 3 * RED, out);
  FGPU_CHECK(hipGetLastError());
}

double dev_dot(Ctx &c, const double *x, const double *y, int n) {
  enqueue_dot(c, x, y, n, c.d_scal + 8);
  FGPU_CHECK(hipMemcpyAsync(c.h_scal + 8, c.d_scal + 8, sizeof(double), hipMemcpyDeviceToHost,
                            c.stream));
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  return c.h_scal[8];
}


}  // namespace fgpu
