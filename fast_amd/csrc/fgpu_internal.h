// fgpu_internal.h - context layout and kernel launchers shared by the
// translation units of libfgpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fgpu.h"

#define FGPU_CHECK(call)                                                     \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      throw fgpu::Error(std::string(#call) + ": " + hipGetErrorString(e_));  \
    }                                                                        \
  } while (0)

namespace fgpu {

struct Error {
  std::string msg;
  explicit Error(std::string m) : msg(std::move(m)) {}
};

// Number of stored symmetric 3x3 node blocks of an element (a <= b).
inline int nsym_blocks(int nn) { return nn * (nn + 1) / 2; }
// Packed index of the upper block (a <= b), row-major upper triangle.
__host__ __device__ inline int sym_index(int a, int b, int nn) {
  return a * nn - (a * (a - 1)) / 2 + (b - a);
}

// Jacobian bookkeeping of the element kernels, status[4] on the device:
// [0] points with det J < 0, [1] failed stress updates, [3] points with
// det J = 0 or not finite.  A mesh whose points all have det J < 0 (a
// left-handed parametrization) is accepted and integrated with |det J|; some
// negative points among positive ones mean inverted elements and are an
// error.  (The reference integrates with the signed det J: its stiffness and
// body forces then both change sign, so its displacements under body forces
// agree with these, and under nodal or face loads they come out negated.)
#ifdef __HIPCC__
template <class T>
__device__ inline void count_jacobian(int *status, T det, int n = 1) {
  if (det < T(0))
    atomicAdd(&status[0], n);
  else if (!(det > T(0)))
    atomicAdd(&status[3], n);
}

// Stress output at one point from L = grad u (L_ij = d u_i / d X_j), Voigt
// order xx yy zz xy xz yz (tensor components).  Linear (p0 = E, p1 = nu):
// sigma = lambda tr(eps) I + 2 mu eps, eps = (L + L^T) / 2.  St.Venant-
// Kirchhoff: second Piola-Kirchhoff S = lambda tr(E) I + 2 mu E with the
// Green strain E = (L + L^T + L^T L) / 2, the reference's s = C e with
// engineering shears.  Neo-Hooke (p0 = mu, p1 = lambda): S = mu (I - C^-1) +
// lambda ln J C^-1, C = F^T F.  Returns false when det C <= 0 (S then 0).
template <int FORM>
__device__ inline bool point_stress(const double L[3][3], double p0, double p1, double s[6]) {
  if (FORM != FGPU_TL_NEOHOOKE) {
    const double lam = p0 * p1 / ((1.0 + p1) * (1.0 - 2.0 * p1)), mu = 0.5 * p0 / (1.0 + p1);
    double E[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        E[i][j] = 0.5 * (L[i][j] + L[j][i]);
        if (FORM == FGPU_TL_SVK)
          E[i][j] += 0.5 * (L[0][i] * L[0][j] + L[1][i] * L[1][j] + L[2][i] * L[2][j]);
      }
    const double tr = E[0][0] + E[1][1] + E[2][2];
    s[0] = lam * tr + 2.0 * mu * E[0][0];
    s[1] = lam * tr + 2.0 * mu * E[1][1];
    s[2] = lam * tr + 2.0 * mu * E[2][2];
    s[3] = 2.0 * mu * E[0][1];
    s[4] = 2.0 * mu * E[0][2];
    s[5] = 2.0 * mu * E[1][2];
    return true;
  }
  double F[3][3], C[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) F[i][j] = L[i][j] + (i == j ? 1.0 : 0.0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
  // adjugate of the symmetric C
  const double a00 = C[1][1] * C[2][2] - C[1][2] * C[2][1], a01 = C[0][2] * C[2][1] - C[0][1] * C[2][2],
               a02 = C[0][1] * C[1][2] - C[0][2] * C[1][1], a11 = C[0][0] * C[2][2] - C[0][2] * C[2][0],
               a12 = C[0][2] * C[1][0] - C[0][0] * C[1][2], a22 = C[0][0] * C[1][1] - C[0][1] * C[1][0];
  const double det = C[0][0] * a00 + C[1][0] * a01 + C[2][0] * a02;
  if (!(det > 0.0)) {
    for (int k = 0; k < 6; k++) s[k] = 0.0;
    return false;
  }
  const double id = 1.0 / det, lnJ = 0.5 * log(det);
  const double ci[6] = {a00 * id, a11 * id, a22 * id, a01 * id, a02 * id, a12 * id};
  for (int k = 0; k < 6; k++) s[k] = (k < 3 ? p0 : 0.0) + (p1 * lnJ - p0) * ci[k];
  return true;
}
#endif

// Element-matrix storage: tiles of T elements; inside a tile, block-major:
//   Ke[((e / T) * nsym + blk) * T * 9 + (e % T) * 9 + (3 i + j)]
// so that one element's 3x3 block is 9 contiguous floats (the assembly
// gather reads it as a unit) and a wave of hex8 elements stores coalesced.
struct KeLayout {
  int T;     // elements per tile (64 for hex8, 1 for B-spline solids)
  int nsym;  // stored blocks per element
  __host__ __device__ size_t offset(long e, int blk) const {
    return ((size_t)(e / T) * nsym + blk) * (size_t)T * 9 + (size_t)(e % T) * 9;
  }
};

// One group of uniform surface loads (fgpu_add_surface_loads), on the device
struct SurfaceGroup {
  int nface = 0, deg[2] = {1, 1}, nip = 2;
  int shape = 0;              // 0 tensor product, 1 8-node serendipity (20-node brick faces)
  long long row0 = 0;         // first face-vector row of the group (faces x nodes)
  int *d_nodes = nullptr;     // [nface * (p1+1)(p2+1)]
  double *d_ext = nullptr;    // [nface * ((p1+1)^2 + (p2+1)^2)] or null (identity)
  int *d_local = nullptr;     // [nface]
  double *d_q = nullptr;      // [nface * 3]
  double *d_gp = nullptr;     // [nip][2] Gauss points and weights (FP64)
  std::vector<int> h_nodes;   // host copy (the gather lists are rebuilt from it)
};

struct Ctx {
  // ---- mesh (host copies)
  int nnode = 0, nelem = 0, shape = 0, nn = 0, form = 0;
  int deg[3] = {1, 1, 1}, nip[3] = {2, 2, 2};
  int neq = 0;
  std::vector<int> eqmap;  // [nnode*3] 0-based equation or -1
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) for f_int beside K, its status counters
  // and the fork event
  hipStream_t stream2 = nullptr;
  int *d_status2 = nullptr;  // d_status + 4 (one allocation)
  int *h_status = nullptr;   // pinned mirror of both status arrays [8]
  hipEvent_t ev_fork = nullptr;

  // ---- device mesh
  double *d_coords = nullptr;   // [nnode*3]
  int *d_conn = nullptr;        // [nelem*nn]
  int *d_eqmap = nullptr;       // [nnode*3]
  int *d_eqnode = nullptr;      // [neq] 3 node + direction of each equation
  double *d_mat = nullptr;      // [nelem*2]
  double *d_ext = nullptr;      // B-spline extraction rows [nelem*3*16]
  double *d_cpw = nullptr;      // [nnode] control-point weights
  float *d_gauss = nullptr;     // [3][nip][2] (point, weight) fp32
  double *d_gauss64 = nullptr;  // same, fp64

  // ---- element outputs
  KeLayout kel{64, 36};
  float *d_Ke = nullptr;        // element matrices (upper node blocks)
  double *d_Ke64 = nullptr;     // the same in FP64, for element kernels that compute in FP64
                                // (B-spline solids outside the tuned shapes); d_Ke unused then
  double *d_ge = nullptr;       // element internal forces [nelem*nn*3]
  double *d_pts = nullptr;      // per-point internal-force records (fused B-spline path)
  int *d_status = nullptr;      // [2]: bad-Jacobian count, stress failure
  double *d_TR = nullptr;       // stress extrapolation points -> nodes [nn*nq]
  // B-spline solids outside the tuned cubic kernels (k_bspg.hip): upper block
  // pairs, per-chunk scratch (gradients, basis values, F g / F^-T g, point data)
  int *d_gsym = nullptr;        // [nsym] int2 (a, b)
  double *d_gG = nullptr, *d_gR = nullptr, *d_gH = nullptr, *d_gPQ = nullptr;
  int gen_chunk = 0;            // elements per scratch chunk
  size_t gen_cap = 0, gen_cap_pq = 0;  // entries the scratch was sized for
  double *d_gnodes = nullptr;   // natural coordinates of the nodes, per direction
  int stress_nodes = 1;         // 1: extrapolated nodal stresses, 0: evaluated at the nodes

  // ---- global matrix: node-block rows of 3x3 FP64 blocks (all nodes; a
  // fixed direction has a unit diagonal and no coupling), the solver's
  // storage.  The equation CSR of the reference (rows/columns = equations,
  // full symmetric pattern) is kept as a pattern for export.
  long long nbf = 0;               // stored blocks
  long long *d_bptr = nullptr;     // [nnode+1]
  int *d_bcol = nullptr;           // [nbf] neighbour node
  int *d_bcoloff = nullptr;        // [nbf] column offset of the block in the equation rows
  long long *d_bdiag = nullptr;    // [nnode] position of the diagonal block
  unsigned char *d_fmask = nullptr;  // [nnode] bit d set = direction d free
  double *d_bval = nullptr;        // [nbf*9]
  // FP32 copy of d_bval for the PCG's products (mat_bits 32), refreshed on
  // the first solve after the matrix changes
  int mat_bits = 64;
  float *d_bval32 = nullptr;
  bool bval32_valid = false;
  long long nnz = 0;
  long long *d_rowptr = nullptr;  // [neq+1] equation CSR pattern
  int *d_col = nullptr;           // [nnz]

  // ---- assembly map: upper node blocks (A <= B)
  long long nblk = 0;
  unsigned char *d_bmask = nullptr;  // [nblk] free directions of A, of B, A == B
  int *d_blkB = nullptr;             // [nblk] column node of each upper block
  int *d_offAB = nullptr;  // [nblk] position of block (A, B) among the node blocks
  int *d_offBA = nullptr;  // [nblk] position of block (B, A)
  int *d_cptr = nullptr;         // [nblk+1] contribution segments
  int *d_contrib = nullptr;      // [ncontrib] (e * nsym + sym) << 1 | transpose
  long long ncontrib = 0;
  // node -> (element, local node) adjacency for the f_int gather
  int *d_nptr = nullptr;         // [nnode+1]
  int *d_nadj = nullptr;         // [nelem*nn] e * nn + a
  // row assembly (elements of at most 64 nodes): upper-block row pointer,
  // and for adjacency entry t and local node b the column of block (a, b)
  // among the row's upper blocks (0xFFFF: node b precedes the row's node)
  long long *d_kptr = nullptr;          // [nnode+1]
  unsigned short *d_rslot = nullptr;    // [nelem*nn*nn]
  int row_maxs = 0;                     // longest row (upper blocks)

  // ---- launch configuration (picked by the autotuner, identical results)
  int hex8_te = 0;       // elements per workgroup of the hex8 K kernel (0: largest)
  int asm_mode = -1;     // assembly: contribution-list loads in batches of 1, 4, 8,
                         // 16 (mode & 3), XCD-contiguous workgroups (mode & 4);
                         // 8 one wave per node row; -1 default (2)
  bool tuned = false;
  int overlap = -1;      // f_int beside the assembly (1) or after K (0); -1 untimed
  std::string tune_report;

  // ---- solver workspace
  double *d_u = nullptr;      // nodal vector scratch [nnode*3]
  // PCG vectors on nodes [nnode*3]; d_b is an equation-vector scratch
  double *d_b = nullptr, *d_x = nullptr, *d_r = nullptr, *d_p = nullptr,
         *d_w = nullptr, *d_z = nullptr, *d_dinv = nullptr, *d_bn = nullptr;
  double *d_part = nullptr;   // partial sums
  int precond = 0;            // 0 point Jacobi (the reference's), 1 node-block Jacobi
  double *d_binv = nullptr;   // [nnode*9] inverses of the diagonal node blocks (precond 1)
  double *d_scal = nullptr;   // device scalars
  int *d_done = nullptr;      // convergence iteration (0 = running)
  int *h_done = nullptr;      // pinned mirror
  double *h_scal = nullptr;   // pinned mirror
  int nblocks_red = 0;        // reduction blocks
  bool use_graphs = true;     // capture PCG iteration chunks into hipGraphs
  bool bsp_tl_valu = false;   // B-spline TL tangents on the vector-ALU kernel
  // hex8 internal force: element vector entry (e, a) written at position
  // epos[e * nn + a] of the node-sorted adjacency (nadj), so that the gather
  // reads each node's entries contiguously; ge_sorted says which layout the
  // last writer of d_ge used
  int *d_epos = nullptr;
  bool ge_sorted = false;
  hipGraphExec_t g_single = nullptr, g_batch = nullptr;
  int g_single_bits = 0, g_batch_bits = 0;  // matrix precision they were captured for

  // ---- batch of independent meshes (contiguous node ranges)
  int nbatch = 0;
  std::vector<long long> h_bnodeptr;  // [nbatch+1]
  int nchunk = 0;                     // solver chunks (never cross a mesh)
  long long *d_chn0 = nullptr, *d_chn1 = nullptr;  // [nchunk] node ranges
  int *d_chmesh = nullptr;            // [nchunk]
  int *d_meshch = nullptr;            // [nbatch+1] chunk range of each mesh
  double *d_mscal = nullptr;          // [nbatch*8]
  int *d_mdone = nullptr;             // [nbatch*4]
  int *d_ndone = nullptr;             // finished meshes

  // ---- direct solve: LDL^T of the equation matrix in FP64 band storage,
  // A(i, j) (j <= i <= j + bw) at d_band[j * (bw + 1) + i - j]
  int band_bw = -1;            // half bandwidth (-1: not computed yet)
  double *d_band = nullptr;    // [neq * (bw + 1)]
  int *d_finfo = nullptr;      // [2] first pivot below the tolerance + 1, negative pivots
  double *d_bpart = nullptr;   // partial sums of the wide backward solve [2][wg][32]
  bool factored = false;       // d_band holds the factors of the current matrix
  int neg_pivots = 0;

  // ---- external loads: surface-load groups, their face vectors [rows * 3]
  // and the node -> face-vector-row lists (entry order) of their gather;
  // prescribed equations with their penalty (node-direction index, value)
  std::vector<SurfaceGroup> sgroups;
  long long nsrow = 0;
  double *d_sfv = nullptr;
  int *d_sptr = nullptr, *d_sadj = nullptr;
  int npen = 0;
  int *d_pnd = nullptr;
  double *d_pval = nullptr;

  // ---- the device's fixed scratch (k_scratch.hip), looked up at creation
  void *scratch = nullptr;

  // ---- timing
  hipEvent_t ev[8] = {};
  double last_ms[4] = {0, 0, 0, 0};
};

// Fixed device scratch of the allocation-free internal force (k_scratch.hip):
// a file-scope device array of kScratchBytes per device.  scratch_setup
// returns its address on `device` (the current device), looking it up once.
constexpr size_t kScratchBytes = size_t(64) << 20;
void *scratch_setup(int device);

// kernel launchers (defined in the .hip files)
void launch_hex8_K(Ctx &c, const double *d_unodal);
void launch_hex8_fint(Ctx &c, const double *d_unodal);
void launch_bsp_K(Ctx &c, const double *d_unodal);
void launch_bsp_fint(Ctx &c, const double *d_unodal);
// linear tuned B-spline kernels: the stiffness kernel also writes the points'
// internal-force records (d_pts), from which launch_bsp_fint_pts finishes f_int
bool bsp_fused_fint(const Ctx &c);
void bsp_alloc_pts(Ctx &c);  // the records, up to 4 GB, at context creation
void launch_bsp_K_pts(Ctx &c, const double *d_unodal);
void launch_bsp_fint_pts(Ctx &c);
void launch_hex8_body(Ctx &c, const double *d_b);   // element vectors into d_ge
// 20-node bricks (k_hex20.hip): the same four entry points
void launch_hex20_K(Ctx &c, const double *d_unodal);
void launch_hex20_fint(Ctx &c, const double *d_unodal);
void launch_hex20_body(Ctx &c, const double *d_b);
void launch_hex20_stress(Ctx &c, const double *d_unodal, double *sig_ip, double *sig_node);
void launch_bsp_body(Ctx &c, const double *d_b);
// stresses at the points (sig_ip [nelem*nq*6]) and, when sig_node is not
// null, extrapolated to the element nodes with d_TR ([nelem*nn*6])
void launch_hex8_stress(Ctx &c, const double *d_unodal, double *sig_ip, double *sig_node);
void launch_bsp_stress(Ctx &c, const double *d_unodal, double *sig_ip, double *sig_node);
void launch_avg_stress(Ctx &c, const double *sig_node, double *sig_avg);
// B-spline solids of any degree / Gauss order (k_bspg.hip); bsp_generic is
// true when the tuned cubic kernels do not cover the element
bool bsp_generic(const Ctx &c);
void gen_prepare(Ctx &c, const double *h_ext);
void gen_launch_K(Ctx &c, const double *d_unodal);
void gen_launch_vec(Ctx &c, const double *d_unodal, const double *d_b);  // f_int, or body if b
void gen_launch_stress(Ctx &c, const double *d_unodal, double *sig_ip, double *sig_node);
void gen_launch_node_stress(Ctx &c, const double *d_unodal, double *sig_node);
void launch_assemble_K(Ctx &c);
std::vector<int> hex8_te_options(const Ctx &c);     // valid hex8 K workgroup sizes
std::vector<int> assembly_options(const Ctx &c);    // valid asm_mode values
void launch_gather_fint(Ctx &c, double *d_g);
// external loads (k_load.hip)
void launch_surface_group(Ctx &c, const SurfaceGroup &g, double *fv);
void launch_gather_surface(Ctx &c, double *f);
void launch_penalty_diag(Ctx &c);                               // after the assembly
void launch_penalty_force(Ctx &c, const double *u, double *g);  // after the f_int gather
void launch_zero_penalty(Ctx &c, double *x);  // x = 0 at the prescribed equations
void launch_spmv_nodal(Ctx &c, const double *x, double *y);  // node vectors
void launch_abs_spmv_nodal(Ctx &c, const double *x, double *y);  // |K| |x|, node vectors
void launch_csr_convert(Ctx &c, double *d_csrval, int to_csr);
// abs_exit: the reference's early return when |b| < rtol (an absolute
// test, mlib/sysmat.cpp PCGSolver); callers that pass a relative tolerance
// computed from the residual (inexact Newton) turn it off.
int run_pcg(Ctx &c, const double *b, double *x, const double *x0, int maxit,
            double rtol, int *iters, double *relres, bool abs_exit = true);
int run_pcg_batch(Ctx &c, const double *b, double *x, int maxit, double rtol, int *iters,
                  double *relres);
void ensure_eq_csr(Ctx &c);
void launch_nodal_to_eq(Ctx &c, const double *nodal, double *eq);
void launch_eq_to_nodal(Ctx &c, const double *eq, double *nodal);
void launch_add_diag(Ctx &c, int n, const int *d_eq, const double *d_v);
void ensure_bval32(Ctx &c);  // d_bval32 = (float) d_bval when stale
// direct solve (k_direct.hip): half bandwidth of the equation matrix; factor
// the current global matrix into d_band (0 + c.err on a tiny pivot); solve
// with the factors (b, x equation vectors, may alias)
int band_width(Ctx &c);
int band_factor(Ctx &c);
void band_solve(Ctx &c, const double *b, double *x);
void launch_expand_Ke(Ctx &c, int e0, int ne, float *d_out);
void bsp_prepare_tables(Ctx &c, const double *h_ext, const double *h_cpw);
void launch_axpby(Ctx &c, int n, double a, const double *x, double b, const double *y,
                  double *z);
double dev_dot(Ctx &c, const double *x, const double *y, int n);
// graph-safe pieces (no host synchronisation, no memset nodes)
void launch_fill(Ctx &c, double *p, long long n, double v);
void launch_fill(Ctx &c, int *p, long long n, int v);
void launch_copy(Ctx &c, const double *x, double *y, long long n);  // y = x
void enqueue_dot(Ctx &c, const double *x, const double *y, int n, double *d_out);
void enqueue_stiffness(Ctx &c, const double *d_unodal);     // status into d_status
void enqueue_int_force(Ctx &c, const double *d_unodal, double *d_g);  // into d_status2
int check_status_values(Ctx &c, const char *what, const int *st);
int band_setup(Ctx &c);                 // width check, buffers; 0 with c.err
void enqueue_band_factor(Ctx &c);       // info into d_finfo
int band_factor_result(Ctx &c, const int *info);  // info read back: 1 ok, 0 singular
void band_solve_setup(Ctx &c);
void enqueue_band_solve(Ctx &c, const double *b, double *x, double *z);  // z: n scratch
void enqueue_pcg_start(Ctx &c, const double *b, const double *d_tol, int abs_exit);
void enqueue_pcg_chunk(Ctx &c);
int pcg_continue(Ctx &c, int maxit, double rtol, int abs_exit, int *iters, double *relres);
// element K + assembly / element f_int + gather on the context stream, with
// the status check (returns 0 and sets c.err on failure)
int stiffness_dev(Ctx &c, const double *d_unodal);
int int_force_dev(Ctx &c, const double *d_unodal, double *d_g);
// both for the same state, f_int on the second stream beside K
int stiffness_fint_dev(Ctx &c, const double *d_unodal, double *d_g);

}  // namespace fgpu
