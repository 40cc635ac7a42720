// fgpu_api.cpp - C-ABI entry points and host-side set-up of libfgpu.
//
// Set-up replaces the reference's equation numbering and matrix creation
// (ctrl/ctrl.cpp:2004-2157 InitDofs/CompEqs/CompProf, mlib/sysmat.cpp:42-70
// CreateMatrix): it builds, once per mesh,
//   * the full symmetric CSR pattern over equations (rows of one node share
//     their column list: the free directions of every neighbour node),
//   * the deterministic assembly map: for every upper node block (A <= B) the
//     segment of element blocks that add into it, sorted by element id
//     (the order in which ctrl.cpp:945-956 adds element matrices), and
//   * the node -> (element, local node) lists of the internal-force gather.
// All integer maps are exact; no floating-point atomics are used anywhere.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "fgpu_internal.h"

using namespace fgpu;

namespace {

template <class T>
T *dalloc(size_t n) {
  T *p = nullptr;
  if (n == 0) n = 1;
  FGPU_CHECK(hipMalloc(&p, n * sizeof(T)));
  return p;
}

template <class T>
T *dupload(const T *h, size_t n, hipStream_t s) {
  T *d = dalloc<T>(n);
  if (n) FGPU_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

template <class F>
void parallel_for(long n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  long chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &f] {
      for (long i = b; i < e; i++) f(i);
    });
  }
  for (auto &x : th) x.join();
}

void gauss_rule(int n, double *r, double *w) {
  // Gauss-Legendre by Newton iteration on P_n, rounded to 15 decimals as in
  // the reference's quadrature table (elc/intpoint.cpp:741-920).
  const long double PI = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; i++) {
    long double x = -cosl(PI * (i + 0.75L) / (n + 0.5L)), dp = 1.0L;
    for (int it = 0; it < 100 && n > 1; it++) {
      long double p0 = 1.0L, p1 = x;
      for (int k = 2; k <= n; k++) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0L);
      long double dx = p1 / dp;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    if (n == 1) { x = 0.0L; dp = 1.0L; }
    char buf[64];
    snprintf(buf, sizeof buf, "%.15Lf", x);
    r[i] = strtod(buf, nullptr);
    if (r[i] == 0.0) r[i] = 0.0;
    snprintf(buf, sizeof buf, "%.15Lf", 2.0L / ((1.0L - x * x) * dp * dp));
    w[i] = strtod(buf, nullptr);
  }
}

void build_maps(Ctx &c, const int *conn) {
  const int nn = c.nn, nnode = c.nnode, nelem = c.nelem;
  const int nsym = nsym_blocks(nn);
  // node -> (element, local) lists, element order
  std::vector<int> nptr(nnode + 1, 0);
  for (long i = 0; i < (long)nelem * nn; i++) nptr[conn[i] + 1]++;
  for (int i = 0; i < nnode; i++) nptr[i + 1] += nptr[i];
  std::vector<int> nadj((size_t)nelem * nn);
  {
    std::vector<int> pos(nptr.begin(), nptr.end() - 1);
    for (int e = 0; e < nelem; e++)
      for (int a = 0; a < nn; a++) nadj[pos[conn[(size_t)e * nn + a]]++] = e * nn + a;
  }
  std::vector<int> nfree(nnode, 0);
  for (int n = 0; n < nnode; n++)
    for (int d = 0; d < 3; d++) nfree[n] += c.eqmap[3 * n + d] >= 0;

  // per node: sorted unique neighbour list (kept, CSR-like), the column
  // offset of each neighbour inside the node's rows, upper-block and
  // contribution counts.
  auto neighbours = [&](int A, std::vector<int> &nb) {
    nb.clear();
    for (int t = nptr[A]; t < nptr[A + 1]; t++) {
      int e = nadj[t] / nn;
      for (int b = 0; b < nn; b++) nb.push_back(conn[(size_t)e * nn + b]);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
  };
  std::vector<long long> nbcnt(nnode + 1, 0), nc_up(nnode);
  parallel_for(nnode, [&](long A) {
    thread_local std::vector<int> nb;
    neighbours((int)A, nb);
    nbcnt[A + 1] = (long long)nb.size();
    long long ncon = 0;
    for (int t = nptr[A]; t < nptr[A + 1]; t++) {
      int e = nadj[t] / nn;
      for (int b = 0; b < nn; b++) ncon += conn[(size_t)e * nn + b] >= A;
    }
    nc_up[A] = ncon;
  });
  for (int n = 0; n < nnode; n++) nbcnt[n + 1] += nbcnt[n];
  std::vector<int> nbidx(nbcnt[nnode]), nbcol(nbcnt[nnode]);
  std::vector<long long> rowlen(nnode), nb_up(nnode);
  parallel_for(nnode, [&](long A) {
    thread_local std::vector<int> nb;
    neighbours((int)A, nb);
    long long base = nbcnt[A];
    int off = 0;
    long long up = 0;
    for (size_t k = 0; k < nb.size(); k++) {
      nbidx[base + k] = nb[k];
      nbcol[base + k] = off;
      off += nfree[nb[k]];
      up += nb[k] >= A;
    }
    rowlen[A] = off;
    nb_up[A] = up;
  });
  auto blk_pos = [&](int A, int B) {  // position of block (A, B) in the node-block rows
    const int *b0 = nbidx.data() + nbcnt[A], *b1 = nbidx.data() + nbcnt[A + 1];
    const int *it = std::lower_bound(b0, b1, B);
    return nbcnt[A] + (long long)(it - b0);
  };
  auto col_offset = [&](int A, int B) {  // column offset of B in the rows of A
    return nbcol[blk_pos(A, B)];
  };
  if (nbcnt[nnode] >= (1LL << 31)) throw Error("mesh too large for 32-bit block positions");
  // CSR row pointers in equation order (equations are node-major)
  std::vector<long long> rowptr(c.neq + 1, 0);
  for (int n = 0; n < nnode; n++)
    for (int d = 0; d < 3; d++) {
      int q = c.eqmap[3 * n + d];
      if (q >= 0) rowptr[q + 1] = rowlen[n];
    }
  for (int q = 0; q < c.neq; q++) rowptr[q + 1] += rowptr[q];
  c.nnz = rowptr[c.neq];
  std::vector<long long> bptr(nnode + 1, 0), cptr0(nnode + 1, 0);
  for (int n = 0; n < nnode; n++) {
    bptr[n + 1] = bptr[n] + nb_up[n];
    cptr0[n + 1] = cptr0[n] + nc_up[n];
  }
  c.nblk = bptr[nnode];
  c.ncontrib = cptr0[nnode];
  if (c.ncontrib >= (1LL << 31) || (long long)nelem * nsym * 2 >= (1LL << 31))
    throw Error("mesh too large for 32-bit assembly codes");
  (void)col_offset;
  std::vector<int> blkA(c.nblk), blkB(c.nblk), offAB(c.nblk), offBA(c.nblk);
  std::vector<int> cptr(c.nblk + 1);  // contribution segments (fewer than 2^31 entries)
  std::vector<int> contrib(c.ncontrib);
  cptr[c.nblk] = (int)c.ncontrib;

  parallel_for(nnode, [&](long Al) {
    int A = (int)Al;
    thread_local std::vector<std::pair<int, int>> items;
    // contributions of upper blocks, sorted by (B, element, local a, local b)
    items.clear();
    for (int t = nptr[A]; t < nptr[A + 1]; t++) {
      int e = nadj[t] / nn, a = nadj[t] % nn;
      for (int b = 0; b < nn; b++) {
        int B = conn[(size_t)e * nn + b];
        if (B < A) continue;
        int lo = std::min(a, b), hi = std::max(a, b);
        int code = ((e * nsym + sym_index(lo, hi, nn)) << 1) | (a > b ? 1 : 0);
        items.emplace_back(B, code);
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                       return x.first < y.first;
                     });
    long long bk = bptr[A], cc = cptr0[A];
    size_t i = 0;
    while (i < items.size()) {
      int B = items[i].first;
      blkA[bk] = A;
      blkB[bk] = B;
      offAB[bk] = (int)blk_pos(A, B);
      offBA[bk] = (int)blk_pos(B, A);
      cptr[bk] = (int)cc;
      while (i < items.size() && items[i].first == B) contrib[cc++] = items[i++].second;
      bk++;
    }
  });

  // node-block (3x3) matrix: rows = nodes, blocks = sorted neighbours
  c.nbf = nbcnt[nnode];
  std::vector<long long> bdiag(nnode);
  std::vector<unsigned char> fmask(nnode);
  for (int n = 0; n < nnode; n++) {
    bdiag[n] = blk_pos(n, n);
    unsigned char mk = 0;
    for (int d = 0; d < 3; d++)
      if (c.eqmap[3 * n + d] >= 0) mk |= (unsigned char)(1 << d);
    fmask[n] = mk;
  }
  // per upper block: the free directions of A (bits 0-2) and of B (bits
  // 3-5), and A == B (bit 6) -- what the assembly needs of A and B
  std::vector<unsigned char> bmask(c.nblk);
  for (long long k = 0; k < c.nblk; k++)
    bmask[k] = (unsigned char)(fmask[blkA[k]] | fmask[blkB[k]] << 3 |
                               (blkA[k] == blkB[k] ? 1 << 6 : 0));
  hipStream_t s = c.stream;
  c.d_rowptr = dupload(rowptr.data(), rowptr.size(), s);  // columns: built on first export
  c.d_bptr = dupload(nbcnt.data(), nbcnt.size(), s);
  c.d_bcol = dupload(nbidx.data(), nbidx.size(), s);
  c.d_bcoloff = dupload(nbcol.data(), nbcol.size(), s);
  c.d_bdiag = dupload(bdiag.data(), bdiag.size(), s);
  c.d_fmask = dupload(fmask.data(), fmask.size(), s);
  c.d_bval = dalloc<double>((size_t)c.nbf * 9);
  c.d_bmask = dupload(bmask.data(), bmask.size(), s);
  c.d_blkB = dupload(blkB.data(), blkB.size(), s);
  c.d_offAB = dupload(offAB.data(), offAB.size(), s);
  c.d_offBA = dupload(offBA.data(), offBA.size(), s);
  c.d_cptr = dupload(cptr.data(), cptr.size(), s);
  c.d_contrib = dupload(contrib.data(), contrib.size(), s);
  c.d_nptr = dupload(nptr.data(), nptr.size(), s);
  c.d_nadj = dupload(nadj.data(), nadj.size(), s);
  if (c.shape == FGPU_HEX8) {  // inverse of nadj: where each element entry goes
    std::vector<int> epos(nadj.size());
    for (size_t t = 0; t < nadj.size(); t++) epos[nadj[t]] = (int)t;
    c.d_epos = dupload(epos.data(), epos.size(), s);
    FGPU_CHECK(hipStreamSynchronize(s));
  }
  // row assembly tables: B-spline solids of at most 64 control points per
  // element whose rows fit the kernel's LDS (ROW_MAXS upper blocks)
  constexpr long long ROW_MAXS = 848;  // 77 bytes of LDS each: 63.8 KB
  long long maxs = 0;
  for (int n = 0; n < nnode; n++) maxs = std::max(maxs, nb_up[n]);
  // the kernel relies on: an element's nodes distinct (one column per lane);
  // a row's upper blocks consecutive in the node-block rows, the diagonal first
  bool fits = true;
  for (int n = 0; n < nnode && fits; n++) {
    for (int t = nptr[n]; t < nptr[n + 1] - 1; t++)
      if (nadj[t] / nn == nadj[t + 1] / nn) fits = false;
    for (long long k = bptr[n]; k < bptr[n + 1]; k++)
      if (offAB[k] != offAB[bptr[n]] + (k - bptr[n])) fits = false;
    if (bptr[n + 1] > bptr[n] && blkB[bptr[n]] != n) fits = false;
  }
  // the table holds 2 bytes per element, local node pair: built up to 4 GB
  const bool small = (double)nadj.size() * nn * 2.0 <= 4.0 * (1 << 30);
  if (c.shape == FGPU_BSP && nn <= 64 && maxs <= ROW_MAXS && fits && small) {
    std::vector<unsigned short> rslot(nadj.size() * (size_t)nn);
    parallel_for(nnode, [&](long Al) {
      const int A = (int)Al;
      const int *r0 = blkB.data() + bptr[A], *r1 = blkB.data() + bptr[A + 1];
      for (int t = nptr[A]; t < nptr[A + 1]; t++) {
        const int e = nadj[t] / nn;
        for (int b = 0; b < nn; b++) {
          const int B = conn[(size_t)e * nn + b];
          rslot[(size_t)t * nn + b] =
              B < A ? (unsigned short)0xFFFF
                    : (unsigned short)(std::lower_bound(r0, r1, B) - r0);
        }
      }
    });
    c.d_kptr = dupload(bptr.data(), bptr.size(), s);
    c.d_rslot = dupload(rslot.data(), rslot.size(), s);
    c.row_maxs = (int)maxs;
  }
  FGPU_CHECK(hipStreamSynchronize(s));
}

// Batch set-up: meshes are contiguous node ranges; the solver works on
// chunks of at most CHN nodes inside one mesh (one workgroup per chunk).
void setup_batch(Ctx &c, const fgpu_mesh &m) {
  const int nb = m.nbatch;
  if (!m.batch_nodeptr) throw Error("batch_nodeptr missing");
  c.h_bnodeptr.assign(m.batch_nodeptr, m.batch_nodeptr + nb + 1);
  if (c.h_bnodeptr[0] != 0 || c.h_bnodeptr[nb] != c.nnode) throw Error("batch_nodeptr range");
  for (int k = 0; k < nb; k++)
    if (c.h_bnodeptr[k + 1] <= c.h_bnodeptr[k]) throw Error("empty mesh in batch");
  // no element may couple two meshes
  std::vector<int> mesh_of(c.nnode);
  for (int k = 0; k < nb; k++)
    for (long long n = c.h_bnodeptr[k]; n < c.h_bnodeptr[k + 1]; n++) mesh_of[n] = k;
  for (long e = 0; e < c.nelem; e++) {
    int k0 = mesh_of[m.conn[(size_t)e * c.nn]];
    for (int a = 1; a < c.nn; a++)
      if (mesh_of[m.conn[(size_t)e * c.nn + a]] != k0) throw Error("element couples two meshes");
  }
  const long long CHN = 1024;
  std::vector<long long> n0, n1;
  std::vector<int> chm, meshch(nb + 1, 0);
  for (int k = 0; k < nb; k++) {
    meshch[k] = (int)n0.size();
    for (long long a = c.h_bnodeptr[k]; a < c.h_bnodeptr[k + 1]; a += CHN) {
      n0.push_back(a);
      n1.push_back(std::min(a + CHN, c.h_bnodeptr[k + 1]));
      chm.push_back(k);
    }
  }
  meshch[nb] = (int)n0.size();
  c.nbatch = nb;
  c.nchunk = (int)n0.size();
  hipStream_t s = c.stream;
  c.d_chn0 = dupload(n0.data(), n0.size(), s);
  c.d_chn1 = dupload(n1.data(), n1.size(), s);
  c.d_chmesh = dupload(chm.data(), chm.size(), s);
  c.d_meshch = dupload(meshch.data(), meshch.size(), s);
  c.d_mscal = dalloc<double>(8 * (size_t)nb);
  c.d_mdone = dalloc<int>(4 * (size_t)nb);
  c.d_ndone = dalloc<int>(1);
  FGPU_CHECK(hipStreamSynchronize(s));
}

void destroy(Ctx *c) {
  if (!c) return;
  void *bufs[] = {c->d_coords, c->d_conn, c->d_eqmap, c->d_mat, c->d_ext,
                  c->d_cpw, c->d_gauss, c->d_gauss64, c->d_Ke, c->d_Ke64, c->d_ge, c->d_pts,
                  c->d_epos,
                  c->d_status, c->d_rowptr, c->d_col, c->d_bptr, c->d_bcol,
                  c->d_bcoloff, c->d_bdiag, c->d_fmask, c->d_bval, c->d_bn,
                  c->d_bmask, c->d_blkB, c->d_offAB, c->d_offBA, c->d_cptr,
                  c->d_contrib, c->d_nptr, c->d_nadj, c->d_kptr, c->d_rslot, c->d_u,
                  c->d_b, c->d_x,
                  c->d_r, c->d_p, c->d_w, c->d_z, c->d_dinv, c->d_part,
                  c->d_scal, c->d_done, c->d_eqnode, c->d_TR, c->d_band, c->d_finfo, c->d_bpart, c->d_binv,
                  c->d_bval32, c->d_gsym, c->d_gG, c->d_gR, c->d_gH, c->d_gPQ,
                  c->d_gnodes};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (c->h_done) (void)hipHostFree(c->h_done);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  void *bbufs[] = {c->d_chn0, c->d_chn1, c->d_chmesh, c->d_meshch, c->d_mscal, c->d_mdone,
                   c->d_ndone};
  for (void *p : bbufs)
    if (p) (void)hipFree(p);
  for (auto &g : c->sgroups) {
    void *gb[] = {g.d_nodes, g.d_ext, g.d_local, g.d_q, g.d_gp};
    for (void *p : gb)
      if (p) (void)hipFree(p);
  }
  void *lbufs[] = {c->d_sfv, c->d_sptr, c->d_sadj, c->d_pnd, c->d_pval};
  for (void *p : lbufs)
    if (p) (void)hipFree(p);
  if (c->g_single) (void)hipGraphExecDestroy(c->g_single);
  if (c->g_batch) (void)hipGraphExecDestroy(c->g_batch);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->h_status) (void)hipHostFree(c->h_status);
  delete c;
}

template <class F>
int guarded(fgpu_ctx *ctx, F f) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return 0;
  try {
    return f(*c);
  } catch (const Error &e) {
    c->err = e.msg;
  } catch (const std::exception &e) {
    c->err = e.what();
  }
  return 0;
}

float elapsed(Ctx &c, int a, int b) {
  float ms = 0;
  FGPU_CHECK(hipEventElapsedTime(&ms, c.ev[a], c.ev[b]));
  return ms;
}

// status[4] of the element kernels (count_jacobian), read into pinned host
// memory behind the kernels that count into it (one synchronization)
void enqueue_status_read(Ctx &c, const int *d_st, int n) {
  FGPU_CHECK(hipMemcpyAsync(c.h_status, d_st, n * sizeof(int), hipMemcpyDeviceToHost, c.stream));
}

int check_status(Ctx &c, const char *what) {
  enqueue_status_read(c, c.d_status, 4);
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  return check_status_values(c, what, c.h_status);
}

}  // namespace

namespace fgpu {

// status[4] of the element kernels (count_jacobian): a zero or non-finite
// Jacobian, or negative Jacobians at some but not all points of the mesh
// (inverted elements), fail; all negative is a left-handed parametrization,
// integrated with |det J|.
int check_status_values(Ctx &c, const char *what, const int *st) {
  const long long npts = (long long)c.nelem * c.nip[0] * c.nip[1] * c.nip[2];
  if (st[3]) {
    c.err = std::string(what) + ": " + std::to_string(st[3]) +
            " integration points with a zero or non-finite Jacobian";
    return 0;
  }
  if (st[0] && st[0] != npts) {
    c.err = std::string(what) + ": " + std::to_string(st[0]) + " of " + std::to_string(npts) +
            " integration points with a negative Jacobian (inverted elements)";
    return 0;
  }
  if (st[1]) {
    c.err = std::string(what) + ": stress update failed at " +
            std::to_string(st[1]) + " integration points (det C <= 0)";
    return 0;
  }
  return 1;
}


// element kernels of the context's shape (the B-spline launchers pick the
// tuned or the generic kernels themselves)
void launch_elem_K(Ctx &c, const double *d_u) {
  if (c.shape == FGPU_HEX8)
    launch_hex8_K(c, d_u);
  else if (c.shape == FGPU_HEX20)
    launch_hex20_K(c, d_u);
  else
    launch_bsp_K(c, d_u);
}

void launch_elem_fint(Ctx &c, const double *d_u) {
  if (c.shape == FGPU_HEX8)
    launch_hex8_fint(c, d_u);
  else if (c.shape == FGPU_HEX20)
    launch_hex20_fint(c, d_u);
  else
    launch_bsp_fint(c, d_u);
}

// Autotuner: times every valid launch configuration of the element-stiffness
// and assembly kernels on this mesh and state (best of 3 runs each, HIP
// events) and keeps the fastest.  All configurations give identical bits.
void autotune(Ctx &c, const double *d_u) {
  c.tuned = true;
  const char *env = std::getenv("FGPU_AUTOTUNE");
  if (env && env[0] == '0') return;
  std::string rep;
  auto time3 = [&](auto &&launch) {
    float best = 1e30f;
    for (int r = 0; r < 3; r++) {
      FGPU_CHECK(hipEventRecord(c.ev[0], c.stream));
      launch();
      FGPU_CHECK(hipEventRecord(c.ev[1], c.stream));
      FGPU_CHECK(hipEventSynchronize(c.ev[1]));
      best = std::min(best, (float)elapsed(c, 0, 1));
    }
    return best;
  };
  char buf[192];
  if (c.shape == FGPU_HEX8) {
    float best = 1e30f;
    int pick = 0;
    for (int te : hex8_te_options(c)) {
      c.hex8_te = te;
      float ms = time3([&] { launch_hex8_K(c, d_u); });
      std::snprintf(buf, sizeof buf, "hex8_K elements/workgroup %d: %.4f ms\n", te, ms);
      rep += buf;
      if (ms < best) best = ms, pick = te;
    }
    c.hex8_te = pick;
    std::snprintf(buf, sizeof buf, "picked hex8_K elements/workgroup %d%s\n", pick,
                  pick == 1 ? " (one element per lane)" : "");
    rep += buf;
    launch_hex8_K(c, d_u);
  } else {
    launch_elem_K(c, d_u);
  }
  float best = 1e30f;
  int pick = 0;
  for (int mode : assembly_options(c)) {
    c.asm_mode = mode;
    float ms = time3([&] { launch_assemble_K(c); });
    const int nb = (mode & 3) == 0 ? 1 : 2 << (mode & 3);
    if (mode == 8)
      std::snprintf(buf, sizeof buf, "assembly mode 8 (one wave per node row): %.4f ms\n", ms);
    else
      std::snprintf(buf, sizeof buf, "assembly mode %d (loads in batches of %d%s): %.4f ms\n",
                    mode, nb, mode & 4 ? ", XCD-contiguous workgroups" : "", ms);
    rep += buf;
    if (ms < best) best = ms, pick = mode;
  }
  c.asm_mode = pick;
  std::snprintf(buf, sizeof buf, "picked assembly mode %d\n", pick);
  rep += buf;
  c.tune_report = rep;
}

// The stiffness of the nodal state d_u, enqueued only: status counters
// cleared by a fill kernel (graph-safe), element kernel, assembly, penalty.
// The autotuner must have run (it synchronises).
void enqueue_stiffness(Ctx &c, const double *d_u) {
  c.factored = false;
  c.bval32_valid = false;
  launch_fill(c, c.d_status, 4, 0);
  launch_elem_K(c, d_u);
  launch_assemble_K(c);
  launch_penalty_diag(c);
}

// The internal force of d_u into d_g, enqueued only, counting into the
// second status array.
void enqueue_int_force(Ctx &c, const double *d_u, double *d_g) {
  std::swap(c.d_status, c.d_status2);
  launch_fill(c, c.d_status, 4, 0);
  launch_elem_fint(c, d_u);
  launch_gather_fint(c, d_g);
  launch_penalty_force(c, d_u, d_g);
  std::swap(c.d_status, c.d_status2);
}

int stiffness_dev(Ctx &c, const double *d_u) {
  if (!c.tuned) autotune(c, d_u);  // its launches count into the status: reset after
  c.factored = false;
  c.bval32_valid = false;
  FGPU_CHECK(hipMemsetAsync(c.d_status, 0, 4 * sizeof(int), c.stream));
  FGPU_CHECK(hipEventRecord(c.ev[0], c.stream));
  launch_elem_K(c, d_u);
  FGPU_CHECK(hipEventRecord(c.ev[1], c.stream));
  launch_assemble_K(c);
  launch_penalty_diag(c);
  FGPU_CHECK(hipEventRecord(c.ev[2], c.stream));
  enqueue_status_read(c, c.d_status, 4);
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  c.last_ms[0] = elapsed(c, 0, 1);
  c.last_ms[1] = elapsed(c, 1, 2);
  return check_status_values(c, "stiffness", c.h_status);
}

int int_force_dev(Ctx &c, const double *d_u, double *d_g) {
  FGPU_CHECK(hipMemsetAsync(c.d_status, 0, 4 * sizeof(int), c.stream));
  FGPU_CHECK(hipEventRecord(c.ev[3], c.stream));
  launch_elem_fint(c, d_u);
  launch_gather_fint(c, d_g);
  launch_penalty_force(c, d_u, d_g);
  FGPU_CHECK(hipEventRecord(c.ev[4], c.stream));
  enqueue_status_read(c, c.d_status, 4);
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  c.last_ms[2] = elapsed(c, 3, 4);
  return check_status_values(c, "internal force", c.h_status);
}

namespace {
// launches issued while it lives go to the second stream and count into the
// second status array (the launchers read c.stream / c.d_status)
struct OnStream2 {
  Ctx &c;
  explicit OnStream2(Ctx &c_) : c(c_) {
    std::swap(c.stream, c.stream2);
    std::swap(c.d_status, c.d_status2);
  }
  ~OnStream2() {
    std::swap(c.stream, c.stream2);
    std::swap(c.d_status, c.d_status2);
  }
};
}  // namespace

int stiffness_fint_overlapped(Ctx &c, const double *d_u, double *d_g, int mode);

// The orders of the combined call, all giving the same bits: 0 = K, then
// f_int, on the context stream; 1 = f_int (element vectors and gather) on the
// second stream beside the assembly; 2 = only f_int's element vectors beside
// the assembly (arithmetic beside a memory-bound kernel), the gather after
// it.  The first call times the three (unless autotuning is off) and keeps
// the fastest.
int stiffness_fint_dev(Ctx &c, const double *d_u, double *d_g) {
  if (!c.tuned) autotune(c, d_u);
  if (c.overlap < 0) {
    const char *env = std::getenv("FGPU_AUTOTUNE");
    if (env && env[0] == '0') {
      c.overlap = 0;
    } else {
      double best[4] = {1e30, 1e30, 1e30, 1e30};
      for (int r = 0; r < 3; r++)
        for (int mode = 0; mode < 4; mode++) {
          FGPU_CHECK(hipStreamSynchronize(c.stream));
          const auto t0 = std::chrono::steady_clock::now();
          const int ok = mode ? stiffness_fint_overlapped(c, d_u, d_g, mode)
                              : (stiffness_dev(c, d_u) && int_force_dev(c, d_u, d_g));
          if (!ok) return 0;
          const double ms =
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                  .count();
          best[mode] = std::min(best[mode], ms);
        }
      c.overlap = 0;
      for (int mode = 1; mode < 4; mode++)
        if (best[mode] < best[c.overlap]) c.overlap = mode;
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "K then f_int: %.4f ms; f_int beside the assembly: %.4f ms; f_int element "
                    "vectors beside the assembly, gather after: %.4f ms; f_int element vectors "
                    "first, gather beside the element K: %.4f ms\n",
                    best[0], best[1], best[2], best[3]);
      c.tune_report += buf;
    }
  }
  if (!c.overlap) return stiffness_dev(c, d_u) && int_force_dev(c, d_u, d_g);
  return stiffness_fint_overlapped(c, d_u, d_g, c.overlap);
}

// Order 3: f_int's element vectors first (arithmetic), then its gather
// (memory-bound) on the second stream beside the stiffness element kernel
// (arithmetic), then the assembly.
int stiffness_fint_fint_first(Ctx &c, const double *d_u, double *d_g) {
  c.factored = false;
  c.bval32_valid = false;
  FGPU_CHECK(hipMemsetAsync(c.d_status, 0, 8 * sizeof(int), c.stream));
  std::swap(c.d_status, c.d_status2);  // f_int counts into the second array
  FGPU_CHECK(hipEventRecord(c.ev[3], c.stream));
  launch_elem_fint(c, d_u);
  std::swap(c.d_status, c.d_status2);
  FGPU_CHECK(hipEventRecord(c.ev[4], c.stream));
  FGPU_CHECK(hipEventRecord(c.ev_fork, c.stream));
  FGPU_CHECK(hipStreamWaitEvent(c.stream2, c.ev_fork, 0));
  {
    OnStream2 on(c);
    launch_gather_fint(c, d_g);
    launch_penalty_force(c, d_u, d_g);
    FGPU_CHECK(hipEventRecord(c.ev[7], c.stream));
  }
  FGPU_CHECK(hipEventRecord(c.ev[0], c.stream));
  launch_elem_K(c, d_u);
  FGPU_CHECK(hipEventRecord(c.ev[1], c.stream));
  launch_assemble_K(c);
  launch_penalty_diag(c);
  FGPU_CHECK(hipEventRecord(c.ev[2], c.stream));
  FGPU_CHECK(hipStreamWaitEvent(c.stream, c.ev[7], 0));  // join
  enqueue_status_read(c, c.d_status, 8);
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  c.last_ms[0] = elapsed(c, 0, 1);
  c.last_ms[1] = elapsed(c, 1, 2);
  c.last_ms[2] = elapsed(c, 3, 4) + elapsed(c, 4, 7);
  if (!check_status_values(c, "stiffness", c.h_status)) return 0;
  return check_status_values(c, "internal force", c.h_status + 4);
}

int stiffness_fint_overlapped(Ctx &c, const double *d_u, double *d_g, int mode) {
  if (mode == 3) return stiffness_fint_fint_first(c, d_u, d_g);
  c.factored = false;
  c.bval32_valid = false;
  // the two status arrays are adjacent (d_status2 = d_status + 4)
  FGPU_CHECK(hipMemsetAsync(c.d_status, 0, 8 * sizeof(int), c.stream));
  FGPU_CHECK(hipEventRecord(c.ev[0], c.stream));
  // linear tuned B-spline kernels: the stiffness kernel also writes the
  // points' internal-force records, and f_int only finishes from them
  const bool fused = bsp_fused_fint(c);
  if (fused)
    launch_bsp_K_pts(c, d_u);
  else
    launch_elem_K(c, d_u);
  FGPU_CHECK(hipEventRecord(c.ev[1], c.stream));
  // the element kernels of K and f_int both fill every CU with arithmetic;
  // f_int starts with the assembly, which waits on memory
  FGPU_CHECK(hipEventRecord(c.ev_fork, c.stream));
  FGPU_CHECK(hipStreamWaitEvent(c.stream2, c.ev_fork, 0));
  {
    OnStream2 on(c);
    FGPU_CHECK(hipEventRecord(c.ev[3], c.stream));
    if (fused)
      launch_bsp_fint_pts(c);
    else
      launch_elem_fint(c, d_u);
    if (mode == 1) {
      launch_gather_fint(c, d_g);
      launch_penalty_force(c, d_u, d_g);
    }
    FGPU_CHECK(hipEventRecord(c.ev[4], c.stream));
  }
  launch_assemble_K(c);
  launch_penalty_diag(c);
  FGPU_CHECK(hipEventRecord(c.ev[2], c.stream));
  FGPU_CHECK(hipStreamWaitEvent(c.stream, c.ev[4], 0));  // join
  if (mode == 2) {
    launch_gather_fint(c, d_g);
    launch_penalty_force(c, d_u, d_g);
    FGPU_CHECK(hipEventRecord(c.ev[7], c.stream));
  }
  enqueue_status_read(c, c.d_status, 8);  // both arrays
  FGPU_CHECK(hipStreamSynchronize(c.stream));
  c.last_ms[0] = elapsed(c, 0, 1);
  c.last_ms[1] = elapsed(c, 1, 2);
  c.last_ms[2] = elapsed(c, 3, 4) + (mode == 2 ? elapsed(c, 2, 7) : 0.0);
  if (!check_status_values(c, "stiffness", c.h_status)) return 0;
  return check_status_values(c, "internal force", c.h_status + 4);
}

}  // namespace fgpu

extern "C" {

int fgpu_create(fgpu_ctx **out, const fgpu_mesh *m) {
  if (!out || !m) return 0;
  *out = nullptr;
  Ctx *c = new Ctx();
  try {
    if (m->shape != FGPU_HEX8 && m->shape != FGPU_BSP && m->shape != FGPU_HEX20)
      throw Error("unknown shape");
    if (m->form < FGPU_LINEAR || m->form > FGPU_TL_NEOHOOKE) throw Error("unknown form");
    if (m->shape == FGPU_HEX8 && m->nn != 8) throw Error("hex8 needs nn = 8");
    if (m->shape == FGPU_HEX20 && m->nn != 20) throw Error("hex20 needs nn = 20");
    if (m->shape == FGPU_BSP) {
      if (m->nn != (m->deg[0] + 1) * (m->deg[1] + 1) * (m->deg[2] + 1))
        throw Error("B-spline solid: nn != prod(p+1)");
      for (int d = 0; d < 3; d++)
        if (m->deg[d] < 1 || m->deg[d] > 7 || m->nip[d] < 1 || m->nip[d] > 16)
          throw Error("B-spline solids: degrees 1..7 and 1..16 Gauss points per direction");
      if (m->nn > 512 || m->nip[0] * m->nip[1] * m->nip[2] > 512)
        throw Error("B-spline solids: at most 512 control points and 512 points per element");
      if (!m->ext) throw Error("B-spline solid needs extraction operators");
    } else {
      for (int d = 0; d < 3; d++)
        if (m->nip[d] != m->nip[0] || m->nip[0] < 1 || m->nip[0] > 3)
          throw Error("hex8 kernels support isotropic Gauss orders 1..3");
    }
    if (m->nnode <= 0 || m->nelem <= 0) throw Error("empty mesh");
    c->nnode = m->nnode;
    c->nelem = m->nelem;
    c->shape = m->shape;
    c->nn = m->nn;
    c->form = m->form;
    for (int d = 0; d < 3; d++) {
      c->deg[d] = m->deg[d];
      c->nip[d] = m->nip[d];
    }
    for (long i = 0; i < (long)m->nelem * m->nn; i++)
      if (m->conn[i] < 0 || m->conn[i] >= m->nnode) throw Error("connectivity out of range");
    c->eqmap.resize((size_t)m->nnode * 3);
    int neq = 0;
    for (long i = 0; i < (long)m->nnode * 3; i++) {
      c->eqmap[i] = m->dofmap[i] - 1;
      neq = std::max(neq, m->dofmap[i]);
    }
    c->neq = neq;
    FGPU_CHECK(hipGetDevice(&c->device));
    c->scratch = scratch_setup(c->device);
    FGPU_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    FGPU_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    for (auto &e : c->ev) FGPU_CHECK(hipEventCreate(&e));
    FGPU_CHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    hipStream_t s = c->stream;
    c->d_coords = dupload(m->coords, (size_t)m->nnode * 3, s);
    c->d_conn = dupload(m->conn, (size_t)m->nelem * m->nn, s);
    c->d_eqmap = dupload(c->eqmap.data(), c->eqmap.size(), s);
    {
      std::vector<int> eqnode(std::max(neq, 1), 0);
      for (long i = 0; i < (long)m->nnode * 3; i++)
        if (c->eqmap[i] >= 0) eqnode[c->eqmap[i]] = (int)i;
      c->d_eqnode = dupload(eqnode.data(), eqnode.size(), s);
      FGPU_CHECK(hipStreamSynchronize(s));
    }
    c->d_mat = dupload(m->mat, (size_t)m->nelem * 2, s);
    double gp[3 * 2 * 16];
    float gpf[3 * 2 * 16];
    for (int d = 0; d < 3; d++) {
      double r[16], w[16];
      gauss_rule(m->nip[d], r, w);
      for (int k = 0; k < m->nip[d]; k++) {
        gp[(d * 16 + k) * 2 + 0] = r[k];
        gp[(d * 16 + k) * 2 + 1] = w[k];
        gpf[(d * 16 + k) * 2 + 0] = (float)r[k];
        gpf[(d * 16 + k) * 2 + 1] = (float)w[k];
      }
    }
    c->d_gauss64 = dupload(gp, 3 * 2 * 16, s);
    c->d_gauss = dupload(gpf, 3 * 2 * 16, s);
    c->d_status = dalloc<int>(8);  // the second stream's counters follow
    c->d_status2 = c->d_status + 4;
    build_maps(*c, m->conn);
    int nsym = nsym_blocks(c->nn);
    c->kel = KeLayout{c->shape == FGPU_BSP ? 1 : 64, nsym};
    size_t ntile = (size_t)(c->nelem + c->kel.T - 1) / c->kel.T;
    if ((c->shape == FGPU_BSP && bsp_generic(*c)) || c->shape == FGPU_HEX20)
      c->d_Ke64 = dalloc<double>(ntile * c->kel.T * (size_t)nsym * 9);
    else
      c->d_Ke = dalloc<float>(ntile * c->kel.T * (size_t)nsym * 9);
    c->d_ge = dalloc<double>((size_t)c->nelem * c->nn * 3);
    if (c->shape == FGPU_BSP) {
      bsp_prepare_tables(*c, m->ext, m->cpw);
      bsp_alloc_pts(*c);
    }
    c->d_u = dalloc<double>((size_t)c->nnode * 3);
    size_t n = (size_t)std::max(3 * c->nnode, std::max(c->neq, 1));
    c->d_bn = dalloc<double>(n);
    c->d_b = dalloc<double>(n);
    c->d_x = dalloc<double>(n);
    c->d_r = dalloc<double>(n);
    c->d_p = dalloc<double>(n);
    c->d_w = dalloc<double>(n);
    c->d_z = dalloc<double>(n);
    c->d_dinv = dalloc<double>(n);
    c->nblocks_red = 1024;
    const char *ng = getenv("FGPU_NO_GRAPH");
    c->use_graphs = !(ng && ng[0] == '1');
    const char *tv = getenv("FGPU_BSP_TL_VALU");  // 1: B-spline TL tangents on the VALU kernel
    c->bsp_tl_valu = tv && tv[0] == '1';
    if (m->nbatch > 0) setup_batch(*c, *m);
    c->d_part = dalloc<double>(std::max<size_t>(4 * c->nblocks_red, 3 * (size_t)c->nchunk + 8));
    c->d_scal = dalloc<double>(16);
    c->d_done = dalloc<int>(4);
    FGPU_CHECK(hipHostMalloc(&c->h_done, 4 * sizeof(int)));
    FGPU_CHECK(hipHostMalloc(&c->h_status, 8 * sizeof(int)));
    FGPU_CHECK(hipHostMalloc(&c->h_scal, 16 * sizeof(double)));
    FGPU_CHECK(hipStreamSynchronize(s));
  } catch (const Error &e) {
    fprintf(stderr, "fgpu_create: %s\n", e.msg.c_str());
    destroy(c);
    return 0;
  }
  *out = reinterpret_cast<fgpu_ctx *>(c);
  return 1;
}

void fgpu_destroy(fgpu_ctx *ctx) { destroy(reinterpret_cast<Ctx *>(ctx)); }

const char *fgpu_last_error(fgpu_ctx *ctx) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  return c ? c->err.c_str() : "null context";
}

int fgpu_neq(fgpu_ctx *ctx) { return ctx ? reinterpret_cast<Ctx *>(ctx)->neq : 0; }

int fgpu_nbatch(fgpu_ctx *ctx) { return ctx ? reinterpret_cast<Ctx *>(ctx)->nbatch : 0; }

int fgpu_solve_batch_dev(fgpu_ctx *ctx, const double *b, double *x, int maxit, double rtol,
                         int *iters, double *relres) {
  return guarded(ctx, [&](Ctx &c) {
    FGPU_CHECK(hipEventRecord(c.ev[5], c.stream));
    int nconv = run_pcg_batch(c, b, x, maxit, rtol, iters, relres);
    FGPU_CHECK(hipEventRecord(c.ev[6], c.stream));
    FGPU_CHECK(hipEventSynchronize(c.ev[6]));
    c.last_ms[3] = elapsed(c, 5, 6);
    if (nconv < c.nbatch)
      c.err = std::to_string(c.nbatch - nconv) + " meshes did not reach the tolerance";
    return nconv;
  });
}

int fgpu_solve_batch(fgpu_ctx *ctx, const double *b, double *x, int maxit, double rtol,
                     int *iters, double *relres) {
  return guarded(ctx, [&](Ctx &c) {
    if (!b || !x) throw Error("null argument");
    size_t nb = sizeof(double) * c.neq;
    double *db = dalloc<double>(c.neq), *dx = dalloc<double>(c.neq);
    FGPU_CHECK(hipMemcpyAsync(db, b, nb, hipMemcpyHostToDevice, c.stream));
    int nconv = fgpu_solve_batch_dev(ctx, db, dx, maxit, rtol, iters, relres);
    FGPU_CHECK(hipMemcpyAsync(x, dx, nb, hipMemcpyDeviceToHost, c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(db);
    (void)hipFree(dx);
    return nconv;
  });
}

long long fgpu_nnz(fgpu_ctx *ctx) { return ctx ? reinterpret_cast<Ctx *>(ctx)->nnz : 0; }

int fgpu_block_counts(fgpu_ctx *ctx, long long *nbf, long long *nblk) {
  return guarded(ctx, [&](Ctx &c) {
    if (nbf) *nbf = c.nbf;
    if (nblk) *nblk = c.nblk;
    return 1;
  });
}

int fgpu_assemble_K_dev(fgpu_ctx *ctx, const double *d_u) {
  return guarded(ctx, [&](Ctx &c) { return stiffness_dev(c, d_u); });
}

int fgpu_assemble_K(fgpu_ctx *ctx, const double *u) {
  return guarded(ctx, [&](Ctx &c) {
    const double *du = nullptr;
    if (u) {
      FGPU_CHECK(hipMemcpyAsync(c.d_u, u, sizeof(double) * 3 * c.nnode,
                                hipMemcpyHostToDevice, c.stream));
      du = c.d_u;
    }
    return fgpu_assemble_K_dev(ctx, du);
  });
}

int fgpu_assemble_K_fint_dev(fgpu_ctx *ctx, const double *d_u, double *d_g) {
  return guarded(ctx, [&](Ctx &c) {
    if (!d_u || !d_g) throw Error("null argument");
    return stiffness_fint_dev(c, d_u, d_g);
  });
}

int fgpu_assemble_K_fint(fgpu_ctx *ctx, const double *u, double *g) {
  return guarded(ctx, [&](Ctx &c) {
    if (!u || !g) throw Error("null argument");
    FGPU_CHECK(hipMemcpyAsync(c.d_u, u, sizeof(double) * 3 * c.nnode,
                              hipMemcpyHostToDevice, c.stream));
    int ok = stiffness_fint_dev(c, c.d_u, c.d_b);
    FGPU_CHECK(hipMemcpyAsync(g, c.d_b, sizeof(double) * c.neq, hipMemcpyDeviceToHost,
                              c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    return ok;
  });
}

int fgpu_int_force_dev(fgpu_ctx *ctx, const double *d_u, double *d_g) {
  return guarded(ctx, [&](Ctx &c) { return int_force_dev(c, d_u, d_g); });
}

int fgpu_int_force_dev_noalloc(fgpu_ctx *ctx, const double *d_u, double *d_g) {
  return guarded(ctx, [&](Ctx &c) {
    if (!d_u || !d_g) throw Error("null argument");
    if (c.shape == FGPU_BSP && bsp_generic(c))
      throw Error("fgpu_int_force_dev_noalloc: B-spline solids of degree (3,3,3) or (3,3,2) on "
                  "4x4x4 Gauss points only (the other shapes allocate their point scratch)");
    const size_t need = sizeof(double) * (size_t)c.nelem * c.nn * 3;
    if (need > kScratchBytes) {
      char msg[200];
      std::snprintf(msg, sizeof msg,
                    "fgpu_int_force_dev_noalloc: the element vectors need %zu bytes, above the "
                    "fixed scratch of %zu bytes (at most %zu elements of %d nodes)",
                    need, kScratchBytes, kScratchBytes / (sizeof(double) * c.nn * 3), c.nn);
      throw Error(msg);
    }
    if (!c.scratch) throw Error("fgpu_int_force_dev_noalloc: no device scratch");
    // the element kernels write every entry of their element vectors before
    // the gather reads them, so nothing in the scratch needs clearing
    struct UseScratch {
      Ctx &c;
      double *keep;
      explicit UseScratch(Ctx &c_) : c(c_), keep(c_.d_ge) {
        c.d_ge = static_cast<double *>(c.scratch);
      }
      ~UseScratch() { c.d_ge = keep; }
    } use(c);
    return int_force_dev(c, d_u, d_g);
  });
}

int fgpu_int_force(fgpu_ctx *ctx, const double *u, double *g) {
  return guarded(ctx, [&](Ctx &c) {
    if (!u || !g) throw Error("null argument");
    FGPU_CHECK(hipMemcpyAsync(c.d_u, u, sizeof(double) * 3 * c.nnode,
                              hipMemcpyHostToDevice, c.stream));
    int ok = fgpu_int_force_dev(ctx, c.d_u, c.d_b);
    FGPU_CHECK(hipMemcpyAsync(g, c.d_b, sizeof(double) * c.neq, hipMemcpyDeviceToHost,
                              c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    return ok;
  });
}

int fgpu_add_diag(fgpu_ctx *ctx, int n, const int *eq, const double *v) {
  return guarded(ctx, [&](Ctx &c) {
    if (n <= 0) return 1;
    c.factored = false;
    c.bval32_valid = false;
    for (int k = 0; k < n; k++)
      if (eq[k] < 1 || eq[k] > c.neq) throw Error("fgpu_add_diag: equation out of range");
    int *d_eq = dupload(eq, n, c.stream);
    double *d_v = dupload(v, n, c.stream);
    launch_add_diag(c, n, d_eq, d_v);
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(d_eq);
    (void)hipFree(d_v);
    return 1;
  });
}

int fgpu_spmv_dev(fgpu_ctx *ctx, const double *x, double *y) {
  return guarded(ctx, [&](Ctx &c) {
    launch_eq_to_nodal(c, x, c.d_bn);
    launch_spmv_nodal(c, c.d_bn, c.d_w);
    launch_nodal_to_eq(c, c.d_w, y);
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    return 1;
  });
}

int fgpu_solve_dev(fgpu_ctx *ctx, const double *b, double *x, const double *x0,
                   int maxit, double rtol, int *iters, double *relres) {
  return guarded(ctx, [&](Ctx &c) {
    FGPU_CHECK(hipEventRecord(c.ev[5], c.stream));
    int conv = run_pcg(c, b, x, x0, maxit, rtol, iters, relres);
    FGPU_CHECK(hipEventRecord(c.ev[6], c.stream));
    FGPU_CHECK(hipEventSynchronize(c.ev[6]));
    c.last_ms[3] = elapsed(c, 5, 6);
    if (!conv) c.err = "PCG did not reach the tolerance";
    return conv;
  });
}

int fgpu_solve(fgpu_ctx *ctx, const double *b, double *x, const double *x0,
               int maxit, double rtol, int *iters, double *relres) {
  return guarded(ctx, [&](Ctx &c) {
    if (!b || !x) throw Error("null argument");
    size_t nb = sizeof(double) * c.neq;
    double *db = dalloc<double>(c.neq), *dx = dalloc<double>(c.neq), *dx0 = nullptr;
    FGPU_CHECK(hipMemcpyAsync(db, b, nb, hipMemcpyHostToDevice, c.stream));
    if (x0) {
      dx0 = dalloc<double>(c.neq);
      FGPU_CHECK(hipMemcpyAsync(dx0, x0, nb, hipMemcpyHostToDevice, c.stream));
    }
    int conv = fgpu_solve_dev(ctx, db, dx, dx0, maxit, rtol, iters, relres);
    FGPU_CHECK(hipMemcpyAsync(x, dx, nb, hipMemcpyDeviceToHost, c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(db);
    (void)hipFree(dx);
    if (dx0) (void)hipFree(dx0);
    return conv;
  });
}

int fgpu_get_csr(fgpu_ctx *ctx, long long *rowptr, int *col, double *val) {
  return guarded(ctx, [&](Ctx &c) {
    if (rowptr)
      FGPU_CHECK(hipMemcpyAsync(rowptr, c.d_rowptr, sizeof(long long) * (c.neq + 1),
                                hipMemcpyDeviceToHost, c.stream));
    if (col) {
      ensure_eq_csr(c);
      FGPU_CHECK(hipMemcpyAsync(col, c.d_col, sizeof(int) * c.nnz, hipMemcpyDeviceToHost,
                                c.stream));
    }
    if (val) {
      double *d = dalloc<double>(c.nnz);
      launch_csr_convert(c, d, 1);
      FGPU_CHECK(hipMemcpyAsync(val, d, sizeof(double) * c.nnz, hipMemcpyDeviceToHost,
                                c.stream));
      FGPU_CHECK(hipStreamSynchronize(c.stream));
      (void)hipFree(d);
    }
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    return 1;
  });
}

int fgpu_set_csr_values(fgpu_ctx *ctx, const double *val) {
  return guarded(ctx, [&](Ctx &c) {
    if (!val) throw Error("null argument");
    c.factored = false;
    c.bval32_valid = false;
    double *d = dupload(val, c.nnz, c.stream);
    launch_csr_convert(c, d, 0);
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(d);
    return 1;
  });
}

int fgpu_get_element_matrices(fgpu_ctx *ctx, int e0, int ne, float *Ke) {
  return guarded(ctx, [&](Ctx &c) {
    if (e0 < 0 || ne < 0 || e0 + ne > c.nelem) throw Error("element range out of bounds");
    size_t nd = 3 * (size_t)c.nn;
    float *d = dalloc<float>((size_t)ne * nd * nd);
    launch_expand_Ke(c, e0, ne, d);
    FGPU_CHECK(hipMemcpyAsync(Ke, d, sizeof(float) * ne * nd * nd, hipMemcpyDeviceToHost,
                              c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(d);
    return 1;
  });
}

// Extrapolation of point values to the element's nodes (cElmParam::
// CompTRMatrix, elc/elmparam.cpp:1444-1535): least-squares fit of
// a + b r + c s + d t over the Gauss points, evaluated at the nodes' natural
// coordinates (bricks: cShapeBrick8::NodeNatCoord, shp/shpsolid.cpp:173-185;
// B-spline solids: a uniform (p+1)^3 grid on [-1, 1]^3, r fastest,
// shp/shpbsp.cpp:906-923); one point along r: every node takes the mean.
static void upload_stress_extrap(Ctx &c) {
  if (c.d_TR) return;
  std::vector<double> pts;  // (r, s, t) per point, r fastest
  double g[3][16], w[16];
  for (int d = 0; d < 3; d++) gauss_rule(c.nip[d], g[d], w);
  for (int k = 0; k < c.nip[2]; k++)
    for (int j = 0; j < c.nip[1]; j++)
      for (int i = 0; i < c.nip[0]; i++) pts.insert(pts.end(), {g[0][i], g[1][j], g[2][k]});
  const int nq = (int)pts.size() / 3, nn = c.nn;
  std::vector<double> nodes;
  if (c.shape == FGPU_HEX8) {
    nodes = {-1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1, -1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1};
  } else if (c.shape == FGPU_HEX20) {  // cShapeBrick20::NodeNatCoord
    nodes = {-1, -1, 1, 0, -1, 1, 1, -1, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1, -1, 1, 1,
             -1, 0, 1, -1, -1, 0, 1, -1, 0, 1, 1, 0, -1, 1, 0, -1, -1, -1, 0, -1, -1,
             1, -1, -1, 1, 0, -1, 1, 1, -1, 0, 1, -1, -1, 1, -1, -1, 0, -1};
  } else {
    const int nb[3] = {c.deg[0] + 1, c.deg[1] + 1, c.deg[2] + 1};
    for (int k = 0; k < nb[2]; k++)
      for (int j = 0; j < nb[1]; j++)
        for (int i = 0; i < nb[0]; i++)
          nodes.insert(nodes.end(), {-1.0 + 2.0 * i / (nb[0] - 1), -1.0 + 2.0 * j / (nb[1] - 1),
                                     -1.0 + 2.0 * k / (nb[2] - 1)});
  }
  if ((int)nodes.size() != 3 * nn) throw Error("stress extrapolation: node count");
  std::vector<double> TR((size_t)nn * nq, 1.0 / nq);
  if (c.nip[0] > 1) {
    // normal matrix A = sum u u^T, u = (1, r, s, t), inverted by Gauss-Jordan
    double A[4][8] = {};
    for (int q = 0; q < nq; q++) {
      const double u[4] = {1.0, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) A[a][b] += u[a] * u[b];
    }
    for (int a = 0; a < 4; a++) A[a][4 + a] = 1.0;
    for (int k = 0; k < 4; k++) {
      const double p = A[k][k];
      if (!(p > 0.0)) throw Error("stress extrapolation: singular point set");
      for (int l = 0; l < 8; l++) A[k][l] /= p;
      for (int r = 0; r < 4; r++)
        if (r != k) {
          const double f = A[r][k];
          for (int l = 0; l < 8; l++) A[r][l] -= f * A[k][l];
        }
    }
    for (int n = 0; n < nn; n++) {
      const double cn[4] = {1.0, nodes[3 * n], nodes[3 * n + 1], nodes[3 * n + 2]};
      double ci[4] = {0, 0, 0, 0};
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) ci[b] += cn[a] * A[a][4 + b];
      for (int q = 0; q < nq; q++)
        TR[(size_t)n * nq + q] =
            ci[0] + ci[1] * pts[3 * q] + ci[2] * pts[3 * q + 1] + ci[3] * pts[3 * q + 2];
    }
  }
  c.d_TR = dupload(TR.data(), TR.size(), c.stream);
  FGPU_CHECK(hipStreamSynchronize(c.stream));
}

int fgpu_set_matrix_precision(fgpu_ctx *ctx, int bits) {
  return guarded(ctx, [&](Ctx &c) {
    if (bits != 32 && bits != 64) throw Error("matrix precision: 32 or 64 bits");
    c.mat_bits = bits;
    return 1;
  });
}

int fgpu_set_preconditioner(fgpu_ctx *ctx, int kind) {
  return guarded(ctx, [&](Ctx &c) {
    if (kind != 0 && kind != 1) throw Error("preconditioner: 0 (point Jacobi) or 1 (node blocks)");
    if (kind == 1 && !c.d_binv) c.d_binv = dalloc<double>((size_t)c.nnode * 9);
    c.precond = kind;
    return 1;
  });
}

int fgpu_factor(fgpu_ctx *ctx) {
  return guarded(ctx, [&](Ctx &c) { return band_factor(c); });
}

int fgpu_factor_solve_dev(fgpu_ctx *ctx, const double *b_dev, double *x_dev) {
  return guarded(ctx, [&](Ctx &c) {
    if (!b_dev || !x_dev) throw Error("null argument");
    if (!c.factored)
      throw Error("no factorization of the current matrix (call fgpu_factor after assembly)");
    FGPU_CHECK(hipEventRecord(c.ev[5], c.stream));
    band_solve(c, b_dev, x_dev);
    FGPU_CHECK(hipEventRecord(c.ev[6], c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    c.last_ms[3] = elapsed(c, 5, 6);
    return 1;
  });
}

int fgpu_factor_solve(fgpu_ctx *ctx, const double *b, double *x) {
  return guarded(ctx, [&](Ctx &c) {
    if (!b || !x) throw Error("null argument");
    double *d = dupload(b, std::max(c.neq, 1), c.stream);
    int ok = fgpu_factor_solve_dev(ctx, d, d);
    if (ok) {
      FGPU_CHECK(hipMemcpyAsync(x, d, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream));
      FGPU_CHECK(hipStreamSynchronize(c.stream));
    }
    (void)hipFree(d);
    return ok;
  });
}

int fgpu_direct_solve_dev(fgpu_ctx *ctx, const double *b_dev, double *x_dev) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.factored && !band_factor(c)) return 0;
    return fgpu_factor_solve_dev(ctx, b_dev, x_dev);
  });
}

int fgpu_direct_solve(fgpu_ctx *ctx, const double *b, double *x) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.factored && !band_factor(c)) return 0;
    return fgpu_factor_solve(ctx, b, x);
  });
}

int fgpu_factor_info(fgpu_ctx *ctx, int *half_band, int *neg_pivots, long long *bytes) {
  return guarded(ctx, [&](Ctx &c) {
    const int bw = band_width(c);
    if (half_band) *half_band = bw;
    if (neg_pivots) *neg_pivots = c.factored ? c.neg_pivots : -1;
    if (bytes) *bytes = (long long)c.neq * (bw + 1) * (long long)sizeof(double);
    return 1;
  });
}

int fgpu_set_stress_nodes(fgpu_ctx *ctx, int extrapolate) {
  return guarded(ctx, [&](Ctx &c) {
    if (!extrapolate && (c.shape != FGPU_BSP || c.form == FGPU_TL_NEOHOOKE))
      throw Error("nodal stresses evaluated at the nodes: B-spline solids with (E, nu) only");
    c.stress_nodes = extrapolate ? 1 : 0;
    return 1;
  });
}

int fgpu_num_points(fgpu_ctx *ctx) {
  return guarded(ctx, [&](Ctx &c) { return c.nip[0] * c.nip[1] * c.nip[2]; });
}

int fgpu_stresses_dev(fgpu_ctx *ctx, const double *u_dev, double *sig_ip_dev,
                      double *sig_node_dev, double *sig_avg_dev) {
  return guarded(ctx, [&](Ctx &c) {
    if (!u_dev || !sig_ip_dev) throw Error("null argument");
    upload_stress_extrap(c);
    double *tmp = nullptr;
    if (sig_avg_dev && !sig_node_dev)
      sig_node_dev = tmp = dalloc<double>((size_t)c.nelem * c.nn * 6);
    FGPU_CHECK(hipMemsetAsync(c.d_status, 0, 4 * sizeof(int), c.stream));
    const bool at_nodes = c.stress_nodes == 0 && sig_node_dev;
    if (c.shape == FGPU_HEX8)
      launch_hex8_stress(c, u_dev, sig_ip_dev, at_nodes ? nullptr : sig_node_dev);
    else if (c.shape == FGPU_HEX20)
      launch_hex20_stress(c, u_dev, sig_ip_dev, sig_node_dev);
    else
      launch_bsp_stress(c, u_dev, sig_ip_dev, at_nodes ? nullptr : sig_node_dev);
    if (at_nodes) gen_launch_node_stress(c, u_dev, sig_node_dev);
    if (sig_avg_dev) launch_avg_stress(c, sig_node_dev, sig_avg_dev);
    const int ok = check_status(c, "stresses");
    if (tmp) (void)hipFree(tmp);
    return ok;
  });
}

int fgpu_stresses(fgpu_ctx *ctx, const double *u, double *sig_ip, double *sig_node,
                  double *sig_avg) {
  return guarded(ctx, [&](Ctx &c) {
    if (!u || !sig_ip) throw Error("null argument");
    const size_t nq = (size_t)c.nip[0] * c.nip[1] * c.nip[2];
    const size_t n_ip = (size_t)c.nelem * nq * 6, n_node = (size_t)c.nelem * c.nn * 6,
                 n_avg = (size_t)c.nnode * 6;
    FGPU_CHECK(hipMemcpyAsync(c.d_u, u, sizeof(double) * 3 * c.nnode, hipMemcpyHostToDevice,
                              c.stream));
    double *d_ip = dalloc<double>(n_ip);
    double *d_node = sig_node || sig_avg ? dalloc<double>(n_node) : nullptr;
    double *d_avg = sig_avg ? dalloc<double>(n_avg) : nullptr;
    int ok = fgpu_stresses_dev(ctx, c.d_u, d_ip, d_node, d_avg);
    FGPU_CHECK(hipMemcpyAsync(sig_ip, d_ip, sizeof(double) * n_ip, hipMemcpyDeviceToHost,
                              c.stream));
    if (sig_node)
      FGPU_CHECK(hipMemcpyAsync(sig_node, d_node, sizeof(double) * n_node,
                                hipMemcpyDeviceToHost, c.stream));
    if (sig_avg)
      FGPU_CHECK(hipMemcpyAsync(sig_avg, d_avg, sizeof(double) * n_avg, hipMemcpyDeviceToHost,
                                c.stream));
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(d_ip);
    if (d_node) (void)hipFree(d_node);
    if (d_avg) (void)hipFree(d_avg);
    return ok;
  });
}

int fgpu_body_force_dev(fgpu_ctx *ctx, const double *b_dev, double *f_dev) {
  return guarded(ctx, [&](Ctx &c) {
    if (c.nbatch > 0) throw Error("body forces of a batch context are not supported");
    FGPU_CHECK(hipMemsetAsync(c.d_status, 0, 4 * sizeof(int), c.stream));
    if (c.shape == FGPU_HEX8)
      launch_hex8_body(c, b_dev);
    else if (c.shape == FGPU_HEX20)
      launch_hex20_body(c, b_dev);
    else
      launch_bsp_body(c, b_dev);
    launch_gather_fint(c, f_dev);
    return check_status(c, "body force");
  });
}

int fgpu_body_force(fgpu_ctx *ctx, const double *b, double *f) {
  return guarded(ctx, [&](Ctx &c) {
    if (!b || !f) throw Error("null argument");
    double *db = dupload(b, (size_t)c.nelem * 3, c.stream);
    double *df = dalloc<double>(std::max(c.neq, 1));
    int ok = fgpu_body_force_dev(ctx, db, df);
    if (ok) {
      FGPU_CHECK(hipMemcpyAsync(f, df, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream));
      FGPU_CHECK(hipStreamSynchronize(c.stream));
    }
    (void)hipFree(db);
    (void)hipFree(df);
    return ok;
  });
}

int fgpu_set_launch(fgpu_ctx *ctx, int hex8_te, int asm_mode) {
  return guarded(ctx, [&](Ctx &c) {
    if (hex8_te != 0) {
      const std::vector<int> ok = hex8_te_options(c);
      if (c.shape != FGPU_HEX8 || std::find(ok.begin(), ok.end(), hex8_te) == ok.end())
        throw Error("elements per workgroup " + std::to_string(hex8_te) +
                    " not available for this mesh");
    }
    if (asm_mode != -1) {
      const std::vector<int> ok = assembly_options(c);
      if (std::find(ok.begin(), ok.end(), asm_mode) == ok.end())
        throw Error("assembly mode " + std::to_string(asm_mode) + " not available for this mesh");
    }
    c.hex8_te = hex8_te;
    c.asm_mode = asm_mode;
    c.tuned = true;  // an explicit choice replaces the autotuner
    c.tune_report = "set by the caller\n";
    return 1;
  });
}

int fgpu_add_surface_loads(fgpu_ctx *ctx, const fgpu_surface *sl) {
  return guarded(ctx, [&](Ctx &c) {
    if (!sl) throw Error("null argument");
    if (sl->nface < 0) throw Error("surface loads: negative count");
    if (sl->nface == 0) return 1;
    for (int d = 0; d < 2; d++)
      if (sl->deg[d] < 1 || sl->deg[d] > 7) throw Error("surface loads: degrees 1..7");
    if (sl->nip < 1 || sl->nip > 16) throw Error("surface loads: 1..16 Gauss points");
    if (!sl->nodes || !sl->q) throw Error("surface loads: nodes and q are required");
    if (sl->shape != 0 && sl->shape != 1) throw Error("surface loads: shape 0 or 1");
    if (sl->shape == 1 && (sl->deg[0] != 2 || sl->deg[1] != 2 || sl->ext))
      throw Error("surface loads: serendipity faces are degree (2, 2) without extraction");
    const int n1 = sl->deg[0] + 1, n2 = sl->deg[1] + 1, nb = sl->shape == 1 ? 8 : n1 * n2;
    const size_t nn = (size_t)sl->nface * nb;
    for (size_t k = 0; k < nn; k++)
      if (sl->nodes[k] < 0 || sl->nodes[k] >= c.nnode)
        throw Error("surface loads: node out of range");
    if (c.nsrow + (long long)nn >= (1LL << 31)) throw Error("surface loads: too many face nodes");
    SurfaceGroup g;
    g.nface = sl->nface;
    g.deg[0] = sl->deg[0];
    g.deg[1] = sl->deg[1];
    g.nip = sl->nip;
    g.shape = sl->shape;
    g.row0 = c.nsrow;
    g.h_nodes.assign(sl->nodes, sl->nodes + nn);
    g.d_nodes = dupload(sl->nodes, nn, c.stream);
    if (sl->ext) g.d_ext = dupload(sl->ext, (size_t)sl->nface * (n1 * n1 + n2 * n2), c.stream);
    if (sl->local) g.d_local = dupload(sl->local, (size_t)sl->nface, c.stream);
    g.d_q = dupload(sl->q, (size_t)sl->nface * 3, c.stream);
    double r[16], w[16], gp[32];
    gauss_rule(sl->nip, r, w);
    for (int k = 0; k < sl->nip; k++) {
      gp[2 * k] = r[k];
      gp[2 * k + 1] = w[k];
    }
    g.d_gp = dupload(gp, 2 * (size_t)sl->nip, c.stream);
    c.sgroups.push_back(std::move(g));
    c.nsrow += (long long)nn;
    // face-vector rows of every group and the gather lists: each node's rows
    // in increasing row order, i.e. in the order the entries were given
    std::vector<int> ptr(c.nnode + 1, 0), adj(c.nsrow);
    for (const auto &gr : c.sgroups)
      for (int n : gr.h_nodes) ptr[n + 1]++;
    for (int n = 0; n < c.nnode; n++) ptr[n + 1] += ptr[n];
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (const auto &gr : c.sgroups)
      for (size_t k = 0; k < gr.h_nodes.size(); k++)
        adj[pos[gr.h_nodes[k]]++] = (int)(gr.row0 + (long long)k);
    for (void *p : {(void *)c.d_sfv, (void *)c.d_sptr, (void *)c.d_sadj})
      if (p) (void)hipFree(p);
    c.d_sfv = dalloc<double>((size_t)c.nsrow * 3);
    c.d_sptr = dupload(ptr.data(), ptr.size(), c.stream);
    c.d_sadj = dupload(adj.data(), adj.size(), c.stream);
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    return 1;
  });
}

int fgpu_clear_surface_loads(fgpu_ctx *ctx) {
  return guarded(ctx, [&](Ctx &c) {
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    for (auto &g : c.sgroups) {
      void *gb[] = {g.d_nodes, g.d_ext, g.d_local, g.d_q, g.d_gp};
      for (void *p : gb)
        if (p) (void)hipFree(p);
    }
    c.sgroups.clear();
    for (void *p : {(void *)c.d_sfv, (void *)c.d_sptr, (void *)c.d_sadj})
      if (p) (void)hipFree(p);
    c.d_sfv = nullptr;
    c.d_sptr = c.d_sadj = nullptr;
    c.nsrow = 0;
    return 1;
  });
}

int fgpu_surface_load_dev(fgpu_ctx *ctx, double *f_dev) {
  return guarded(ctx, [&](Ctx &c) {
    if (!f_dev) throw Error("null argument");
    if (c.sgroups.empty()) {
      FGPU_CHECK(hipMemsetAsync(f_dev, 0, sizeof(double) * c.neq, c.stream));
    } else {
      for (const auto &g : c.sgroups) launch_surface_group(c, g, c.d_sfv + 3 * g.row0);
      launch_gather_surface(c, f_dev);
    }
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    return 1;
  });
}

int fgpu_surface_load(fgpu_ctx *ctx, double *f) {
  return guarded(ctx, [&](Ctx &c) {
    if (!f) throw Error("null argument");
    if (!fgpu_surface_load_dev(ctx, c.d_b)) return 0;
    FGPU_CHECK(hipMemcpy(f, c.d_b, sizeof(double) * c.neq, hipMemcpyDeviceToHost));
    return 1;
  });
}

int fgpu_set_penalty(fgpu_ctx *ctx, int n, const int *eq, const double *v) {
  return guarded(ctx, [&](Ctx &c) {
    if (n < 0 || (n > 0 && (!eq || !v))) throw Error("fgpu_set_penalty: bad arguments");
    std::vector<int> eqnode(c.neq, -1);
    for (int i = 0; i < 3 * c.nnode; i++)
      if (c.eqmap[i] >= 0) eqnode[c.eqmap[i]] = i;
    // one entry per equation, repeated equations summed in the given order
    std::vector<int> nd;
    std::vector<double> val;
    std::vector<int> slot(3 * (size_t)c.nnode, -1);
    for (int k = 0; k < n; k++) {
      if (eq[k] < 1 || eq[k] > c.neq) throw Error("fgpu_set_penalty: equation out of range");
      const int i = eqnode[eq[k] - 1];
      if (slot[i] < 0) {
        slot[i] = (int)nd.size();
        nd.push_back(i);
        val.push_back(v[k]);
      } else {
        val[slot[i]] += v[k];
      }
    }
    FGPU_CHECK(hipStreamSynchronize(c.stream));
    if (c.d_pnd) (void)hipFree(c.d_pnd);
    if (c.d_pval) (void)hipFree(c.d_pval);
    c.d_pnd = nullptr;
    c.d_pval = nullptr;
    c.npen = (int)nd.size();
    if (c.npen) {
      c.d_pnd = dupload(nd.data(), nd.size(), c.stream);
      c.d_pval = dupload(val.data(), val.size(), c.stream);
      FGPU_CHECK(hipStreamSynchronize(c.stream));
    }
    c.factored = false;
    c.bval32_valid = false;
    return 1;
  });
}

int fgpu_set_overlap(fgpu_ctx *ctx, int mode) {
  return guarded(ctx, [&](Ctx &c) {
    if (mode < -1 || mode > 3) throw Error("overlap mode: -1, 0, 1, 2 or 3");
    c.overlap = mode;
    return 1;
  });
}

const char *fgpu_tuning_report(fgpu_ctx *ctx) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  return c ? c->tune_report.c_str() : "";
}

double fgpu_last_time_ms(fgpu_ctx *ctx, int phase) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || phase < 0 || phase > 3) return -1.0;
  return c->last_ms[phase];
}

void *fgpu_device_buffer(fgpu_ctx *ctx, int which) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return nullptr;
  switch (which) {
    case 0: return c->d_bval;
    case 1: return c->d_Ke;
    case 2: return c->d_u;
    default: return nullptr;
  }
}

}  // extern "C"
