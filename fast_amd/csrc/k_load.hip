// k_load.hip - external loads on the device: uniform surface loads on element
// faces, and the prescribed-equation penalty of the stiffness and the
// internal force.
//
// Surface loads (cFaceUnifLoad, load/load.cpp:1440-1500, and
// cFaceUnifLoadIGA, load/loadiga.cpp:305-354, integrated by
// cElemLoad::ExtForce, load/load.cpp:618-691): each entry is one face, a
// tensor-product surface of degree (p1, p2) over (p1+1)(p2+1) nodes (first
// parameter fastest) with Bezier extraction rows per parameter; the basis is
// N_ij = (C1 B(s'))_i (C2 B(t'))_j on the face's parent square, made rational
// with the nodes' weights.  At each of nip x nip Gauss points (the first
// parameter fastest, as cIntPoint::CreateQuadPoints orders them,
// elc/intpoint.cpp:266-300)
//   f_a += w_s w_t |x_,s x x_,t| R_a load,
// load = q in global axes, or in the face's local frame (local = 1;
// cShapeSurf::LocalSys, shp/shpsurf.cpp:121-164): u = x_,s / |x_,s|,
// w = u x x_,t / |x_,t| (not normalised), v = w x u, load = q0 u + q1 v +
// q2 w.  A brick face is the degree-(1,1) surface over its four corners with
// identity extraction; a face of a 20-node brick is the 8-node serendipity
// quadrilateral (surface_q8_kernel).  One thread per face accumulates its nodes' vectors in
// point order; a gather sums every node's entries in entry order (FP64, no
// atomics).
#include "fgpu_dev.h"

namespace fgpu {
namespace {

constexpr int SMAXB = 8;  // basis functions per parameter (degree <= 7)

// Bernstein polynomials of degree p at u in [0, 1] and their derivatives
// with respect to u (de Casteljau-style recurrence)
__device__ inline void bernstein(int p, double u, double *B, double *dB) {
  double lo[SMAXB];
  for (int j = 0; j <= p; j++) B[j] = 0.0;
  B[0] = 1.0;
  for (int k = 1; k <= p; k++) {
    if (k == p)
      for (int j = 0; j < p; j++) lo[j] = B[j];  // degree p - 1
    double prev = 0.0;
    for (int j = 0; j <= k; j++) {
      const double cur = j < k ? B[j] : 0.0;
      B[j] = (1.0 - u) * cur + u * prev;
      prev = cur;
    }
  }
  if (p == 0) {
    dB[0] = 0.0;
    return;
  }
  for (int j = 0; j <= p; j++)
    dB[j] = p * ((j > 0 ? lo[j - 1] : 0.0) - (j < p ? lo[j] : 0.0));
}

__global__ __launch_bounds__(64) void surface_load_kernel(
    int nface, int p1, int p2, int nip, const int *__restrict__ nodes,
    const double *__restrict__ ext, const int *__restrict__ local, const double *__restrict__ q,
    const double *__restrict__ coords, const double *__restrict__ cpw,
    const double *__restrict__ gp, double *__restrict__ fv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nface) return;
  const int n1 = p1 + 1, n2 = p2 + 1, nb = n1 * n2;
  const int *nd = nodes + (size_t)e * nb;
  const double *C1 = ext ? ext + (size_t)e * (n1 * n1 + n2 * n2) : nullptr;
  const double *C2 = C1 ? C1 + n1 * n1 : nullptr;
  double *out = fv + (size_t)e * nb * 3;
  for (int k = 0; k < 3 * nb; k++) out[k] = 0.0;
  const double qx = q[3 * e], qy = q[3 * e + 1], qz = q[3 * e + 2];
  const bool loc = local && local[e];
  for (int b = 0; b < nip; b++) {  // points of the reference's quad rule: r fastest
    double B2[SMAXB], dB2[SMAXB], N2[SMAXB], dN2[SMAXB];
    bernstein(p2, 0.5 * (gp[2 * b] + 1.0), B2, dB2);
    for (int i = 0; i < n2; i++) {
      double s = 0.0, ds = 0.0;
      for (int j = 0; j < n2; j++) {
        const double c = C2 ? C2[i * n2 + j] : (i == j ? 1.0 : 0.0);
        s += c * B2[j];
        ds += c * dB2[j];
      }
      N2[i] = s;
      dN2[i] = 0.5 * ds;  // d/dt of the parent coordinate t in [-1, 1]
    }
    for (int a = 0; a < nip; a++) {
      double B1[SMAXB], dB1[SMAXB], N1[SMAXB], dN1[SMAXB];
      bernstein(p1, 0.5 * (gp[2 * a] + 1.0), B1, dB1);
      for (int i = 0; i < n1; i++) {
        double s = 0.0, ds = 0.0;
        for (int j = 0; j < n1; j++) {
          const double c = C1 ? C1[i * n1 + j] : (i == j ? 1.0 : 0.0);
          s += c * B1[j];
          ds += c * dB1[j];
        }
        N1[i] = s;
        dN1[i] = 0.5 * ds;
      }
      // rational basis sums: W = sum N W_k, W1 / W2 its derivatives
      double sw = 0.0, sw1 = 0.0, sw2 = 0.0;
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          const double w = cpw ? cpw[nd[i + n1 * j]] : 1.0;
          sw += N1[i] * N2[j] * w;
          sw1 += dN1[i] * N2[j] * w;
          sw2 += N1[i] * dN2[j] * w;
        }
      double x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0};
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          const int n = nd[i + n1 * j];
          const double w = cpw ? cpw[n] : 1.0;
          const double R = N1[i] * N2[j] * w / sw;
          const double R1 = (dN1[i] * N2[j] * w - R * sw1) / sw;
          const double R2 = (N1[i] * dN2[j] * w - R * sw2) / sw;
          for (int k = 0; k < 3; k++) {
            x1[k] += R1 * coords[3 * (size_t)n + k];
            x2[k] += R2 * coords[3 * (size_t)n + k];
          }
        }
      const double cx = x1[1] * x2[2] - x1[2] * x2[1], cy = x1[2] * x2[0] - x1[0] * x2[2],
                   cz = x1[0] * x2[1] - x1[1] * x2[0];
      // coefficient of cElemLoad::ExtForce: weight x det J (|x_,s x x_,t|)
      const double dA = (gp[2 * a + 1] * gp[2 * b + 1]) * sqrt(cx * cx + cy * cy + cz * cz);
      double L[3] = {qx, qy, qz};
      if (loc) {
        const double l1 = sqrt(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
        const double l2 = sqrt(x2[0] * x2[0] + x2[1] * x2[1] + x2[2] * x2[2]);
        const double u[3] = {x1[0] / l1, x1[1] / l1, x1[2] / l1};
        const double t[3] = {x2[0] / l2, x2[1] / l2, x2[2] / l2};
        const double w[3] = {u[1] * t[2] - u[2] * t[1], u[2] * t[0] - u[0] * t[2],
                             u[0] * t[1] - u[1] * t[0]};
        const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2],
                             w[0] * u[1] - w[1] * u[0]};
        for (int k = 0; k < 3; k++) L[k] = qx * u[k] + qy * v[k] + qz * w[k];
      }
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          const int n = nd[i + n1 * j];
          const double w = cpw ? cpw[n] : 1.0;
          const double R = N1[i] * N2[j] * w / sw;
          if (R == 0.0) continue;
          double *o = out + 3 * (i + n1 * j);
          for (int k = 0; k < 3; k++) o[k] += R * (dA * L[k]);  // MultTAcc's order
        }
    }
  }
}

// Faces of 20-node bricks: the 8-node serendipity quadrilateral of cShapeQ8
// (shp/shpplane.cpp:428-486), nodes counterclockwise from (-1, -1), corners
// and mid-edge nodes alternating.  Corner c: N = (1 + r cr)(1 + s cs)
// (r cr + s cs - 1) / 4; mid-edge node with coordinate k = 0:
// N = (1 - x_k^2)(1 + x_j c_j) / 2.  Points, dA and the local frame as in
// surface_load_kernel.
__constant__ signed char kQ8[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0},
                                      {1, 1},   {0, 1},  {-1, 1}, {-1, 0}};

__device__ inline double q8_shape(double r, double s, int a, double &dr, double &ds) {
  const double cr = kQ8[a][0], cs = kQ8[a][1];
  const double fr = 1.0 + r * cr, fs = 1.0 + s * cs;
  if (cr != 0.0 && cs != 0.0) {
    const double g = r * cr + s * cs - 1.0;
    dr = 0.25 * cr * fs * (g + fr);
    ds = 0.25 * cs * fr * (g + fs);
    return 0.25 * fr * fs * g;
  }
  if (cr == 0.0) {
    dr = -r * fs;
    ds = 0.5 * (1.0 - r * r) * cs;
    return 0.5 * (1.0 - r * r) * fs;
  }
  dr = 0.5 * (1.0 - s * s) * cr;
  ds = -s * fr;
  return 0.5 * (1.0 - s * s) * fr;
}

__global__ __launch_bounds__(64) void surface_q8_kernel(
    int nface, int nip, const int *__restrict__ nodes, const int *__restrict__ local,
    const double *__restrict__ q, const double *__restrict__ coords,
    const double *__restrict__ gp, double *__restrict__ fv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nface) return;
  const int *nd = nodes + (size_t)e * 8;
  double *out = fv + (size_t)e * 24;
  for (int k = 0; k < 24; k++) out[k] = 0.0;
  const double qx = q[3 * e], qy = q[3 * e + 1], qz = q[3 * e + 2];
  const bool loc = local && local[e];
  for (int b = 0; b < nip; b++)
    for (int a = 0; a < nip; a++) {
      const double r = gp[2 * a], s = gp[2 * b];
      double N[8], x1[3] = {0, 0, 0}, x2[3] = {0, 0, 0};
      for (int k = 0; k < 8; k++) {
        double dr, ds;
        N[k] = q8_shape(r, s, k, dr, ds);
        const double *X = coords + 3 * (size_t)nd[k];
        for (int d = 0; d < 3; d++) {
          x1[d] += dr * X[d];
          x2[d] += ds * X[d];
        }
      }
      const double cx = x1[1] * x2[2] - x1[2] * x2[1], cy = x1[2] * x2[0] - x1[0] * x2[2],
                   cz = x1[0] * x2[1] - x1[1] * x2[0];
      const double dA = (gp[2 * a + 1] * gp[2 * b + 1]) * sqrt(cx * cx + cy * cy + cz * cz);
      double L[3] = {qx, qy, qz};
      if (loc) {
        const double l1 = sqrt(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
        const double l2 = sqrt(x2[0] * x2[0] + x2[1] * x2[1] + x2[2] * x2[2]);
        const double u[3] = {x1[0] / l1, x1[1] / l1, x1[2] / l1};
        const double t[3] = {x2[0] / l2, x2[1] / l2, x2[2] / l2};
        const double w[3] = {u[1] * t[2] - u[2] * t[1], u[2] * t[0] - u[0] * t[2],
                             u[0] * t[1] - u[1] * t[0]};
        const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2],
                             w[0] * u[1] - w[1] * u[0]};
        for (int k = 0; k < 3; k++) L[k] = qx * u[k] + qy * v[k] + qz * w[k];
      }
      for (int k = 0; k < 8; k++)
        for (int d = 0; d < 3; d++) out[3 * k + d] += N[k] * (dA * L[d]);
    }
}

// One thread per node direction: the node's face-vector entries summed in
// entry order (the adjacency lists them so), written to its equation.
__global__ __launch_bounds__(256) void gather_surface_kernel(
    int nnode, const int *__restrict__ sptr, const int *__restrict__ sadj,
    const double *__restrict__ fv, const int *__restrict__ eqmap, double *__restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nnode) return;
  const int A = i / 3, d = i - 3 * A;
  double s = 0.0;
  for (int t = sptr[A]; t < sptr[A + 1]; t++) s += fv[(size_t)sadj[t] * 3 + d];
  const int eq = eqmap[i];
  if (eq >= 0) f[eq] = s;
}

// penalty stiffness on the diagonal of the prescribed equations (after the
// element contributions, as cControl::StiffnessMatrix adds it)
__global__ __launch_bounds__(256) void penalty_diag_kernel(int npen, const int *__restrict__ pnd,
                                                           const double *__restrict__ pval,
                                                           const long long *__restrict__ bdiag,
                                                           double *__restrict__ bval) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npen) return;
  const int i = pnd[k], A = i / 3, d = i - 3 * A;
  bval[9 * bdiag[A] + 4 * d] += pval[k];
}

// penalty force of the prescribed equations, after the element gather
// (cControl::IntForceVector, ctrl/ctrl.cpp:794-810): g += penalty * u
__global__ __launch_bounds__(256) void penalty_force_kernel(int npen, const int *__restrict__ pnd,
                                                            const double *__restrict__ pval,
                                                            const int *__restrict__ eqmap,
                                                            const double *__restrict__ u,
                                                            double *__restrict__ g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npen) return;
  const int i = pnd[k];
  g[eqmap[i]] += pval[k] * u[i];
}

// x[eq] = 0 at the prescribed equations (the norms of the Newton test run
// over the free equations, cControl::Convergence, ctrl/ctrl.cpp:1391-1413)
__global__ __launch_bounds__(256) void zero_penalty_kernel(int npen, const int *__restrict__ pnd,
                                                           const int *__restrict__ eqmap,
                                                           double *__restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npen) return;
  x[eqmap[pnd[k]]] = 0.0;
}

}  // namespace

void launch_zero_penalty(Ctx &c, double *x) {
  if (c.npen == 0) return;
  hipLaunchKernelGGL(zero_penalty_kernel, dim3((c.npen + 255) / 256), dim3(256), 0, c.stream,
                     c.npen, c.d_pnd, c.d_eqmap, x);
  FGPU_CHECK(hipGetLastError());
}

void launch_penalty_force(Ctx &c, const double *u, double *g) {
  if (c.npen == 0) return;
  hipLaunchKernelGGL(penalty_force_kernel, dim3((c.npen + 255) / 256), dim3(256), 0, c.stream,
                     c.npen, c.d_pnd, c.d_pval, c.d_eqmap, u, g);
  FGPU_CHECK(hipGetLastError());
}

void launch_surface_group(Ctx &c, const SurfaceGroup &g, double *fv) {
  if (g.nface == 0) return;
  if (g.shape == 1) {
    hipLaunchKernelGGL(surface_q8_kernel, dim3((g.nface + 63) / 64), dim3(64), 0, c.stream,
                       g.nface, g.nip, g.d_nodes, g.d_local, g.d_q, c.d_coords, g.d_gp, fv);
    FGPU_CHECK(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(surface_load_kernel, dim3((g.nface + 63) / 64), dim3(64), 0, c.stream,
                     g.nface, g.deg[0], g.deg[1], g.nip, g.d_nodes, g.d_ext, g.d_local, g.d_q,
                     c.d_coords, c.d_cpw, g.d_gp, fv);
  FGPU_CHECK(hipGetLastError());
}

void launch_gather_surface(Ctx &c, double *f) {
  hipLaunchKernelGGL(gather_surface_kernel, dim3((3 * c.nnode + 255) / 256), dim3(256), 0,
                     c.stream, c.nnode, c.d_sptr, c.d_sadj, c.d_sfv, c.d_eqmap, f);
  FGPU_CHECK(hipGetLastError());
}

void launch_penalty_diag(Ctx &c) {
  if (c.npen == 0) return;
  hipLaunchKernelGGL(penalty_diag_kernel, dim3((c.npen + 255) / 256), dim3(256), 0, c.stream,
                     c.npen, c.d_pnd, c.d_pval, c.d_bdiag, c.d_bval);
  FGPU_CHECK(hipGetLastError());
}

}  // namespace fgpu
