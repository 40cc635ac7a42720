/* fast_oracle.c - CPU FP64 restatement of FAST's element-integration,
 * assembly and linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path (tests, smoke check,
 * CPU-baseline leg of the benchmark).  The product library never links it.
 *
 * Operation order follows the reference routines cited at each function so
 * that, built with -ffp-contract=off, element matrices and vectors agree bit
 * for bit with the reference build.  Neo-Hookean is not in the reference; its
 * routines are this project's own and are pinned by finite differences and by
 * their small-strain limit (which equals the reference's linear model).
 */
#define _GNU_SOURCE
#include "fast_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNN 512   /* nodes per element upper bound (p <= 7 B-splines) */
#define MAXDOF (3 * MAXNN)

/* ------------------------------------------------------------------------
 * Quadrature.  Reference: elc/intpoint.cpp:741-920 tabulates Gauss-Legendre
 * abscissae and weights to 15 decimals.  We compute the rule by Newton
 * iteration on P_n in long double and round each value to 15 decimals through
 * its decimal text, which yields the same doubles as the reference's literals.
 * ---------------------------------------------------------------------- */
static double round15(long double x)
{
  char buf[64];
  snprintf(buf, sizeof buf, "%.15Lf", x);
  if (strcmp(buf, "-0.000000000000000") == 0) return 0.0;
  return strtod(buf, NULL);
}

int orc_gauss(int n, double *r, double *w)
{
  if (n < 1 || n > 32) return 0;
  const long double PI = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; i++) {
    /* root i in ascending order */
    long double x = -cosl(PI * (i + 0.75L) / (n + 0.5L));
    long double dp = 1.0L;
    for (int it = 0; it < 100; it++) {
      long double p0 = 1.0L, p1 = x;
      if (n == 1) { p1 = x; p0 = 1.0L; }
      for (int k = 2; k <= n; k++) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1; p1 = p2;
      }
      /* p1 = P_n(x), p0 = P_{n-1}(x) */
      dp = n * (x * p1 - p0) / (x * x - 1.0L);
      long double dx = p1 / dp;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    if (n == 1) { x = 0.0L; dp = 1.0L; }
    r[i] = round15(x);
    w[i] = round15(2.0L / ((1.0L - x * x) * dp * dp));
  }
  return 1;
}

/* Reference: elc/intpoint.cpp:579-617 (CreateBrickPoints): r fastest,
 * weight = w_r * w_s * w_t in that multiplication order. */
int orc_brick_points(const int nip[3], double *rst, double *w)
{
  double gr[3][32], gw[3][32];
  for (int d = 0; d < 3; d++)
    if (!orc_gauss(nip[d], gr[d], gw[d])) return 0;
  int k = 0;
  for (int i = 0; i < nip[2]; i++)
    for (int j = 0; j < nip[1]; j++)
      for (int l = 0; l < nip[0]; l++) {
        rst[3 * k + 0] = gr[0][l];
        rst[3 * k + 1] = gr[1][j];
        rst[3 * k + 2] = gr[2][i];
        w[k] = gw[0][l] * gw[1][j] * gw[2][i];
        k++;
      }
  return k;
}

/* ------------------------------------------------------------------------
 * Hex8 shape functions.  Reference: shp/shpsolid.cpp:173-238 (node order:
 * t = +1 face first, counter-clockwise in (r,s); then t = -1 face).
 * dN is [8][3] = (d/dr, d/ds, d/dt).
 * ---------------------------------------------------------------------- */
void orc_hex8_shape(double r, double s, double t, double *N, double *dN)
{
  static const double nr[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
  static const double ns[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  static const double nt[8] = {1, 1, 1, 1, -1, -1, -1, -1};
  for (int a = 0; a < 8; a++) {
    double fr = 1.0 + nr[a] * r, fs = 1.0 + ns[a] * s, ft = 1.0 + nt[a] * t;
    /* (1 +- r) written as (1.0 - r) or (1.0 + r): same double either way */
    if (N) N[a] = 0.125 * fr * fs * ft;
    if (dN) {
      dN[3 * a + 0] = (nr[a] > 0 ? 0.125 : -0.125) * fs * ft;
      dN[3 * a + 1] = (ns[a] > 0 ? 0.125 : -0.125) * fr * ft;
      dN[3 * a + 2] = (nt[a] > 0 ? 0.125 : -0.125) * fr * fs;
    }
  }
}

/* ------------------------------------------------------------------------
 * 20-node serendipity brick (cShapeBrick20, shp/shpsolid.cpp:301-435): the
 * node order of NodeNatCoord (t = +1 face first, counterclockwise from
 * (-1,-1), corners and mid-edges alternating; then the four mid-height
 * nodes; then the t = -1 face alike).  Functions from the serendipity
 * definition: a corner node c = (cr, cs, ct) has
 *   N = (1 + r cr)(1 + s cs)(1 + t ct)(r cr + s cs + t ct - 2) / 8,
 * a mid-edge node with coordinate k = 0
 *   N = (1 - x_k^2) prod_{j != k} (1 + x_j c_j) / 4.
 * ---------------------------------------------------------------------- */
static const double kHex20Nat[20][3] = {
    {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {-1, 1, 1},
    {-1, 0, 1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, -1, -1},
    {0, -1, -1}, {1, -1, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}, {-1, 1, -1}, {-1, 0, -1}};

void orc_hex20_shape(double r, double s, double t, double *N, double *dN)
{
  const double x[3] = {r, s, t};
  for (int a = 0; a < 20; a++) {
    const double *c = kHex20Nat[a];
    int k = -1;
    for (int j = 0; j < 3; j++)
      if (c[j] == 0.0) k = j;
    double f[3];
    for (int j = 0; j < 3; j++) f[j] = 1.0 + x[j] * c[j];
    if (k < 0) {  /* corner */
      const double g = x[0] * c[0] + x[1] * c[1] + x[2] * c[2] - 2.0;
      if (N) N[a] = 0.125 * f[0] * f[1] * f[2] * g;
      if (dN)
        for (int d = 0; d < 3; d++) {
          /* d/dx_d of f_d g, times the other two factors */
          const double fo = f[(d + 1) % 3] * f[(d + 2) % 3];
          dN[3 * a + d] = 0.125 * c[d] * fo * (g + f[d]);
        }
    } else {  /* mid-edge, coordinate k zero */
      const int j1 = (k + 1) % 3, j2 = (k + 2) % 3;
      const double q = 1.0 - x[k] * x[k];
      if (N) N[a] = 0.25 * q * f[j1] * f[j2];
      if (dN) {
        dN[3 * a + k] = -0.5 * x[k] * f[j1] * f[j2];
        dN[3 * a + j1] = 0.25 * q * c[j1] * f[j2];
        dN[3 * a + j2] = 0.25 * q * f[j1] * c[j2];
      }
    }
  }
}

/* 8-node serendipity quadrilateral (cShapeQ8, shp/shpplane.cpp:428-486):
 * nodes counterclockwise from (-1,-1), corners and mid-edges alternating;
 * corner N = (1 + r cr)(1 + s cs)(r cr + s cs - 1) / 4, mid-edge
 * N = (1 - x_k^2)(1 + x_j c_j) / 2.  dN: [8][2]. */
void orc_q8_shape(double r, double s, double *N, double *dN)
{
  static const double c8[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0},
                                  {1, 1},   {0, 1},  {-1, 1}, {-1, 0}};
  const double x[2] = {r, s};
  for (int a = 0; a < 8; a++) {
    const double *c = c8[a];
    if (c[0] != 0.0 && c[1] != 0.0) {
      const double f0 = 1.0 + r * c[0], f1 = 1.0 + s * c[1];
      const double g = r * c[0] + s * c[1] - 1.0;
      N[a] = 0.25 * f0 * f1 * g;
      dN[2 * a] = 0.25 * c[0] * f1 * (g + f0);
      dN[2 * a + 1] = 0.25 * c[1] * f0 * (g + f1);
    } else {
      const int k = c[0] == 0.0 ? 0 : 1, j = 1 - k;
      const double q = 1.0 - x[k] * x[k], fj = 1.0 + x[j] * c[j];
      N[a] = 0.5 * q * fj;
      dN[2 * a + k] = -x[k] * fj;
      dN[2 * a + j] = 0.5 * q * c[j];
    }
  }
}

/* ------------------------------------------------------------------------
 * Bernstein basis on [0,1].  Reference: geom/bernbasis.cpp:91-106
 * (CompPols, de Casteljau-style triangle) and 110-124 (CompDervPols:
 * n * (B_{i-1,n-1} - B_{i,n-1})).
 * ---------------------------------------------------------------------- */
void orc_bernstein(int n, double u, double *B)
{
  B[0] = 1.0;
  double u1 = 1.0 - u;
  for (int j = 1; j <= n; j++) {
    double saved = 0.0;
    for (int k = 0; k < j; k++) {
      double temp = B[k];
      B[k] = saved + u1 * temp;
      saved = u * temp;
    }
    B[j] = saved;
  }
}

void orc_bernstein_deriv(int n, double u, double *dB)
{
  orc_bernstein(n - 1, u, dB);
  dB[n] = 0.0;
  double prev = 0.0;
  for (int i = 0; i <= n; i++) {
    double curr = dB[i];
    dB[i] = n * (prev - curr);
    prev = curr;
  }
}

/* ------------------------------------------------------------------------
 * Bezier extraction (Borden, Scott, Evans, Hughes 2011, Algorithm 1).
 * Reference: geom/knotvec.cpp:415-493.  C is [nspan][p+1][p+1] row-major:
 * row = local B-spline basis, column = Bernstein polynomial.
 * Returns the number of non-zero spans.
 * ---------------------------------------------------------------------- */
int orc_bezier_extraction(int p, int m, const double *U, double *C)
{
  int ns = 0;
  for (int i = 1; i < m; i++)
    if (fabs(U[i] - U[i - 1]) > 1e-6) ns++;
  int q = p + 1;
  memset(C, 0, sizeof(double) * (size_t)ns * q * q);
#define CE(e, i, j) C[((size_t)(e) * q + (i)) * q + (j)]
  for (int i = 0; i < q; i++) CE(0, i, i) = 1.0;
  int a = p + 1, b = a + 1, nb = 1;
  double alphas[64];
  while (b < m) {
    if (nb < ns)
      for (int i = 0; i < q; i++) CE(nb, i, i) = 1.0;
    int i = b;
    while (b < m && U[b] == U[b - 1]) b++;
    int mult = b - i + 1;
    if (mult < p) {
      double numer = U[b - 1] - U[a - 1];
      for (int j = p; j >= mult + 1; j--)
        alphas[j - mult - 1] = numer / (U[a + j - 1] - U[a - 1]);
      int r = p - mult;
      for (int j = 1; j <= r; j++) {
        int save = r - j + 1;
        int s = mult + j;
        for (int k = p + 1; k >= s + 1; k--) {
          double alpha = alphas[k - s - 1];
          for (int w = 0; w < q; w++)
            CE(nb - 1, w, k - 1) =
                alpha * CE(nb - 1, w, k - 1) + (1.0 - alpha) * CE(nb - 1, w, k - 2);
        }
        if (b < m && nb < ns)
          for (int w = save, v = p - j + 1; w <= j + save; w++, v++)
            CE(nb, w - 1, save - 1) = CE(nb - 1, v - 1, p);
      }
    }
    nb++;
    if (b < m) a = b++;
  }
#undef CE
  return ns;
}

/* ------------------------------------------------------------------------
 * Natural derivatives dN/d(r,s,t) for one element.
 * Hex8: shp/shpsolid.cpp:211-238.
 * B-spline solid (rational branch; the non-rational branch of the reference
 * never advances its basis counter, shp/shpbsp.cpp:1015-1031, so decks always
 * declare rational patches and unit weights reduce to plain B-splines):
 * shp/shpbsp.cpp:254-300 (EvalBspBas/Derv: t' = (t+1)/2, derivative / 2,
 * N_i += C[i][j] B_j) and 985-1067 (tensor product k(r) fastest, quotient).
 * ---------------------------------------------------------------------- */
static void bsp_univariate(int p, const double *C, double t, double *N,
                           double *dN)
{
  double B[16], dB[16];
  double u = (t + 1.0) / 2.0;
  orc_bernstein(p, u, B);
  orc_bernstein_deriv(p, u, dB);
  for (int i = 0; i <= p; i++) dB[i] /= 2.0;
  for (int i = 0; i <= p; i++) {
    N[i] = 0.0;
    dN[i] = 0.0;
  }
  for (int i = 0; i <= p; i++)
    for (int j = 0; j <= p; j++) N[i] += C[i * (p + 1) + j] * B[j];
  for (int i = 0; i <= p; i++)
    for (int j = 0; j <= p; j++) dN[i] += C[i * (p + 1) + j] * dB[j];
}

/* dN: natural derivatives [3 nn]; N (optional): shape values [nn] */
static void natural_derivs(const orc_etype *et, const double *ext,
                           const double *w, const double rst[3], double *dN,
                           double *Nval)
{
  if (et->shape == ORC_HEX8) {
    orc_hex8_shape(rst[0], rst[1], rst[2], Nval, dN);
    return;
  }
  if (et->shape == ORC_HEX20) {
    orc_hex20_shape(rst[0], rst[1], rst[2], Nval, dN);
    return;
  }
  const int *p = et->deg;
  double Nr[16], Ns[16], Nt[16], dNr[16], dNs[16], dNt[16];
  size_t o0 = 0, o1 = (size_t)(p[0] + 1) * (p[0] + 1);
  size_t o2 = o1 + (size_t)(p[1] + 1) * (p[1] + 1);
  bsp_univariate(p[0], ext + o0, rst[0], Nr, dNr);
  bsp_univariate(p[1], ext + o1, rst[1], Ns, dNs);
  bsp_univariate(p[2], ext + o2, rst[2], Nt, dNt);
  double N[MAXNN];
  int a = 0;
  if (!et->rational) {
    /* plain tensor product (what the reference intends for this branch) */
    for (int i = 0; i <= p[2]; i++)
      for (int j = 0; j <= p[1]; j++)
        for (int k = 0; k <= p[0]; k++, a++) {
          dN[3 * a + 0] = dNr[k] * Ns[j] * Nt[i];
          dN[3 * a + 1] = Nr[k] * dNs[j] * Nt[i];
          dN[3 * a + 2] = Nr[k] * Ns[j] * dNt[i];
          if (Nval) Nval[a] = Nr[k] * Ns[j] * Nt[i];
        }
    return;
  }
  double W = 0, dwr = 0, dws = 0, dwt = 0;
  for (int i = 0; i <= p[2]; i++)
    for (int j = 0; j <= p[1]; j++)
      for (int k = 0; k <= p[0]; k++) {
        double wa = w ? w[a] : 1.0;
        N[a] = Nr[k] * Ns[j] * Nt[i] * wa;
        W += N[a];
        dN[3 * a + 0] = dNr[k] * Ns[j] * Nt[i] * wa;
        dN[3 * a + 1] = Nr[k] * dNs[j] * Nt[i] * wa;
        dN[3 * a + 2] = Nr[k] * Ns[j] * dNt[i] * wa;
        dwr += dN[3 * a + 0];
        dws += dN[3 * a + 1];
        dwt += dN[3 * a + 2];
        a++;
      }
  for (int i = 0; i < a; i++) {
    N[i] /= W;
    dN[3 * i + 0] = (dN[3 * i + 0] - N[i] * dwr) / W;
    dN[3 * i + 1] = (dN[3 * i + 1] - N[i] * dws) / W;
    dN[3 * i + 2] = (dN[3 * i + 2] - N[i] * dwt) / W;
  }
  if (Nval)
    for (int i = 0; i < a; i++) Nval[i] = N[i];
}

/* Reference: shp/shpsolid.cpp:58-126 (J = dN^T X with rows r,s,t) and
 * shp/shape.cpp:360-391 (cofactor inverse; fails on det == 0). */
int orc_shape_dxyz(const orc_etype *et, const double *X, const double *ext,
                   const double *w, const double rst[3], double *detJ,
                   double *dNxyz)
{
  double dN[3 * MAXNN];
  natural_derivs(et, ext, w, rst, dN, NULL);
  double A[3][3] = {{0}};
  int nn = et->nn;
  for (int i = 0; i < nn; i++) {
    A[0][0] += dN[3 * i + 0] * X[3 * i + 0];
    A[0][1] += dN[3 * i + 0] * X[3 * i + 1];
    A[0][2] += dN[3 * i + 0] * X[3 * i + 2];
    A[1][0] += dN[3 * i + 1] * X[3 * i + 0];
    A[1][1] += dN[3 * i + 1] * X[3 * i + 1];
    A[1][2] += dN[3 * i + 1] * X[3 * i + 2];
    A[2][0] += dN[3 * i + 2] * X[3 * i + 0];
    A[2][1] += dN[3 * i + 2] * X[3 * i + 1];
    A[2][2] += dN[3 * i + 2] * X[3 * i + 2];
  }
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  *detJ = det;
  if (fabs(det) == 0.0) return 0;
  double B[3][3];
  B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  B[0][1] = -(A[0][1] * A[2][2] - A[0][2] * A[2][1]) / det;
  B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  B[1][0] = -(A[1][0] * A[2][2] - A[1][2] * A[2][0]) / det;
  B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  B[1][2] = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]) / det;
  B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  B[2][1] = -(A[0][0] * A[2][1] - A[0][1] * A[2][0]) / det;
  B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  for (int i = 0; i < nn; i++) {
    double r = dN[3 * i + 0], s = dN[3 * i + 1], t = dN[3 * i + 2];
    dNxyz[3 * i + 0] = B[0][0] * r + B[0][1] * s + B[0][2] * t;
    dNxyz[3 * i + 1] = B[1][0] * r + B[1][1] * s + B[1][2] * t;
    dNxyz[3 * i + 2] = B[2][0] * r + B[2][1] * s + B[2][2] * t;
  }
  return 1;
}

/* ------------------------------------------------------------------------
 * Kinematic matrices of the 3D solid model.  Reference: elc/anmodel.cpp
 * 1418-1447 (B: Voigt xx,yy,zz,xy,xz,yz with engineering shear),
 * 1451-1510 (Bl), 1514-1532 (Bnl, 9 rows), 1536-1568 (S block diagonal),
 * 1286-1298 (isotropic Q).
 * ---------------------------------------------------------------------- */
static void bmatrix(int nn, const double *d, double *B /* [6][3nn] */)
{
  int nd = 3 * nn;
  memset(B, 0, sizeof(double) * 6 * nd);
  for (int i = 0; i < nn; i++) {
    double x = d[3 * i], y = d[3 * i + 1], z = d[3 * i + 2];
    B[0 * nd + 3 * i] = x;
    B[1 * nd + 3 * i + 1] = y;
    B[2 * nd + 3 * i + 2] = z;
    B[3 * nd + 3 * i] = y;
    B[3 * nd + 3 * i + 1] = x;
    B[4 * nd + 3 * i] = z;
    B[4 * nd + 3 * i + 2] = x;
    B[5 * nd + 3 * i + 1] = z;
    B[5 * nd + 3 * i + 2] = y;
  }
}

static void blmatrix(int nn, const double *d, const double *u, double *Bl)
{
  int nd = 3 * nn;
  double l11 = 0, l12 = 0, l13 = 0, l21 = 0, l22 = 0, l23 = 0, l31 = 0,
         l32 = 0, l33 = 0;
  for (int i = 0; i < nn; i++) {
    double x = d[3 * i], y = d[3 * i + 1], z = d[3 * i + 2];
    l11 += u[3 * i] * x;     l12 += u[3 * i] * y;     l13 += u[3 * i] * z;
    l21 += u[3 * i + 1] * x; l22 += u[3 * i + 1] * y; l23 += u[3 * i + 1] * z;
    l31 += u[3 * i + 2] * x; l32 += u[3 * i + 2] * y; l33 += u[3 * i + 2] * z;
  }
  for (int i = 0; i < nn; i++) {
    double x = d[3 * i], y = d[3 * i + 1], z = d[3 * i + 2];
    double *r0 = Bl, *r1 = Bl + nd, *r2 = Bl + 2 * nd, *r3 = Bl + 3 * nd,
           *r4 = Bl + 4 * nd, *r5 = Bl + 5 * nd;
    r0[3 * i] = l11 * x; r0[3 * i + 1] = l21 * x; r0[3 * i + 2] = l31 * x;
    r1[3 * i] = l12 * y; r1[3 * i + 1] = l22 * y; r1[3 * i + 2] = l32 * y;
    r2[3 * i] = l13 * z; r2[3 * i + 1] = l23 * z; r2[3 * i + 2] = l33 * z;
    r3[3 * i] = l11 * y + l12 * x;
    r3[3 * i + 1] = l21 * y + l22 * x;
    r3[3 * i + 2] = l31 * y + l32 * x;
    r4[3 * i] = l11 * z + l13 * x;
    r4[3 * i + 1] = l21 * z + l23 * x;
    r4[3 * i + 2] = l31 * z + l33 * x;
    r5[3 * i] = l12 * z + l13 * y;
    r5[3 * i + 1] = l22 * z + l23 * y;
    r5[3 * i + 2] = l32 * z + l33 * y;
  }
}

static void bnlmatrix(int nn, const double *d, double *Bnl /* [9][3nn] */)
{
  int nd = 3 * nn;
  memset(Bnl, 0, sizeof(double) * 9 * nd);
  for (int i = 0; i < nn; i++)
    for (int c = 0; c < 3; c++)
      for (int k = 0; k < 3; k++) Bnl[(3 * c + k) * nd + 3 * i + c] = d[3 * i + k];
}

static void smatrix(const double *s, double *S /* [9][9] */)
{
  memset(S, 0, sizeof(double) * 81);
  for (int b = 0; b < 3; b++) {
    double *M = S + (3 * b) * 9 + 3 * b;
    M[0 * 9 + 0] = s[0]; M[1 * 9 + 0] = s[3]; M[2 * 9 + 0] = s[4];
    M[0 * 9 + 1] = s[3]; M[1 * 9 + 1] = s[1]; M[2 * 9 + 1] = s[5];
    M[0 * 9 + 2] = s[4]; M[1 * 9 + 2] = s[5]; M[2 * 9 + 2] = s[2];
  }
}

static void qmatrix(double E, double nu, double *Q /* [6][6] */)
{
  double d = (1.0 + nu) * (1.0 - 2.0 * nu);
  memset(Q, 0, sizeof(double) * 36);
  Q[0] = Q[7] = Q[14] = E * (1.0 - nu) / d;
  Q[1] = Q[6] = Q[2] = Q[12] = Q[8] = Q[13] = E * nu / d;
  Q[21] = Q[28] = Q[35] = 0.5 * E / (1.0 + nu);
}

/* Compressible neo-Hookean (this project's model; absent from the reference):
 * W = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2.  S = mu (I - C^-1) +
 * lambda ln J C^-1; D_IJKL = lambda c_IJ c_KL + (mu - lambda ln J)
 * (c_IK c_JL + c_IL c_JK), c = C^-1, in Voigt form for engineering shear.
 * Returns 0 when det C <= 0. */
static const int VI[6] = {0, 1, 2, 0, 0, 1}, VJ[6] = {0, 1, 2, 1, 2, 2};

static int neohooke(double mu, double lam, const double *e, double *s,
                    double *D)
{
  double C[3][3];
  C[0][0] = 1.0 + 2.0 * e[0];
  C[1][1] = 1.0 + 2.0 * e[1];
  C[2][2] = 1.0 + 2.0 * e[2];
  C[0][1] = C[1][0] = e[3];
  C[0][2] = C[2][0] = e[4];
  C[1][2] = C[2][1] = e[5];
  double det = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) -
               C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
               C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
  if (!(det > 0.0)) return 0;
  double c[3][3];
  c[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) / det;
  c[0][1] = -(C[0][1] * C[2][2] - C[0][2] * C[2][1]) / det;
  c[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) / det;
  c[1][0] = c[0][1];
  c[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) / det;
  c[1][2] = -(C[0][0] * C[1][2] - C[0][2] * C[1][0]) / det;
  c[2][0] = c[0][2];
  c[2][1] = c[1][2];
  c[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) / det;
  double lnJ = 0.5 * log(det);
  for (int a = 0; a < 6; a++) {
    int I = VI[a], J = VJ[a];
    s[a] = mu * ((I == J ? 1.0 : 0.0) - c[I][J]) + lam * lnJ * c[I][J];
  }
  if (D) {
    double m2 = mu - lam * lnJ;
    for (int a = 0; a < 6; a++)
      for (int b = 0; b < 6; b++) {
        int I = VI[a], J = VJ[a], K = VI[b], L = VJ[b];
        D[a * 6 + b] = lam * c[I][J] * c[K][L] +
                       m2 * (c[I][K] * c[J][L] + c[I][L] * c[J][K]);
      }
  }
  return 1;
}

/* Reference: mlib/matvec.cpp:372-390 (K += w B^T C B, skipping zero B and C
 * entries; operation order aux1 = w*B, aux2 = aux1*C, K += aux2*B). */
static void trip_btcb(int m, int n, const double *B, const double *C,
                      double w, double *K)
{
  for (int i = 0; i < n; i++)
    for (int k = 0; k < m; k++) {
      double bki = B[k * n + i];
      if (bki != 0.0) {
        double aux1 = w * bki;
        for (int l = 0; l < m; l++) {
          double ckl = C[k * m + l];
          if (ckl != 0.0) {
            double aux2 = aux1 * ckl;
            const double *Bl = B + (size_t)l * n;
            double *Ki = K + (size_t)i * n;
            for (int j = 0; j < n; j++) Ki[j] += aux2 * Bl[j];
          }
        }
      }
    }
}

/* Reference: mlib/matvec.cpp:324-335 (row dot products from 0.0). */
static void matvec(int m, int n, const double *A, const double *x, double *y)
{
  for (int i = 0; i < m; i++) {
    double d = 0.0;
    for (int j = 0; j < n; j++) d += A[(size_t)i * n + j] * x[j];
    y[i] = d;
  }
}

/* Reference: mlib/matvec.cpp:359-368 (g += a B^T s, column by column). */
static void multtacc(double a, int m, int n, const double *B, const double *s,
                     double *g)
{
  for (int j = 0; j < m; j++) {
    double auj = a * s[j];
    const double *Bj = B + (size_t)j * n;
    for (int i = 0; i < n; i++) g[i] += Bj[i] * auj;
  }
}

/* Material response at one point.  Linear / SVK: mat/cmodel.cpp:203-216 and
 * 273-288 (sigma = Q eps, Q rebuilt per call). */
static int material(int form, const double *mat, const double *e, double *s,
                    double *D)
{
  if (form == ORC_TL_NEOHOOKE) return neohooke(mat[0], mat[1], e, s, D);
  double Q[36];
  qmatrix(mat[0], mat[1], Q);
  if (s) matvec(6, 6, Q, e, s);
  if (D) memcpy(D, Q, sizeof Q);
  return 1;
}

/* Green-Lagrange strain: elc/elmparam.cpp:1899-1919 (B u + 0.5 (Bl u)). */
static void green_strain(int nd, const double *B, const double *Bl,
                         const double *u, double *e)
{
  double s1[6], s2[6];
  matvec(6, nd, B, u, s1);
  matvec(6, nd, Bl, u, s2);
  for (int i = 0; i < 6; i++) s2[i] *= 0.5;
  for (int i = 0; i < 6; i++) e[i] = s1[i] + s2[i];
}

/* Linear: elc/elmparam.cpp:539-597 (StiffMat).  TL: elc/elmparam.cpp
 * 1711-1807 (K_t = (B+Bl)^T C (B+Bl) + Bnl^T S Bnl per point; a failed
 * stress update ends the point loop, as in the reference). */
int orc_elem_stiff(const orc_etype *et, const double *X, const double *ext,
                   const double *w, const double *mat, const double *u,
                   double *K)
{
  int nn = et->nn, nd = 3 * nn;
  double rst[3 * 512], wq[512];
  int nq = orc_brick_points(et->nip, rst, wq);
  double *B = malloc(sizeof(double) * 6 * nd);
  double *Bl = malloc(sizeof(double) * 6 * nd);
  double *Bnl = malloc(sizeof(double) * 9 * nd);
  double *d = malloc(sizeof(double) * 3 * nn);
  double *uz = calloc(nd, sizeof(double));
  const double *ue = u ? u : uz;
  memset(K, 0, sizeof(double) * nd * nd);
  int ok = 1;
  for (int q = 0; q < nq; q++) {
    double detJ;
    if (!orc_shape_dxyz(et, X, ext, w, rst + 3 * q, &detJ, d)) { ok = 0; }
    bmatrix(nn, d, B);
    double coeff = wq[q] * detJ * 1.0;
    if (et->form == ORC_LINEAR) {
      double C[36];
      material(ORC_LINEAR, mat, NULL, NULL, C);
      trip_btcb(6, nd, B, C, coeff, K);
    } else {
      double C[36], e[6], s[6], S[81];
      blmatrix(nn, d, ue, Bl);
      bnlmatrix(nn, d, Bnl);
      green_strain(nd, B, Bl, ue, e);
      if (et->form == ORC_TL_SVK) {
        material(ORC_LINEAR, mat, NULL, NULL, C);
        material(ORC_LINEAR, mat, e, s, NULL);
      } else if (!material(et->form, mat, e, s, C)) {
        ok = 0;
        break;
      }
      smatrix(s, S);
      for (int i = 0; i < 6 * nd; i++) B[i] += Bl[i];
      trip_btcb(6, nd, B, C, coeff, K);
      trip_btcb(9, nd, Bnl, S, coeff, K);
    }
  }
  free(B); free(Bl); free(Bnl); free(d); free(uz);
  return ok;
}

/* Linear: elc/elmparam.cpp:449-535.  TL: elc/elmparam.cpp:1615-1707. */
int orc_elem_intforce(const orc_etype *et, const double *X,
                      const double *ext, const double *w, const double *mat,
                      const double *u, double *g)
{
  int nn = et->nn, nd = 3 * nn;
  double rst[3 * 512], wq[512];
  int nq = orc_brick_points(et->nip, rst, wq);
  double *B = malloc(sizeof(double) * 6 * nd);
  double *Bl = malloc(sizeof(double) * 6 * nd);
  double *d = malloc(sizeof(double) * 3 * nn);
  memset(g, 0, sizeof(double) * nd);
  int ok = 1;
  for (int q = 0; q < nq; q++) {
    double detJ, e[6], s[6];
    if (!orc_shape_dxyz(et, X, ext, w, rst + 3 * q, &detJ, d)) ok = 0;
    bmatrix(nn, d, B);
    if (et->form == ORC_LINEAR) {
      matvec(6, nd, B, u, e);
    } else {
      blmatrix(nn, d, u, Bl);
      green_strain(nd, B, Bl, u, e);
    }
    if (!material(et->form == ORC_TL_SVK ? ORC_LINEAR : et->form, mat, e, s,
                  NULL)) {
      ok = 0;
      break;
    }
    if (et->form != ORC_LINEAR)
      for (int i = 0; i < 6 * nd; i++) B[i] += Bl[i];
    double coeff = wq[q] * detJ * 1.0;
    multtacc(coeff, 6, nd, B, s, g);
  }
  free(B); free(Bl); free(d);
  return ok;
}

/* Element body-force vector: cElemLoad::ExtForce (load/load.cpp:618-691)
 * with cBodyLoad::GetVal (load.cpp:969-974): per point, coeff = w detJ (the
 * element's Gauss rule), f += coeff N^T b. */
int orc_elem_bodyforce(const orc_etype *et, const double *X, const double *ext,
                       const double *w, const double *b, double *f)
{
  int nn = et->nn;
  double rst[3 * 512], wq[512];
  int nq = orc_brick_points(et->nip, rst, wq);
  double Nv[MAXNN], dN[3 * MAXNN], dxyz[3 * MAXNN];
  memset(f, 0, sizeof(double) * 3 * nn);
  int ok = 1;
  for (int q = 0; q < nq; q++) {
    double detJ;
    natural_derivs(et, ext, w, rst + 3 * q, dN, Nv);
    if (!orc_shape_dxyz(et, X, ext, w, rst + 3 * q, &detJ, dxyz)) ok = 0;
    double coeff = wq[q] * detJ * 1.0;
    for (int d = 0; d < 3; d++) {
      double ab = coeff * b[d];
      for (int a = 0; a < nn; a++) f[3 * a + d] += Nv[a] * ab;
    }
  }
  return ok;
}

/* Integration-point stresses.  Linear: cElmParam::IntPntStress
 * (elc/elmparam.cpp:859-937): e = B u, s = C e per point.  Total Lagrangian:
 * cElmParamTL::IntPntStress (elmparam.cpp:1811-1885): Green strain
 * e = B u + 0.5 Bl u (GreenStrain, 1899-1919), s = C e (second
 * Piola-Kirchhoff); neo-Hooke: the second Piola-Kirchhoff stress of the
 * project's model.  Voigt order xx yy zz xy xz yz; points in the brick
 * rule's order (r fastest).  sig: [nq*6]. */
int orc_elem_stress(const orc_etype *et, const double *X, const double *ext,
                    const double *w, const double *mat, const double *u,
                    double *sig)
{
  int nn = et->nn, nd = 3 * nn;
  double rst[3 * 512], wq[512];
  int nq = orc_brick_points(et->nip, rst, wq);
  double *B = malloc(sizeof(double) * 6 * nd);
  double *Bl = malloc(sizeof(double) * 6 * nd);
  double *d = malloc(sizeof(double) * 3 * nn);
  int ok = 1;
  for (int q = 0; q < nq; q++) {
    double detJ, e[6];
    if (!orc_shape_dxyz(et, X, ext, w, rst + 3 * q, &detJ, d)) ok = 0;
    bmatrix(nn, d, B);
    if (et->form == ORC_LINEAR) {
      matvec(6, nd, B, u, e);
    } else {
      blmatrix(nn, d, u, Bl);
      green_strain(nd, B, Bl, u, e);
    }
    if (!material(et->form == ORC_TL_SVK ? ORC_LINEAR : et->form, mat, e,
                  sig + 6 * q, NULL))
      ok = 0;
  }
  free(B); free(Bl); free(d);
  return ok;
}

/* Natural coordinates of the element's nodes: cShapeBrick8::NodeNatCoord
 * (shp/shpsolid.cpp:173-185), cShapeBrick20::NodeNatCoord (301-325) and cShapeBspSolid::NodeNatCoord
 * (shp/shpbsp.cpp:906-923: a uniform (p+1)^3 grid on [-1, 1]^3, r fastest). */
static void node_nat_coords(const orc_etype *et, double *c)
{
  if (et->shape == ORC_HEX8) {
    static const double h[8][3] = {{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
                                   {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}};
    memcpy(c, h, sizeof h);
    return;
  }
  if (et->shape == ORC_HEX20) {
    memcpy(c, kHex20Nat, sizeof kHex20Nat);
    return;
  }
  int nb[3] = {et->deg[0] + 1, et->deg[1] + 1, et->deg[2] + 1}, n = 0;
  for (int i = 0; i < nb[2]; i++)
    for (int j = 0; j < nb[1]; j++)
      for (int k = 0; k < nb[0]; k++, n++) {
        c[3 * n] = -1.0 + (2.0 / (nb[0] - 1)) * k;
        c[3 * n + 1] = -1.0 + (2.0 / (nb[1] - 1)) * j;
        c[3 * n + 2] = -1.0 + (2.0 / (nb[2] - 1)) * i;
      }
}

/* Extrapolation of integration-point values to the nodes:
 * cElmParam::CompTRMatrix (elc/elmparam.cpp:1444-1535).  Least-squares fit
 * of a + b r + c s + d t over the points (normal matrix A = sum u u^T,
 * u = (1, r, s, t)), evaluated at the nodes: TR = CN A^-1 [u_1 .. u_nq].
 * One point along r: every node takes the mean.  TR: [nn*nq] row-major. */
int orc_stress_extrap(const orc_etype *et, double *TR)
{
  int nn = et->nn;
  double rst[3 * 512], wq[512];
  int nq = orc_brick_points(et->nip, rst, wq);
  if (et->nip[0] == 1) {
    for (int i = 0; i < nn * nq; i++) TR[i] = 1.0 / (double)nq;
    return nq;
  }
  double A[4][4] = {{0}};
  for (int q = 0; q < nq; q++) {
    double v[4] = {1.0, rst[3 * q], rst[3 * q + 1], rst[3 * q + 2]};
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++) A[k][l] += v[k] * v[l];
  }
  /* A^-1 by Gauss-Jordan elimination (A is symmetric positive definite) */
  double I[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  for (int k = 0; k < 4; k++) {
    double p = A[k][k];
    if (!(p > 0.0)) return 0;
    for (int l = 0; l < 4; l++) { A[k][l] /= p; I[k][l] /= p; }
    for (int r = 0; r < 4; r++) {
      if (r == k) continue;
      double f = A[r][k];
      for (int l = 0; l < 4; l++) { A[r][l] -= f * A[k][l]; I[r][l] -= f * I[k][l]; }
    }
  }
  double cn[3 * MAXNN];
  node_nat_coords(et, cn);
  for (int a = 0; a < nn; a++) {
    double c[4] = {1.0, cn[3 * a], cn[3 * a + 1], cn[3 * a + 2]}, ci[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++) ci[l] += c[k] * I[k][l];
    for (int q = 0; q < nq; q++) {
      double v[4] = {1.0, rst[3 * q], rst[3 * q + 1], rst[3 * q + 2]};
      TR[(size_t)a * nq + q] = ci[0] * v[0] + ci[1] * v[1] + ci[2] * v[2] + ci[3] * v[3];
    }
  }
  return nq;
}

/* ------------------------------------------------------------------------
 * Mesh-level loops.
 * ---------------------------------------------------------------------- */
static void gather_elem(const orc_mesh *m, int e, const double *u_nodal,
                        double *X, double *ue, double *we, int *edof)
{
  int nn = m->et.nn;
  const int *cn = m->conn + (size_t)e * nn;
  for (int a = 0; a < nn; a++) {
    int n = cn[a];
    for (int c = 0; c < 3; c++) {
      X[3 * a + c] = m->coords[3 * (size_t)n + c];
      if (ue) ue[3 * a + c] = u_nodal ? u_nodal[3 * (size_t)n + c] : 0.0;
      if (edof) edof[3 * a + c] = m->dofmap[3 * (size_t)n + c];
    }
    if (we) we[a] = m->cpw ? m->cpw[n] : 1.0;
  }
}

static size_t ext_stride(const orc_etype *et)
{
  if (et->shape != ORC_BSP) return 0;
  size_t s = 0;
  for (int d = 0; d < 3; d++) s += (size_t)(et->deg[d] + 1) * (et->deg[d] + 1);
  return s;
}

static long long csr_find(const long long *rowptr, const int *col, int r,
                          int c)
{
  long long lo = rowptr[r], hi = rowptr[r + 1] - 1;
  while (lo <= hi) {
    long long mid = (lo + hi) >> 1;
    if (col[mid] == c) return mid;
    if (col[mid] < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* Reference: ctrl/ctrl.cpp:886-975 (element loop, element order) and
 * elc/element.cpp:2975-3002 (AddGlobMat: only dof[i] >= dof[j] > 0 for the
 * symmetric matrix).  The lower entry is mirrored to the upper position, so
 * the full CSR's lower triangle equals the reference's stored values.
 * With nthreads > 1 the element loop is split into element ranges whose
 * element matrices are computed in parallel and added in element order. */
int orc_assemble_K(const orc_mesh *m, const double *u_nodal, int neq,
                   const long long *rowptr, const int *col, double *val,
                   int nthreads, int eb, int ee)
{
  (void)neq;
  int nn = m->et.nn, nd = 3 * nn;
  size_t es = ext_stride(&m->et);
  if (ee < 0 || ee > m->nelem) ee = m->nelem;
  int ok = 1;
  const int CH = 256;
  double *Kbuf = malloc(sizeof(double) * (size_t)CH * nd * nd);
  int *dbuf = malloc(sizeof(int) * (size_t)CH * nd);
  int *okb = malloc(sizeof(int) * CH);
  if (nthreads < 1) nthreads = 1;
  for (int c0 = eb; c0 < ee; c0 += CH) {
    int c1 = c0 + CH < ee ? c0 + CH : ee;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int e = c0; e < c1; e++) {
      double X[3 * MAXNN], ue[3 * MAXNN], we[MAXNN];
      int *edof = dbuf + (size_t)(e - c0) * nd;
      gather_elem(m, e, u_nodal, X, ue, we, edof);
      okb[e - c0] = orc_elem_stiff(&m->et, X, m->ext ? m->ext + es * e : NULL,
                                   we, m->mat + 2 * (size_t)e,
                                   u_nodal ? ue : NULL,
                                   Kbuf + (size_t)(e - c0) * nd * nd);
    }
    for (int e = c0; e < c1; e++) {
      const double *Ke = Kbuf + (size_t)(e - c0) * nd * nd;
      const int *edof = dbuf + (size_t)(e - c0) * nd;
      if (!okb[e - c0]) ok = 0;
      for (int i = 0; i < nd; i++) {
        int di = edof[i];
        if (!di) continue;
        for (int j = 0; j < nd; j++) {
          int dj = edof[j];
          if (!dj || di < dj) continue;
          long long p = csr_find(rowptr, col, di - 1, dj - 1);
          if (p < 0) { ok = 0; continue; }
          val[p] += Ke[(size_t)i * nd + j];
          if (di != dj) {
            long long p2 = csr_find(rowptr, col, dj - 1, di - 1);
            if (p2 < 0) { ok = 0; continue; }
            val[p2] += Ke[(size_t)i * nd + j];
          }
        }
      }
    }
  }
  free(Kbuf); free(dbuf); free(okb);
  return ok;
}

/* Reference: ctrl/ctrl.cpp:719-813 and elc/element.cpp:2953-2971. */
/* Body-force load vector: element loop in element order, AddGlobVec
 * (load.cpp:695-705). b: [nelem * 3]. */
int orc_assemble_body(const orc_mesh *m, const double *b, int neq, double *f)
{
  int nn = m->et.nn;
  size_t es = ext_stride(&m->et);
  double X[3 * MAXNN], we[MAXNN], fe[3 * MAXNN];
  int edof[3 * MAXNN];
  memset(f, 0, sizeof(double) * neq);
  int ok = 1;
  for (int e = 0; e < m->nelem; e++) {
    gather_elem(m, e, NULL, X, NULL, we, edof);
    const double *ext = m->ext ? m->ext + (size_t)e * es : NULL;
    if (!orc_elem_bodyforce(&m->et, X, ext, we, b + 3 * (size_t)e, fe)) ok = 0;
    for (int i = 0; i < 3 * nn; i++)
      if (edof[i] > 0) f[edof[i] - 1] += fe[i];
  }
  return ok;
}

/* Uniform surface loads, one group, added into f in entry order.
 * cElemLoad::ExtForce (load/load.cpp:618-691) over the face shape of
 * cFaceUnifLoad (load.cpp:1440-1500: cShapeQ4Surf over the brick face's
 * nodes, shp/shpsolid.cpp:264-276) or of cFaceUnifLoadIGA (loadiga.cpp
 * 305-354: the patch's boundary surface over the element's spans):
 *   points of cIntPoint::CreateQuadPoints (elc/intpoint.cpp:266-300: r
 *   fastest, weight w_r w_s); coeff = weight * detJ * 1 with detJ =
 *   |x_,r x x_,s| (cShapeSurf::DrvShpXYZ, shp/shpsurf.cpp:56-116);
 *   load = q, or R q with R = [u v w] of cShapeSurf::LocalSys
 *   (shpsurf.cpp:121-164: u, v the unit tangents, w = u x v, v = w x u)
 *   when the entry is local (cFaceUnifLoad::GetVal, load.cpp:1487-1500);
 *   fe += N^T (coeff load) (MultTAcc), then f[eq] += fe (AddGlobVec).
 * A group without extraction rows and of degree (1, 1) is a brick face:
 * its nodes are given first parameter fastest (n0 n1 n3 n2) and the face
 * is integrated with the bilinear functions of cShapeQ4 (shp/shpplane.cpp
 * 363-390) over n0 n1 n2 n3.  Otherwise the basis is the rational
 * Bezier-extracted tensor product of the element's two parameters (the
 * first fastest), as for the solid (shp/shpbsp.cpp:254-300, 985-1067). */
int orc_surface_load(const orc_mesh *m, const orc_surface *s, double *f)
{
  const int q8 = s->shape == 1;
  if (q8 && (s->ext || s->deg[0] != 2 || s->deg[1] != 2)) return 0;
  const int p1 = s->deg[0], p2 = s->deg[1], n1 = p1 + 1, n2 = p2 + 1, nb = q8 ? 8 : n1 * n2;
  const int q4 = !s->ext && p1 == 1 && p2 == 1;
  double r[16], wr[16];
  if (nb > 64 || s->nip < 1 || s->nip > 16 || !orc_gauss(s->nip, r, wr)) return 0;
  for (int e = 0; e < s->nface; e++) {
    const int *nd = s->nodes + (size_t)e * nb;
    const double *C1 = s->ext ? s->ext + (size_t)e * (n1 * n1 + n2 * n2) : NULL;
    const double *C2 = C1 ? C1 + n1 * n1 : NULL;
    const double *qv = s->q + 3 * (size_t)e;
    const int loc = s->local && s->local[e];
    double fe[3 * 64], X[3 * 64], W[64];
    int perm[64];  /* shape-function index -> entry node index */
    for (int a = 0; a < nb; a++) perm[a] = a;
    if (q4) { perm[2] = 3; perm[3] = 2; }
    for (int a = 0; a < nb; a++) {
      const int n = nd[perm[a]];
      for (int k = 0; k < 3; k++) X[3 * a + k] = m->coords[3 * (size_t)n + k];
      W[a] = m->cpw ? m->cpw[n] : 1.0;
    }
    memset(fe, 0, sizeof(double) * 3 * nb);
    for (int i = 0; i < s->nip; i++)
      for (int j = 0; j < s->nip; j++) {
        const double pr = r[j], ps = r[i], wgt = wr[j] * wr[i];
        double N[64], dr[64], ds[64];
        if (q8) {
          double d8[16];
          orc_q8_shape(pr, ps, N, d8);
          for (int a = 0; a < 8; a++) { dr[a] = d8[2 * a]; ds[a] = d8[2 * a + 1]; }
        } else if (q4) {
          N[0] = 0.25 * (1.0 - pr) * (1.0 - ps);
          N[1] = 0.25 * (1.0 + pr) * (1.0 - ps);
          N[2] = 0.25 * (1.0 + pr) * (1.0 + ps);
          N[3] = 0.25 * (1.0 - pr) * (1.0 + ps);
          dr[0] = -0.25 * (1.0 - ps); dr[1] = 0.25 * (1.0 - ps);
          dr[2] = 0.25 * (1.0 + ps);  dr[3] = -0.25 * (1.0 + ps);
          ds[0] = -0.25 * (1.0 - pr); ds[1] = -0.25 * (1.0 + pr);
          ds[2] = 0.25 * (1.0 + pr);  ds[3] = 0.25 * (1.0 - pr);
        } else {
          static const double I[64] = {0};
          double Nr[16], dNr[16], Ns[16], dNs[16], Id1[64], Id2[64];
          const double *c1 = C1, *c2 = C2;
          if (!c1) {  /* identity extraction: the Bernstein polynomials */
            memcpy(Id1, I, sizeof(Id1));
            memcpy(Id2, I, sizeof(Id2));
            for (int a = 0; a < n1; a++) Id1[a * n1 + a] = 1.0;
            for (int a = 0; a < n2; a++) Id2[a * n2 + a] = 1.0;
            c1 = Id1;
            c2 = Id2;
          }
          bsp_univariate(p1, c1, pr, Nr, dNr);
          bsp_univariate(p2, c2, ps, Ns, dNs);
          double sw = 0.0, swr = 0.0, sws = 0.0;
          for (int b = 0; b < n2; b++)
            for (int a = 0; a < n1; a++) {
              const int k = a + n1 * b;
              N[k] = Nr[a] * Ns[b] * W[k];
              dr[k] = dNr[a] * Ns[b] * W[k];
              ds[k] = Nr[a] * dNs[b] * W[k];
              sw += N[k];
              swr += dr[k];
              sws += ds[k];
            }
          for (int k = 0; k < nb; k++) {
            N[k] /= sw;
            dr[k] = (dr[k] - N[k] * swr) / sw;
            ds[k] = (ds[k] - N[k] * sws) / sw;
          }
        }
        double J0[3] = {0, 0, 0}, J1[3] = {0, 0, 0};
        for (int a = 0; a < nb; a++)
          for (int k = 0; k < 3; k++) {
            J0[k] += dr[a] * X[3 * a + k];
            J1[k] += ds[a] * X[3 * a + k];
          }
        const double nrm[3] = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2],
                               J0[0] * J1[1] - J0[1] * J1[0]};
        const double detJ = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        double load[3] = {qv[0], qv[1], qv[2]};
        if (loc) {
          double u[3], v[3], w[3];
          const double lu = sqrt(J0[0] * J0[0] + J0[1] * J0[1] + J0[2] * J0[2]);
          const double lv = sqrt(J1[0] * J1[0] + J1[1] * J1[1] + J1[2] * J1[2]);
          for (int k = 0; k < 3; k++) { u[k] = J0[k] / lu; v[k] = J1[k] / lv; }
          w[0] = u[1] * v[2] - u[2] * v[1];
          w[1] = u[2] * v[0] - u[0] * v[2];
          w[2] = u[0] * v[1] - u[1] * v[0];
          v[0] = w[1] * u[2] - w[2] * u[1];
          v[1] = w[2] * u[0] - w[0] * u[2];
          v[2] = w[0] * u[1] - w[1] * u[0];
          for (int k = 0; k < 3; k++) load[k] = u[k] * qv[0] + v[k] * qv[1] + w[k] * qv[2];
        }
        const double coeff = wgt * detJ * 1.0;
        for (int d = 0; d < 3; d++) {
          const double cl = coeff * load[d];
          for (int a = 0; a < nb; a++) fe[3 * a + d] += N[a] * cl;
        }
      }
    for (int a = 0; a < nb; a++) {
      const int n = nd[perm[a]];
      for (int d = 0; d < 3; d++) {
        const int eq = m->dofmap[3 * (size_t)n + d];
        if (eq > 0) f[eq - 1] += fe[3 * a + d];
      }
    }
  }
  return 1;
}

/* Nodal stresses without extrapolation (HEADER.OUTPUT.STRESS.EXTRAPOLATION
 * 0): cElmParam::NodalStress (elc/elmparam.cpp:1313-1333) through PntStress
 * (1206-1290) at the nodes' natural coordinates: e = B u (the linear strain,
 * also for total-Lagrangian elements, which do not override PntStress) and
 * s = C e of the linear elastic material.  sig_node [nelem*nn*6]. */
int orc_node_stresses(const orc_mesh *m, const double *u_nodal, double *sig_node)
{
  int nn = m->et.nn, nd = 3 * nn;
  size_t es = ext_stride(&m->et);
  double cn[3 * MAXNN];
  node_nat_coords(&m->et, cn);
  double X[3 * MAXNN], ue[3 * MAXNN], we[MAXNN];
  double *B = malloc(sizeof(double) * 6 * nd);
  double *d = malloc(sizeof(double) * 3 * nn);
  int ok = 1;
  for (int e = 0; e < m->nelem; e++) {
    gather_elem(m, e, u_nodal, X, ue, we, NULL);
    const double *ext = m->ext ? m->ext + (size_t)e * es : NULL;
    for (int a = 0; a < nn; a++) {
      double detJ, eps[6];
      orc_shape_dxyz(&m->et, X, ext, we, cn + 3 * a, &detJ, d);
      bmatrix(nn, d, B);
      matvec(6, nd, B, ue, eps);
      if (!material(ORC_LINEAR, m->mat + 2 * (size_t)e, eps,
                    sig_node + ((size_t)e * nn + a) * 6, NULL))
        ok = 0;
    }
  }
  free(B); free(d);
  return ok;
}

/* Stress output of the whole mesh: cControl::PrintIntPntStress
 * (ctrl/ctrl.cpp:2189-2259: IntPntStress per element), PrintNodalStress
 * (2264-2330: cElmParam::NodalStress, elmparam.cpp:1304-1333, with the
 * default extrapolation [S]node = TR [S]ipt) and PrintAvrNodalStress
 * (2397-2470: element-nodal values summed per node in element order, then
 * divided by their count).  sig_ip [nelem*nq*6], sig_node [nelem*nn*6],
 * sig_avg [nnode*6]; sig_node / sig_avg may be NULL. */
int orc_stresses(const orc_mesh *m, const double *u_nodal, double *sig_ip,
                 double *sig_node, double *sig_avg)
{
  int nn = m->et.nn;
  size_t es = ext_stride(&m->et);
  double rst[3 * 512], wq[512];
  int nq = orc_brick_points(m->et.nip, rst, wq);
  double *TR = malloc(sizeof(double) * (size_t)nn * nq);
  int ok = orc_stress_extrap(&m->et, TR) > 0;
  double X[3 * MAXNN], ue[3 * MAXNN], we[MAXNN];
  double *cnt = NULL;
  if (sig_avg) {
    memset(sig_avg, 0, sizeof(double) * 6 * (size_t)m->nnode);
    cnt = calloc((size_t)m->nnode, sizeof(double));
  }
  double *sn = malloc(sizeof(double) * 6 * (size_t)nn);
  for (int e = 0; e < m->nelem; e++) {
    gather_elem(m, e, u_nodal, X, ue, we, NULL);
    const double *ext = m->ext ? m->ext + (size_t)e * es : NULL;
    double *si = sig_ip + (size_t)e * nq * 6;
    if (!orc_elem_stress(&m->et, X, ext, we, m->mat + 2 * (size_t)e, ue, si)) ok = 0;
    for (int a = 0; a < nn; a++)
      for (int c = 0; c < 6; c++) {
        double v = 0.0;
        for (int q = 0; q < nq; q++) v += TR[(size_t)a * nq + q] * si[6 * q + c];
        sn[6 * a + c] = v;
      }
    if (sig_node) memcpy(sig_node + (size_t)e * nn * 6, sn, sizeof(double) * 6 * nn);
    if (sig_avg)
      for (int a = 0; a < nn; a++) {
        int n = m->conn[(size_t)e * nn + a];
        for (int c = 0; c < 6; c++) sig_avg[6 * (size_t)n + c] += sn[6 * a + c];
        cnt[n] += 1.0;
      }
  }
  if (sig_avg)
    for (int n = 0; n < m->nnode; n++)
      if (cnt[n] > 0.0)
        for (int c = 0; c < 6; c++) sig_avg[6 * (size_t)n + c] /= cnt[n];
  free(cnt); free(sn); free(TR);
  return ok;
}

int orc_assemble_fint(const orc_mesh *m, const double *u_nodal, int neq,
                      double *g, int nthreads, int eb, int ee)
{
  (void)neq;
  int nn = m->et.nn, nd = 3 * nn;
  size_t es = ext_stride(&m->et);
  if (ee < 0 || ee > m->nelem) ee = m->nelem;
  int ok = 1;
  const int CH = 1024;
  double *gbuf = malloc(sizeof(double) * (size_t)CH * nd);
  int *dbuf = malloc(sizeof(int) * (size_t)CH * nd);
  int *okb = malloc(sizeof(int) * CH);
  if (nthreads < 1) nthreads = 1;
  for (int c0 = eb; c0 < ee && ok; c0 += CH) {
    int c1 = c0 + CH < ee ? c0 + CH : ee;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int e = c0; e < c1; e++) {
      double X[3 * MAXNN], ue[3 * MAXNN], we[MAXNN];
      int *edof = dbuf + (size_t)(e - c0) * nd;
      gather_elem(m, e, u_nodal, X, ue, we, edof);
      okb[e - c0] = orc_elem_intforce(&m->et, X,
                                      m->ext ? m->ext + es * e : NULL, we,
                                      m->mat + 2 * (size_t)e, ue,
                                      gbuf + (size_t)(e - c0) * nd);
    }
    for (int e = c0; e < c1; e++) {
      if (!okb[e - c0]) { ok = 0; break; }
      const double *ge = gbuf + (size_t)(e - c0) * nd;
      const int *edof = dbuf + (size_t)(e - c0) * nd;
      for (int i = 0; i < nd; i++)
        if (edof[i]) g[edof[i] - 1] += ge[i];
    }
  }
  free(gbuf); free(dbuf); free(okb);
  return ok;
}

/* Brute-force pattern: per row, the set of equations of all elements that
 * touch the row's equation (sorted, unique). */
static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static void pattern_rows(const orc_mesh *m, int neq, long long *rowptr,
                         int *col)
{
  int nn = m->et.nn;
  /* node -> element adjacency */
  int *cnt = calloc((size_t)m->nnode + 1, sizeof(int));
  for (size_t i = 0; i < (size_t)m->nelem * nn; i++) cnt[m->conn[i] + 1]++;
  for (int i = 0; i < m->nnode; i++) cnt[i + 1] += cnt[i];
  int *adj = malloc(sizeof(int) * ((size_t)m->nelem * nn + 1));
  int *pos = malloc(sizeof(int) * ((size_t)m->nnode + 1));
  memcpy(pos, cnt, sizeof(int) * m->nnode);
  int maxadj = 0;
  for (int i = 0; i < m->nnode; i++)
    if (cnt[i + 1] - cnt[i] > maxadj) maxadj = cnt[i + 1] - cnt[i];
  for (int e = 0; e < m->nelem; e++)
    for (int a = 0; a < nn; a++) adj[pos[m->conn[(size_t)e * nn + a]]++] = e;
  int *buf = malloc(sizeof(int) * ((size_t)maxadj * nn * 3 + 1));
  int *eqnode = malloc(sizeof(int) * (size_t)(neq > 0 ? neq : 1));
  for (int n = 0; n < m->nnode; n++)
    for (int c = 0; c < 3; c++) {
      int d = m->dofmap[3 * n + c];
      if (d) eqnode[d - 1] = n;
    }
  long long total = 0, prev_start = 0;
  int prev_node = -1, prev_len = 0;
  rowptr[0] = 0;
  for (int r = 0; r < neq; r++) {
    int n = eqnode[r];
    if (n == prev_node) { /* rows of one node share their columns */
      if (col) memcpy(col + total, col + prev_start, sizeof(int) * prev_len);
      total += prev_len;
      rowptr[r + 1] = total;
      continue;
    }
    /* neighbour nodes (sorted, unique), then their active equations,
     * sorted (a solver node order numbers nodes out of label order) */
    int k = 0;
    for (int t = cnt[n]; t < cnt[n + 1]; t++) {
      int e = adj[t];
      for (int a = 0; a < nn; a++) buf[k++] = m->conn[(size_t)e * nn + a];
    }
    qsort(buf, k, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < k; i++)
      if (i == 0 || buf[i] != buf[i - 1])
        for (int c = 0; c < 3; c++) {
          int d = m->dofmap[3 * buf[i] + c];
          if (!d) continue;
          if (col) col[total + u] = d - 1;
          u++;
        }
    if (col) qsort(col + total, u, sizeof(int), cmp_int);
    prev_node = n;
    prev_start = total;
    prev_len = u;
    total += u;
    rowptr[r + 1] = total;
  }
  free(cnt); free(adj); free(pos); free(buf); free(eqnode);
}

long long orc_csr_count(const orc_mesh *m, int neq, long long *rowptr)
{
  pattern_rows(m, neq, rowptr, NULL);
  return rowptr[neq];
}

void orc_csr_fill(const orc_mesh *m, int neq, const long long *rowptr,
                  int *col)
{
  long long *rp = malloc(sizeof(long long) * (size_t)(neq + 1));
  pattern_rows(m, neq, rp, col);
  (void)rowptr;
  free(rp);
}

/* ------------------------------------------------------------------------
 * Solvers.
 * ---------------------------------------------------------------------- */
void orc_spmv(int n, const long long *rowptr, const int *col,
              const double *val, const double *x, double *y)
{
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (long long p = rowptr[i]; p < rowptr[i + 1]; p++) s += val[p] * x[col[p]];
    y[i] = s;
  }
}

static double vlen(int n, const double *u)
{
  double d = 0.0;
  for (int i = 0; i < n; i++) d += u[i] * u[i];
  return sqrt(d);
}

/* Reference: mlib/sysmat.cpp:1061-1125 (PCGSolver): Jacobi preconditioner,
 * u0 = 0, convergence on |r| / |f| < tol with |f| > 1e11 replaced by 1,
 * immediate return when |f| < tol.  Returns 1 when converged. */
int orc_pcg(int n, const long long *rowptr, const int *col, const double *val,
            const double *f, double *u, int maxit, double tol, int *iters,
            double *relres)
{
  double lenf = vlen(n, f);
  for (int i = 0; i < n; i++) u[i] = 0.0;
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (lenf < tol) return 1;
  if (lenf > 10e+10) lenf = 1.0;
  double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
  double *d = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n);
  double *z = malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    d[i] = 0.0;
    for (long long k = rowptr[i]; k < rowptr[i + 1]; k++)
      if (col[k] == i) d[i] = val[k];
  }
  for (int i = 0; i < n; i++) { r[i] = f[i]; z[i] = r[i] / d[i]; p[i] = z[i]; }
  double dot0 = 0.0;
  for (int i = 0; i < n; i++) dot0 += z[i] * r[i];
  int conv = 0, k;
  double err = 1.0;
  for (k = 1; k <= maxit; k++) {
    orc_spmv(n, rowptr, col, val, p, w);
    double pw = 0.0;
    for (int i = 0; i < n; i++) pw += p[i] * w[i];
    double alpha = dot0 / pw;
    for (int i = 0; i < n; i++) u[i] = u[i] + alpha * p[i];
    for (int i = 0; i < n; i++) r[i] = r[i] + (-alpha) * w[i];
    for (int i = 0; i < n; i++) z[i] = r[i] / d[i];
    err = vlen(n, r) / lenf;
    if (err < tol) { conv = 1; break; }
    double dot1 = 0.0;
    for (int i = 0; i < n; i++) dot1 += z[i] * r[i];
    double beta = dot1 / dot0;
    dot0 = dot1;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  if (iters) *iters = conv ? k : maxit;
  if (relres) *relres = err;
  free(r); free(p); free(d); free(w); free(z);
  return conv;
}

/* Reference: ctrl/ctrl.cpp:2112-2157 (profile = lowest equation coupled to
 * each row) and mlib/matvec.cpp:609-664 (CroutSolver: L D L^T on the skyline,
 * forward reduction, diagonal scaling, back substitution; fails on a pivot
 * below 1e-12).  Skyline rows are built from the CSR lower part. */
int orc_skyline_solve(int n, const long long *rowptr, const int *col,
                      const double *val, const double *f, double *u)
{
  int *p = malloc(sizeof(int) * n);
  size_t total = 0;
  for (int i = 0; i < n; i++) {
    int mn = i;
    for (long long k = rowptr[i]; k < rowptr[i + 1]; k++)
      if (col[k] < mn) mn = col[k];
    p[i] = mn;
    total += (size_t)(i - mn + 1);
  }
  double **A = malloc(sizeof(double *) * n);
  double *store = calloc(total, sizeof(double));
  size_t off = 0;
  for (int i = 0; i < n; i++) {
    A[i] = store + off - p[i]; /* A[i][j] valid for p[i] <= j <= i */
    off += (size_t)(i - p[i] + 1);
    for (long long k = rowptr[i]; k < rowptr[i + 1]; k++)
      if (col[k] <= i) A[i][col[k]] = val[k];
  }
  int ok = 1;
  for (int j = 1; j < n && ok; j++) {
    for (int i = p[j] + 1; i < j; i++) {
      int k0 = p[i] > p[j] ? p[i] : p[j];
      for (int k = k0; k < i; k++) A[j][i] -= A[i][k] * A[j][k];
    }
    for (int k = p[j]; k < j; k++) {
      if (fabs(A[k][k]) < 1e-12) { ok = 0; break; }
      A[j][j] -= A[j][k] / A[k][k] * A[j][k];
      A[j][k] /= A[k][k];
    }
  }
  if (ok) {
    for (int i = 0; i < n; i++) u[i] = f[i];
    for (int i = 1; i < n; i++)
      for (int k = p[i]; k < i; k++) u[i] -= A[i][k] * u[k];
    for (int i = 0; i < n; i++) u[i] /= A[i][i];
    for (int i = n - 1; i > 0; i--)
      for (int k = p[i]; k < i; k++) u[k] -= A[i][k] * u[i];
  }
  free(p); free(A); free(store);
  return ok;
}
