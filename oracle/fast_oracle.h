/* fast_oracle.h - CPU FP64 restatement of the FAST element-integration,
 * assembly and linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the HIP path: it
 * is loaded by tests, by the smoke check and by the CPU-baseline leg of the
 * benchmark, and never by the product library.
 *
 * Every routine cites the reference file:line whose algorithm it restates.
 * Arithmetic is FP64 with the reference's operation order (compile with
 * -ffp-contract=off) so that element matrices agree bit for bit with the
 * reference build on the same inputs.
 */
#ifndef FAST_ORACLE_H
#define FAST_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_HEX8 = 0, ORC_BSP = 1, ORC_HEX20 = 2 };
enum { ORC_LINEAR = 0, ORC_TL_SVK = 1, ORC_TL_NEOHOOKE = 2 };

/* Element family description (one per mesh). */
typedef struct {
  int shape;      /* ORC_HEX8 | ORC_BSP | ORC_HEX20 */
  int nn;         /* nodes / control points per element */
  int deg[3];     /* B-spline degree per parametric direction (ORC_BSP) */
  int nip[3];     /* Gauss points per direction */
  int rational;   /* ORC_BSP: apply control-point weights (NURBS quotient) */
  int form;       /* ORC_LINEAR | ORC_TL_SVK | ORC_TL_NEOHOOKE */
} orc_etype;

/* Mesh: one element family, per-element material. */
typedef struct {
  int nnode, nelem;
  const double *coords;  /* [nnode*3] */
  const int    *conn;    /* [nelem*nn], 0-based node ids, reference node order */
  const int    *dofmap;  /* [nnode*3], 1-based equation or 0 = fixed */
  const double *ext;     /* ORC_BSP: [nelem*3*(p+1)^2] extraction rows per dir */
  const double *cpw;     /* ORC_BSP: [nnode] control-point weights */
  const double *mat;     /* [nelem*2]: (E, nu) or (mu, lambda) for neo-Hooke */
  orc_etype et;
} orc_mesh;

/* Gauss-Legendre rule with the reference's 15-decimal table values. */
int  orc_gauss(int n, double *r, double *w);
/* Brick rule: r fastest, then s, then t. */
int  orc_brick_points(const int nip[3], double *rst, double *w);

void orc_hex8_shape(double r, double s, double t, double *N, double *dN);
/* 20-node serendipity brick, node order of cShapeBrick20; dN [20][3] */
void orc_hex20_shape(double r, double s, double t, double *N, double *dN);
/* 8-node serendipity quadrilateral, node order of cShapeQ8; dN [8][2] */
void orc_q8_shape(double r, double s, double *N, double *dN);
void orc_bernstein(int p, double u, double *B);
void orc_bernstein_deriv(int p, double u, double *dB);
/* Bezier extraction operators for a knot vector (nspan*(p+1)^2). */
int  orc_bezier_extraction(int p, int nknot, const double *knots, double *C);

/* Shape-function derivatives w.r.t. x for one element at natural point. */
int  orc_shape_dxyz(const orc_etype *et, const double *X, const double *ext,
                    const double *w, const double rst[3], double *detJ,
                    double *dNxyz);

/* Element routines; u may be NULL (zero displacement). Return 1 ok / 0 fail. */
int  orc_elem_stiff(const orc_etype *et, const double *X, const double *ext,
                    const double *w, const double *mat, const double *u,
                    double *K);
int  orc_elem_intforce(const orc_etype *et, const double *X,
                       const double *ext, const double *w, const double *mat,
                       const double *u, double *g);

/* Mesh-level loops (ctrl.cpp StiffnessMatrix / IntForceVector), on a
 * caller-given full symmetric CSR pattern over equations (0-based rows). */
int  orc_assemble_K(const orc_mesh *m, const double *u_nodal, int neq,
                    const long long *rowptr, const int *col, double *val,
                    int nthreads, int elem_begin, int elem_end);
int  orc_elem_bodyforce(const orc_etype *et, const double *X, const double *ext,
                        const double *w, const double *b, double *f);
int  orc_assemble_body(const orc_mesh *m, const double *b, int neq, double *f);
int  orc_assemble_fint(const orc_mesh *m, const double *u_nodal, int neq,
                       double *g, int nthreads, int elem_begin, int elem_end);
/* Uniform surface loads on faces (cElemLoad::ExtForce over cFaceUnifLoad /
 * cFaceUnifLoadIGA faces): one group of faces of degree deg, nip x nip
 * Gauss points, nodes first surface parameter fastest, per-face extraction
 * rows of the two parameters (NULL: a brick face of degree (1, 1)), local
 * flags (NULL: global) and loads q [nface*3].  shape 1: 8-node serendipity
 * faces of 20-node bricks (deg (2, 2), ext NULL, nodes in cShapeQ8's order).
 * Adds into f [neq]. */
typedef struct {
  int nface;
  int deg[2];
  int nip;
  const int *nodes;
  const double *ext;
  const int *local;
  const double *q;
  int shape;
} orc_surface;
int  orc_surface_load(const orc_mesh *m, const orc_surface *s, double *f);
/* Stresses: integration-point values (Voigt xx yy zz xy xz yz, points in
 * the brick rule's order), the extrapolation matrix to the element's nodes
 * [nn*nq], and the whole-mesh output (element-nodal and node-averaged values
 * may be NULL). */
int  orc_elem_stress(const orc_etype *et, const double *X, const double *ext,
                     const double *w, const double *mat, const double *u,
                     double *sig);
int  orc_stress_extrap(const orc_etype *et, double *TR);
int  orc_stresses(const orc_mesh *m, const double *u_nodal, double *sig_ip,
                  double *sig_node, double *sig_avg);
int  orc_node_stresses(const orc_mesh *m, const double *u_nodal, double *sig_node);
/* Full CSR sparsity pattern of the equations (two-pass). */
long long orc_csr_count(const orc_mesh *m, int neq, long long *rowptr);
void orc_csr_fill(const orc_mesh *m, int neq, const long long *rowptr, int *col);

/* Reference PCG (Jacobi, x0 = 0) on a full symmetric CSR. */
int  orc_pcg(int n, const long long *rowptr, const int *col, const double *val,
             const double *f, double *u, int maxit, double tol, int *iters,
             double *relres);
void orc_spmv(int n, const long long *rowptr, const int *col,
              const double *val, const double *x, double *y);
/* Skyline Crout LDL^T direct solve from the CSR lower part. */
int  orc_skyline_solve(int n, const long long *rowptr, const int *col,
                       const double *val, const double *f, double *u);

#ifdef __cplusplus
}
#endif
#endif
