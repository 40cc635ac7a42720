// ctrlgpu.cpp - FAST side of the MI355X path (see ctrlgpu.h).
//
// GpuInit reads the model the reference has set up (after cControl::Init:
// nodes with their VecDof equations, elements with their shapes, sections
// and integration points, B-spline patches with their extraction operators)
// into an fgpu_mesh and creates the device context.  The reference's own
// tables are handed over unchanged: the element's local node order
// (cShape::GetNode), VecDof (1-based equations, 0 = fixed), the control-point
// weights (cNode::GetW) and the Bezier extraction rows of the element's spans
// (cBspPat::getExtMat, the C of cShapeBsp::EvalBspBas, shp/shpbsp.cpp:
// 252-271).
#include "ctrlgpu.h"

#include <cstdio>
#include <iostream>
#include <string>
#include <unordered_map>
#include <vector>

#include "element.h"
#include "elmparam.h"
#include "anmodel.h"
#include "bspbasis.h"
#include "gbldef.h"
#include "knotvec.h"
#include "material.h"
#include "node.h"
#include "patch.h"
#include "prescdof.h"
#include "section.h"
#include "shape.h"
#include "shpbsp.h"
#include "vec.h"

namespace {

fgpu_ctx *Ctx = 0;
std::string Err;
std::vector<int> PrDof;  // prescribed equations (1-based), cPrescDof order

int Fail(const std::string &msg)
{
  Err = msg;
  return 0;
}

int FailCtx(const char *what)
{
  Err = std::string(what) + ": " + fgpu_last_error(Ctx);
  return 0;
}

// Current nodal displacements, node-major [nnode][3] (cNode::VecDispl).
std::vector<double> NodalDispl(void)
{
  int nnode = cNode :: GetNumNode( );
  std::vector<double> u(3 * (size_t)nnode);
  for (int i = 0; i < nnode; i++)
    for (int d = 0; d < 3; d++) u[3 * (size_t)i + d] = cNode :: GetNode(i)->GetDispl(d);
  return u;
}

// Gauss points per direction of an element's stiffness rule: the distinct
// natural coordinates of its integration points (a tensor brick rule,
// elc/intpoint.cpp:579-617).
void PointsPerDirection(cElement *elm, int nip[3])
{
  std::vector<double> c[3];
  int n = elm->GetNumIntPnt( );
  for (int k = 0; k < n; k++)
  {
    sNatCoord p;
    double w;
    elm->GetIntPnt(k, p, w);
    double v[3] = {p.r, p.s, p.t};
    for (int d = 0; d < 3; d++)
    {
      bool seen = false;
      for (double x : c[d]) seen = seen || x == v[d];
      if (!seen) c[d].push_back(v[d]);
    }
  }
  for (int d = 0; d < 3; d++) nip[d] = (int)c[d].size( );
}

// The element's span in direction d of its patch: the span whose first
// basis function is the element's first (cShapeBsp::CompBasisIndex).
int SpanOf(cShapeBsp *shp, cBspPat *pat, int d)
{
  int nb = 0, *bi = 0;
  shp->CompBasisIndex(d, nb, bi);
  int first = bi[0];
  delete[] bi;
  const cBSplineBasis &bas = pat->getBasis(d);
  int nspan = bas.getKnotVec( ).getNumSpan( );
  for (int s = 0; s < nspan; s++)
    if (bas.getBasInd(s) == first) return s;
  return -1;
}

}  // namespace

// ============================== GpuInit ===================================

int GpuInit(void)
{
  GpuDestroy( );
  int nnode = cNode :: GetNumNode( ), nelm = cElement :: GetNumElm( );
  if (nnode <= 0 || nelm <= 0) return Fail("GpuInit: empty model");

  std::unordered_map<const cNode *, int> index;
  std::vector<double> X(3 * (size_t)nnode), cpw(nnode);
  std::vector<int> dof(3 * (size_t)nnode);
  for (int i = 0; i < nnode; i++)
  {
    cNode *n = cNode :: GetNode(i);
    index[n] = i;
    sNodeCoord c = n->GetCoord( );
    X[3 * (size_t)i] = c.x;
    X[3 * (size_t)i + 1] = c.y;
    X[3 * (size_t)i + 2] = c.z;
    for (int d = 0; d < 3; d++) dof[3 * (size_t)i + d] = n->GetDof(d);
    cpw[i] = n->GetW( );
  }

  // one element family: linear or total-Lagrangian parametric solids, all
  // 8-node bricks, all 20-node bricks or all B-spline solids of one degree
  // and one integration rule
  cElement *e0 = cElement :: GetElm(0);
  eShpType st = e0->GetShape( )->GetType( );
  if (st != SHAPE_BRICK8 && st != SHAPE_BRICK20 && st != SHAPE_BSP_SOLID)
    return Fail("GpuInit: the device path takes BRICK8, BRICK20 and B-spline solid elements");
  fgpu_mesh m = fgpu_mesh( );
  m.nnode = nnode;
  m.nelem = nelm;
  m.shape = st == SHAPE_BRICK8 ? FGPU_HEX8 : st == SHAPE_BRICK20 ? FGPU_HEX20 : FGPU_BSP;
  m.nn = e0->GetShape( )->GetNumNode( );
  m.form = e0->GetType( ) == PARAMETRICTL ? FGPU_TL_SVK : FGPU_LINEAR;
  PointsPerDirection(e0, m.nip);
  m.deg[0] = m.deg[1] = m.deg[2] = 1;

  std::vector<int> conn((size_t)nelm * m.nn);
  std::vector<double> mat(2 * (size_t)nelm), ext;
  for (int e = 0; e < nelm; e++)
  {
    cElement *elm = cElement :: GetElm(e);
    cShape *shp = elm->GetShape( );
    int nip[3];
    PointsPerDirection(elm, nip);
    if (shp->GetType( ) != st || shp->GetNumNode( ) != m.nn || elm->GetType( ) != e0->GetType( ) ||
        nip[0] != m.nip[0] || nip[1] != m.nip[1] || nip[2] != m.nip[2])
      return Fail("GpuInit: elements of more than one kind, size or integration order");
    if (elm->GetType( ) != PARAMETRIC && elm->GetType( ) != PARAMETRICTL)
      return Fail("GpuInit: parametric (cElmParam) elements only");
    cAnModel *mdl = elm->GetAnModel( );
    if (!mdl || mdl->GetType( ) != SOLID) return Fail("GpuInit: 3D solid models only");
    cSection *sec = elm->GetSection( );
    cMaterial *mt = sec ? sec->GetMaterial( ) : 0;
    cMatMec *mec = mt ? mt->GetMecProp( ) : 0;
    if (!mec || mec->GetType( ) != MAT_ELASTIC_ISOTROPIC)
      return Fail("GpuInit: isotropic elastic materials only");
    double prm[3];
    mec->GetParam(prm);  // cMatElastIso: E, nu, alpha (mat/material.cpp:846-851)
    mat[2 * (size_t)e] = prm[0];
    mat[2 * (size_t)e + 1] = prm[1];
    for (int a = 0; a < m.nn; a++)
      conn[(size_t)e * m.nn + a] = index[shp->GetNode(a)];
    if (m.shape == FGPU_BSP)
    {
      cShapeBsp *bsp = dynamic_cast<cShapeBsp *>(shp);
      cBspPat *pat = bsp ? dynamic_cast<cBspPat *>(bsp->GetPatch( )) : 0;
      if (!pat) return Fail("GpuInit: B-spline solid without its patch");
      for (int d = 0; d < 3; d++)
      {
        int p = pat->getBasis(d).getDegree( );
        if (e == 0) m.deg[d] = p;
        if (p != m.deg[d]) return Fail("GpuInit: B-spline solids of more than one degree");
        int s = SpanOf(bsp, pat, d);
        double **C = s >= 0 ? pat->getExtMat(d, s) : 0;
        if (!C) return Fail("GpuInit: no extraction operator for an element span");
        for (int i = 0; i <= p; i++)
          for (int j = 0; j <= p; j++) ext.push_back(C[i][j]);  // N_i = sum_j C_ij B_j
      }
    }
  }
  m.coords = X.data( );
  m.conn = conn.data( );
  m.dofmap = dof.data( );
  m.mat = mat.data( );
  m.ext = ext.empty( ) ? 0 : ext.data( );
  m.cpw = m.shape == FGPU_BSP ? cpw.data( ) : 0;
  if (!fgpu_create(&Ctx, &m))
  {
    Ctx = 0;
    return Fail("GpuInit: fgpu_create failed (see stderr)");
  }

  // the penalty of the prescribed equations stays on the device: every
  // stiffness adds it to the diagonal, every internal force adds
  // PENALTY_STIFFNESS times the current value (ctrl/ctrl.cpp:958-974,
  // 794-810)
  int npd = cPrescDof :: GetNumPrescDofs( );
  PrDof.assign(npd, 0);
  if (npd > 0)
  {
    cPrescDof :: GetPrescDofs(PrDof.data( ));
    std::vector<double> pen(npd, PENALTY_STIFFNESS);
    if (!fgpu_set_penalty(Ctx, npd, PrDof.data( ), pen.data( ))) return FailCtx("GpuInit");
  }
  return 1;
}

void GpuDestroy(void)
{
  if (Ctx) fgpu_destroy(Ctx);
  Ctx = 0;
  PrDof.clear( );
}

fgpu_ctx *GpuContext(void) { return Ctx; }

const char *GpuLastError(void) { return Err.c_str( ); }

// ========================== GpuStiffnessMatrix ============================

int GpuStiffnessMatrix(void)
{
  if (!Ctx) return Fail("GpuStiffnessMatrix: no context (GpuInit)");
  std::vector<double> u = NodalDispl( );
  if (!fgpu_assemble_K(Ctx, u.data( ))) return FailCtx("GpuStiffnessMatrix");
  return 1;
}

// ========================== GpuIntForceVector =============================

int GpuIntForceVector(cVector &g)
{
  if (!Ctx) return Fail("GpuIntForceVector: no context (GpuInit)");
  std::vector<double> u = NodalDispl( );
  if (!fgpu_int_force(Ctx, u.data( ), &g[0])) return FailCtx("GpuIntForceVector");
  return 1;
}

// ============================ cGpuSysMatrix ===============================

int cGpuSysMatrix :: Solve(cVector &b, cVector &x)
{
  int it = 0;
  double rr = 0.0;
  return fgpu_solve(Ctx, &b[0], &x[0], 0, maxit > 0 ? maxit : dim, tol, &it, &rr);
}

double cGpuSysMatrix :: Get(int i, int j)
{
  std::vector<long long> rp(dim + 1);
  std::vector<int> col((size_t)fgpu_nnz(Ctx));
  std::vector<double> val(col.size( ));
  if (!fgpu_get_csr(Ctx, rp.data( ), col.data( ), val.data( ))) return 0.0;
  for (long long k = rp[i]; k < rp[i + 1]; k++)
    if (col[k] == j) return val[k];
  return 0.0;
}

void cGpuSysMatrix :: Add(int i, int j, double v)
{
  if (i == j)
  {
    int eq = i + 1;
    fgpu_add_diag(Ctx, 1, &eq, &v);
    return;
  }
  std::vector<long long> rp(dim + 1);
  std::vector<int> col((size_t)fgpu_nnz(Ctx));
  std::vector<double> val(col.size( ));
  if (!fgpu_get_csr(Ctx, rp.data( ), col.data( ), val.data( ))) return;
  for (int r = 0; r < 2; r++)  // the pattern is stored in full: both (i, j) and (j, i)
  {
    int a = r ? j : i, b = r ? i : j;
    for (long long k = rp[a]; k < rp[a + 1]; k++)
      if (col[k] == b) val[k] += v;
  }
  fgpu_set_csr_values(Ctx, val.data( ));
}

void cGpuSysMatrix :: Print(void)
{
  std::cout << "cGpuSysMatrix: " << dim << " equations, " << fgpu_nnz(Ctx)
            << " stored entries on the device" << std::endl;
}

void cGpuSysMatrix :: MultVect(double *x, double *y)
{
  std::vector<long long> rp(dim + 1);
  std::vector<int> col((size_t)fgpu_nnz(Ctx));
  std::vector<double> val(col.size( ));
  if (!fgpu_get_csr(Ctx, rp.data( ), col.data( ), val.data( ))) return;
  for (int i = 0; i < dim; i++)
  {
    double s = 0.0;
    for (long long k = rp[i]; k < rp[i + 1]; k++) s += val[k] * x[col[k]];
    y[i] = s;
  }
}
