// ctrlgpu.h - FAST side of the MI355X path: the binding a FAST maintainer adds
// (ctrl/ctrlgpu.h, ctrl/ctrlgpu.cpp) to run the element loops, the assembly
// and the linear solve of cControl on the device through include/fgpu.h.
//
// It uses only public reference interfaces: cNode (node/node.h:341-360),
// cElement / cElmParam (elc/element.h:640-700, elc/elmparam.h:50-110),
// cShapeBsp (shp/shpbsp.h:21-36), cBspPat (geom/patch.h:569-592),
// cMatElastIso (mat/material.h), cPrescDof (load/prescdof.h:90-111) and
// cSysMatrix (mlib/sysmat.h:60-99).  The analysis classes call:
//
//   cControl::Init      -> GpuInit() once the equations are numbered
//   StiffnessMatrix(K)  -> GpuStiffnessMatrix()    (ctrl/ctrl.cpp:886-975)
//   IntForceVector(g)   -> GpuIntForceVector(g)    (ctrl/ctrl.cpp:719-813)
//   K->Solve(r, du)     -> a cGpuSysMatrix created for NumEq equations
//
// Every function returns 1 on success and 0 on failure, with the reason on
// GpuLastError() (the reference's int returns; nothing calls exit()).
#ifndef CTRLGPU_H
#define CTRLGPU_H

#include "fgpu.h"
#include "sysmat.h"

class cVector;

int         GpuInit(void);
void        GpuDestroy(void);
fgpu_ctx   *GpuContext(void);
const char *GpuLastError(void);

// global stiffness of the current nodal displacements, with the penalty
// stiffness on the prescribed equations (ctrl/ctrl.cpp:958-974)
int GpuStiffnessMatrix(void);

// internal force of the current nodal displacements into g (NumEq
// entries), with the prescribed-equation penalty force (ctrl/ctrl.cpp:
// 794-810); 0 when a stress update fails or a Jacobian is not positive
int GpuIntForceVector(cVector &g);

// A system matrix whose Solve is the device Jacobi-PCG of the context's
// assembled matrix (cSysMatrix::Solve with PCGSolver, mlib/sysmat.cpp:
// 1061-1125).  Get, Add and MultVect act on the device matrix through its
// equation CSR (fgpu_get_csr / fgpu_set_csr_values / fgpu_add_diag).
class cGpuSysMatrix : public cSysMatrix
{
 public:
  explicit cGpuSysMatrix(int n) : cSysMatrix(n) { }
  int    Symmetric(void) { return 1; }
  void   Zero(void) { }
  int    Solve(cVector &b, cVector &x);
  double Get(int i, int j);
  void   Add(int i, int j, double v);
  void   Print(void);
  void   MultVect(double *x, double *y);
};

#endif
