// ref_gpu.cpp - test driver of the FAST-side binding (fast_amd/binding):
// the reference's own model set-up with the element loops, the assembly and
// the linear solve on the device.
//
// Linked against the reference's object files (built from its sources by
// oracle/Makefile), fast_amd/binding/ctrlgpu.cpp and fast_amd/libfgpu.so.
// It reads a neutral-file deck exactly as the reference's main() does
// (main/main.cpp:55-124: InitIGA, InpReadData, InpCopyData, cControl::Init),
// calls GpuInit, and runs the deck's analysis with the binding in the places
// the reference's analysis classes call cControl:
//   linear static (ctrl/ctrllins.cpp:46-99): f = ExtLoadVector, K = device
//     stiffness with the prescribed-equation penalty, u from refinements
//     du = K^-1 (f - g(u)) with the device PCG (cGpuSysMatrix) and the
//     device internal force g, until |f - g| <= 1e-10 |f|;
//   load control (ctrl/ctrlnr.cpp:47-267): f = lambda q, full Newton with the
//     device tangent, PCG corrections and internal forces, convergence
//     |r| / max(1, |f|) < Tol over the free equations (ctrl/ctrl.cpp
//     1366-1413).
// Every converged step is written as "step <k> factor <lambda> iters <n>"
// followed by "<node label> ux uy uz" per node (%.17g).
//
// usage: ref_gpu <deck-basename> <out.txt> [pcg|direct]
//   pcg (default): corrections by cGpuSysMatrix::Solve (device Jacobi-PCG to
//   1e-12); direct: by the device LDL^T of the same matrix (fgpu_direct_solve,
//   the reference's default skyline solver).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace std;

#include "input.h"
#include "ctrl.h"
#include "element.h"
#include "node.h"
#include "prescdof.h"
#include "shpiga.h"
#include "vec.h"
#include "ctrlgpu.h"

// the analysis parameters cControl keeps for its drivers
struct cGpuCtrl : public cControl
{
  static int    Neq(void) { return NumEq; }
  static int    Steps(void) { return MaxStep; }
  static int    Iters(void) { return MaxIter; }
  static double Factor(void) { return StepFactor; }
  static double Tolerance(void) { return Tol; }
};

ifstream in;
ofstream out, outem;
string   fname;
bool Feedback = false;
bool Printing = false;

static void Write(FILE *fo, int step, double lam, int iters)
{
  fprintf(fo, "step %d factor %.17g iters %d\n", step, lam, iters);
  for (int i = 0; i < cNode :: GetNumNode( ); i++)
  {
    cNode *n = cNode :: GetNode(i);
    fprintf(fo, "%d %.17g %.17g %.17g\n", n->GetLabel( ), n->GetDispl(0), n->GetDispl(1),
            n->GetDispl(2));
  }
}

// |v| over the free equations (the prescribed ones excluded)
static double FreeNorm(cVector &v, const vector<int> &pd)
{
  double s = 0.0;
  vector<char> skip(v.Dim( ), 0);
  for (int k : pd) skip[k - 1] = 1;
  for (int i = 0; i < v.Dim( ); i++)
    if (!skip[i]) s += v[i] * v[i];
  return sqrt(s);
}

int main(int argc, char **argv)
{
  if (argc < 3) { fprintf(stderr, "usage: ref_gpu deck out.txt [pcg|direct]\n"); return 2; }
  const bool direct = argc > 3 && string(argv[3]) == "direct";
  InitIGA( );
  fname = string(argv[1]);
  in.open((fname + ".dat").c_str( ));
  if (!in.is_open( )) { fprintf(stderr, "cannot open deck\n"); return 1; }
  out.open((fname + ".gpu.pos").c_str( ));
  InpReadData( );
  InpCopyData( );
  in.close( );
  cControl :: Init( );

  if (!GpuInit( )) { fprintf(stderr, "%s\n", GpuLastError( )); return 1; }
  const int neq = cGpuCtrl :: Neq( );
  vector<int> pd(cPrescDof :: GetNumPrescDofs( ));
  if (!pd.empty( )) cPrescDof :: GetPrescDofs(pd.data( ));
  FILE *fo = fopen(argv[2], "w");
  cGpuSysMatrix K(neq);
  K.SetParam(50 * neq, 1e-12);
  cVector u(neq), du(neq), g(neq), r(neq), f(neq), q(neq);
  u.Zero( );
  cControl :: AssignDispl(u);

  if (cControl :: GetType( ) == LINEAR_STATIC)
  {
    cControl :: ExtLoadVector(f);
    if (!GpuStiffnessMatrix( )) { fprintf(stderr, "%s\n", GpuLastError( )); return 1; }
    const double flen = f.Length( );
    int k = 0;
    for (; k < 50; k++)
    {
      if (!GpuIntForceVector(g)) { fprintf(stderr, "%s\n", GpuLastError( )); return 1; }
      for (int i = 0; i < neq; i++) r[i] = f[i] - g[i];
      if (r.Length( ) <= 1e-10 * flen) break;
      du.Zero( );
      if (direct ? !fgpu_direct_solve(GpuContext( ), &r[0], &du[0]) : !K.Solve(r, du))
        fprintf(stderr, "solve: %s\n", fgpu_last_error(GpuContext( )));
      for (int i = 0; i < neq; i++) u[i] += du[i];
      cControl :: AssignDispl(u);
    }
    Write(fo, 1, 1.0, k);
  }
  else if (cControl :: GetType( ) == LOAD_CONTROL)
  {
    cControl :: ExtLoadVector(q);
    if (!GpuIntForceVector(g)) { fprintf(stderr, "%s\n", GpuLastError( )); return 1; }
    double lam = 0.0;
    for (int step = 1; step <= cGpuCtrl :: Steps( ); step++)
    {
      lam += cGpuCtrl :: Factor( );
      for (int i = 0; i < neq; i++) { f[i] = lam * q[i]; r[i] = f[i] - g[i]; }
      const double flen = FreeNorm(f, pd);
      int conv = 0, it = 1;
      for (; it <= cGpuCtrl :: Iters( ); it++)
      {
        if (!GpuStiffnessMatrix( )) { fprintf(stderr, "%s\n", GpuLastError( )); return 1; }
        du.Zero( );
        if (direct ? !fgpu_direct_solve(GpuContext( ), &r[0], &du[0]) : !K.Solve(r, du))
          fprintf(stderr, "solve: %s\n", fgpu_last_error(GpuContext( )));
        for (int i = 0; i < neq; i++) u[i] += du[i];
        cControl :: AssignDispl(u);
        if (!GpuIntForceVector(g)) { fprintf(stderr, "%s\n", GpuLastError( )); return 1; }
        for (int i = 0; i < neq; i++) r[i] = f[i] - g[i];
        const double ref = flen > cGpuCtrl :: Tolerance( ) ? flen : FreeNorm(g, pd);
        if (FreeNorm(r, pd) / (ref > 1.0 ? ref : 1.0) < cGpuCtrl :: Tolerance( )) { conv = 1; break; }
      }
      if (!conv) { fprintf(stderr, "no convergence at step %d\n", step); fclose(fo); return 1; }
      Write(fo, step, lam, it);
    }
  }
  else
  {
    fprintf(stderr, "ref_gpu: linear static and load-control decks only\n");
    return 1;
  }
  fclose(fo);
  GpuDestroy( );
  return 0;
}
