// ref_probe.cpp - test-infrastructure driver linked against the reference's
// own object files (built from the reference sources by oracle/Makefile).
//
// It reads a neutral-file deck exactly as the reference's main() does
// (main/main.cpp:55-124: InitIGA, InpReadData, InpCopyData, cControl::Init),
// then calls the reference element plugins directly and dumps, per element,
// the DOF list (cElement::GetElmDofs, elc/element.cpp:2907-2926), the
// stiffness matrix (cElement::StiffMat, elc/elmparam.cpp:539 / 1711) and the
// internal force vector (cElement::IntForce, elc/elmparam.cpp:449 / 1615)
// for a prescribed nodal displacement field.
//
// usage: ref_probe <deck-basename> <nodal-displ-file|-> <out.bin>
//        ref_probe <deck-basename> <nodal-displ-file|-> --time <reps>
//   nodal-displ-file: text, 3 doubles per node in node order (or "-" = zero)
//   --time: instead of the dump, times the reference's own assembly loops
//   cControl::StiffnessMatrix (element StiffMat + AddGlobMat into the system
//   matrix the deck names, ctrl/ctrl.cpp:886-975) and
//   cControl::IntForceVector (ctrl/ctrl.cpp:719-813), best of <reps> calls
//   each, and prints "nelem <n> stiffness_s <t> int_force_s <t>".
// out.bin layout (little endian):
//   int32 nnode, int32 numeq, int32 nelem
//   int32 dof[nnode*3]                 (node VecDof[0..2], 1-based, 0 = fixed)
//   per element: int32 ndof, int32 edof[ndof], f64 K[ndof*ndof] (row major),
//                f64 g[ndof], int32 ok
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <fstream>
#include <iostream>
#include <vector>

using namespace std;

#include "input.h"
#include "ctrl.h"
#include "element.h"
#include "node.h"
#include "vec.h"
#include "mat.h"
#include "shpiga.h"

// access to the analysis set-up (system-matrix type, equations, profile)
// that cControl keeps for its drivers
struct cProbeCtrl : public cControl
{
  static cSysMatrix *NewMatrix(void) { return cSysMatrix :: CreateMatrix(SysMatType, NumEq, Profile); }
  static int Neq(void) { return NumEq; }
};

static double Seconds(chrono::steady_clock::time_point t0)
{
  return chrono::duration<double>(chrono::steady_clock::now( ) - t0).count( );
}

ifstream in;
ofstream out, outem;
string   fname;
bool Feedback = false;
bool Printing = false;

int main(int argc, char **argv)
{
  if (argc < 4) { fprintf(stderr, "usage: ref_probe deck displ out\n"); return 2; }
  InitIGA( );
  fname = string(argv[1]);
  in.open((fname + ".dat").c_str());
  if (!in.is_open()) { fprintf(stderr, "cannot open deck\n"); return 1; }
  out.open((fname + ".probe.pos").c_str());
  InpReadData( );
  InpCopyData( );
  in.close( );
  cControl :: Init( );

  int nnode = cNode :: GetNumNode( );
  vector<double> ud(3*(size_t)nnode, 0.0);
  if (string(argv[2]) != "-")
  {
    ifstream uf(argv[2]);
    for (size_t i = 0; i < ud.size(); i++) uf >> ud[i];
  }
  for (int i = 0; i < nnode; i++)
  {
    cNode *nd = cNode :: GetNode(i);
    for (int j = 0; j < 6; j++) nd->SetDispl(j, j < 3 ? ud[3*i+j] : 0.0);
  }

  if (string(argv[3]) == "--time")
  {
    int reps = argc > 4 ? atoi(argv[4]) : 1;
    if (reps < 1) reps = 1;
    cSysMatrix *K = cProbeCtrl :: NewMatrix( );
    cVector g(cProbeCtrl :: Neq( ));
    double tk = 1e30, tf = 1e30;
    for (int r = 0; r < reps; r++)
    {
      chrono::steady_clock::time_point t0 = chrono::steady_clock::now( );
      cControl :: StiffnessMatrix(K);
      double t = Seconds(t0);
      if (t < tk) tk = t;
      t0 = chrono::steady_clock::now( );
      cControl :: IntForceVector(g);
      t = Seconds(t0);
      if (t < tf) tf = t;
    }
    printf("nelem %d stiffness_s %.6f int_force_s %.6f\n", cElement :: GetNumElm( ), tk, tf);
    delete K;
    return 0;
  }

  FILE *fo = fopen(argv[3], "wb");
  int nelm = cElement :: GetNumElm( );
  int numeq = 0;
  vector<int> dofs(3*(size_t)nnode);
  for (int i = 0; i < nnode; i++)
    for (int j = 0; j < 3; j++)
    {
      dofs[3*i+j] = cNode :: GetNode(i)->GetDof(j);
      if (dofs[3*i+j] > numeq) numeq = dofs[3*i+j];
    }
  fwrite(&nnode, 4, 1, fo);
  fwrite(&numeq, 4, 1, fo);
  fwrite(&nelm, 4, 1, fo);
  fwrite(dofs.data(), 4, dofs.size(), fo);

  for (int e = 0; e < nelm; e++)
  {
    cElement *elm = cElement :: GetElm(e);
    int ndof = elm->GetNumElmDof( );
    vector<int> edof(ndof);
    elm->GetElmDofs(edof.data());
    cMatrix K(ndof, ndof);
    cVector g(ndof);
    elm->StiffMat(K);
    int ok = elm->IntForce(g);
    fwrite(&ndof, 4, 1, fo);
    fwrite(edof.data(), 4, ndof, fo);
    for (int r = 0; r < ndof; r++) fwrite(K[r], 8, ndof, fo);
    for (int r = 0; r < ndof; r++) { double v = g[r]; fwrite(&v, 8, 1, fo); }
    fwrite(&ok, 4, 1, fo);
  }
  fclose(fo);
  return 0;
}
