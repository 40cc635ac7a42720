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
//   nodal-displ-file: text, 3 doubles per node in node order (or "-" = zero)
// out.bin layout (little endian):
//   int32 nnode, int32 numeq, int32 nelem
//   int32 dof[nnode*3]                 (node VecDof[0..2], 1-based, 0 = fixed)
//   per element: int32 ndof, int32 edof[ndof], f64 K[ndof*ndof] (row major),
//                f64 g[ndof], int32 ok
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
