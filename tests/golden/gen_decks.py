"""Fixtures of decks run by the reference program.

Decks in tests/golden/decks/:
  TestB8.dat, ScordelisLoDeg3.dat, PinCylDeg4.dat, ThinCylinderDeg5.dat,
  CurvedBeamDeg4.dat  the reference's own
      example decks (dat/continuum/TestB8.dat: one brick under a uniform end
      traction; dat/iga/ScordelisLoDeg3.dat: the Scordelis-Lo roof, 64 NURBS
      solids of degree (3,3,2) under self weight; dat/iga/PinCylDeg4.dat: a
      pinched cylinder, 64 NURBS solids of degree (4,2,4) on 5x5x5 Gauss
      points, nodal stresses evaluated at the nodes; dat/iga/ThinCylinderDeg5:
      a thin cylinder under a normal pressure on its inner faces, 80 NURBS
      solids of degree (5,2,5); dat/iga/CurvedBeamDeg4: a curved
      total-Lagrangian beam under an end-face load, 200 load steps;
      dat/iga/PinSemiCylDeg4.dat: the pinched semi-cylinder of Sze et al.,
      256 total-Lagrangian NURBS solids of degree (4,4,2) on 5x5x5 Gauss
      points under a point load in load-control steps of 0.025 -- run for
      the 40 steps of the deck's header table (load factors 0.05 to 1.0;
      the deck itself goes on to 100 steps) with every second step printed,
      keeping the stresses of its first 32 elements);
  BeamB8.dat, BeamB8TL.dat  written by this script: a 4x2x2 cantilever of
      distorted bricks under tip forces, linear and total-Lagrangian
      (the latter under load control, 4 steps to 4 times the load);
  BeamB8Face.dat  the same beam under a uniform load on its distorted top
      faces in the faces' local axes (tangential and normal components);
  ScordelisLoDeg3TL.dat  the Scordelis-Lo deck with total-Lagrangian
      B-spline solids under load control (2 steps);
  BeamB20.dat  the reference's dat/continuum/BeamB20.dat: a 2x2x50
      cantilever of 20-node bricks under uniform face loads on its tip faces;
  BeamB20TL.dat  the same beam with total-Lagrangian 20-node bricks under
      load control, 2 steps of 20 times the load.
This script runs the reference program built from its sources
(``make -C oracle ref`` -> oracle/_ref/fast) on a temporary copy of each deck
with ``%HEADER.OUTPUT.PRECISION 16`` added, and stores what it prints for the
last step as tests/golden/deck_<name>.npz: node labels, displacements u, the
load factor, every step's displacements u_steps and load factors, element
labels, integration-point stresses sig_ip [nelem, npt, 6]
and element-nodal stresses sig_node [nelem, nen, 6] (Voigt xx yy zz xy xz
yz), and the number of steps.

usage: python tests/golden/gen_decks.py [deck names (default: all)]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fast_amd import neutral  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "fast")
DECKS = ("TestB8", "ScordelisLoDeg3", "PinCylDeg4", "ThinCylinderDeg5", "CurvedBeamDeg4",
         "BeamB8", "BeamB8TL", "BeamB8Face", "ScordelisLoDeg3TL", "PinSemiCylDeg4", "BeamB20",
         "BeamB20TL")
PRINT_STEP = {"PinSemiCylDeg4": 2}   # HEADER.ANALYSIS.PRINT.STEPS of the run
RUN_STEPS = {"PinSemiCylDeg4": 40}   # HEADER.ANALYSIS.MAXIMUM.STEPS of the run
SIG_ELEMS = {"PinSemiCylDeg4": 32}   # stresses kept for the first elements only
LCM = "%HEADER.ANALYSIS.ALGORITHM\n'lcm'\n\n%HEADER.ANALYSIS.MAXIMUM.STEPS\n{steps}\n\n" \
      "%HEADER.ANALYSIS.STEP.FACTOR\n{factor}\n\n%HEADER.ANALYSIS.MAXIMUM.ITERATIONS\n30\n\n" \
      "%HEADER.ANALYSIS.TOLERANCE\n{tol}\n\n"


def beam_deck(tl: bool, face: bool = False) -> str:
    """4x2x2 bricks on [0,4]x[0,1]x[0,1], interior nodes moved by a seeded
    +-0.1 h, clamped at x = 0, tip forces (y and z) at x = 4, or (face) a
    uniform load in local axes on the top faces."""
    nx, ny, nz = 4, 2, 2
    rng = np.random.default_rng(11)
    nid = lambda i, j, k: 1 + i + (nx + 1) * (j + (ny + 1) * k)
    X = {}
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                x = np.array([i * 1.0, j * 0.5, k * 0.5])
                if 0 < i < nx:
                    x += rng.uniform(-0.1, 0.1, 3) * np.array([1.0, 0.5, 0.5])
                X[nid(i, j, k)] = x
    out = ["%HEADER", "Cantilever of distorted bricks under tip forces", ""]
    if tl:
        out.append(LCM.format(steps=4, factor=1.0, tol="1.0e-10"))  # 4x the linear load
    out += ["%NODE", str(len(X)), "", "%NODE.COORD", str(len(X))]
    out += ["%d  %.17g  %.17g  %.17g" % (n, *X[n]) for n in sorted(X)]
    sup = [nid(0, j, k) for k in range(nz + 1) for j in range(ny + 1)]
    out += ["", "%NODE.SUPPORT", str(len(sup))] + ["%d  1  1  1  0  0  0" % n for n in sup]
    out += ["", "%MATERIAL", "1", "", "%MATERIAL.ISOTROPIC", "1", "1  1000.0  0.3", "",
            "%SECTION", "1", "", "%SECTION.HOMOGENEOUS.ISOTROPIC.3D", "1", "1  1", "",
            "%INTEGRATION.ORDER", "1", "1  2  2  2  2  2  2", "",
            "%ELEMENT", str(nx * ny * nz), "", "%ELEMENT.BRICK8TL" if tl else "%ELEMENT.BRICK8",
            str(nx * ny * nz)]
    e = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                e += 1
                c = [nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                     nid(i, j + 1, k + 1), nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k),
                     nid(i, j + 1, k)]
                out.append("%d  1  1  " % e + "  ".join(str(n) for n in c))
    out += ["", "%LOAD", "1", "1  'tip'", "%LOAD.CASE", "1"]
    if face:
        top = [(1 + i + nx * (j + ny * (nz - 1)), nid(i, j, nz), nid(i + 1, j, nz),
                nid(i + 1, j + 1, nz)) for j in range(ny) for i in range(nx)]
        out += ["%LOAD.CASE.AREA.UNIFORM", str(len(top))]
        out += ["%d  %d  %d  %d  1  0.02  -0.01  -0.3" % t for t in top]
    else:
        tip = [nid(nx, j, k) for k in range(nz + 1) for j in range(ny + 1)]
        out += ["%LOAD.CASE.NODAL.FORCE", str(len(tip))]
        out += ["%d  0.0  %.6g  %.6g  0.0  0.0  0.0" % (n, 0.05, -0.25) for n in tip]
    out += ["", "%END", ""]
    return "\n".join(out)


def deck_text(name: str) -> str:
    if name == "BeamB8":
        return beam_deck(False)
    if name == "BeamB8TL":
        return beam_deck(True)
    if name == "BeamB8Face":
        return beam_deck(False, face=True)
    if name == "ScordelisLoDeg3TL":
        text = open(os.path.join(HERE, "decks", "ScordelisLoDeg3.dat")).read()
        text = text.replace("%ELEMENT.BSPLINE.SOLID\n", "%ELEMENT.BSPLINE.SOLIDTL\n", 1)
        return text.replace("%NODE\n", LCM.format(steps=2, factor=0.5, tol="1.0e-8") + "%NODE\n", 1)
    if name == "BeamB20TL":
        text = open(os.path.join(HERE, "decks", "BeamB20.dat")).read()
        text = text.replace("%ELEMENT.BRICK20\n", "%ELEMENT.BRICK20TL\n", 1)
        return text.replace("%NODE\n", LCM.format(steps=2, factor=20.0, tol="1.0e-8") + "%NODE\n",
                            1)
    return open(os.path.join(HERE, "decks", name + ".dat")).read()


def main(names=None):
    if not os.path.exists(REF):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"])
    with tempfile.TemporaryDirectory() as wd:
        for name in names or DECKS:
            text = deck_text(name)
            with open(os.path.join(HERE, "decks", name + ".dat"), "w") as f:
                f.write(text)
            # full precision in the result file (ctrl/ctrl.cpp ReadOutPrec)
            extra = "%HEADER.OUTPUT.PRECISION\n16\n\n"
            pstep = PRINT_STEP.get(name, 1)
            if pstep > 1:
                extra += "%%HEADER.ANALYSIS.PRINT.STEPS\n%d\n\n" % pstep
            text = text.replace("%NODE\n", extra + "%NODE\n", 1)
            if name in RUN_STEPS:
                k = text.index("%HEADER.ANALYSIS.MAXIMUM.STEPS")
                e = text.index("%", k + 1)
                text = text[:k] + "%%HEADER.ANALYSIS.MAXIMUM.STEPS\n%d\n\n" % RUN_STEPS[name] + \
                    text[e:]
            with open(os.path.join(wd, name + ".dat"), "w") as f:
                f.write(text)
            subprocess.check_call([REF, name], cwd=wd, stdout=subprocess.DEVNULL)
            steps = neutral.read_pos(os.path.join(wd, name + ".pos"))
            st = steps[-1]
            nsteps = len(steps) * pstep
            ne = SIG_ELEMS.get(name, len(st["sig_ip"]))
            np.savez_compressed(os.path.join(HERE, "deck_%s.npz" % name), labels=st["labels"],
                                u=st["u"], factor=st["factor"], nsteps=nsteps,
                                u_steps=np.stack([x["u"] for x in steps]),
                                factors=np.array([x["factor"] for x in steps]),
                                elem_labels=st["elem_labels"],
                                sig_ip=np.stack(st["sig_ip"][:ne]),
                                sig_node=np.stack(st["sig_node"][:ne]))
            print(name, len(st["labels"]), "nodes,", len(steps), "steps, max |u|",
                  np.abs(st["u"]).max(), "max |sig|", np.abs(np.stack(st["sig_ip"])).max())


if __name__ == "__main__":
    main(sys.argv[1:])
