"""Fixtures of the reference's own example decks run by the reference.

The decks in tests/golden/decks/ are data files of the reference
(dat/continuum/TestB8.dat: one brick under a uniform end traction;
dat/iga/ScordelisLoDeg3.dat: the Scordelis-Lo roof, 64 NURBS solids of degree
(3,3,2) under self weight).  This script runs the reference program built
from its sources (``make -C oracle ref`` -> oracle/_ref/fast) on a temporary
copy of each deck with ``%HEADER.OUTPUT.PRECISION 16`` added and stores the
nodal displacements it prints as tests/golden/deck_<name>.npz (labels, u).

usage: python tests/golden/gen_decks.py
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
DECKS = ("TestB8", "ScordelisLoDeg3")


def main():
    if not os.path.exists(REF):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"])
    with tempfile.TemporaryDirectory() as wd:
        for name in DECKS:
            text = open(os.path.join(HERE, "decks", name + ".dat")).read()
            # full precision in the result file (ctrl/ctrl.cpp ReadOutPrec)
            text = text.replace("%NODE\n", "%HEADER.OUTPUT.PRECISION\n16\n\n%NODE\n", 1)
            with open(os.path.join(wd, name + ".dat"), "w") as f:
                f.write(text)
            subprocess.check_call([REF, name], cwd=wd, stdout=subprocess.DEVNULL)
            labels, u = neutral.read_pos_displacements(os.path.join(wd, name + ".pos"))
            np.savez_compressed(os.path.join(HERE, "deck_%s.npz" % name), labels=labels, u=u)
            print(name, len(labels), "nodes, max |u|", np.abs(u).max())


if __name__ == "__main__":
    main()
