"""Run a FAST neutral-file deck on the device: read (fast_amd.neutral), form
the load vector (nodal and face loads on the host, body forces on the
device), solve (linear static, or load-control Newton-Raphson), write the
nodal displacements in the reference's result layout.

usage: python -m fast_amd.deck <deck.dat> [--out <file.pos>] [--tol 1e-9]
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import neutral
from .fgpu import FgpuError, GpuModel


def run(path: str, out: str | None = None, tol: float = 1e-9):
    """Returns (deck, nodal displacements [nnode, 3] of the last step)."""
    d = neutral.read(path)
    model = GpuModel(d.mesh)
    f = d.load.copy()
    if d.body is not None:
        f += model.body_force(d.body)
    if d.analysis == "linear":
        u, it, rr = model.linear_static(f, maxit=max(100000, 200 * d.mesh.neq), tol=tol)
        factor = 1.0
    else:
        # ctrl/ctrlnr.cpp: f = lambda q per step, Newton to TOLERANCE
        us, st, err = model.newton(f, steps=d.steps, factor=d.factor, maxit=d.maxit, tol=d.tol,
                                   pcg_tol=tol)
        if err:
            raise FgpuError(err)
        u = us[-1]
        factor = d.steps * d.factor
    un = d.mesh.scatter(u)
    if out:
        neutral.write_pos(out, d.labels, un, factor)
    model.close()
    return d, un


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("deck")
    ap.add_argument("--out", default=None)
    ap.add_argument("--tol", type=float, default=1e-9)
    a = ap.parse_args(argv)
    d, un = run(a.deck, a.out, a.tol)
    k = int(np.argmax(np.abs(un).max(axis=1)))
    print("%s: %d nodes, %d equations; largest displacement at node %d: %s"
          % (a.deck, d.mesh.nnode, d.mesh.neq, d.labels[k], un[k]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
