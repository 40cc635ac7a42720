"""Run a FAST neutral-file deck on the device: read (fast_amd.neutral), form
the load vector (nodal forces from the deck, surface and body loads
integrated on the device, the penalty force of prescribed displacements),
solve (linear static, or load-control Newton-Raphson), compute the
stresses on the device, and write displacements and stresses in the
reference's result layout.

usage: python -m fast_amd.deck <deck.dat> [--out <file.pos>] [--tol 1e-9]
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import neutral
from .fgpu import FgpuError, GpuModel


MAX_DIRECT_BAND = 16384  # widest half band of the device direct solve
PENALTY_STIFFNESS = 1.0e30  # main/gbldef.h:148


def prepare(d, model):
    """Hand the deck's surface loads and prescribed equations to the model:
    the surface groups are uploaded for device integration and every
    prescribed equation gets the penalty stiffness on its diagonal and the
    penalty force in the internal force (ctrl/ctrl.cpp:958-974, 794-810)."""
    model.clear_surface_loads()
    for g in d.surfaces:
        model.add_surface_loads(g)
    if d.prescribed is not None:
        eq, _ = d.prescribed
        model.set_penalty(eq, np.full(len(eq), PENALTY_STIFFNESS))


def load_vector(d, model):
    """The deck's external load vector, cControl::ExtLoadVector
    (ctrl/ctrl.cpp:666-716) at t = 1: nodal forces, the surface loads and
    body forces integrated on the device, and the penalty force
    PENALTY_STIFFNESS * value of each prescribed equation.  ``prepare`` must
    have been called on the model."""
    f = d.load.copy()
    if d.surfaces:
        f += model.surface_load()
    if d.body is not None:
        f += model.body_force(d.body)
    if d.prescribed is not None:
        for eq, v in zip(*d.prescribed):
            f[eq - 1] += PENALTY_STIFFNESS * v
    return f


def run(path: str, out: str | None = None, tol: float = 1e-9, stresses: bool = True,
        solver: str | None = None):
    """Returns (deck, nodal displacements [nnode, 3] of the last step, stresses)
    with stresses = (integration points [nelem, nq, 6], element nodes
    [nelem, nn, 6], node averages [nnode, 6] or None) of that step, or None.

    solver: "direct" (the device LDL^T, the deck's default skyline matrix),
    "pcg" (the deck's sparse matrix types), or None to follow the deck; a
    band wider than the direct solve allows falls back to PCG.  Decks with
    prescribed displacements use the direct solve (the penalty diagonal
    leaves the PCG's residual test to the prescribed equations)."""
    d = neutral.read(path)
    model = GpuModel(d.mesh)
    prepare(d, model)
    if stresses and not d.stress_extrapolation:
        model.set_stress_nodes(False)  # HEADER.OUTPUT.STRESS.EXTRAPOLATION 0
    if solver is None:
        solver = "direct" if d.matrix_type.endswith("skyline") else "pcg"
    if d.prescribed is not None:
        solver = "direct"
        if model.factor_info()[0] > MAX_DIRECT_BAND:
            raise FgpuError("prescribed displacements need the direct solve; the band is "
                            "wider than %d" % MAX_DIRECT_BAND)
    elif solver == "direct" and model.factor_info()[0] > MAX_DIRECT_BAND:
        solver = "pcg"
    f = load_vector(d, model)
    if d.analysis == "linear":
        if solver == "direct":
            u, it, rr = model.linear_static_direct(f, tol=tol)
        else:
            u, it, rr = model.linear_static(f, maxit=max(100000, 200 * d.mesh.neq), tol=tol)
        us, factors = [u], [1.0]
    else:
        # ctrl/ctrlnr.cpp: f = lambda q per step, Newton to TOLERANCE.  With
        # PCG for the corrections its iteration limit is that of a
        # direct-quality solve.
        us, st, err = model.newton(f, steps=d.steps, factor=d.factor, maxit=d.maxit, tol=d.tol,
                                   pcg_tol=tol, pcg_maxit=max(100000, 200 * d.mesh.neq),
                                   direct=solver == "direct")
        if err:
            raise FgpuError(err)
        factors = [s["lam"] for s in st]
    # every step is printed, as the reference does at its default print step
    steps = []
    for u, lam in zip(us, factors):
        un = d.mesh.scatter(u)
        sig = model.stresses(un, nodal=True, average=d.average_stress) if stresses else None
        steps.append(dict(factor=lam, u=un, sig_ip=sig[0] if sig else None,
                          sig_node=sig[1] if sig else None, sig_avg=sig[2] if sig else None))
    if out:
        neutral.write_pos_steps(out, d.labels, steps, elem_labels=d.elem_labels)
    model.close()
    last = steps[-1]
    sig = (last["sig_ip"], last["sig_node"], last["sig_avg"]) if stresses else None
    return d, last["u"], sig


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("deck")
    ap.add_argument("--out", default=None)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--solver", choices=["direct", "pcg"], default=None,
                    help="default: the deck's matrix type (skyline: direct)")
    a = ap.parse_args(argv)
    d, un, sig = run(a.deck, a.out, a.tol, solver=a.solver)
    k = int(np.argmax(np.abs(un).max(axis=1)))
    print("%s: %d nodes, %d equations; largest displacement at node %d: %s"
          % (a.deck, d.mesh.nnode, d.mesh.neq, d.labels[k], un[k]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
