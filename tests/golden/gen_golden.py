"""Write the golden fixtures of tests/golden/ by running the reference.

Runs only where the reference build exists (``make -C oracle ref`` compiles
it from the reference sources into oracle/_ref/).  For each generated deck it
runs

* ``oracle/_ref/ref_probe``: the reference's own element plugins
  (cElmParam[TL]::StiffMat / IntForce) on a prescribed nodal displacement
  field -> element DOF lists, K_e and g_e;
* ``oracle/_ref/fast``: the whole reference program (linear static with the
  skyline Crout solver, or load-control Newton-Raphson) with
  ``%HEADER.OUTPUT.PRECISION 16`` -> nodal displacements per step.

The fixtures hold only inputs and outputs (numpy .npz, no pickles).

usage: python tests/golden/gen_golden.py
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fast_amd import mesh as M  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def write_deck(path, mesh, algo="linear", steps=1, factor=1.0, maxit=20,
               tol=1e-8, prec=16, matrix="skyline", nodal_loads=None):
    """Neutral-file deck for a Mesh (main/input.cpp keyword table)."""
    L = []
    w = L.append
    w("%HEADER\ngenerated deck\n")
    w("%%HEADER.ANALYSIS.ALGORITHM\n'%s'\n" % algo)
    w("%%HEADER.ANALYSIS.MATRIX.TYPE\n'%s'\n" % matrix)
    w("%%HEADER.ANALYSIS.MAXIMUM.STEPS\n%d\n" % steps)
    w("%%HEADER.ANALYSIS.STEP.FACTOR\n%.17g\n" % factor)
    w("%%HEADER.ANALYSIS.MAXIMUM.ITERATIONS\n%d\n" % maxit)
    w("%%HEADER.ANALYSIS.TOLERANCE\n%.17g\n" % tol)
    w("%%HEADER.OUTPUT.PRECISION\n%d\n" % prec)
    w("%%NODE\n%d\n" % mesh.nnode)
    w("%%NODE.COORD\n%d\n" % mesh.nnode)
    for i, (x, y, z) in enumerate(mesh.coords):
        w("%d %.17g %.17g %.17g" % (i + 1, x, y, z))
    fx = np.where(mesh.fixed.any(axis=1))[0]
    if len(fx):
        w("\n%%NODE.SUPPORT\n%d" % len(fx))
        for i in fx:
            f = mesh.fixed[i].astype(int)
            w("%d %d %d %d 0 0 0" % (i + 1, f[0], f[1], f[2]))
    if mesh.shape == "bsp":
        w("\n%%CONTROL.POINT.WEIGHT\n%d" % mesh.nnode)
        for i in range(mesh.nnode):
            w("%d %.17g" % (i + 1, mesh.cpw[i]))
        w("\n%PATCH\n1\n1 'bspsol' 1")
        for d in range(3):
            U = mesh.knots[d]
            uniq, mult = np.unique(U, return_counts=True)
            w("%d 'General' %d %s %s" % (
                mesh.deg[d], len(uniq), " ".join("%.17g" % v for v in uniq),
                " ".join(str(m) for m in mult)))
        w(" ".join(str(i + 1) for i in range(mesh.nnode)))
    mats = {}
    for e in range(mesh.nelem):
        key = tuple(mesh.mat[e])
        mats.setdefault(key, len(mats) + 1)
    w("\n%%MATERIAL\n%d\n\n%%MATERIAL.ISOTROPIC\n%d" % (len(mats), len(mats)))
    for (E, nu), k in mats.items():
        w("%d %.17g %.17g" % (k, E, nu))
    w("\n%%SECTION\n%d\n\n%%SECTION.HOMOGENEOUS.ISOTROPIC.3D\n%d" % (len(mats), len(mats)))
    for _, k in mats.items():
        w("%d %d" % (k, k))
    n = mesh.nip
    w("\n%%INTEGRATION.ORDER\n1\n1 %d %d %d %d %d %d" % (n[0], n[1], n[2], n[0], n[1], n[2]))
    w("\n%%ELEMENT\n%d\n" % mesh.nelem)
    tl = mesh.form != "linear"
    if mesh.shape == "hex8":
        w("%%ELEMENT.BRICK8%s\n%d" % ("TL" if tl else "", mesh.nelem))
        for e in range(mesh.nelem):
            sec = mats[tuple(mesh.mat[e])]
            w("%d %d 1 %s" % (e + 1, sec, " ".join(str(c + 1) for c in mesh.conn[e])))
    else:
        w("%%ELEMENT.BSPLINE.SOLID%s\n%d" % ("TL" if tl else "", mesh.nelem))
        for e in range(mesh.nelem):
            sec = mats[tuple(mesh.mat[e])]
            s = mesh.span[e] + 1
            w("%d %d 1 1 %d %d %d" % (e + 1, sec, s[0], s[1], s[2]))
    if nodal_loads is not None:
        idx = np.where(np.abs(nodal_loads).sum(axis=1) > 0)[0]
        w("\n%LOAD\n1\n1 'case1'\n%LOAD.CASE\n1\n")
        w("%%LOAD.CASE.NODAL.FORCE\n%d" % len(idx))
        for i in idx:
            f = nodal_loads[i]
            w("%d %.17g %.17g %.17g 0 0 0" % (i + 1, f[0], f[1], f[2]))
    w("\n%END\n")
    with open(path, "w") as fh:
        fh.write("\n".join(L))


def run_probe(mesh, u_nodal, workdir, name):
    deck = os.path.join(workdir, name)
    write_deck(deck + ".dat", mesh)
    ufile = os.path.join(workdir, name + ".u")
    np.savetxt(ufile, u_nodal.reshape(-1, 3), fmt="%.17g")
    out = os.path.join(workdir, name + ".bin")
    subprocess.check_call([os.path.join(REF, "ref_probe"), deck, ufile, out],
                          stdout=subprocess.DEVNULL)
    raw = open(out, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a

    nnode, numeq, nelem = take(np.int32, 3)
    dofs = take(np.int32, 3 * nnode).reshape(nnode, 3)
    K, g, ed, ok = [], [], [], []
    for _ in range(nelem):
        nd = int(take(np.int32, 1)[0])
        ed.append(take(np.int32, nd).copy())
        K.append(take(np.float64, nd * nd).reshape(nd, nd).copy())
        g.append(take(np.float64, nd).copy())
        ok.append(int(take(np.int32, 1)[0]))
    return dict(dofs=dofs.copy(), numeq=int(numeq), edof=np.array(ed),
                K=np.array(K), g=np.array(g), ok=np.array(ok))


def run_fast(mesh, workdir, name, **kw):
    deck = os.path.join(workdir, name)
    write_deck(deck + ".dat", mesh, **kw)
    res = subprocess.run([os.path.join(REF, "fast"), deck], capture_output=True,
                         text=True, cwd=workdir)
    pos = open(deck + ".pos").read()
    steps = []
    for blk in re.finditer(r"%RESULT\.CASE\.STEP\.NODAL\.DISPLACEMENT\n(\d+)[^\n]*\n", pos):
        n = int(blk.group(1))
        lines = pos[blk.end():].split("\n")[:n]
        u = np.array([[float(v) for v in ln.split()[1:4]] for ln in lines])
        steps.append(u)
    iters = [int(m) for m in re.findall(r"Iter:\s+(\d+)", res.stdout)]
    return np.array(steps), iters, res.stdout


def mesh_arrays(mesh):
    d = dict(coords=mesh.coords, conn=mesh.conn, mat=mesh.mat, fixed=mesh.fixed,
             nip=np.array(mesh.nip), deg=np.array(mesh.deg),
             shape=np.array(mesh.shape), form=np.array(mesh.form))
    if mesh.ext is not None:
        d.update(ext=mesh.ext, cpw=mesh.cpw, span=mesh.span)
    return d


def main():
    if not os.path.exists(os.path.join(REF, "ref_probe")):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"])
    rng = np.random.default_rng(1234)
    with tempfile.TemporaryDirectory() as wd:
        # --- element level: hex8 linear and TL (SVK), perturbed 2x2x2
        for form, E, scale in (("linear", 200e3, 1e-3), ("tl_svk", 2.6, 0.05)):
            m = M.hex8_block(2, 2, 2, E=E, nu=0.3, form=form, perturb=0.25, seed=7)
            u = rng.uniform(-scale, scale, size=(m.nnode, 3))
            r = run_probe(m, u, wd, "h8_" + form)
            np.savez_compressed(os.path.join(HERE, "elem_hex8_%s.npz" % form),
                                u=u, **mesh_arrays(m), **r)
            print("hex8", form, "K", r["K"].shape, "ok", r["ok"].min())
        # --- element level: cubic B-spline solid patch 2x2x2, linear and TL
        for form, scale in (("linear", 1e-3), ("tl_svk", 0.05)):
            m = M.bsp_block(ne=(2, 2, 2), p=3, size=(1.0, 0.8, 0.6), E=2.6, nu=0.3,
                            form=form, perturb=0.2, seed=11)
            u = rng.uniform(-scale, scale, size=(m.nnode, 3))
            r = run_probe(m, u, wd, "bsp_" + form)
            v = rng.standard_normal(r["K"].shape[1])
            np.savez_compressed(
                os.path.join(HERE, "elem_bsp3_%s.npz" % form), u=u, v=v,
                K0=r["K"][0], Kv=np.einsum("eij,j->ei", r["K"], v),
                Kdiag=np.einsum("eii->ei", r["K"]), g=r["g"], edof=r["edof"],
                dofs=r["dofs"], numeq=r["numeq"], ok=r["ok"], **mesh_arrays(m))
            print("bsp", form, "K", r["K"].shape)
        # --- whole program: configuration-1 linear static (skyline Crout)
        m = M.hex8_block(10, 10, 10, E=200e3, nu=0.3, traction=(-1.0, 0.0, 0.0))
        steps, _, _ = run_fast(m, wd, "cfg1", algo="linear", nodal_loads=m.fext)
        np.savez_compressed(os.path.join(HERE, "cfg1_linear_u.npz"),
                            u=steps[-1], fext=m.fext, **mesh_arrays(m))
        print("cfg1 u max", np.abs(steps[-1]).max())
        # --- whole program: TL SVK cantilever under NR, 5 load steps
        m = M.hex8_block(4, 4, 4, E=2.6, nu=0.3, form="tl_svk",
                         traction=(0.0, 0.0, -0.2))
        steps, iters, log = run_fast(m, wd, "nr", algo="lcm", steps=5, factor=0.2,
                                     maxit=30, tol=1e-9, nodal_loads=m.fext)
        np.savez_compressed(os.path.join(HERE, "nr_tl_hex8_4x4x4.npz"),
                            u_steps=steps, iters=np.array(iters), fext=m.fext,
                            **mesh_arrays(m))
        print("nr steps", steps.shape, "iters", iters)
        # --- whole program: cubic B-spline solid, linear static
        m = M.bsp_block(ne=(4, 3, 2), p=3, size=(2.0, 1.0, 0.5), E=1.0, nu=0.3,
                        perturb=0.1, seed=5)
        fext = np.zeros((m.nnode, 3))
        tip = np.arange(m.nnode) % m.nbas[0] == m.nbas[0] - 1
        fext[tip, 2] = -1e-3
        steps, _, _ = run_fast(m, wd, "bspl", algo="linear", nodal_loads=fext)
        np.savez_compressed(os.path.join(HERE, "bsp3_linear_u.npz"),
                            u=steps[-1], fext=fext, **mesh_arrays(m))
        print("bsp linear u max", np.abs(steps[-1]).max())


if __name__ == "__main__":
    main()
