"""Reader for the part of FAST's neutral input file (``.dat``) that this path
runs, and writer of its nodal-displacement result section (``.pos``).

Supported: solid 8-node and 20-node bricks (``ELEMENT.BRICK8[TL]``,
``ELEMENT.BRICK20[TL]``) and B-spline solids of degree up to 7 per direction (``ELEMENT.BSPLINE.SOLID`` /
``...SOLIDTL``) with isotropic elasticity (``MATERIAL.ISOTROPIC`` through
``SECTION.HOMOGENEOUS.ISOTROPIC.3D``), supports, the solver node order,
uniform face loads on bricks and on B-spline patch faces in global or local
axes (read into surface groups that the device integrates), body forces,
nodal forces, prescribed displacements (the penalty method), linear static
and load-control Newton analyses, and the stress output of the result file.
Any other keyword that changes the analysis raises ``NotImplementedError``
naming it; keywords that only steer printing are skipped.

Formats follow the reference's readers (main/input.cpp:70-330 keyword table):
NODE.COORD / NODE.SUPPORT / CONTROL.POINT.WEIGHT / NODE.SOLVER.ORDER
(node/node.cpp:60-280), MATERIAL.ISOTROPIC (mat/material.cpp:824-831),
INTEGRATION.ORDER (elc/element.cpp:194-228), element lines
(elc/elmparam.cpp:221-245: label, section, order, then the shape's incidence,
elc/element.cpp:2043-2100 for 20-node bricks;
shp/shpbsp.cpp:57-110 for B-spline patch + spans), PATCH
(geom/patch.cpp:229-264, 1460-1484; knot vectors geom/knotvec.cpp:93-175;
control points shp/shpiga.cpp:174-215), face loads (load/load.cpp:1440-1480),
body forces (load/load.cpp:944-966).  Equation numbering follows the solver
node order (ctrl/ctrl.cpp:2004-2104, node/node.cpp:248-285).
"""
from __future__ import annotations

import dataclasses
import re
from typing import Dict, List, Optional

import numpy as np

from .mesh import HEX20_FACES, Mesh, SurfaceGroup, bezier_extraction

# keywords that only control printing / output; their data are skipped
_OUTPUT_ONLY = {
    "HEADER", "END", "MATERIAL.LABEL", "PATCH.OUTPUT.POINT", "LOAD", "LOAD.CASE",
    "HEADER.OUTPUT.PRECISION", "HEADER.PRINT.SUPPORT.REACTIONS",
    "HEADER.OUTPUT.STRESS.EXTRAPOLATION", "HEADER.PRINT.EQUIVALENT.MESH",
    "HEADER.EQUIVALENT.MESH.DIVISION", "HEADER.OUTPUT.PATCH.STRESS.SMOOTH", "NODE",
    "ELEMENT", "MATERIAL", "SECTION", "HEADER.ANALYSIS.PRINT.STEPS",
    # read below, with their data, for the runner's choices
    "HEADER.ANALYSIS.MATRIX.TYPE", "HEADER.PRINT.AVERAGE.NODAL.STRESS",
}

# faces of the reference's Brick8 in its table order and node order
# (cShapeBrick8::GetFace, shp/shpsolid.cpp:264-276): the node order fixes the
# face's natural axes r, s and with them the local load axes
_HEX8_FACES = ((0, 1, 2, 3), (3, 2, 6, 7), (1, 5, 6, 2), (0, 4, 5, 1), (0, 3, 7, 4),
               (4, 7, 6, 5))


@dataclasses.dataclass
class Deck:
    """A parsed deck: the mesh handed to the device, its loads and the node
    labels (row i of every nodal array is the node labelled labels[i]).  The
    loads are the nodal forces in equation order (``load``), the surface
    loads (``surfaces``, integrated on the device) and the body forces
    (``body``, integrated on the device); fast_amd.deck.load_vector sums
    them."""
    mesh: Mesh
    load: np.ndarray                      # nodal forces, [neq]
    labels: np.ndarray
    title: str = ""
    analysis: str = "linear"              # "linear" or "newton" (load control)
    body: Optional[np.ndarray] = None     # [nelem, 3] body forces (force / volume)
    steps: int = 1                        # Newton: MAXIMUM.STEPS, STEP.FACTOR,
    factor: float = 1.0                   # MAXIMUM.ITERATIONS, TOLERANCE
    maxit: int = 20
    tol: float = 1e-5
    elem_labels: Optional[np.ndarray] = None  # row e of every element array
    stress_extrapolation: bool = True     # HEADER.OUTPUT.STRESS.EXTRAPOLATION
    average_stress: bool = False          # HEADER.PRINT.AVERAGE.NODAL.STRESS
    matrix_type: str = "skyline"          # HEADER.ANALYSIS.MATRIX.TYPE
    surfaces: List[SurfaceGroup] = dataclasses.field(default_factory=list)
    # LOAD.CASE.PRESCRIBED.DOF: (equations, 1-based, and values) in deck order
    prescribed: Optional[tuple] = None


def _sections(text: str) -> Dict[str, List[str]]:
    """Keyword -> whitespace tokens of its data (a quoted label is one token)."""
    out: Dict[str, List[str]] = {}
    cur = None
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("%"):
            cur = s[1:].split()[0].upper()
            if cur in out and cur not in _OUTPUT_ONLY:
                raise ValueError("keyword %%%s given twice" % cur)
            out[cur] = []
            continue
        if cur is not None:
            out[cur] += re.findall(r"'[^']*'|\S+", s)
    return out


class _Toks:
    def __init__(self, toks: List[str], key: str):
        self.t, self.i, self.key = toks, 0, key

    def next(self):
        if self.i >= len(self.t):
            raise ValueError("%%%s: data ended early" % self.key)
        self.i += 1
        return self.t[self.i - 1]

    def int(self):
        return int(self.next())

    def float(self):
        return float(self.next())


def _knot_vector(tk: _Toks) -> np.ndarray:
    kind = tk.next().upper()
    n = tk.int()
    if kind in ("'GENERAL'", "'G'"):
        vals = [tk.float() for _ in range(n)]
    elif kind in ("'UNIFORM'", "'U'"):
        v0, dv = tk.float(), tk.float()
        vals = [v0 + dv * k for k in range(n)]
    else:
        raise ValueError("unknown knot vector type %s" % kind)
    mult = [tk.int() for _ in range(n)]
    return np.repeat(np.asarray(vals, dtype=np.float64), mult)


def _span_first_basis(U: np.ndarray, p: int, sid: int) -> int:
    """First basis function of the sid-th non-empty knot span
    (geom/knotvec.cpp:303-313, geom/bspbasis.cpp:59-62)."""
    cur = -1
    for k in range(1, len(U)):
        if U[k] - U[k - 1] > 1e-5:
            cur += 1
            if cur == sid:
                return (k - 1) - p
    raise ValueError("knot span %d does not exist" % (sid + 1))


def read(path: str) -> Deck:
    text = open(path, errors="replace").read()
    sec = _sections(text)
    unknown = [k for k in sec if k not in _OUTPUT_ONLY and k not in _HANDLED]
    if unknown:
        raise NotImplementedError("keywords outside this path: " + ", ".join(sorted(unknown)))
    title = " ".join(sec.get("HEADER", []))[:200]

    # nodes
    tk = _Toks(sec["NODE.COORD"], "NODE.COORD")
    n = tk.int()
    lab, xyz = [], []
    for _ in range(n):
        lab.append(tk.int())
        xyz.append([tk.float(), tk.float(), tk.float()])
    order = np.argsort(lab)
    labels = np.asarray(lab)[order]
    coords = np.asarray(xyz, dtype=np.float64)[order]
    idx = {int(l): i for i, l in enumerate(labels)}
    nnode = len(labels)

    cpw = np.ones(nnode)
    if "CONTROL.POINT.WEIGHT" in sec:
        tk = _Toks(sec["CONTROL.POINT.WEIGHT"], "CONTROL.POINT.WEIGHT")
        for _ in range(tk.int()):
            i = idx[tk.int()]
            cpw[i] = tk.float()

    fixed = np.zeros((nnode, 3), dtype=bool)
    if "NODE.SUPPORT" in sec:
        tk = _Toks(sec["NODE.SUPPORT"], "NODE.SUPPORT")
        for _ in range(tk.int()):
            i = idx[tk.int()]
            flags = [tk.int() for _ in range(6)]
            fixed[i] = np.asarray(flags[:3]) != 0

    # materials, sections, integration orders
    mats = {}
    tk = _Toks(sec["MATERIAL.ISOTROPIC"], "MATERIAL.ISOTROPIC")
    for _ in range(tk.int()):
        m = tk.int()
        mats[m] = (tk.float(), tk.float())
    secs = {}
    tk = _Toks(sec["SECTION.HOMOGENEOUS.ISOTROPIC.3D"], "SECTION.HOMOGENEOUS.ISOTROPIC.3D")
    for _ in range(tk.int()):
        s_id = tk.int()
        secs[s_id] = tk.int()
    ords = {}
    tk = _Toks(sec["INTEGRATION.ORDER"], "INTEGRATION.ORDER")
    for _ in range(tk.int()):
        o = tk.int()
        k = [tk.int() for _ in range(3)]
        [tk.int() for _ in range(3)]
        ords[o] = tuple(k)

    # elements (labels 1..nelem)
    kinds = [k for k in ("ELEMENT.BRICK8", "ELEMENT.BRICK8TL", "ELEMENT.BRICK20",
                         "ELEMENT.BRICK20TL", "ELEMENT.BSPLINE.SOLID",
                         "ELEMENT.BSPLINE.SOLIDTL") if k in sec]
    if len(kinds) != 1:
        raise NotImplementedError("one element type per deck: found %s" % kinds)
    kind = kinds[0]
    form = "tl_svk" if kind.endswith("TL") else "linear"
    tk = _Toks(sec[kind], kind)
    ne = tk.int()
    rows = []
    for _ in range(ne):
        lab_e, s_id, o_id = tk.int(), tk.int(), tk.int()
        if kind.startswith("ELEMENT.BRICK"):
            inc = [idx[tk.int()] for _ in range(20 if "BRICK20" in kind else 8)]
        else:
            inc = [tk.int() for _ in range(4)]  # patch, span r, s, t (1-based)
        rows.append((lab_e, s_id, o_id, inc))
    rows.sort(key=lambda r: r[0])
    mat = np.asarray([mats[secs[r[1]]] for r in rows], dtype=np.float64)
    nips = {ords[r[2]] for r in rows}
    if len(nips) != 1:
        raise NotImplementedError("one integration order per deck")
    nip = nips.pop()

    if kind.startswith("ELEMENT.BRICK"):
        conn = np.asarray([r[3] for r in rows], dtype=np.int32)
        mesh_kw = dict(shape="hex20" if "BRICK20" in kind else "hex8")
        ext = None
        deg = (1, 1, 1)
    else:
        conn, ext, deg, iga = _bspline_elements(sec, rows, idx)
        mesh_kw = dict(shape="bsp", cpw=cpw)

    # equation numbering in solver order: nodes sorted by their rank
    rank = np.arange(nnode)
    if "NODE.SOLVER.ORDER" in sec:
        tk = _Toks(sec["NODE.SOLVER.ORDER"], "NODE.SOLVER.ORDER")
        if tk.int() != nnode:
            raise ValueError("NODE.SOLVER.ORDER must list every node")
        rank = np.asarray([tk.int() for _ in range(nnode)])  # rank of node i (label order)
    dofmap = _number_dofs(conn, fixed, nnode, np.argsort(rank, kind="stable"))
    mesh = Mesh(coords=coords, conn=conn, nip=nip, form=form, mat=mat, fixed=fixed, deg=deg,
                ext=ext, dofmap=dofmap, **mesh_kw)

    # loads
    f = np.zeros(mesh.neq)
    body = None
    if "LOAD.CASE.NODAL.FORCE" in sec:
        tk = _Toks(sec["LOAD.CASE.NODAL.FORCE"], "LOAD.CASE.NODAL.FORCE")
        for _ in range(tk.int()):
            i = idx[tk.int()]
            v = [tk.float() for _ in range(6)]
            for d in range(3):
                q = dofmap[i, d]
                if q > 0:
                    f[q - 1] += v[d]
    surfaces = []
    if "LOAD.CASE.AREA.UNIFORM" in sec:
        if mesh.shape == "bsp":
            raise NotImplementedError("face loads on B-spline solids")
        surfaces.append(_face_group(mesh, sec["LOAD.CASE.AREA.UNIFORM"], idx))
    if "LOAD.CASE.AREA.IGA.UNIFORM" in sec:
        if mesh.shape != "bsp":
            raise NotImplementedError("IGA face loads need B-spline solids")
        elab = {r[0]: k for k, r in enumerate(rows)}
        surfaces.append(_iga_face_group(mesh, sec["LOAD.CASE.AREA.IGA.UNIFORM"], elab, iga))
    prescribed = None
    if "LOAD.CASE.PRESCRIBED.DOF" in sec:
        # cPrescDof::ReadPrescDofs (load/prescdof.cpp:35-70): node, direction
        # (1-based), value; the direction must hold an equation
        # (GetPrescDofs, prescdof.cpp:113-128)
        tk = _Toks(sec["LOAD.CASE.PRESCRIBED.DOF"], "LOAD.CASE.PRESCRIBED.DOF")
        eqs, vals = [], []
        for _ in range(tk.int()):
            lab, dr, v = tk.int(), tk.int(), tk.float()
            if not 1 <= dr <= 3:
                raise NotImplementedError("prescribed direction %d on a solid" % dr)
            eq = int(dofmap[idx[lab], dr - 1])
            if eq <= 0:
                raise ValueError("invalid prescribed dof at node %d" % lab)
            eqs.append(eq)
            vals.append(v)
        prescribed = (np.asarray(eqs, dtype=np.int32), np.asarray(vals))
    if "LOAD.CASE.BODY.FORCES" in sec:
        tk = _Toks(sec["LOAD.CASE.BODY.FORCES"], "LOAD.CASE.BODY.FORCES")
        body = np.zeros((mesh.nelem, 3))
        for _ in range(tk.int()):
            e = tk.int() - 1
            body[e] += [tk.float(), tk.float(), tk.float()]
    deck = Deck(mesh=mesh, load=f, labels=labels, title=title, body=body,
                elem_labels=np.asarray([r[0] for r in rows]), surfaces=surfaces,
                prescribed=prescribed)
    # stress output flags (ctrl/ctrl.cpp:354-385 ReadPrintAverageStress /
    # ReadStrExtFlag; extrapolation on and averages off by default)
    if "HEADER.OUTPUT.STRESS.EXTRAPOLATION" in sec:
        deck.stress_extrapolation = int(sec["HEADER.OUTPUT.STRESS.EXTRAPOLATION"][0].split()[0]) != 0
    # system matrix (ctrl/ctrl.cpp:213-238 ReadMatrixType: skyline, the
    # direct solver, unless the deck names a sparse one)
    if "HEADER.ANALYSIS.MATRIX.TYPE" in sec:
        mt = sec["HEADER.ANALYSIS.MATRIX.TYPE"][0].split()[0].strip("'")
        deck.matrix_type = mt if mt in ("skyline", "unsym_skyline", "sparse", "unsym_sparse",
                                        "eigen_unsym_sparse") else "skyline"
    if "HEADER.PRINT.AVERAGE.NODAL.STRESS" in sec:
        deck.average_stress = int(sec["HEADER.PRINT.AVERAGE.NODAL.STRESS"][0].split()[0]) != 0
    # analysis (ctrl/ctrl.cpp:148-190 ReadAlg and the HEADER.ANALYSIS readers)
    alg = sec.get("HEADER.ANALYSIS.ALGORITHM", ["'Linear'"])[0].strip("'")
    if alg in ("Linear", "linear"):
        deck.analysis = "linear"
    elif alg in ("LoadControl", "lcm", "NewtonRaphson", "snr"):
        deck.analysis = "newton"
        deck.steps = int(sec["HEADER.ANALYSIS.MAXIMUM.STEPS"][0])
        deck.factor = float(sec["HEADER.ANALYSIS.STEP.FACTOR"][0])
        deck.maxit = int(sec["HEADER.ANALYSIS.MAXIMUM.ITERATIONS"][0])
        deck.tol = float(sec["HEADER.ANALYSIS.TOLERANCE"][0])
    else:
        raise NotImplementedError("analysis algorithm %s" % alg)
    return deck


# HEADER.ANALYSIS.TARGET.ITERATIONS is read by the reference
# (cControl::ReadTargetIter, ctrl/ctrl.cpp:264-271) but only its arc-length
# controls use it (ctrl/ctrlarclen.cpp:125-281); load control ignores it, and
# the other algorithms are refused below
_HANDLED = {"HEADER.ANALYSIS.ALGORITHM", "HEADER.ANALYSIS.MAXIMUM.STEPS",
            "HEADER.ANALYSIS.TARGET.ITERATIONS",
            "HEADER.ANALYSIS.STEP.FACTOR", "HEADER.ANALYSIS.MAXIMUM.ITERATIONS",
            "HEADER.ANALYSIS.TOLERANCE",
            "NODE.COORD", "CONTROL.POINT.WEIGHT", "NODE.SUPPORT", "NODE.SOLVER.ORDER",
            "MATERIAL.ISOTROPIC", "SECTION.HOMOGENEOUS.ISOTROPIC.3D", "INTEGRATION.ORDER",
            "ELEMENT.BRICK8", "ELEMENT.BRICK8TL", "ELEMENT.BRICK20", "ELEMENT.BRICK20TL",
            "ELEMENT.BSPLINE.SOLID",
            "ELEMENT.BSPLINE.SOLIDTL", "PATCH", "LOAD.CASE.AREA.UNIFORM",
            "LOAD.CASE.AREA.IGA.UNIFORM",
            "LOAD.CASE.BODY.FORCES", "LOAD.CASE.NODAL.FORCE", "LOAD.CASE.PRESCRIBED.DOF"}


def _number_dofs(conn, fixed, nnode, node_order):
    """VecDof numbering: active, unsupported directions numbered node by node
    in solver order, x, y, z (ctrl/ctrl.cpp:2004-2104)."""
    active = np.zeros((nnode, 3), dtype=bool)
    active[np.unique(conn.ravel())] = True
    active &= ~fixed
    dofmap = np.zeros((nnode, 3), dtype=np.int32)
    q = 0
    for i in node_order:
        for d in range(3):
            if active[i, d]:
                q += 1
                dofmap[i, d] = q
    return dofmap


def _bspline_elements(sec, rows, idx):
    tk = _Toks(sec["PATCH"], "PATCH")
    pats = {}
    for _ in range(tk.int()):
        pid, ptype = tk.int(), tk.next()
        if ptype.lower() != "'bspsol'":
            raise NotImplementedError("patch type %s" % ptype)
        if tk.int() != 1:
            raise NotImplementedError("non-rational B-spline patches (the reference's "
                                      "non-rational solid basis is not evaluated correctly)")
        bases = []
        for _d in range(3):
            p = tk.int()
            bases.append((p, _knot_vector(tk)))
        nb = [len(U) - p - 1 for p, U in bases]
        cps = [idx[tk.int()] for _ in range(nb[0] * nb[1] * nb[2])]
        pats[pid] = (bases, nb, cps)
    degs = {tuple(p for p, _ in pats[r[3][0]][0]) for r in rows}
    if len(degs) != 1:
        raise NotImplementedError("one B-spline degree per deck")
    deg = degs.pop()
    conn, ext, einfo = [], [], []
    cache = {}
    for r in rows:
        pid, span = r[3][0], [s - 1 for s in r[3][1:]]
        bases, nb, cps = pats[pid]
        first, cs = [], []
        for d in range(3):
            p, U = bases[d]
            key = (pid, d)
            if key not in cache:
                cache[key] = bezier_extraction(p, U)
            cs.append(cache[key][span[d]])
            first.append(_span_first_basis(U, p, span[d]))
        q = [p + 1 for p in deg]
        c = []
        for k in range(q[2]):
            for j in range(q[1]):
                for i in range(q[0]):
                    cp = (first[0] + i) + nb[0] * ((first[1] + j) + nb[1] * (first[2] + k))
                    c.append(cps[cp])
        conn.append(c)
        ext.append(np.concatenate([x.ravel() for x in cs]))
        einfo.append((pid, first, [np.asarray(x) for x in cs]))
    return np.asarray(conn, dtype=np.int32), np.asarray(ext), deg, (pats, einfo)


# IGA faces (shp/shpiga.h:169-176 eFaceType; geom/subpat.cpp:123-190
# nSurfAlg::CreateSubPatch): the fixed parametric direction and whether the
# face lies at its start (0) or end (1); the surface's parameters are the
# other two directions in increasing order
_IGA_FACES = {"Front": (2, 0), "Back": (2, 1), "Left": (0, 0), "Right": (0, 1),
              "Above": (1, 1), "Below": (1, 0)}


def _iga_face_group(mesh, toks, elab, iga):
    """Uniform loads on faces of a B-spline solid patch (cFaceUnifLoadIGA,
    load/loadiga.cpp:305-354) as one surface group: each face is the patch's
    boundary surface (geom/subpat.cpp CreateSubPatch) over the element's
    knot spans in its two other parameters (the lower one first), with
    those parameters' extraction rows, integrated on max(K) x max(K) Gauss
    points (cElmParam::GetIntOrd, elc/elmparam.cpp:349-360)."""
    pats, einfo = iga
    tk = _Toks(toks, "LOAD.CASE.AREA.IGA.UNIFORM")
    nodes, ext, local, q, deg = [], [], [], [], None
    for _ in range(tk.int()):
        e = elab[tk.int()]
        face = tk.next().strip("'")
        local.append(tk.int())
        q.append([tk.float(), tk.float(), tk.float()])
        if face not in _IGA_FACES:
            raise ValueError("unknown face %s" % face)
        fd, end = _IGA_FACES[face]
        pid, first, cs = einfo[e]
        bases, nb, cps = pats[pid]
        d1, d2 = [d for d in range(3) if d != fd]
        p1, p2 = bases[d1][0], bases[d2][0]
        if deg is None:
            deg = (p1, p2)
        elif deg != (p1, p2):
            raise NotImplementedError("IGA face loads of different degrees in one deck")
        fixed = 0 if end == 0 else nb[fd] - 1
        fn = []
        for j in range(p2 + 1):      # first surface parameter fastest
            for i in range(p1 + 1):
                ijk = [0, 0, 0]
                ijk[fd], ijk[d1], ijk[d2] = fixed, first[d1] + i, first[d2] + j
                fn.append(cps[ijk[0] + nb[0] * (ijk[1] + nb[1] * ijk[2])])
        nodes.append(fn)
        ext.append(np.concatenate([np.ravel(cs[d1]), np.ravel(cs[d2])]))
    return SurfaceGroup(deg=deg, nip=max(mesh.nip), nodes=np.asarray(nodes), q=np.asarray(q),
                        ext=np.asarray(ext), local=np.asarray(local))


def _face_group(mesh, toks, idx):
    """Uniform face loads on bricks (cFaceUnifLoad, load/load.cpp:1440-1498)
    as one surface group: the face is the first of the brick's faces in
    table order whose corners hold the three given nodes (cShape::
    GetShapeFace, shp/shape.cpp:461-489).  8-node bricks: its nodes n0 n1 n2
    n3 given as n0 n1 n3 n2 (the bilinear surface, first parameter fastest);
    20-node bricks: the face's eight nodes in cShapeBrick20::GetFace order
    (the 8-node serendipity quadrilateral).  On the element's Gauss order per
    direction (cElemLoad::ExtForce, load.cpp:618-691)."""
    tk = _Toks(toks, "LOAD.CASE.AREA.UNIFORM")
    nodes, local, q = [], [], []
    hex20 = mesh.shape == "hex20"
    for _ in range(tk.int()):
        e = tk.int() - 1
        corners = {idx[tk.int()] for _ in range(3)}
        local.append(tk.int())
        q.append([tk.float(), tk.float(), tk.float()])
        cn = mesh.conn[e]
        if hex20:
            face = [fc for fc in HEX20_FACES
                    if corners <= {int(cn[fc[k]]) for k in (0, 2, 4, 6)}]
        else:
            face = [fc for fc in _HEX8_FACES if corners <= {int(cn[a]) for a in fc}]
        if not face:
            raise ValueError("face of element %d not found" % (e + 1))
        if hex20:
            nodes.append([int(cn[a]) for a in face[0]])
        else:
            n0, n1, n2, n3 = (int(cn[a]) for a in face[0])
            nodes.append([n0, n1, n3, n2])
    if hex20:
        return SurfaceGroup(deg=(2, 2), nip=mesh.nip[0], nodes=np.asarray(nodes),
                            q=np.asarray(q), local=np.asarray(local), shape="q8")
    return SurfaceGroup(deg=(1, 1), nip=mesh.nip[0], nodes=np.asarray(nodes), q=np.asarray(q),
                        local=np.asarray(local))


STRESS_LABELS = ("STRESS_XX", "STRESS_YY", "STRESS_ZZ", "STRESS_XY", "STRESS_XZ", "STRESS_YZ")


def _fmt(v, digits):
    return "%+.*e" % (digits, v)


def write_pos(path: str, labels: np.ndarray, u_nodal: np.ndarray, factor: float = 1.0,
              digits: int = 16, elem_labels=None, sig_ip=None, sig_node=None, sig_avg=None):
    """One step in the result layout of the reference's .pos file (see
    write_pos_steps)."""
    write_pos_steps(path, labels, [dict(factor=factor, u=u_nodal, sig_ip=sig_ip,
                                        sig_node=sig_node, sig_avg=sig_avg)],
                    elem_labels=elem_labels, digits=digits)


def write_pos_steps(path: str, labels: np.ndarray, steps, elem_labels=None, digits: int = 16):
    """Result steps in the layout of the reference's .pos file: the case
    header listing the steps (ctrl/ctrl.cpp:1117-1175), then per step its
    load factor, the nodal displacements (label and six components per node)
    and, when given, the stresses of cControl::PrintIntPntStress
    (ctrl.cpp:2189-2259: per element its label, point count and one line per
    point), PrintNodalStress (2264-2330: per element its label and one line
    per element node) and PrintAvrNodalStress (2397-2470: one line per
    node).  steps: dicts with factor, u [nnode, 3] and optional sig_ip,
    sig_node, sig_avg; stress components in Voigt order xx yy zz xy xz yz."""
    head = "6\n" + "  ".join("'%s'" % k for k in STRESS_LABELS) + "  \n\n"
    with open(path, "w") as out:
        out.write("%%RESULT\n1\n1  'result1'\n\n%%RESULT.CASE\n1   %d\n" % len(steps))
        for k in range(len(steps)):
            out.write("%d   'step%d'\n" % (k + 1, k + 1))
        out.write("\n")
        for k, st in enumerate(steps):
            out.write("%%RESULT.CASE.STEP\n%d\n\n%%RESULT.CASE.STEP.FACTOR\n%.*e\n\n"
                      % (k + 1, digits, st["factor"]))
            out.write("%%RESULT.CASE.STEP.NODAL.DISPLACEMENT\n%d  'displacement'\n" % len(labels))
            for lab, u in zip(labels, st["u"]):
                out.write("%-5d " % lab + "  ".join(_fmt(v, digits) for v in
                                                      list(u) + [0.0, 0.0, 0.0]) + "  \n")
            out.write("\n")
            sig_ip, sig_node, sig_avg = st.get("sig_ip"), st.get("sig_node"), st.get("sig_avg")
            el = elem_labels
            if el is None and (sig_ip is not None or sig_node is not None):
                el = np.arange(1, len(sig_ip if sig_ip is not None else sig_node) + 1)
            if sig_ip is not None:
                out.write("%RESULT.CASE.STEP.ELEMENT.GAUSS.SCALAR\n" + head)
                out.write("%%RESULT.CASE.STEP.ELEMENT.GAUSS.SCALAR.DATA\n%d\n" % len(sig_ip))
                for lab, sg in zip(el, sig_ip):
                    out.write("%d\n%d\n" % (lab, len(sg)))
                    for row in sg:
                        out.write(" ".join(_fmt(v, digits) for v in row) + " \n")
                out.write("\n")
            if sig_node is not None:
                out.write("%RESULT.CASE.STEP.ELEMENT.NODAL.SCALAR\n" + head)
                out.write("%%RESULT.CASE.STEP.ELEMENT.NODAL.SCALAR.DATA\n%d\n" % len(sig_node))
                for lab, sg in zip(el, sig_node):
                    out.write("%d\n" % lab)
                    for row in sg:
                        out.write(" ".join(_fmt(v, digits) for v in row) + " \n")
                out.write("\n")
            if sig_avg is not None:
                out.write("%RESULT.CASE.STEP.NODAL.SCALAR\n" + head)
                out.write("%%RESULT.CASE.STEP.NODAL.SCALAR.DATA\n%d\n" % len(sig_avg))
                for lab, row in zip(labels, sig_avg):
                    out.write("%-5d " % lab + "  ".join(_fmt(v, digits) for v in row) + "  \n")
                out.write("\n")
        out.write("%END\n")


def read_pos(path: str):
    """Every result step of a .pos file: a list of dicts with the step's
    ``factor``, nodal ``labels`` and displacements ``u`` (x, y, z), and where
    present ``elem_labels`` with ``sig_ip`` (list of [npt, 6] arrays),
    ``sig_node`` (list of [nen, 6] arrays) and ``sig_avg`` ([nnode, 6])."""
    lines = [l.strip() for l in open(path).read().splitlines()]
    steps = []
    cur = None
    i = 0

    def vals(l):
        return [float(x) for x in l.split()]

    while i < len(lines):
        l = lines[i]
        if l == "%RESULT.CASE.STEP.FACTOR":
            cur = {"factor": float(lines[i + 1].split()[0])}
            steps.append(cur)
            i += 2
            continue
        if l == "%RESULT.CASE.STEP.NODAL.DISPLACEMENT" and cur is not None:
            n = int(lines[i + 1].split()[0])
            rows = [lines[i + 2 + k].split() for k in range(n)]
            cur["labels"] = np.asarray([int(r[0]) for r in rows])
            cur["u"] = np.asarray([[float(x) for x in r[1:4]] for r in rows])
            i += 2 + n
            continue
        if l == "%RESULT.CASE.STEP.ELEMENT.GAUSS.SCALAR.DATA" and cur is not None:
            ne = int(lines[i + 1])
            j = i + 2
            labs, sig = [], []
            for _ in range(ne):
                labs.append(int(lines[j]))
                npt = int(lines[j + 1])
                sig.append(np.asarray([vals(lines[j + 2 + k]) for k in range(npt)]))
                j += 2 + npt
            cur["elem_labels"] = np.asarray(labs)
            cur["sig_ip"] = sig
            i = j
            continue
        if l == "%RESULT.CASE.STEP.ELEMENT.NODAL.SCALAR.DATA" and cur is not None:
            ne = int(lines[i + 1])
            j = i + 2
            labs, sig = [], []
            for _ in range(ne):
                labs.append(int(lines[j]))
                j += 1
                rows = []
                while j < len(lines) and lines[j] and (lines[j][0] in "+-"):
                    rows.append(vals(lines[j]))
                    j += 1
                sig.append(np.asarray(rows))
            cur["elem_labels"] = np.asarray(labs)
            cur["sig_node"] = sig
            i = j
            continue
        if l == "%RESULT.CASE.STEP.NODAL.SCALAR.DATA" and cur is not None:
            n = int(lines[i + 1])
            rows = [vals(lines[i + 2 + k])[1:] for k in range(n)]
            cur["sig_avg"] = np.asarray(rows)
            i += 2 + n
            continue
        i += 1
    return steps


def read_pos_displacements(path: str, step: int = 0):
    """Labels and displacements (x, y, z) of one result step of a .pos file
    (the first by default; -1 the last)."""
    st = read_pos(path)[step]
    return st["labels"], st["u"]
