"""Mesh data model and seeded generators for the element/assembly/solve path.

The data model mirrors what FAST's element plugins read from a neutral file:

* nodes (control points) with coordinates and supports
  (node/node.h:295-369, ``%NODE.COORD`` / ``%NODE.SUPPORT``);
* one element family per mesh: 8-node hexahedra (``%ELEMENT.BRICK8[TL]``,
  elc/element.cpp:1985-2041), 20-node serendipity hexahedra
  (``%ELEMENT.BRICK20[TL]``, elc/element.cpp:2043-2100) or B-spline solids
  over one patch (``%ELEMENT.BSPLINE.SOLID[TL]``, elc/elmparam.cpp:152-190);
* per-element isotropic material (E, nu) -- or (mu, lambda) for the
  neo-Hookean total-Lagrangian form, which the reference does not have;
* the equation map ``dofmap[node, dir]`` with the semantics of
  ``cControl::InitDofs`` / ``CompEqs`` (ctrl/ctrl.cpp:2004-2104): a direction is
  active when an element uses the node, supports deactivate it, and active
  directions are numbered 1.. in node order; 0 marks a fixed direction.

Node order of a hex8 follows shp/shpsolid.cpp:173-184 (t = +1 face first),
of a hex20 shp/shpsolid.cpp:301-325 (t = +1 face counterclockwise, corners
and mid-edge nodes alternating, the four mid-height nodes, the t = -1 face).
Control points of a B-spline element follow shp/shpbsp.cpp:873-902
(k(t) slowest, i(r) fastest) and patch indexing geom/patch.cpp:1221-1231.
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Sequence

import numpy as np

FORMS = {"linear": 0, "tl_svk": 1, "tl_neohooke": 2}
SHAPES = {"hex8": 0, "bsp": 1, "hex20": 2}


@dataclasses.dataclass
class Mesh:
    coords: np.ndarray                 # (nnode, 3) float64
    conn: np.ndarray                   # (nelem, nn) int32, 0-based
    shape: str                         # "hex8" | "bsp" | "hex20"
    nip: tuple                         # Gauss points per direction
    form: str                          # "linear" | "tl_svk" | "tl_neohooke"
    mat: np.ndarray                    # (nelem, 2) float64
    fixed: np.ndarray                  # (nnode, 3) bool
    deg: tuple = (1, 1, 1)
    ext: Optional[np.ndarray] = None   # (nelem, 3, p+1, p+1) float64
    cpw: Optional[np.ndarray] = None   # (nnode,) float64
    span: Optional[np.ndarray] = None  # (nelem, 3) int32 knot-span ids
    knots: Optional[list] = None       # per-direction knot vectors
    nbas: Optional[tuple] = None       # basis functions per direction
    fext: Optional[np.ndarray] = None  # (nnode, 3) reference load vector
    grid: Optional[tuple] = None       # structured element counts
    dofmap: Optional[np.ndarray] = None
    batch_nodeptr: Optional[np.ndarray] = None  # independent meshes: node ranges

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(self.conn, dtype=np.int32)
        self.mat = np.ascontiguousarray(self.mat, dtype=np.float64)
        self.fixed = np.ascontiguousarray(self.fixed, dtype=bool)
        if self.dofmap is None:
            self.dofmap = number_dofs(self.conn, self.fixed, len(self.coords))

    @property
    def nnode(self) -> int:
        return int(self.coords.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.conn.shape[0])

    @property
    def nn(self) -> int:
        return int(self.conn.shape[1])

    @property
    def neq(self) -> int:
        return int(self.dofmap.max(initial=0))

    def load_vector(self, scale: float = 1.0) -> np.ndarray:
        """Global external load vector (equation order) from nodal loads."""
        f = np.zeros(self.neq)
        if self.fext is None:
            return f
        d = self.dofmap.ravel()
        m = d > 0
        np.add.at(f, d[m] - 1, scale * self.fext.ravel()[m])
        return f

    def scatter(self, u_eq: np.ndarray) -> np.ndarray:
        """Equation vector -> nodal (nnode, 3) field (ctrl.cpp:1094-1113)."""
        out = np.zeros((self.nnode, 3))
        d = self.dofmap
        m = d > 0
        out[m] = np.asarray(u_eq)[d[m] - 1]
        return out

    def gather(self, u_nodal: np.ndarray) -> np.ndarray:
        u = np.zeros(self.neq)
        d = self.dofmap
        m = d > 0
        u[d[m] - 1] = np.asarray(u_nodal)[m]
        return u


@dataclasses.dataclass
class SurfaceGroup:
    """Uniform surface loads on faces of one degree and Gauss rule
    (``%LOAD.CASE.AREA.UNIFORM``, load/load.cpp:1440-1500, and
    ``%LOAD.CASE.AREA.IGA.UNIFORM``, load/loadiga.cpp:305-354).  Each face is
    a tensor-product surface of degree ``deg`` over (p1+1)(p2+1) nodes, the
    first surface parameter fastest; ``ext`` holds its Bezier extraction
    rows of the two parameters, (p1+1)^2 then (p2+1)^2 per face, or is None
    for brick faces (degree (1, 1), nodes n0 n1 n3 n2 of the face).
    ``shape`` "q8": faces of 20-node bricks, the 8-node serendipity
    quadrilateral (deg (2, 2), the face's nodes in cShapeBrick20::GetFace
    order).  ``local[f]`` = 1 gives the load in the face's local axes, ``q``
    is the load per unit area."""
    deg: tuple
    nip: int
    nodes: np.ndarray                  # (nface, nb) int32
    q: np.ndarray                      # (nface, 3) float64
    ext: Optional[np.ndarray] = None   # (nface, (p1+1)^2 + (p2+1)^2) float64
    local: Optional[np.ndarray] = None  # (nface,) int32
    shape: str = "tensor"              # "tensor" | "q8" (faces of 20-node bricks)

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.int32)
        self.q = np.ascontiguousarray(self.q, dtype=np.float64).reshape(-1, 3)
        if self.ext is not None:
            self.ext = np.ascontiguousarray(self.ext, dtype=np.float64).reshape(len(self.q), -1)
        if self.local is not None:
            self.local = np.ascontiguousarray(self.local, dtype=np.int32)

    @property
    def nface(self) -> int:
        return len(self.q)


def number_dofs(conn: np.ndarray, fixed: np.ndarray, nnode: int) -> np.ndarray:
    """Equation numbering of ctrl/ctrl.cpp:2004-2104 (InitDofs + CompEqs)."""
    active = np.zeros((nnode, 3), dtype=bool)
    active[np.unique(conn.ravel())] = True
    active &= ~fixed
    dofmap = np.zeros((nnode, 3), dtype=np.int32)
    flat = active.ravel()
    dofmap.ravel()[flat] = np.arange(1, int(flat.sum()) + 1, dtype=np.int32)
    return dofmap


# ---------------------------------------------------------------- hex8 ----

def hex8_block(nx: int, ny: int, nz: int, size=(1.0, 1.0, 1.0), E=1.0,
               nu=0.3, form="linear", nip=2, perturb=0.0, seed=0,
               clamp_x0=True, traction=None, mat=None) -> Mesh:
    """Structured block of nx*ny*nz 8-node bricks, nodes numbered x fastest.

    ``perturb`` moves interior nodes by U(-perturb, perturb) * h per
    direction.  ``traction`` (3-vector) is a uniform surface traction on
    the face x = size[0], integrated exactly for flat faces (q*A/4 per
    corner, what a 2x2 Gauss face rule gives on a parallelogram face).
    """
    lx, ly, lz = size
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if perturb:
        rng = np.random.default_rng(seed)
        h = np.array([lx / nx, ly / ny, lz / nz])
        i = np.tile(np.arange(nx + 1), (nz + 1) * (ny + 1))
        j = np.tile(np.repeat(np.arange(ny + 1), nx + 1), nz + 1)
        k = np.repeat(np.arange(nz + 1), (nx + 1) * (ny + 1))
        inner = ((i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz))
        d = rng.uniform(-perturb, perturb, size=coords.shape) * h
        coords[inner] += d[inner]

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    ei = ei.transpose(2, 1, 0).ravel()
    ej = ej.transpose(2, 1, 0).ravel()
    ek = ek.transpose(2, 1, 0).ravel()
    conn = np.stack([
        nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1),
        nid(ei + 1, ej + 1, ek + 1), nid(ei, ej + 1, ek + 1),
        nid(ei, ej, ek), nid(ei + 1, ej, ek),
        nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek)], axis=1).astype(np.int32)
    nnode = coords.shape[0]
    fixed = np.zeros((nnode, 3), dtype=bool)
    if clamp_x0:
        fixed[np.isclose(coords[:, 0], 0.0)] = True
    if mat is None:
        mat = np.tile([E, nu], (conn.shape[0], 1))
    fext = None
    if traction is not None:
        fext = np.zeros((nnode, 3))
        t = np.asarray(traction, dtype=np.float64)
        area = (ly / ny) * (lz / nz)
        for k in range(nz):
            for j in range(ny):
                for (jj, kk) in ((j, k), (j + 1, k), (j + 1, k + 1), (j, k + 1)):
                    fext[nid(nx, jj, kk)] += t * area / 4.0
    return Mesh(coords=coords, conn=conn, shape="hex8", nip=(nip, nip, nip),
                form=form, mat=mat, fixed=fixed, fext=fext, grid=(nx, ny, nz))


# ---------------------------------------------------------------- hex20 ---

# natural coordinates of a 20-node brick's nodes (cShapeBrick20 node order)
HEX20_NAT = np.array([
    (-1, -1, 1), (0, -1, 1), (1, -1, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1), (-1, 1, 1),
    (-1, 0, 1), (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, -1),
    (0, -1, -1), (1, -1, -1), (1, 0, -1), (1, 1, -1), (0, 1, -1), (-1, 1, -1),
    (-1, 0, -1)], dtype=np.int64)

# the eight nodes of each face, in the order of cShapeBrick20::GetFace
# (shp/shpsolid.cpp:459-470; corners at positions 0, 2, 4, 6)
HEX20_FACES = ((0, 1, 2, 3, 4, 5, 6, 7), (2, 1, 0, 8, 12, 13, 14, 9),
               (4, 3, 2, 9, 14, 15, 16, 10), (6, 5, 4, 10, 16, 17, 18, 11),
               (12, 19, 18, 17, 16, 15, 14, 13), (18, 11, 6, 7, 0, 8, 12, 19))


def hex20_block(nx: int, ny: int, nz: int, size=(1.0, 1.0, 1.0), E=1.0, nu=0.3,
                form="linear", nip=3, perturb=0.0, seed=0, clamp_x0=True,
                mat=None) -> Mesh:
    """Structured block of nx*ny*nz 20-node bricks.  Nodes: the points of the
    (2nx+1)(2ny+1)(2nz+1) half-step lattice that a serendipity brick uses
    (at most one odd lattice index), numbered x fastest.  ``perturb`` moves
    interior nodes by U(-perturb, perturb) * h / 2 per direction."""
    lx, ly, lz = size
    mx, my, mz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    keep = (i % 2) + (j % 2) + (k % 2) <= 1
    lattice_id = np.full(mx * my * mz, -1, dtype=np.int64)
    lattice_id[keep] = np.arange(int(keep.sum()))
    coords = np.stack([i[keep] * (lx / (2 * nx)), j[keep] * (ly / (2 * ny)),
                       k[keep] * (lz / (2 * nz))], axis=1).astype(np.float64)
    if perturb:
        rng = np.random.default_rng(seed)
        ii, jj, kk = i[keep], j[keep], k[keep]
        inner = (ii > 0) & (ii < mx - 1) & (jj > 0) & (jj < my - 1) & (kk > 0) & (kk < mz - 1)
        h = np.array([lx / nx, ly / ny, lz / nz]) / 2.0
        d = rng.uniform(-perturb, perturb, size=coords.shape) * h
        coords[inner] += d[inner]
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    conn = np.empty((nx * ny * nz, 20), dtype=np.int32)
    for a, (cr, cs, ct) in enumerate(HEX20_NAT):
        li = (2 * ex + 1 + cr) + mx * ((2 * ey + 1 + cs) + my * (2 * ez + 1 + ct))
        conn[:, a] = lattice_id[li]
    nnode = coords.shape[0]
    fixed = np.zeros((nnode, 3), dtype=bool)
    if clamp_x0:
        fixed[np.isclose(coords[:, 0], 0.0)] = True
    if mat is None:
        mat = np.tile([E, nu], (conn.shape[0], 1))
    return Mesh(coords=coords, conn=conn, shape="hex20", nip=(nip, nip, nip), form=form,
                mat=mat, fixed=fixed, grid=(nx, ny, nz))


# ---------------------------------------------------------------- B-spline --

def open_uniform_knots(p: int, nspan: int) -> np.ndarray:
    inner = np.arange(1, nspan) / nspan
    return np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])


def bernstein(p: int, u: float) -> np.ndarray:
    B = np.zeros(p + 1)
    B[0] = 1.0
    for j in range(1, p + 1):
        saved = 0.0
        for k in range(j):
            temp = B[k]
            B[k] = saved + (1.0 - u) * temp
            saved = u * temp
        B[j] = saved
    return B


def bezier_extraction(p: int, U: Sequence[float]) -> np.ndarray:
    """Bezier extraction operators, one (p+1)x(p+1) matrix per span.

    Algorithm of Borden, Scott, Evans & Hughes (2011), the one used by the
    reference at geom/knotvec.cpp:415-493: successive knot insertion, rows =
    local B-spline functions, columns = Bernstein polynomials of the span.
    """
    U = np.asarray(U, dtype=np.float64)
    m = len(U)
    ns = int(np.sum(np.abs(np.diff(U)) > 1e-6))
    q = p + 1
    C = np.zeros((ns, q, q))
    C[0] = np.eye(q)
    a, b, nb = p + 1, p + 2, 1
    while b < m:
        if nb < ns:
            C[nb] = np.eye(q)
        i = b
        while b < m and U[b] == U[b - 1]:
            b += 1
        mult = b - i + 1
        if mult < p:
            numer = U[b - 1] - U[a - 1]
            alphas = np.zeros(p)
            for j in range(p, mult, -1):
                alphas[j - mult - 1] = numer / (U[a + j - 1] - U[a - 1])
            r = p - mult
            for j in range(1, r + 1):
                save = r - j + 1
                s = mult + j
                for k in range(p + 1, s, -1):
                    al = alphas[k - s - 1]
                    C[nb - 1][:, k - 1] = (al * C[nb - 1][:, k - 1]
                                           + (1.0 - al) * C[nb - 1][:, k - 2])
                if b < m and nb < ns:
                    for w in range(save, j + save + 1):
                        v = p - j + 1 + (w - save)
                        C[nb][w - 1, save - 1] = C[nb - 1][v - 1, p]
        nb += 1
        if b < m:
            a = b
            b += 1
    return C


def greville(p: int, U: np.ndarray) -> np.ndarray:
    n = len(U) - p - 1
    return np.array([U[i + 1:i + p + 1].mean() for i in range(n)])


def bsp_block(ne=(32, 32, 8), p=3, size=(1.0, 1.0, 1.0), E=1.0, nu=0.3,
              form="linear", nip=None, perturb=0.05, seed=0,
              clamp_x0=True, mat=None) -> Mesh:
    """Single B-spline solid patch over a box, open uniform knots.

    Control points sit at the Greville abscissae of the box, interior ones
    moved by U(-perturb, perturb) * h_dir (h_dir = size/ne).  Elements are
    ordered with the r-span fastest; each element's control points follow the
    reference order (k slowest, i fastest).  All weights are 1 and the patch
    is rational, the case the reference evaluates correctly.
    """
    deg = tuple(p) if np.ndim(p) else (p, p, p)
    if nip is None:
        nip = max(deg) + 1
    knots = [open_uniform_knots(deg[d], ne[d]) for d in range(3)]
    nb = tuple(len(knots[d]) - deg[d] - 1 for d in range(3))
    gx = greville(deg[0], knots[0]) * size[0]
    gy = greville(deg[1], knots[1]) * size[1]
    gz = greville(deg[2], knots[2]) * size[2]
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if perturb:
        rng = np.random.default_rng(seed)
        h = np.array([size[0] / ne[0], size[1] / ne[1], size[2] / ne[2]])
        i = np.tile(np.arange(nb[0]), nb[1] * nb[2])
        j = np.tile(np.repeat(np.arange(nb[1]), nb[0]), nb[2])
        k = np.repeat(np.arange(nb[2]), nb[0] * nb[1])
        inner = ((i > 0) & (i < nb[0] - 1) & (j > 0) & (j < nb[1] - 1)
                 & (k > 0) & (k < nb[2] - 1))
        d = rng.uniform(-perturb, perturb, size=coords.shape) * h
        coords[inner] += d[inner]
    exts = [bezier_extraction(deg[d], knots[d]) for d in range(3)]
    nel = ne[0] * ne[1] * ne[2]
    span = np.zeros((nel, 3), dtype=np.int32)
    q = [d + 1 for d in deg]
    conn = np.zeros((nel, q[0] * q[1] * q[2]), dtype=np.int32)
    # packed per direction: (p_d + 1)^2 row-major each
    ext = np.zeros((nel, sum(x * x for x in q)))
    e = 0
    for st in range(ne[2]):
        for ss in range(ne[1]):
            for sr in range(ne[0]):
                span[e] = (sr, ss, st)
                # basis index of span s for open knots with simple interior
                # knots is s (bspbasis.cpp:59-62: FindSpan(s) - p)
                kk, jj, ii = np.meshgrid(st + np.arange(q[2]), ss + np.arange(q[1]),
                                         sr + np.arange(q[0]), indexing="ij")
                conn[e] = (ii + nb[0] * (jj + nb[1] * kk)).ravel()
                ext[e] = np.concatenate([exts[0][sr].ravel(), exts[1][ss].ravel(),
                                         exts[2][st].ravel()])
                e += 1
    nnode = coords.shape[0]
    fixed = np.zeros((nnode, 3), dtype=bool)
    if clamp_x0:
        fixed[np.arange(nnode) % nb[0] == 0] = True
    if mat is None:
        mat = np.tile([E, nu], (nel, 1))
    return Mesh(coords=coords, conn=conn, shape="bsp", nip=(nip, nip, nip),
                form=form, mat=mat, fixed=fixed, deg=deg, ext=ext,
                cpw=np.ones(nnode), span=span, knots=knots, nbas=nb,
                grid=tuple(ne))


# ---------------------------------------------------------------- batches --

def hex8_cube(n: int, E=1.0, nu=0.3, clamp_face=0, load_face=1, pressure=1.0) -> Mesh:
    """Unit cube of n^3 bricks; one face clamped, uniform pressure on another.

    Faces are numbered 2*axis + side (side 0: coordinate 0, side 1:
    coordinate 1).  The pressure acts against the outward normal of the
    loaded face; nodal forces are the exact q*A/4 corner shares of each face.
    """
    m = hex8_block(n, n, n, E=E, nu=nu, clamp_x0=False)
    X = m.coords
    ax, side = divmod(int(clamp_face), 2)
    m.fixed[np.isclose(X[:, ax], float(side))] = True
    lax, lside = divmod(int(load_face), 2)
    on = np.isclose(X[:, lax], float(lside))
    # corner shares: interior face nodes 4 quarters, edges 2, corners 1
    others = [d for d in range(3) if d != lax]
    cnt = np.ones(len(X))
    for d in others:
        edge = np.isclose(X[:, d], 0.0) | np.isclose(X[:, d], 1.0)
        cnt *= np.where(edge, 1.0, 2.0)
    f = np.zeros((len(X), 3))
    normal = 1.0 if lside == 1 else -1.0
    f[on, lax] = -normal * pressure * cnt[on] * (1.0 / n) ** 2 / 4.0
    m.fext = f
    m.dofmap = number_dofs(m.conn, m.fixed, m.nnode)
    return m


def union(meshes) -> Mesh:
    """Independent meshes of one element family as one block-diagonal model.

    Nodes, elements and equations of mesh k follow those of mesh k-1, so each
    mesh owns a contiguous node range (``batch_nodeptr``) and a contiguous
    equation range."""
    m0 = meshes[0]
    nodeptr = np.zeros(len(meshes) + 1, dtype=np.int64)
    nodeptr[1:] = np.cumsum([m.nnode for m in meshes])
    conn = np.concatenate([m.conn + nodeptr[k] for k, m in enumerate(meshes)]).astype(np.int32)
    fext = None
    if all(m.fext is not None for m in meshes):
        fext = np.concatenate([m.fext for m in meshes])
    out = Mesh(coords=np.concatenate([m.coords for m in meshes]), conn=conn,
               shape=m0.shape, nip=m0.nip, form=m0.form,
               mat=np.concatenate([m.mat for m in meshes]),
               fixed=np.concatenate([m.fixed for m in meshes]), deg=m0.deg, fext=fext)
    out.batch_nodeptr = nodeptr
    return out


def cube_batch(nmesh=1024, nmin=4, nmax=32, seed=0, E=1.0):
    """Batch of independent elastic cubes: edge n ~ U{nmin..nmax},
    nu ~ U(0.2, 0.45), a random clamped face and pressure 1 on a random other
    face, all from ``seed``.  Returns (union mesh, per-mesh parameters)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=nmesh)
    nus = rng.uniform(0.2, 0.45, size=nmesh)
    cf = rng.integers(0, 6, size=nmesh)
    lf = (cf + rng.integers(1, 6, size=nmesh)) % 6
    meshes = [hex8_cube(int(n), E=E, nu=float(nu), clamp_face=int(c), load_face=int(l))
              for n, nu, c, l in zip(ns, nus, cf, lf)]
    return union(meshes), {"n": ns, "nu": nus, "clamp_face": cf, "load_face": lf}
