"""Unstructured triangle meshes for the unstructured DVM (solver_evolution.jl:872-1035,
solver_update.jl:359-471): the fields of KitBase's UnstructPSpace (which KitBase fills from a mesh
file through FiniteMesh), built here for a triangulated rectangle, and the geometry init_fvm derives
from them (solver_init.jl:387-640).

Indices are 0-based here (KitBase's are 1-based); -1 marks "no cell" in faceCells and
cellNeighbors, as in KitBase.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class UnstructPSpace:
    points: np.ndarray         # [np, 2]
    cellid: np.ndarray         # [nc, 3] point indices
    facePoints: np.ndarray     # [nf, 2]
    faceCells: np.ndarray      # [nf, 2] (-1 outside)
    cellFaces: np.ndarray      # [nc, 3]
    cellNeighbors: np.ndarray  # [nc, 3] (-1 outside)
    cellCenter: np.ndarray     # [nc, 2]
    faceCenter: np.ndarray     # [nf, 2]
    cellArea: np.ndarray       # [nc]
    cellType: np.ndarray       # [nc]: 0 inner, 2 next to a wall (both stepped), 3 mean of two neighbours
    x0: float = 0.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0

    @property
    def ncell(self) -> int:
        return self.cellid.shape[0]

    @property
    def nface(self) -> int:
        return self.facePoints.shape[0]


def tri_rectangle(x0: float, x1: float, nx: int, y0: float, y1: float, ny: int) -> UnstructPSpace:
    """nx x ny rectangles on [x0, x1] x [y0, y1], each cut into two triangles along a diagonal that
    alternates between neighbouring rectangles; cells with a boundary face are cellType 2 (walls)."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    pts = np.array([[x, y] for y in ys for x in xs])
    pid = lambda i, j: j * (nx + 1) + i
    cells = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = pid(i, j), pid(i + 1, j), pid(i + 1, j + 1), pid(i, j + 1)
            if (i + j) % 2 == 0:
                cells += [(a, b, c), (a, c, d)]
            else:
                cells += [(a, b, d), (b, c, d)]
    cellid = np.array(cells, dtype=np.int64)
    edges: dict[tuple[int, int], int] = {}
    face_pts, face_cells = [], []
    cell_faces = np.zeros_like(cellid)
    for c, tri in enumerate(cellid):
        for k in range(3):
            p, q = int(tri[k]), int(tri[(k + 1) % 3])
            key = (min(p, q), max(p, q))
            if key not in edges:
                edges[key] = len(face_pts)
                face_pts.append((p, q))
                face_cells.append([c, -1])
            else:
                face_cells[edges[key]][1] = c
            cell_faces[c, k] = edges[key]
    facePoints = np.array(face_pts, dtype=np.int64)
    faceCells = np.array(face_cells, dtype=np.int64)
    nb = -np.ones_like(cellid)
    for c in range(len(cellid)):
        for k in range(3):
            f = cell_faces[c, k]
            a, b = faceCells[f]
            nb[c, k] = b if a == c else a
    cc = pts[cellid].mean(axis=1)
    fc = pts[facePoints].mean(axis=1)
    p0, p1, p2 = pts[cellid[:, 0]], pts[cellid[:, 1]], pts[cellid[:, 2]]
    area = 0.5 * np.abs((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1]))
    ctype = np.where((nb == -1).any(axis=1), 2, 0).astype(np.int64)
    return UnstructPSpace(pts, cellid, facePoints, faceCells, cell_faces, nb, cc, fc, area, ctype, x0, x1, y0, y1)


def unit_normal(p1: np.ndarray, p2: np.ndarray) -> np.ndarray:
    """geo_general.jl:230-235, with its 1e-6 added to the length (so |n| is slightly below 1)."""
    d = p2 - p1
    return np.array([-d[1], d[0]]) / (np.linalg.norm(d) + 1e-6)


def face_geometry(ps: UnstructPSpace) -> tuple[np.ndarray, np.ndarray]:
    """Face normals and lengths as init_fvm sets them (solver_init.jl:447-466): the normal points
    from faceCells[0] to faceCells[1], or from the face into its cell on a boundary face."""
    n = np.zeros((ps.nface, 2))
    ln = np.zeros(ps.nface)
    for f in range(ps.nface):
        a, b = ps.points[ps.facePoints[f, 0]], ps.points[ps.facePoints[f, 1]]
        ln[f] = np.linalg.norm(a - b)
        nf = unit_normal(a, b)
        c1, c2 = ps.faceCells[f]
        if c1 != -1 and c2 != -1:
            n0 = ps.cellCenter[c2] - ps.cellCenter[c1]
        else:
            n0 = ps.cellCenter[c1 if c1 != -1 else c2] - ps.faceCenter[f]
        if np.dot(nf, n0) < 0:
            nf = -nf
        n[f] = nf
    return n, ln


def cell_dirc(ps: UnstructPSpace, fn: np.ndarray) -> np.ndarray:
    """sign(dot(ctr[i].n[j], face[cellFaces[i, j]].n)) of update! (solver_update.jl:375), with the
    cell's outward face normals of init_fvm (solver_init.jl:474-492)."""
    d = np.zeros((ps.ncell, 3))
    for c in range(ps.ncell):
        for j in range(3):
            f = ps.cellFaces[c, j]
            nc = unit_normal(ps.points[ps.facePoints[f, 0]], ps.points[ps.facePoints[f, 1]])
            if np.dot(ps.faceCenter[f] - ps.cellCenter[c], nc) < 0:
                nc = -nc
            d[c, j] = np.sign(np.dot(nc, fn[f]))
    return d


def face_kinds(ps: UnstructPSpace) -> np.ndarray:
    """0 interior (KFVS), 1 Maxwell wall (its cell is cellType 2), 2 boundary face whose stored flux
    stays (solver_evolution.jl:883-935)."""
    k = np.zeros(ps.nface, dtype=np.int64)
    for f in range(ps.nface):
        c1, c2 = ps.faceCells[f]
        if c1 == -1 or c2 == -1:
            c = c1 if c1 != -1 else c2
            k[f] = 1 if ps.cellType[c] == 2 else 2
    return k


def point_distance(p: np.ndarray, p1: np.ndarray, p2: np.ndarray) -> float:
    """geo_general.jl:260-267 (its 1e-6 inside the square root kept)."""
    (x0, y0), (x1, y1), (x2, y2) = p, p1, p2
    return abs((x2 - x1) * (y1 - y0) - (x1 - x0) * (y2 - y1)) / np.sqrt((x2 - x1) ** 2 + (y2 - y1) ** 2 + 1e-6)


def cell_dx(ps: UnstructPSpace) -> np.ndarray:
    """ctr[i].dx of init_fvm (solver_init.jl:494-514): the cell centre's distance to its three edges."""
    d = np.zeros((ps.ncell, 3))
    for c in range(ps.ncell):
        a, b, e = (ps.points[k] for k in ps.cellid[c])
        d[c] = [point_distance(ps.cellCenter[c], a, b), point_distance(ps.cellCenter[c], b, e),
                point_distance(ps.cellCenter[c], e, a)]
    return d


@dataclass
class UnstructState:
    """ctr (w, prim, f or h and b per cell) and face (fw, ff) arrays of an unstructured solve."""
    w: np.ndarray     # [nc, 4]
    prim: np.ndarray  # [nc, 4]
    f: np.ndarray     # [nc, nf, nv]
    fw: np.ndarray    # [nface, 4]
    ff: np.ndarray    # [nface, nf, nv]


def init_fvm_us(ks, ps: UnstructPSpace) -> UnstructState:
    """init_fvm for an UnstructPSpace (solver_init.jl:387-640): every cell from KS.ib at its centre,
    zero face fluxes."""
    nf, nv = ks.nf, ks.vs.nv
    nc = ps.ncell
    w = np.zeros((nc, 4))
    prim = np.zeros((nc, 4))
    f = np.zeros((nc, nf, nv))
    from .theory import conserve_prim
    for c in range(nc):
        x, y = ps.cellCenter[c]
        w[c] = ks.ib.fw(x, y)
        prim[c] = conserve_prim(w[c], ks.gas.gamma)
        g = ks.ib.ff(x, y)
        if nf == 1:
            f[c, 0] = g
        else:
            f[c, 0], f[c, 1] = g
    return UnstructState(w, prim, f, np.zeros((ps.nface, 4)), np.zeros((ps.nface, nf, nv)))


def wall_prims(ks, ps: UnstructPSpace) -> np.ndarray:
    """Global-frame wall primitive vector of every boundary face: KS.ib.bc at the face centre
    (solver_evolution.jl:906-910); the lid is the top edge, chosen by index."""
    out = np.zeros((ps.nface, 4))
    for f in range(ps.nface):
        if -1 in ps.faceCells[f]:
            x, y = ps.faceCenter[f]
            top = bool(np.all(ps.points[ps.facePoints[f], 1] == ps.y1))
            out[f] = ks.ib.bc(x, y, top)
    return out


def timestep_us(ks, ps: UnstructPSpace, st: UnstructState, simTime: float = 0.0) -> float:
    """timestep for unstructured control volumes (solver_main.jl:400-426)."""
    from .theory import sound_speed
    dx = cell_dx(ps)
    sos = sound_speed(st.prim[:, -1], ks.gas.gamma)
    umax, vmax = ks.vs.u1 + sos, ks.vs.v1 + sos
    tmax = float(np.max(np.sqrt(umax ** 2 + vmax ** 2) / dx.min(axis=1)))
    dt = ks.set.cfl / tmax
    return dt if dt < ks.set.maxTime - simTime else ks.set.maxTime - simTime


def make_udesc(ks, ps: UnstructPSpace, precision: str = "fp64", fwall: np.ndarray | None = None,
               fkind: np.ndarray | None = None):
    """kbg_udesc for an unstructured solve; returns the struct and the arrays it points into (keep
    them alive through kbg_usolve_create)."""
    from . import _lib

    keep: list = []

    def d(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _lib.dptr(a)

    def i(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return _lib.iptr(a)

    u = _lib.KbgUDesc()
    u.nf, u.nv, u.ncell, u.nface = ks.nf, ks.vs.nv, ps.ncell, ps.nface
    u.precision = _lib.KBG_FP64 if precision == "fp64" else _lib.KBG_FP32
    u.collision = _lib.COLL[ks.set.collision]
    u.interp_order = ks.set.interpOrder
    g = ks.gas
    u.K, u.gamma, u.mu_ref, u.omega, u.Pr = g.K, g.gamma, g.mu_ref, g.omega, g.Pr
    u.u, u.v, u.weights = d(ks.vs.u.ravel()), d(ks.vs.v.ravel()), d(ks.vs.weights.ravel())
    fn, fl = face_geometry(ps)
    u.face_cells = i(ps.faceCells)
    u.face_kind = i(face_kinds(ps) if fkind is None else fkind)
    u.face_geo = d(np.column_stack([fn, fl]))
    u.face_wall = d(wall_prims(ks, ps) if fwall is None else fwall)
    u.cell_faces = i(ps.cellFaces)
    u.cell_dirc = d(cell_dirc(ps, fn))
    u.cell_area = d(ps.cellArea)
    u.cell_type = i(ps.cellType)
    u.cell_nb = i(ps.cellNeighbors)
    u.cell_dmin = d(cell_dx(ps).min(axis=1))
    return u, keep


class DeviceSolveUS:
    """An unstructured triangle solve resident in HBM (kbg_usolve_create): the device counterpart of
    the ctr / face arrays init_fvm builds for an UnstructPSpace (solver_init.jl:387-640), stepped by
    evolve! (solver_evolution.jl:872-1035) and update! (solver_update.jl:359-471)."""

    def __init__(self, ks, ps: UnstructPSpace, st: UnstructState, precision: str = "fp64", ctx=None,
                 fwall: np.ndarray | None = None, fkind: np.ndarray | None = None):
        import ctypes as C

        from . import _lib
        from .solver import default_context

        self.ks, self.ps = ks, ps
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.precision = precision
        desc, keep = make_udesc(ks, ps, precision, fwall, fkind)
        h = C.c_void_p()
        _lib.check(self.lib.kbg_usolve_create(self.ctx.handle, C.byref(desc), C.byref(h)))
        self.handle = h
        self.upload(st, faces=True)

    def upload(self, st: UnstructState, faces: bool = False):
        from . import _lib

        _lib.check(self.lib.kbg_upload(self.handle, _lib.dptr(st.w), _lib.dptr(st.prim), _lib.dptr(st.f)))
        if faces:
            _lib.check(self.lib.kbg_upload_faces(self.handle, _lib.dptr(st.fw), _lib.dptr(st.ff)))

    def download(self, st: UnstructState, faces: bool = False):
        from . import _lib

        _lib.check(self.lib.kbg_download(self.handle, _lib.dptr(st.w), _lib.dptr(st.prim), _lib.dptr(st.f)))
        if faces:
            _lib.check(self.lib.kbg_download_faces(self.handle, _lib.dptr(st.fw), _lib.dptr(st.ff)))

    def evolve(self, dt: float):
        from . import _lib

        _lib.check(self.lib.kbg_evolve(self.handle, dt))

    def update(self, dt: float, residual: np.ndarray | None = None):
        from . import _lib

        _lib.check(self.lib.kbg_update(self.handle, dt, _lib.dptr(residual) if residual is not None else None))

    def step(self, dt: float, nsteps: int = 1, residual: np.ndarray | None = None):
        from . import _lib

        r = None if residual is None else np.zeros(5)
        _lib.check(self.lib.kbg_step(self.handle, dt, nsteps, _lib.dptr(r)))
        if residual is not None:
            residual[...] = r[: residual.size]

    def timestep(self, simTime: float = 0.0) -> float:
        """timestep for unstructured control volumes (solver_main.jl:400-426) from the device prim."""
        import ctypes as C

        from . import _lib

        ks = self.ks
        dt = C.c_double()
        _lib.check(self.lib.kbg_timestep(self.handle, ks.set.cfl, ks.vs.u1, ks.vs.v1, simTime, ks.set.maxTime, C.byref(dt)))
        return dt.value

    def health(self, clear: bool = False) -> int:
        import ctypes as C

        from . import _lib

        flags = C.c_int(0)
        _lib.check(self.lib.kbg_solve_health(self.handle, C.byref(flags), 1 if clear else 0))
        return flags.value

    def close(self):
        if self.handle:
            self.lib.kbg_solve_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def usolve_bytes(ks, ps: UnstructPSpace, precision: str = "fp32") -> int:
    import ctypes as C

    from . import _lib

    d, _keep = make_udesc(ks, ps, precision)
    return int(_lib.load().kbg_usolve_bytes(C.byref(d)))
