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
