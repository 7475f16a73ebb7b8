"""Physical spaces and frame rotations (KitBase src/Geometry).

PSpace1D arrays include the ghost layers: index k = i - 1 + ng for the Julia index i in
[1-ng, nx+ng].  PSpace2D has no ghost layers here; cell (i, j) is at (j-1, i-1) of the
(ny, nx)-shaped arrays, i.e. flat index (i-1) + nx (j-1).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class PSpace1D:
    """1D structured mesh (geo_struct.jl:10-32)."""

    x0: float
    x1: float
    nx: int
    ng: int
    x: np.ndarray
    dx: np.ndarray


def pspace1d(X0: float, X1: float, NX: int, NG: int = 0) -> PSpace1D:
    d = (X1 - X0) / NX
    i = np.arange(1 - NG, NX + NG + 1)
    return PSpace1D(X0, X1, NX, NG, X0 + (i - 0.5) * d, np.full(i.size, d))


def _unit_normal(p1, p2):
    """unit_normal (geo_general.jl:230-235), including its +1e-6 in the length."""
    d0, d1 = p2[0] - p1[0], p2[1] - p1[1]
    length = np.sqrt(d0 * d0 + d1 * d1) + 1e-6
    return np.array([-d1 / length, d0 / length])


@dataclass
class PSpace2D:
    """2D structured mesh (geo_struct.jl:47-175).

    areas[j, i, k] and n[j, i, k, :] for the edges k = 0..3 (KitBase 1..4: y-, x+, y+, x-).
    """

    x0: float
    x1: float
    nx: int
    y0: float
    y1: float
    ny: int
    x: np.ndarray
    y: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    vertices: np.ndarray
    areas: np.ndarray
    n: np.ndarray


def pspace2d(X0, X1, NX, Y0, Y1, NY) -> PSpace2D:
    dx_, dy_ = (X1 - X0) / NX, (Y1 - Y0) / NY
    I, J = np.meshgrid(np.arange(1, NX + 1), np.arange(1, NY + 1))  # (ny, nx)
    x = X0 + (I - 0.5) * dx_
    y = Y0 + (J - 0.5) * dy_
    dx = np.full(x.shape, dx_)
    dy = np.full(x.shape, dy_)
    vt = np.empty(x.shape + (4, 2))
    vt[..., 0, 0] = x - 0.5 * dx
    vt[..., 1, 0] = x + 0.5 * dx
    vt[..., 2, 0] = x + 0.5 * dx
    vt[..., 3, 0] = x - 0.5 * dx
    vt[..., 0, 1] = y - 0.5 * dy
    vt[..., 1, 1] = y - 0.5 * dy
    vt[..., 2, 1] = y + 0.5 * dy
    vt[..., 3, 1] = y + 0.5 * dy
    areas = np.empty(x.shape + (4,))
    n = np.empty(x.shape + (4, 2))
    # edge lengths and outward unit normals of every cell (geo_struct.jl:98-175), elementwise over
    # the mesh: the same operations per edge as unit_normal (_unit_normal) and the orientation test
    for k in range(4):
        p1, p2 = vt[..., k, :], vt[..., (k + 1) % 4, :]
        e0, e1 = p1[..., 0] - p2[..., 0], p1[..., 1] - p2[..., 1]
        areas[..., k] = np.sqrt(e0 * e0 + e1 * e1)
        d0, d1 = p2[..., 0] - p1[..., 0], p2[..., 1] - p1[..., 1]
        length = np.sqrt(d0 * d0 + d1 * d1) + 1e-6
        nk0, nk1 = -d1 / length, d0 / length
        mid0, mid1 = x - (p1[..., 0] + p2[..., 0]) / 2, y - (p1[..., 1] + p2[..., 1]) / 2
        flip = ~(nk0 * mid0 + nk1 * mid1 <= 0)
        n[..., k, 0] = np.where(flip, -nk0, nk0)
        n[..., k, 1] = np.where(flip, -nk1, nk1)
    return PSpace2D(X0, X1, NX, Y0, Y1, NY, x, y, dx, dy, vt, areas, n)


def local_frame(w: np.ndarray, c: float, s: float) -> np.ndarray:
    """local_frame (geo_general.jl:130-150) for 2- and 4-vectors."""
    w = np.asarray(w, dtype=float)
    if w.size == 2:
        return np.array([w[0] * c + w[1] * s, w[1] * c - w[0] * s])
    return np.array([w[0], w[1] * c + w[2] * s, w[2] * c - w[1] * s, w[3]])


def global_frame(w: np.ndarray, c: float, s: float) -> np.ndarray:
    """global_frame (geo_general.jl:6-26) for 2- and 4-vectors."""
    w = np.asarray(w, dtype=float)
    if w.size == 2:
        return np.array([w[0] * c - w[1] * s, w[0] * s + w[1] * c])
    return np.array([w[0], w[1] * c - w[2] * s, w[1] * s + w[2] * c, w[3]])


def face_geometry_2d(ps: PSpace2D, nf: int = 2):
    """Per-face normals and lengths the 2D evolve! uses (solver_evolution.jl:643-674,
    evolve_boundary.jl:71-175; 2D1F2V walls solver_evolution.jl:500-604 take the edge length
    ps.areas instead of dy / dx).

    Returns xgeo (ny, nx+1, 3) and ygeo (ny+1, nx, 3) with [n1, n2, len]:
      interior x-face (i, j), i = 2..nx: n = ps.n[i-1, j, 2], len = ps.areas[i-1, j, 2]
      interior y-face (i, j), j = 2..ny: n = ps.n[i, j-1, 3], len = ps.areas[i, j-1, 3]
      walls: left n = -ps.n[1, j, 4], len = dy[1, j]; right n = ps.n[nx, j, 2], len = dy[nx, j];
             down n = -ps.n[i, 1, 1], len = dx[i, 1]; up n = ps.n[i, ny, 3], len = dx[i, ny].
    """
    nx, ny = ps.nx, ps.ny
    xg = np.zeros((ny, nx + 1, 3))
    yg = np.zeros((ny + 1, nx, 3))
    # Julia edge numbers 1..4 are 0..3 here
    xg[:, 1:nx, 0:2] = ps.n[:, 0 : nx - 1, 1, :]
    xg[:, 1:nx, 2] = ps.areas[:, 0 : nx - 1, 1]
    yg[1:ny, :, 0:2] = ps.n[0 : ny - 1, :, 2, :]
    yg[1:ny, :, 2] = ps.areas[0 : ny - 1, :, 2]
    xg[:, 0, 0:2] = -ps.n[:, 0, 3, :]
    xg[:, 0, 2] = ps.dy[:, 0]
    xg[:, nx, 0:2] = ps.n[:, nx - 1, 1, :]
    xg[:, nx, 2] = ps.dy[:, nx - 1]
    yg[0, :, 0:2] = -ps.n[0, :, 0, :]
    yg[0, :, 2] = ps.dx[0, :]
    yg[ny, :, 0:2] = ps.n[ny - 1, :, 2, :]
    yg[ny, :, 2] = ps.dx[ny - 1, :]
    if nf == 1:
        xg[:, 0, 2] = ps.areas[:, 0, 3]
        xg[:, nx, 2] = ps.areas[:, nx - 1, 1]
        yg[0, :, 2] = ps.areas[0, :, 0]
        yg[ny, :, 2] = ps.areas[ny - 1, :, 2]
    return np.ascontiguousarray(xg), np.ascontiguousarray(yg)
