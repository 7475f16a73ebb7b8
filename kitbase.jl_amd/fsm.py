"""The fast spectral method's set-up for the 1D1F3V collision :fsm: KitBase's Gauss-Legendre rule
`lgwt` (math_general.jl:58-98), `hs_boltz_kn` (theory_fsm.jl:16-17), the kernel modes of
`kernel_mode` (theory_fsm.jl:24-90) and `fsm_kernel` (theory_fsm.jl:196-215), which
`gas_property(vs; ..., collision = "fsm")` (struct_property.jl:516-567) stores as Gas.fsm.

Arrays are [direction][w][v][u] with u fastest (KitBase's column-major phi[i, j, k, idx]); the
collision operator itself runs on the device (libkbgpu, kbg_fsm.hip).
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass

import numpy as np


def lgwt(N: int, a: float, b: float) -> tuple[np.ndarray, np.ndarray]:
    """N Gauss-Legendre nodes and weights on [a, b], KitBase's damped Newton iteration (its
    derivative carries N + 1 for N, undone in the weights); nodes in descending order."""
    N1, N2 = N, N + 1
    i = np.arange(1, N1 + 1)
    y = np.cos((2.0 * (i - 1.0) + 1.0) * 4.0 * math.atan(1.0) / (2.0 * (N - 1.0) + 2.0)) + \
        0.27 / N1 * np.sin(4.0 * math.atan(1.0) * (-1.0 + i * 2.0 / (N1 - 1.0)) * (N - 1.0) / N2)
    y0 = np.full(N1, 2.0)
    Lp = np.zeros(N1)
    L = np.zeros((N1, N2))
    while np.max(np.abs(y - y0)) > 0.0000000000001:
        L[:, 0] = 1.0
        L[:, 1] = y
        for k in range(2, N1 + 1):
            L[:, k] = ((2.0 * k - 1.0) * y * L[:, k - 1] - (k - 1) * L[:, k - 2]) / k
        Lp = N2 * (L[:, N1 - 1] - y * L[:, N2 - 1]) / (1.0 - y ** 2)
        y0 = y
        y = y0 - L[:, N2 - 1] / Lp
    x = (a * (1.0 - y) + b * (1.0 + y)) / 2.0
    w = N2 ** 2 * (b - a) / ((1.0 - y ** 2) * Lp ** 2) / N1 ** 2
    return x, w


def hs_boltz_kn(mu_ref: float, alpha: float) -> float:
    """Effective Knudsen number of the hard-sphere fast spectral method (Gamma(2) = 1)."""
    return 64 * math.sqrt(2.0) ** alpha / 5.0 * math.gamma((alpha + 3) / 2) * math.gamma(2.0) * math.sqrt(math.pi) * mu_ref


def kernel_mode(M: int, umax: float, vmax: float, wmax: float, du: float, dv: float, dw: float, unum: int, vnum: int,
                wnum: int, alpha: float, quad_num: int = 64):
    """phi, psi [M(M-1)][wnum][vnum][unum] and phipsi [wnum][vnum][unum]: the kernel modes of the
    variant with grid spacings (the one fsm_kernel calls)."""
    from scipy.special import j1  # besselj(1, .)

    supp = math.sqrt(2.0) * 2.0 * max(umax, vmax, wmax) / (3.0 + math.sqrt(2.0))

    def freq(n, d):
        stop = (n // 2 - 1) * 2.0 * math.pi / n / d
        return -math.pi / d + np.arange(n) * ((stop - (-math.pi / d)) / (n - 1))

    fx = freq(unum, du)[None, None, :]
    fy = freq(vnum, dv)[None, :, None]
    fz = freq(wnum, dw)[:, None, None]
    absc, gw = lgwt(quad_num, 0.0, supp)
    wq = 2.0 * gw * absc ** alpha
    nd = M * (M - 1)
    phi = np.zeros((nd, wnum, vnum, unum))
    psi = np.zeros((nd, wnum, vnum, unum))
    f2 = fx ** 2 + fy ** 2 + fz ** 2
    for loop in range(1, M):
        theta = math.pi / M * loop
        for loop2 in range(1, M + 1):
            theta2 = math.pi / M * loop2
            idx = (loop - 1) * M + loop2 - 1
            s = fx * math.sin(theta) * math.cos(theta2) + fy * math.sin(theta) * math.sin(theta2) + fz * math.cos(theta)
            phi[idx] = np.tensordot(np.cos(s[..., None] * absc), wq, axes=([3], [0])) * math.sin(theta)
            r2 = f2 - s ** 2
            r = np.sqrt(np.where(r2 > 0.0, r2, 1.0))
            psi[idx] = np.where(r2 <= 0.0, math.pi * supp ** 2, 2.0 * math.pi * supp * j1(supp * r) / r)
    phipsi = np.einsum("dijk,dijk->ijk", phi, psi)
    return phi, psi, phipsi


@dataclass
class FsmKernel:
    """Gas.fsm = (Kn, nm, phi, psi, chi) of KitBase."""

    Kn: float
    nm: int
    phi: np.ndarray
    psi: np.ndarray
    phipsi: np.ndarray


@functools.lru_cache(maxsize=16)
def _modes(nm: int, u1: float, v1: float, w1: float, du: float, dv: float, dw: float, nu: int, nv: int, nw: int,
           alpha: float):
    """kernel_mode of one grid, computed once per process (0.6 s for 24^3 at M = 5): a pure function
    of the grid, so solve set-ups on the same grid (a serving loop's requests) share it read-only."""
    out = kernel_mode(nm, u1, v1, w1, du, dv, dw, nu, nv, nw, alpha)
    for a in out:
        a.flags.writeable = False
    return out


def fsm_kernel(vs, mu: float, nm: int = 5, alpha: float = 1.0) -> FsmKernel:
    """fsm_kernel(vs, mu, nm, alpha) for a uniform 3V grid (phase.VSpace3D)."""
    nu, nv, nw = vs.shape
    du, dv, dw = vs.spacing
    if nu * nv * nw != vs.nv or min(nu, nv, nw) < 2:
        raise ValueError("the fast spectral method needs a uniform 3V grid (vMeshType rectangle)")
    phi, psi, phipsi = _modes(int(nm), float(vs.u1), float(vs.v1), float(vs.w1), float(du), float(dv), float(dw),
                              int(nu), int(nv), int(nw), float(alpha))
    return FsmKernel(hs_boltz_kn(mu, alpha), nm, phi, psi, phipsi)
