"""Macroscopic conversions and equilibria on the host (KitBase src/Theory), used by the setup
(initial conditions, time step) of the host mirror.  The time step itself runs in libkbgpu.so."""
from __future__ import annotations

import numpy as np


def prim_conserve(prim, gamma: float) -> np.ndarray:
    """prim_conserve! (theory_macro.jl:10-33)."""
    p = np.asarray(prim, dtype=float)
    W = np.empty_like(p)
    W[0] = p[0]
    if p.size == 3:
        W[1] = p[0] * p[1]
        W[2] = 0.5 * p[0] / p[2] / (gamma - 1.0) + 0.5 * p[0] * p[1] ** 2
    elif p.size == 4:
        W[1] = p[0] * p[1]
        W[2] = p[0] * p[2]
        W[3] = 0.5 * p[0] / p[3] / (gamma - 1.0) + 0.5 * p[0] * (p[1] ** 2 + p[2] ** 2)
    elif p.size == 5:
        W[1:4] = p[0] * p[1:4]
        W[4] = 0.5 * p[0] / p[4] / (gamma - 1.0) + 0.5 * p[0] * (p[1] ** 2 + p[2] ** 2 + p[3] ** 2)
    else:
        raise ValueError("prim -> w : dimension error")
    return W


def conserve_prim(W, gamma: float) -> np.ndarray:
    """conserve_prim! (theory_macro.jl:185-206); works on [..., n] arrays."""
    W = np.asarray(W, dtype=float)
    p = np.empty_like(W)
    n = W.shape[-1]
    p[..., 0] = W[..., 0]
    kin = np.zeros(W.shape[:-1])
    for m in range(1, n - 1):
        p[..., m] = W[..., m] / W[..., 0]
        kin = kin + W[..., m] ** 2
    p[..., n - 1] = 0.5 * W[..., 0] / (gamma - 1.0) / (W[..., n - 1] - 0.5 * kin / W[..., 0])
    return p


def heat_capacity_ratio(K: float, D: int) -> float:
    """heat_capacity_ratio (theory_macro.jl:376-384)."""
    return {1: (K + 3.0) / (K + 1.0), 2: (K + 4.0) / (K + 2.0), 3: (K + 5.0) / (K + 3.0)}[D]


def sound_speed(lam: float, gamma: float) -> float:
    """sound_speed (theory_macro.jl:421)."""
    return (0.5 * gamma / lam) ** 0.5


def ref_vhs_vis(Kn: float, alpha: float, omega: float) -> float:
    """ref_vhs_vis (theory_macro.jl:447-449)."""
    return 5.0 * (alpha + 1.0) * (alpha + 2.0) * np.sqrt(np.pi) / (4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega)) * Kn


def vhs_collision_time(rho: float, lam: float, mu: float, omega: float) -> float:
    """vhs_collision_time (theory_relaxation.jl:28)."""
    return mu * 2.0 * lam ** (1.0 - omega) / rho


def maxwellian_1v(u, rho, U, lam):
    """maxwellian 1V (theory_equilibrium.jl:8)."""
    return rho * np.sqrt(lam / np.pi) * np.exp(-lam * (u - U) ** 2)


def maxwellian_2v(u, v, rho, U, V, lam):
    """maxwellian 2V (theory_equilibrium.jl:20)."""
    return rho * (lam / np.pi) * np.exp(-lam * ((u - U) ** 2 + (v - V) ** 2))


def maxwellian_3v(u, v, w, rho, U, V, W, lam):
    """maxwellian 3V (theory_equilibrium.jl:32-33)."""
    return rho * np.sqrt((lam / np.pi) ** 3) * np.exp(-lam * ((u - U) ** 2 + (v - V) ** 2 + (w - W) ** 2))
