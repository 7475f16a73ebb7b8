"""Velocity spaces and quadratures (KitBase src/Phase).

Arrays follow KitBase's layout flattened in Julia (column-major) order: for 2V and 3V spaces the
first velocity index is the fastest, so u.ravel()[i + nu*j] is the node (i, j).  Velocity spaces
here have no ghost nodes in the solver paths (KitBase's `nug` is only used by the Newton-Cotes
weights, which this module honours).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def newton_cotes(idx: int, num: int) -> float:
    """Newton-Cotes weight coefficient (phase_quadrature.jl:276-288).

    Julia's `%` keeps the sign of the dividend; idx >= 1 here so Python's `%` agrees.
    """
    if idx == 1 or idx == num:
        return 14.0 / 45.0
    if (idx - 5) % 4 == 0:
        return 28.0 / 45.0
    if (idx - 3) % 4 == 0:
        return 24.0 / 45.0
    return 64.0 / 45.0


def maxwell_quadrature(n2: int, C: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """Half-range Gauss rule for exp(-x^2) mirrored to the full line (phase_quadrature.jl:13-44).

    Golub-Welsch on the three-term recurrence coefficients a, b; the weights are then multiplied
    by exp(x^2) * C and the nodes by C.  KitBase documents n2 <= 32.
    """
    N = n2 // 2
    a = np.zeros(N)
    b = np.zeros(N)
    a[0] = 1.0 / np.sqrt(np.pi)
    a[1] = 2.0 / np.sqrt(np.pi) / (np.pi - 2.0)
    b[1] = a[0] / (a[0] + a[1]) / 2.0
    for i in range(3, N + 1):  # Julia indices
        b[i - 1] = (i - 2) + 1.0 / 2.0 - b[i - 2] - a[i - 2] ** 2
        a[i - 1] = ((i - 1) ** 2 / (4.0 * b[i - 1]) - b[i - 2] - 1.0 / 2) / a[i - 2] - a[i - 2]
    J = np.diag(a) + np.diag(np.sqrt(b[1:]), 1) + np.diag(np.sqrt(b[1:]), -1)
    v, V = np.linalg.eigh(J)
    w = V[0, :] ** 2 * np.sqrt(np.pi) / 2.0
    order = np.argsort(v, kind="stable")
    v, w = v[order], w[order]
    xis = np.concatenate([-v[::-1], v])
    weights = np.concatenate([w[::-1], w])
    weights = weights * (np.exp(xis**2) * C)
    return xis * C, weights


@dataclass
class VSpace1D:
    """1D velocity space (phase_velocity.jl:14-65)."""

    u0: float
    u1: float
    nu: int
    u: np.ndarray
    du: np.ndarray
    weights: np.ndarray

    @property
    def nv(self) -> int:
        return self.u.size


def vspace1d(U0: float, U1: float, NU: int, type: str = "rectangle", ng: int = 0) -> VSpace1D:
    d = (U1 - U0) / NU
    idx = np.arange(1 - ng, NU + ng + 1)
    if type == "algebra":
        n1 = NU + 1
        uu = np.array([U1 / (n1 - 1) ** 3 * (-n1 + 1 + 2 * (i - 1)) ** 3 for i in range(1, n1 + 1)], dtype=float)
        u = (uu[:-1] + uu[1:]) / 2
        du = uu[1:] - uu[:-1]
        return VSpace1D(U0, U1, NU, u, du, du.copy())
    u = U0 + (idx - 0.5) * d
    du = np.full(idx.size, d)
    if type == "rectangle":
        w = np.full(idx.size, d)
    elif type == "newton":
        w = np.array([newton_cotes(int(i) + ng, NU + ng * 2) * d for i in idx])
    else:
        raise ValueError("No velocity quadrature available")
    return VSpace1D(U0, U1, NU, u, du, w)


@dataclass
class VSpace2D:
    """2D velocity space (phase_velocity.jl:128-258); arrays are flattened u-fastest."""

    u0: float
    u1: float
    nu: int
    v0: float
    v1: float
    nv_: int
    u: np.ndarray
    v: np.ndarray
    weights: np.ndarray

    @property
    def nv(self) -> int:
        return self.u.size


def vspace2d(U0, U1, NU, V0, V1, NV, type: str = "rectangle") -> VSpace2D:
    if type == "maxwell":
        xu, wu = maxwell_quadrature(NU, U1 / 5.76)
        xv, wv = maxwell_quadrature(NV, V1 / 5.76)
    else:
        du, dv = (U1 - U0) / NU, (V1 - V0) / NV
        xu = U0 + (np.arange(1, NU + 1) - 0.5) * du
        xv = V0 + (np.arange(1, NV + 1) - 0.5) * dv
        if type == "rectangle":
            wu, wv = np.full(NU, du), np.full(NV, dv)
        elif type == "newton":
            wu = np.array([newton_cotes(i, NU) * du for i in range(1, NU + 1)])
            wv = np.array([newton_cotes(j, NV) * dv for j in range(1, NV + 1)])
        else:
            raise ValueError("No velocity quadrature available")
    # u[i, j] = xu[i], v[i, j] = xv[j], weights = wu[i] wv[j]; flatten with i fastest
    U = np.tile(xu, NV)
    V = np.repeat(xv, NU)
    if type in ("rectangle", "newton"):
        # KitBase multiplies the two 1D factors in the order (wu_i * du) * (wv_j * dv)
        W = np.tile(wu, NV) * np.repeat(wv, NU) if type == "newton" else np.full(NU * NV, du * dv)
    else:
        W = np.tile(wu, NV) * np.repeat(wv, NU)
    return VSpace2D(U0, U1, NU, V0, V1, NV, U, V, W)


@dataclass
class VSpace3D:
    """3D velocity space (phase_velocity.jl:341-420); arrays flattened u-fastest."""

    u0: float
    u1: float
    nu: int
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    weights: np.ndarray
    v0: float = 0.0
    v1: float = 0.0
    w0: float = 0.0
    w1: float = 0.0
    shape: tuple = (0, 0, 0)        # (nu, nv, nw): KitBase's per-axis counts
    spacing: tuple = (0.0, 0.0, 0.0)  # (du, dv, dw) of a uniform grid (du[1, 1, 1], ...)

    @property
    def nv(self) -> int:
        return self.u.size


def vspace3d(U0, U1, NU, V0, V1, NV, W0, W1, NW, type: str = "rectangle") -> VSpace3D:
    du, dv, dw = (U1 - U0) / NU, (V1 - V0) / NV, (W1 - W0) / NW
    xu = U0 + (np.arange(1, NU + 1) - 0.5) * du
    xv = V0 + (np.arange(1, NV + 1) - 0.5) * dv
    xw = W0 + (np.arange(1, NW + 1) - 0.5) * dw
    U = np.tile(xu, NV * NW)
    V = np.tile(np.repeat(xv, NU), NW)
    Wv = np.repeat(xw, NU * NV)
    if type == "rectangle":
        Wt = np.full(NU * NV * NW, du * dv * dw)
    elif type == "newton":
        cu = np.array([newton_cotes(i, NU) for i in range(1, NU + 1)])
        cv = np.array([newton_cotes(j, NV) for j in range(1, NV + 1)])
        cw = np.array([newton_cotes(k, NW) for k in range(1, NW + 1)])
        Wt = (np.tile(cu, NV * NW) * du) * np.tile(np.repeat(cv, NU), NW) * dv * np.repeat(cw, NU * NV) * dw
    else:
        raise ValueError("No velocity quadrature available")
    return VSpace3D(U0, U1, NU, U, V, Wv, Wt, V0, V1, W0, W1, (NU, NV, NW), (du, dv, dw))
