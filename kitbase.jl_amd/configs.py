"""The benchmark and test workloads as KitBase config dicts (the keys of KitBase's config files,
read_dict io_read.jl:95-128).

sod_1d1f1v      Sod shock tube, 1D1F1V BGK + KFVS, 100 cells x 64 Newton-Cotes velocities on [-5, 5],
                Kn = 1e-4, omega = 0.81, CFL 0.5, van Leer, fixed boundaries, t_end = 0.2.
shock_1d2f1v    Ma = 2 normal shock, 1D2F1V Shakhov (Pr = 2/3) + UGKS, 400 cells on [-25, 25],
                101 velocities on [-8, 8], Kn = 1, 2000 steps.
cavity_2d2f2v   lid-driven cavity, 2D2F2V UGKS, 64 x 64 cells, 32 x 32 Maxwell-quadrature velocities
                (|v| <= 5), gamma = 5/3 (K = 1), lid U = 0.15, Kn in {0.075, 1}, Maxwell walls, 5000 steps.
batch_1d        512 independent shock tubes drawn from a seed: nx ~ U{64..2048}, nv ~ U{32..256},
                Kn = 10^U(-4, 1), rho, T ~ U(0.1, 1) left and right, BGK or Shakhov by coin flip,
                1D1F1V KFVS, u in [-5, 5] rectangle, t_end = 0.2.
"""
from __future__ import annotations

import numpy as np

BATCH_SEED = 20260317


def sod_1d1f1v(**over) -> dict:
    cfg = dict(matter="gas", case="sod", space="1d1f1v", flux="kfvs", collision="bgk", nSpecies=1,
               interpOrder=2, limiter="vanleer", boundary="fix", cfl=0.5, maxTime=0.2, hasForce=False,
               x0=0.0, x1=1.0, nx=100, pMeshType="uniform", nxg=0,
               umin=-5.0, umax=5.0, nu=64, vMeshType="newton", nug=0,
               knudsen=1e-4, mach=0.0, prandtl=1.0, inK=0.0, omega=0.81, alphaRef=1.0, omegaRef=0.5)
    cfg.update(over)
    return cfg


def shock_1d2f1v(**over) -> dict:
    cfg = dict(matter="gas", case="shock", space="1d2f1v", flux="ugks", collision="shakhov", nSpecies=1,
               interpOrder=2, limiter="vanleer", boundary="fix", cfl=0.95, maxTime=1e9, hasForce=False,
               x0=-25.0, x1=25.0, nx=400, pMeshType="uniform", nxg=0,
               umin=-8.0, umax=8.0, nu=101, vMeshType="rectangle", nug=0,
               knudsen=1.0, mach=2.0, prandtl=2.0 / 3.0, inK=2.0, omega=0.72, alphaRef=1.0, omegaRef=0.5)
    cfg.update(over)
    return cfg


SHOCK_STEPS = 2000


def cavity_2d2f2v(**over) -> dict:
    cfg = dict(matter="gas", case="cavity", space="2d2f2v", flux="ugks", collision="bgk", nSpecies=1,
               interpOrder=2, limiter="vanleer", boundary="maxwell", cfl=0.8, maxTime=1e9, hasForce=False,
               x0=0.0, x1=1.0, nx=64, y0=0.0, y1=1.0, ny=64, pMeshType="uniform", nxg=0, nyg=0,
               umin=-5.0, umax=5.0, nu=32, vmin=-5.0, vmax=5.0, nv=32, vMeshType="maxwell", nug=0, nvg=0,
               knudsen=0.075, mach=0.0, prandtl=2.0 / 3.0, inK=1.0, omega=0.72, alphaRef=1.0, omegaRef=0.5,
               uLid=0.15, vLid=0.0, tLid=1.0)
    cfg.update(over)
    return cfg


CAVITY_STEPS = 5000


def cavity_2d1f2v(**over) -> dict:
    """The lid-driven cavity with one distribution (2D1F2V: no internal degrees of freedom, gamma = 2),
    KFVS + BGK (KitBase's 2D1F2V Shakhov step reads an undefined b, step_1f.jl:194)."""
    cfg = cavity_2d2f2v(space="2d1f2v", flux="kfvs", inK=0.0)
    cfg.update(over)
    return cfg


def shock_1d1f3v(**over) -> dict:
    cfg = dict(matter="gas", case="shock", space="1d1f3v", flux="kfvs", collision="bgk", nSpecies=1,
               interpOrder=2, limiter="vanleer", boundary="fix", cfl=0.95, maxTime=1e9, hasForce=False,
               x0=-35.0, x1=35.0, nx=50, pMeshType="uniform", nxg=0,
               umin=-10.0, umax=10.0, nu=24, vmin=-10.0, vmax=10.0, nv=24, wmin=-10.0, wmax=10.0, nw=24,
               vMeshType="rectangle", nug=0, nvg=0, nwg=0,
               knudsen=1.0, mach=2.0, prandtl=1.0, inK=0.0, omega=0.72, alphaRef=1.0, omegaRef=0.5)
    cfg.update(over)
    return cfg


def batch_instances(n: int = 512, seed: int = BATCH_SEED) -> list[dict]:
    """Instance descriptions of the batched workload (a synthetic Sod-type Riemann problem each)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(64, 2049))
        nv = int(rng.integers(32, 257))
        kn = float(10.0 ** rng.uniform(-4.0, 1.0))
        rl, rr, tl, tr = (float(x) for x in rng.uniform(0.1, 1.0, 4))
        coll = "shakhov" if rng.integers(0, 2) == 1 else "bgk"
        out.append(dict(nx=nx, nu=nv, knudsen=kn, rhoL=rl, rhoR=rr, TL=tl, TR=tr, collision=coll))
    return out


def riemann_1d1f1v(inst: dict) -> dict:
    """KitBase config of one batched instance (1D1F1V KFVS, rectangle grid on [-5, 5], t_end 0.2);
    the left/right states are set by the caller through `primL` / `primR` (lambda = 1 / (2 T))."""
    return dict(matter="gas", case="riemann", space="1d1f1v", flux="kfvs", collision=inst["collision"], nSpecies=1,
                interpOrder=2, limiter="vanleer", boundary="fix", cfl=0.5, maxTime=0.2, hasForce=False,
                x0=0.0, x1=1.0, nx=inst["nx"], pMeshType="uniform", nxg=0,
                umin=-5.0, umax=5.0, nu=inst["nu"], vMeshType="rectangle", nug=0,
                knudsen=inst["knudsen"], mach=0.0, prandtl=2.0 / 3.0, inK=0.0, omega=0.81, alphaRef=1.0,
                omegaRef=0.5,
                primL=[inst["rhoL"], 0.0, 1.0 / (2.0 * inst["TL"])],
                primR=[inst["rhoR"], 0.0, 1.0 / (2.0 * inst["TR"])])


SERVE_SEED = 20260318


def serving_requests(n: int = 48, rate: float = 100.0, seed: int = SERVE_SEED, mix=(0.75, 0.15, 0.10),
                     cavity_cells=(32, 48, 64, 96, 128), cavity_vel=(16, 24, 32, 48), v3_cells: int = 24,
                     cavity_steps: int = 100, v3_steps: int = 40, v3_collision: str = "fsm"):
    """The serving stream: Poisson arrivals at `rate` requests/s mixing
      * 1D shock tubes of the batched workload (run to t_end = 0.2, KitBase's nt = floor(T/dt) + 1),
      * lid-driven cavities (2D2F2V UGKS) at cavity_cells^2 cells x cavity_vel^2 velocities
        (Maxwell quadrature up to 32 per axis, a rectangle grid above),
      * 1D1F3V shock structures (24^3 velocities, Rankine-Hugoniot states, the fast spectral
        Boltzmann collision by default, v3_collision = "bgk" for BGK).
    Returns (kind, cfg, nsteps or None, arrival) tuples; nsteps None means "to maxTime"."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    inst = batch_instances(n, seed)
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        kind = rng.choice(3, p=np.asarray(mix) / np.sum(mix))
        if kind == 0:
            out.append(("1d", riemann_1d1f1v(inst[k]), None, t))
        elif kind == 1:
            nc = int(rng.choice(cavity_cells))
            nvel = int(rng.choice(cavity_vel))
            out.append(("2d", cavity_2d2f2v(nx=nc, ny=nc, nu=nvel, nv=nvel,
                                            vMeshType="maxwell" if nvel <= 32 else "rectangle"), cavity_steps, t))
        else:
            out.append(("3v", shock_1d1f3v(nx=v3_cells, nu=24, nv=24, nw=24, collision=v3_collision), v3_steps, t))
    return out
