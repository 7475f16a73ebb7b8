"""Failure detection (kbg_solve_health): a state that goes bad -- a NaN, a non-positive density or
temperature -- raises the solve's flag bits in the kernels that update the cells, on every path (the
fused 1D sweep in FP32 and FP64, the fused 2D step, the phase-level update), while every kernel runs
to completion.  The bad values are planted in the host state before the upload."""
import copy

import numpy as np
import pytest

from conftest import kb

pytestmark = pytest.mark.gpu

NONFINITE, RHO, LAM = kb._lib.KBG_HEALTH_NONFINITE, kb._lib.KBG_HEALTH_RHO, kb._lib.KBG_HEALTH_LAMBDA


def _solve(cfg, precision, ctx, plant=None):
    ks = kb.SolverSet(cfg)
    ctr, face = kb.init_fvm(ks)
    dt = kb.timestep(ks, ctr)
    if plant:
        plant(ks, ctr)
    return ks, ctr, dt, kb.DeviceSolve(ks, ctr, face, precision=precision, ctx=ctx)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_healthy_runs_raise_nothing(gpu_ctx, precision):
    for cfg in (kb.configs.sod_1d1f1v(), kb.configs.cavity_2d2f2v(nx=10, ny=10, nu=8, nv=8)):
        ks, ctr, dt, dev = _solve(cfg, precision, gpu_ctx)
        dev.step(dt, 6)
        assert dev.health() == 0
        dev.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_nan_in_the_sweep(gpu_ctx, precision):
    def plant(ks, ctr):
        ctr.f[50, 0, 10] = np.nan

    ks, ctr, dt, dev = _solve(kb.configs.sod_1d1f1v(), precision, gpu_ctx, plant)
    dev.step(dt, 3)
    h = dev.health(clear=True)
    assert h & NONFINITE, h
    assert dev.health() == 0  # cleared
    dev.close()


def test_negative_temperature_in_the_sweep(gpu_ctx):
    """energy below the kinetic energy: lambda = rho / (2 (gamma - 1) (E - m^2 / (2 rho))) < 0 in the
    planted cell (its Maxwellian, with lambda^(1/2), then spreads NaNs to its neighbours)"""
    def plant(ks, ctr):
        w = ctr.w[50]
        w[2] = 0.25 * w[1] ** 2 / w[0] - 0.1
        ctr.prim[50] = kb.theory.conserve_prim(w, ks.gas.gamma)

    ks, ctr, dt, dev = _solve(kb.configs.sod_1d1f1v(), "fp64", gpu_ctx, plant)
    dev.step(dt, 1)
    h = dev.health()
    assert h & LAM, h
    dev.close()


def test_negative_density_in_the_phase_update(gpu_ctx):
    def plant(ks, ctr):
        ctr.w[30, 0] = -0.5

    ks, ctr, dt, dev = _solve(kb.configs.shock_1d2f1v(nx=60, nu=24), "fp32", gpu_ctx, plant)
    dev.reconstruct()
    dev.evolve(dt)
    dev.update(dt)
    assert dev.health() & RHO
    dev.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_nan_in_the_2d_step(gpu_ctx, precision):
    def plant(ks, ctr):
        ctr.f[4, 5, 1, 7] = np.nan

    ks, ctr, dt, dev = _solve(kb.configs.cavity_2d2f2v(nx=10, ny=10, nu=8, nv=8), precision, gpu_ctx, plant)
    dev.step(dt, 3)
    assert dev.health() & NONFINITE
    dev.close()


def test_a_bad_solve_does_not_flag_its_batch_neighbours(gpu_ctx):
    good = _solve(kb.configs.sod_1d1f1v(), "fp32", gpu_ctx)
    bad = _solve(kb.configs.sod_1d1f1v(), "fp32", gpu_ctx, lambda ks, ctr: ctr.f.__setitem__((20, 0, 3), np.inf))
    kb.batch_step([good[3], bad[3]], [good[2], bad[2]], 4, ctx=gpu_ctx)
    assert good[3].health() == 0 and bad[3].health() & NONFINITE
    good[3].close()
    bad[3].close()
