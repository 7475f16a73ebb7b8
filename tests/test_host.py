"""The host mirror's own arithmetic on the CPU (no GPU): the CFL time step in closed form."""
import numpy as np
import pytest

from conftest import kb


def test_timestep_closed_form_sod():
    """timestep (solver_main.jl:240-288) on the Sod state: u1 = 5, c = 3^(1/2) (gamma = 3, T = 1),
    dx = 0.01: dt = 0.5 / ((5 + 3^(1/2)) / 0.01); nt = floor(0.2 / dt) + 1 = 270."""
    ks = kb.SolverSet(kb.configs.sod_1d1f1v())
    ctr, _ = kb.init_fvm(ks)
    dt = kb.timestep(ks, ctr)
    assert dt == pytest.approx(0.5 / ((5.0 + np.sqrt(3.0)) / 0.01), rel=1e-15)
    assert int(np.floor(ks.set.maxTime / dt)) + 1 == 270
    # clipped to maxTime - simTime near the end
    assert kb.timestep(ks, ctr, ks.set.maxTime - 0.25 * dt) == pytest.approx(0.25 * dt, rel=1e-12)


def test_timestep_closed_form_cavity():
    """2D (solver_main.jl:290-342): the x and y rates add; gamma = 5/3, lambda = 1, u1 = v1 = 5."""
    ks = kb.SolverSet(kb.configs.cavity_2d2f2v())
    ctr, _ = kb.init_fvm(ks)
    assert kb.timestep(ks, ctr) == pytest.approx(0.8 / (2 * 64 * (5.0 + np.sqrt(5.0 / 6.0))), rel=1e-15)


def test_timestep_uses_the_flow_speed_when_it_exceeds_u1():
    """max(u1, |U|): a cell moving faster than the velocity-space bound sets the rate."""
    ks = kb.SolverSet(kb.configs.sod_1d1f1v())
    ctr, _ = kb.init_fvm(ks)
    ctr.prim[40] = [1.0, 7.0, 0.5]
    c = np.sqrt(3.0)
    assert kb.timestep(ks, ctr) == pytest.approx(0.5 / ((7.0 + c) / 0.01), rel=1e-15)
