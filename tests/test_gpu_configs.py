"""The configurations at their full sizes and step counts, against the FP64 oracle, plus the closed
forms and invariants that hold over runs too long for the oracle:

  C1  Sod 1D1F1V 100 x 64: the CFL time step in closed form
  C2  Ma = 2 shock 1D2F1V UGKS + Shakhov 400 x 101: all 2000 steps
  C3  lid-driven cavity 2D2F2V UGKS 64^2 x 32^2 at Kn 0.075 and 1: 500 steps against the oracle,
      5000 steps (the configuration's length) for mass conservation and FP32 against FP64
  checkpoint -> resume against an uninterrupted oracle run of the same total steps

Errors are the plain relative ones, ||a - b|| / ||b|| (L2) and max |a - b| / max |b| (L-inf), for
rho, each velocity component and T."""
import copy

import numpy as np
import pytest

from conftest import kb, oracle

pytestmark = pytest.mark.gpu


def macro(prim):
    p = np.asarray(prim)
    return {"rho": p[..., 0], "U": p[..., 1:-1], "T": 1.0 / (2.0 * p[..., -1])}


def errs(pg, po):
    out = {}
    for k, a in macro(pg).items():
        b = macro(po)[k]
        out[k + "_l2"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out[k + "_linf"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return out


def assert_close(pg, po, l2, linf, label=""):
    e = errs(pg, po)
    print(label, {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    bad = {k: v for k, v in e.items() if not v <= (l2 if k.endswith("_l2") else linf)}  # NaN fails
    assert not bad, bad


def gpu_run(ks, ctr, face, dt, nsteps, precision, ctx, chunk=None):
    dev = kb.DeviceSolve(ks, copy.deepcopy(ctr), face, precision=precision, ctx=ctx)
    left = nsteps
    while left > 0:
        n = left if chunk is None else min(chunk, left)
        dev.step(dt, n)
        left -= n
    c = copy.deepcopy(ctr)
    dev.download(c)
    dev.close()
    return c


def oracle_run(ks, ctr, face, dt, nsteps):
    c, f = copy.deepcopy(ctr), copy.deepcopy(face)
    oracle.OracleSolve(ks, c, f).run(dt, nsteps)
    return c


# ---------------------------------------------------------------- C1: the time step in closed form

def test_sod_timestep_closed_form(gpu_ctx):
    """timestep (solver_main.jl:240-288): dt = cfl / max_i (max(u1, |U|) + c) / dx; on the Sod state
    the largest rate is the left state's: u1 = 5, c = (gamma / (2 lambda))^(1/2) = 3^(1/2) (gamma = 3,
    T = 1), dx = 0.01, so dt = 0.5 / ((5 + 3^(1/2)) / 0.01) = 7.4272e-4."""
    ks = kb.SolverSet(kb.configs.sod_1d1f1v())
    ctr, face = kb.init_fvm(ks)
    dev = kb.DeviceSolve(ks, ctr, face, "fp64", gpu_ctx)
    expect = 0.5 / ((5.0 + np.sqrt(3.0)) / 0.01)
    assert dev.timestep() == pytest.approx(expect, rel=1e-14)
    assert kb.timestep(ks, ctr) == pytest.approx(expect, rel=1e-14)
    dev.close()


def test_cavity_timestep_closed_form(gpu_ctx):
    """2D: dt = cfl / max ((max(u1, |U|) + c) / dx + (max(v1, |V|) + c) / dy); the cavity at rest,
    lambda = 1, gamma = 5/3: c = (5/6)^(1/2), u1 = v1 = 5, dx = dy = 1/64, cfl 0.8."""
    ks = kb.SolverSet(kb.configs.cavity_2d2f2v())
    ctr, face = kb.init_fvm(ks)
    dev = kb.DeviceSolve(ks, ctr, face, "fp32", gpu_ctx)
    expect = 0.8 / (2 * 64 * (5.0 + np.sqrt(5.0 / 6.0)))
    assert dev.timestep() == pytest.approx(expect, rel=1e-14)
    dev.close()


# ---------------------------------------------------------------- C2: all 2000 steps

@pytest.mark.parametrize("precision,l2,linf", [("fp64", 1e-6, 1e-6), ("fp32", 1e-4, 1e-4)])
def test_shock_config_all_steps(gpu_ctx, precision, l2, linf):
    ks = kb.SolverSet(kb.configs.shock_1d2f1v())
    ctr, face = kb.init_fvm(ks)
    dt = kb.timestep(ks, ctr)
    n = kb.configs.SHOCK_STEPS
    co = oracle_run(ks, ctr, face, dt, n)
    cg = gpu_run(ks, ctr, face, dt, n, precision, gpu_ctx)
    assert_close(cg.prim, co.prim, l2, linf, f"C2 {precision} {n} steps")


# ---------------------------------------------------------------- C3: both Knudsen numbers

@pytest.mark.parametrize("kn", [0.075, 1.0])
def test_cavity_config_against_oracle(gpu_ctx, kn):
    """64^2 cells x 32^2 velocities, UGKS + BGK, Maxwell walls with the moving lid: 500 steps."""
    ks = kb.SolverSet(kb.configs.cavity_2d2f2v(knudsen=kn))
    ctr, face = kb.init_fvm(ks)
    dt = kb.timestep(ks, ctr)
    co = oracle_run(ks, ctr, face, dt, 500)
    c64 = gpu_run(ks, ctr, face, dt, 500, "fp64", gpu_ctx)
    assert_close(c64.prim, co.prim, 1e-6, 1e-6, f"C3 Kn {kn} fp64")
    c32 = gpu_run(ks, ctr, face, dt, 500, "fp32", gpu_ctx)
    assert_close(c32.prim, co.prim, 1e-4, 3e-4, f"C3 Kn {kn} fp32")


@pytest.mark.parametrize("kn", [0.075, 1.0])
def test_cavity_config_all_steps(gpu_ctx, kn):
    """All 5000 steps: the walls carry no mass, so the total mass stays at its initial value (the
    interior face fluxes cancel in pairs); the FP32 state stays within 1e-4 of the FP64 one (FP64 is
    held to the oracle above)."""
    ks = kb.SolverSet(kb.configs.cavity_2d2f2v(knudsen=kn))
    ctr, face = kb.init_fvm(ks)
    dt = kb.timestep(ks, ctr)
    n = kb.configs.CAVITY_STEPS
    area = ks.ps.dx * ks.ps.dy
    m0 = float(np.sum(ctr.w[..., 0] * area))
    out = {}
    for prec in ("fp64", "fp32"):
        c = gpu_run(ks, ctr, face, dt, n, prec, gpu_ctx, chunk=1000)
        assert np.all(np.isfinite(c.prim))
        assert float(np.sum(c.w[..., 0] * area)) == pytest.approx(m0, rel=1e-11)
        out[prec] = c
    # the lid has set the gas moving: a velocity field well above round-off
    assert np.max(np.abs(out["fp64"].prim[..., 1])) > 0.02
    assert_close(out["fp32"].prim, out["fp64"].prim, 1e-4, 3e-4, f"C3 Kn {kn} 5000 steps fp32 vs fp64")


# ---------------------------------------------------------------- checkpoint -> resume vs the oracle

_CK = {
    "sod_fp32": (lambda: kb.configs.sod_1d1f1v(), "fp32", 1e-4),
    "walls_fp64": (lambda: kb.configs.shock_1d2f1v(boundary="maxwell", nx=60, nu=40, flux="kfvs"), "fp64", 1e-9),
    "cavity_fp64": (lambda: kb.configs.cavity_2d2f2v(nx=12, ny=10, nu=12, nv=12, collision="shakhov"), "fp64", 1e-9),
}


@pytest.mark.parametrize("case", sorted(_CK))
def test_checkpoint_resume_against_oracle(gpu_ctx, case, tmp_path):
    """9 steps, download -> file -> upload into a new solve -> 13 more steps, against the oracle's
    uninterrupted 22 steps."""
    cfgf, prec, tol = _CK[case]
    ks = kb.SolverSet(cfgf())
    ctr, face = kb.init_fvm(ks)
    dt = kb.timestep(ks, ctr)
    a = kb.DeviceSolve(ks, copy.deepcopy(ctr), face, prec, gpu_ctx)
    a.step(dt, 9)
    snap = copy.deepcopy(ctr)
    a.download(snap)
    a.close()
    path = tmp_path / "checkpoint.npz"
    np.savez(path, w=snap.w, prim=snap.prim, f=snap.f)
    z = np.load(path)
    back = copy.deepcopy(ctr)
    back.w[...], back.prim[...], back.f[...] = z["w"], z["prim"], z["f"]
    b = kb.DeviceSolve(ks, back, None, prec, gpu_ctx)
    b.step(dt, 13)
    cb = copy.deepcopy(ctr)
    b.download(cb)
    b.close()
    co = oracle_run(ks, ctr, face, dt, 22)
    assert_close(cb.prim, co.prim, tol, tol, case)
    if prec == "fp64":
        assert np.max(np.abs(cb.f - co.f)) <= 1e-9 * np.max(np.abs(co.f))


# ---------------------------------------------------------------- per-instance adaptive dt in a batch

def test_batch_timestep_and_adaptive_batch_step(gpu_ctx):
    """kbg_batch_timestep gives every solve the host timestep formula (solver_main.jl:240-342) of its
    current state, and kbg_batch_step_cfl steps each solve with its own dt: the same as stepping it
    alone with that dt, bit for bit."""
    cfgs = [kb.configs.riemann_1d1f1v(i) for i in kb.configs.batch_instances(5, seed=3)]
    cfgs += [kb.configs.sod_1d1f1v(), kb.configs.cavity_2d2f2v(nx=10, ny=9, nu=8, nv=8)]
    runs = []
    for cfg in cfgs:
        ks = kb.SolverSet(cfg)
        ctr, face = kb.init_fvm(ks)
        runs.append((ks, ctr, kb.DeviceSolve(ks, ctr, face, precision="fp32", ctx=gpu_ctx),
                     kb.DeviceSolve(ks, copy.deepcopy(ctr), face, precision="fp32", ctx=gpu_ctx)))
    devs = [r[2] for r in runs]
    # the first call is KitBase's dt of the initial state (the sod case: 7.427e-4)
    dts0 = kb.batch_timestep(devs)
    for (ks, ctr, _, _), d in zip(runs, dts0):
        assert d == pytest.approx(kb.timestep(ks, ctr), rel=1e-14)
    assert dts0[5] == pytest.approx(7.427e-4, rel=1e-3)
    for call in range(3):
        used = kb.batch_step_cfl(devs, 4)
        # the state moved: the dt of each solve follows its own state
        for (ks, ctr, a, b), d in zip(runs, used):
            b.step(d, 4)
            ca, cb = copy.deepcopy(ctr), copy.deepcopy(ctr)
            a.download(ca)
            b.download(cb)
            assert np.array_equal(ca.f, cb.f) and np.array_equal(ca.w, cb.w), (ks.set.space, call)
        after = kb.batch_timestep(devs)
        for (ks, ctr, a, _), d in zip(runs, after):
            c = copy.deepcopy(ctr)
            a.download(c)
            assert d == pytest.approx(kb.timestep(ks, c), rel=1e-12)
    # max_time clips the last step
    clipped = kb.batch_timestep(devs, sim_time=0.05, max_time=0.05 + 1e-6)
    assert np.allclose(clipped, 1e-6, rtol=1e-9)
    for r in runs:
        r[2].close()
        r[3].close()
