"""GPU parity of KitBase's fast spectral collision (1D1F3V step!(:fsm), kbg_fsm.hip: hipFFT batched
3D transforms plus the mode products) against the oracle's orc_step_1f3v_fsm, through the C ABI.

Tolerances: FP64 storage <= 1e-10 relative (max norm of f and w after a few steps); FP32 storage
(f and the transforms in FP32, Q summed in FP64) <= 1e-4 relative L2 in density, velocity and
temperature, and in the dt Q increment of a homogeneous gas.
"""
import copy
import ctypes as C

import numpy as np
import pytest

from conftest import kb, oracle

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_l2(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fsm_case(n=16, nx=8, kn=1.0, shape=None):
    nu, nv, nw = shape or (n, n, n)
    cfg = kb.configs.shock_1d1f3v(collision="fsm", nu=nu, nv=nv, nw=nw, umin=-8.0, umax=8.0, vmin=-8.0, vmax=8.0,
                                  wmin=-8.0, wmax=8.0, nx=nx, knudsen=kn)
    ks = kb.SolverSet(cfg)
    ctr, face = kb.init_fvm(ks)
    return ks, ctr, face, kb.timestep(ks, ctr)


def run_both(ks, ctr, face, dt, nsteps, precision):
    cg, fg = copy.deepcopy(ctr), copy.deepcopy(face)
    dev = kb.DeviceSolve(ks, cg, fg, precision=precision)
    rg = np.zeros(5)
    dev.step(dt, nsteps, residual=rg)
    dev.download(cg)
    assert dev.health() == 0
    dev.close()
    co, fo = copy.deepcopy(ctr), copy.deepcopy(face)
    oracle.OracleSolve(ks, co, fo).run(dt, nsteps)
    return cg, co


def test_fsm_fp64_matches_the_oracle():
    ks, ctr, face, dt = fsm_case()
    cg, co = run_both(ks, ctr, face, dt, 3, "fp64")
    assert rel_max(co.f, ctr.f) > 1e-6  # the shock and the collisions have moved f
    assert rel_max(cg.f, co.f) <= 1e-10, rel_max(cg.f, co.f)
    assert rel_max(cg.w, co.w) <= 1e-10, rel_max(cg.w, co.w)
    assert rel_max(cg.prim, co.prim) <= 1e-10


def test_fsm_fp32_matches_the_oracle():
    ks, ctr, face, dt = fsm_case()
    cg, co = run_both(ks, ctr, face, dt, 3, "fp32")
    pg, po = cg.prim, co.prim
    assert rel_l2(pg[:, 0], po[:, 0]) <= 1e-4
    assert rel_l2(1 / pg[:, 4], 1 / po[:, 4]) <= 1e-4
    assert float(np.linalg.norm(pg[:, 1:4] - po[:, 1:4]) / np.linalg.norm(po[:, 1:4])) <= 1e-4


def homogeneous_bimodal(ks, ctr):
    """Every cell holds the same bimodal f (two Maxwellians at u = +-1.2): the face fluxes cancel, so
    a step moves f by exactly dt Q(f) (step_1f.jl:146-151) and w stays."""
    vs = ks.vs

    def maxw(U, lam):
        return (lam / np.pi) ** 1.5 * np.exp(-lam * ((vs.u - U) ** 2 + vs.v ** 2 + vs.w ** 2))

    fb = 0.5 * maxw(1.2, 0.6) + 0.5 * maxw(-1.2, 0.6)
    wt = vs.weights
    w = np.array([np.sum(fb * wt), np.sum(fb * vs.u * wt), np.sum(fb * vs.v * wt), np.sum(fb * vs.w * wt),
                  0.5 * np.sum(fb * (vs.u ** 2 + vs.v ** 2 + vs.w ** 2) * wt)])
    rho = w[0]
    lam = 0.75 * rho / (w[4] - 0.5 * np.sum(w[1:4] ** 2) / rho)  # monatomic, K = 0
    ctr.f[:, 0, :] = fb
    ctr.w[:] = w
    ctr.prim[:] = [rho, w[1] / rho, w[2] / rho, w[3] / rho, lam]


@pytest.mark.parametrize("precision,tol,shape", [
    # cubes of side 8, 12, 16, 18, 24 take the library's own forward transforms (one kernel instance
    # per side and precision: FP64 through HBM in three passes, FP32 in LDS); other shapes hipFFT
    ("fp64", 1e-10, (8, 8, 8)), ("fp64", 1e-10, (12, 12, 12)), ("fp64", 1e-10, (16, 16, 16)),
    ("fp64", 1e-10, (18, 18, 18)), ("fp64", 1e-10, (24, 24, 24)), ("fp64", 1e-10, (16, 12, 12)),
    ("fp32", 1e-4, (8, 8, 8)), ("fp32", 1e-4, (12, 12, 12)), ("fp32", 1e-4, (16, 16, 16)),
    ("fp32", 1e-4, (18, 18, 18)), ("fp32", 1e-4, (24, 24, 24)), ("fp32", 1e-4, (16, 12, 12))])
def test_fsm_homogeneous_step_is_dt_q(precision, tol, shape):
    """The collision operator alone: in a homogeneous gas one step moves f by dt Q(f) (64 % of f on
    16^3).  FP64 storage runs the transforms in FP64 (bar 1e-10), FP32 storage in FP32 (bar 1e-4
    relative L2 of the increment; the spectra, modes and transforms round at 6e-8)."""
    ks, ctr, face, dt = fsm_case(nx=6, shape=shape)
    homogeneous_bimodal(ks, ctr)
    cg, co = run_both(ks, ctr, face, dt, 1, precision)
    dg, do = cg.f[1:-1] - ctr.f[1:-1], co.f[1:-1] - ctr.f[1:-1]
    assert np.linalg.norm(do) > 0.02 * np.linalg.norm(ctr.f[1:-1])
    err = rel_l2(dg, do)
    print(f"{precision}: relative L2 of the dt Q increment {err:.3e}")
    assert err <= tol, err
    np.testing.assert_allclose(cg.w, co.w, rtol=1e-12, atol=0)


def test_fsm_phase_calls_match_the_oracle():
    """kbg_reconstruct / kbg_evolve / kbg_update one phase at a time (update! computes Q from f
    before it moves f) against the oracle's phases, with the residual."""
    ks, ctr, face, dt = fsm_case(nx=6)
    cg, fg = copy.deepcopy(ctr), copy.deepcopy(face)
    co, fo = copy.deepcopy(ctr), copy.deepcopy(face)
    dev = kb.DeviceSolve(ks, cg, fg, precision="fp64")
    orc = oracle.OracleSolve(ks, co, fo)
    for _ in range(2):
        rg, ro = np.zeros(5), np.zeros(5)
        dev.reconstruct()
        dev.evolve(dt)
        dev.update(dt, rg)
        orc.reconstruct()
        orc.evolve(dt)
        orc.update(dt, ro)
        dev.download(cg)
        assert rel_max(cg.f, co.f) <= 1e-10
        assert rel_max(cg.w, co.w) <= 1e-10
        # mass, x-momentum, energy; the transverse momenta of a shock along x are rounding noise
        # (divided by sum |w| + 1e-7, itself noise), compared on the scale of the largest residual
        np.testing.assert_allclose(rg[[0, 1, 4]], ro[[0, 1, 4]], rtol=1e-8)
        np.testing.assert_allclose(rg[2:4], ro[2:4], rtol=0, atol=1e-6 * np.max(np.abs(ro)))
    dev.close()


def test_fsm_in_a_batch_steps_as_alone():
    """A :fsm solve batched with a 1D sweep solve and a BGK 1D1F3V solve advances exactly as alone
    (bit for bit), with graph replay switched on (batches with :fsm solves launch eagerly)."""
    ks, ctr, face, dt = fsm_case()
    k1 = kb.SolverSet(kb.configs.sod_1d1f1v())
    c1, f1 = kb.init_fvm(k1)
    k3 = kb.SolverSet(kb.configs.shock_1d1f3v(nu=12, nv=12, nw=12, nx=10))
    c3, f3 = kb.init_fvm(k3)
    dts = [dt, kb.timestep(k1, c1), kb.timestep(k3, c3)]
    alone = kb.DeviceSolve(ks, ctr, face, precision="fp64")
    alone.step(dt, 4)
    ctx = kb.Context(0)
    ctx.set_graphs(True)
    batch = [kb.DeviceSolve(ks, ctr, face, precision="fp64", ctx=ctx), kb.DeviceSolve(k1, c1, f1, precision="fp32", ctx=ctx),
             kb.DeviceSolve(k3, c3, f3, precision="fp32", ctx=ctx)]
    kb.batch_step(batch, dts, 4, ctx=ctx)
    a, b = copy.deepcopy(ctr), copy.deepcopy(ctr)
    alone.download(a)
    batch[0].download(b)
    np.testing.assert_array_equal(a.f, b.f)
    np.testing.assert_array_equal(a.w, b.w)
    for d in [alone] + batch:
        d.close()
    ctx.close()


def test_fsm_refusals():
    """:fsm is the 1D1F3V step!: other spaces are refused, and so is the allocation-free batch
    step (hipFFT is outside its contract)."""
    k1 = kb.SolverSet(kb.configs.sod_1d1f1v())
    d, keep = kb.solver.make_desc(k1, "fp64")
    d.collision = kb._lib.KBG_FSM
    lib = kb._lib.load()
    ctx = kb.Context(0)
    h = C.c_void_p()
    assert lib.kbg_solve_create(ctx.handle, C.byref(d), C.byref(h)) == -3  # KBG_ERR_UNSUPPORTED
    assert b"1D1F3V" in lib.kbg_last_error()
    ks, ctr, face, dt = fsm_case(nx=5)
    dev = kb.DeviceSolve(ks, ctr, face, precision="fp64", ctx=ctx)
    hs = (C.c_void_p * 1)(dev.handle.value)
    dts = np.array([dt])
    rc = lib.kbg_batch_step_noalloc(ctx.handle, hs, 1, kb._lib.dptr(dts), 1, None)
    assert rc == -4 and b"fast spectral" in lib.kbg_last_error()  # KBG_ERR_STATE
    dev.close()
    ctx.close()
