"""The serving loop on the GPU: requests of different kinds arrive while others are being stepped,
share batched launches in chunks, and each result equals the same request run alone."""
import copy

import numpy as np
import pytest

from conftest import kb

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _full_width_sweep_after(gpu_ctx):
    """DeviceBackend sets its velocity cap on the shared context; later tests expect the default."""
    yield
    gpu_ctx.set_sweep_vw(0)


def _mix():
    inst = kb.configs.batch_instances(4, seed=7)
    reqs = []
    for k, i in enumerate(inst):
        i = dict(i, nx=60 + 37 * k, nu=33 + 29 * k)
        reqs.append(kb.serving.Request(rid=k, cfg=kb.configs.riemann_1d1f1v(i), nsteps=45 + 11 * k, arrival=0.002 * k))
    reqs.append(kb.serving.Request(rid=10, cfg=kb.configs.cavity_2d2f2v(nx=12, ny=12, nu=8, nv=8), nsteps=23, arrival=0.001))
    reqs.append(kb.serving.Request(rid=11, cfg=kb.configs.cavity_2d2f2v(nx=9, ny=11, nu=6, nv=6, collision="shakhov"),
                                   nsteps=17, arrival=0.004))
    reqs.append(kb.serving.Request(rid=12, cfg=kb.configs.shock_1d1f3v(nx=14, nu=6, nv=6, nw=6), nsteps=9, arrival=0.003))
    reqs.append(kb.serving.Request(rid=13, cfg=kb.configs.shock_1d2f1v(nx=50, nu=30), nsteps=31, arrival=0.0))
    # the fast spectral collision, as the bench mix's 24^3 requests take it (8^3: the LDS transforms)
    reqs.append(kb.serving.Request(rid=14, cfg=kb.configs.shock_1d1f3v(nx=10, nu=8, nv=8, nw=8, umin=-8.0, umax=8.0, vmin=-8.0,
                                                                       vmax=8.0, wmin=-8.0, wmax=8.0, collision="fsm"),
                                   nsteps=11, arrival=0.0025))
    return reqs


def test_each_result_is_the_same_alone_and_in_the_mix(gpu_ctx):
    reqs = _mix()
    backend = kb.serving.DeviceBackend(kb, ctx=gpu_ctx)
    srv = kb.serving.Server(backend, budget_bytes=1 << 32, chunk_steps=7)
    st = srv.run(reqs)
    assert st.summary()["requests"] == len(reqs) and st.max_live >= 2
    for r in reqs:
        ks = kb.SolverSet(r.cfg)
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        dev = kb.DeviceSolve(ks, ctr, face, precision=r.precision, ctx=gpu_ctx)
        dev.step(dt, r.nsteps)
        alone = copy.deepcopy(ctr)
        dev.download(alone)
        dev.close()
        mix = r.result
        assert np.array_equal(alone.f, mix.f), (r.rid, np.max(np.abs(alone.f - mix.f)))
        assert np.array_equal(alone.w, mix.w) and np.array_equal(alone.prim, mix.prim), r.rid


def test_transfers_on_a_second_context_give_the_same_results(gpu_ctx):
    """Uploads and downloads through a second context while the chunks run on the first
    (kbg_batch_step takes solves of any context of its device): each result still equals the
    request run alone, bit for bit."""
    reqs = _mix()
    xctx = kb.Context(0)
    step_be = kb.serving.DeviceBackend(kb, ctx=gpu_ctx)
    xfer_be = kb.serving.DeviceBackend(kb, ctx=xctx)
    srv = kb.serving.Server(step_be, budget_bytes=1 << 32, chunk_steps=7, overlap=True, transfer=xfer_be)
    st = srv.run(reqs)
    assert st.summary()["requests"] == len(reqs) and st.max_live >= 2
    for r in reqs:
        ks = kb.SolverSet(r.cfg)
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        dev = kb.DeviceSolve(ks, ctr, face, precision=r.precision, ctx=gpu_ctx)
        dev.step(dt, r.nsteps)
        alone = copy.deepcopy(ctr)
        dev.download(alone)
        dev.close()
        mix = r.result
        assert np.array_equal(alone.f, mix.f), (r.rid, np.max(np.abs(alone.f - mix.f)))
        assert np.array_equal(alone.w, mix.w) and np.array_equal(alone.prim, mix.prim), r.rid
    xctx.close()


def test_admission_waits_for_memory(gpu_ctx):
    """A budget that holds one request at a time still serves them all, one after another."""
    reqs = _mix()[:3]
    backend = kb.serving.DeviceBackend(kb, ctx=gpu_ctx)
    big = max(kb.solve_bytes(kb.SolverSet(r.cfg), "fp32") for r in reqs)
    srv = kb.serving.Server(backend, budget_bytes=big, chunk_steps=10)
    st = srv.run(reqs)
    assert st.max_live == 1 and st.summary()["requests"] == 3
