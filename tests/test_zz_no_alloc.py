"""kbg_batch_step_noalloc (the allocation-free variant of kbg_batch_step) against kbg_batch_step:
the same batch stepped by both, bit for bit (states and residuals), on two calls in a row, the second
with a smaller batch; and its refusals (a phase-stepped solve without buffers, a batch larger than
the staging array)."""
import copy
import ctypes as C

import numpy as np
import pytest

from conftest import kb

pytestmark = pytest.mark.gpu


def _solves(cfgs, ctx, precision="fp32", faces=True):
    out = []
    for cfg in cfgs:
        ks = kb.SolverSet(cfg)
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        dev = kb.DeviceSolve(ks, copy.deepcopy(ctr), face if faces else None, precision=precision, ctx=ctx)
        out.append((ks, ctr, face, dt, dev))
    return out


def _call(ctx, name, runs, nsteps, with_res):
    lib = ctx.lib
    hs = (C.c_void_p * len(runs))(*[r[4].handle.value for r in runs])
    dts = np.ascontiguousarray([r[3] for r in runs], dtype=np.float64)
    res = np.zeros((len(runs), 5)) if with_res else None
    kb._lib.check(getattr(lib, name)(ctx.handle, hs, len(runs), kb._lib.dptr(dts), nsteps, kb._lib.dptr(res)))
    return res


def _state(run):
    ks, ctr, face, dt, dev = run
    c = copy.deepcopy(ctr)
    dev.download(c)
    return c


def _mix(k):
    inst = kb.configs.batch_instances(5, seed=11 + k)
    cfgs = [kb.configs.riemann_1d1f1v(dict(i, nx=90 + 61 * j, nu=40 + 23 * j)) for j, i in enumerate(inst)]
    cfgs.append(kb.configs.cavity_2d2f2v(nx=10, ny=9, nu=8, nv=8))
    cfgs.append(kb.configs.shock_1d2f1v(nx=40, nu=24))
    return cfgs


def test_noalloc_equals_batch_step_on_two_calls(gpu_ctx):
    cfgs = _mix(0)
    # call 1: the whole mix; call 2: a smaller batch (the first three solves), both with residuals
    for ncall, take, nsteps in ((1, len(cfgs), 9), (2, 3, 6)):
        print(f"noalloc call {ncall}: {take} solves, {nsteps} steps", flush=True)
        ref = _solves(cfgs[:take], gpu_ctx)
        new = _solves(cfgs[:take], gpu_ctx)
        r_ref = _call(gpu_ctx, "kbg_batch_step", ref, nsteps, True)
        r_new = _call(gpu_ctx, "kbg_batch_step_noalloc", new, nsteps, True)
        assert np.array_equal(r_ref, r_new), (ncall, r_ref, r_new)
        for a, b in zip(ref, new):
            sa, sb = _state(a), _state(b)
            assert np.array_equal(sa.f, sb.f) and np.array_equal(sa.w, sb.w) and np.array_equal(sa.prim, sb.prim)
        # and once more without residuals (the two-steps-per-pass sweep)
        _call(gpu_ctx, "kbg_batch_step", ref, nsteps + 1, False)
        _call(gpu_ctx, "kbg_batch_step_noalloc", new, nsteps + 1, False)
        for a, b in zip(ref, new):
            sa, sb = _state(a), _state(b)
            assert np.array_equal(sa.f, sb.f) and np.array_equal(sa.w, sb.w)
        for r in ref + new:
            r[4].close()


def test_noalloc_refusals(gpu_ctx):
    # a 1D1F3V solve is stepped by the phase kernels, whose slope and face buffers the first
    # allocating call (a step, a phase-level call or a face upload) creates
    runs = _solves([kb.configs.shock_1d1f3v(nx=8, nu=4, nv=4, nw=4)], gpu_ctx, faces=False)
    with pytest.raises(kb._lib.KbgError, match="phase kernels"):
        _call(gpu_ctx, "kbg_batch_step_noalloc", runs, 2, False)
    ref = _solves([kb.configs.shock_1d1f3v(nx=8, nu=4, nv=4, nw=4)], gpu_ctx, faces=False)
    _call(gpu_ctx, "kbg_batch_step", runs, 2, False)       # allocates the phase buffers
    _call(gpu_ctx, "kbg_batch_step_noalloc", runs, 3, False)
    _call(gpu_ctx, "kbg_batch_step", ref, 5, False)
    assert np.array_equal(_state(runs[0]).f, _state(ref[0]).f)
    # more staging than the array holds: one cell per work item over many cells
    big = _solves([kb.configs.riemann_1d1f1v(dict(kb.configs.batch_instances(1)[0], nx=2048, nu=32))] * 40, gpu_ctx)
    seg, streams = gpu_ctx.schedule()
    gpu_ctx.set_schedule(1, streams)
    try:
        with pytest.raises(kb._lib.KbgError, match="staging"):
            _call(gpu_ctx, "kbg_batch_step_noalloc", big, 1, False)
    finally:
        gpu_ctx.set_schedule(seg, streams)
    for r in runs + ref + big:
        r[4].close()
