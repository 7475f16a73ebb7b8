"""Throughput benchmark of the batched 1D discrete-velocity time step on one MI355X.

Workload: 512 independent 1D1F1V shock tubes (KFVS flux, BGK or Shakhov by coin flip, van Leer
slopes, fixed boundaries) drawn from a fixed seed: nx ~ U{64..2048}, nv ~ U{32..256} on a
[-5, 5] rectangle grid, Kn = 10^U(-4, 1), left/right density and temperature ~ U(0.1, 1).
All 512 solves are resident in HBM in FP32 (conservative variables in FP64) and advanced together
by kbg_batch_step.  One "step" advances every instance by one time step (its own CFL dt).

Printed (one JSON line on stdout):
  value          cell x velocity updates/s of the FP32 batched path (wall clock around the K steps)
  l2_vs_fp64     accuracy of the FP32 batch at the workload's end time t = 0.2 (nt = floor(0.2 / dt) + 1
                 steps per instance, up to ~5500): plain relative L2 and L-inf errors in rho, U, T of
                 every instance against the FP64 single-solve GPU path, the largest over instances,
                 and the instances above the 1e-4 L2 bar (none expected)
  fp64_plain     the FP64 single-solve path's updates/s (kbg_step per solve, one at a time) over the
                 same runs to t = 0.2
  roofline       HBM roofline of the sweep kernels: algorithmic bytes per step (8 B per
                 cell x velocity + 48 B per cell) / device time per step (HIP events on the
                 library's stream around the step loop); the sweep advances two steps per pass
                 over f, so its measured HBM traffic ("traffic", per step) is below that figure
  cpu_baseline   the FP64 CPU oracle (OpenMP) on a bounded sample of the same instances run to
                 t = 0.2; the same instances' GPU results (FP32 batch, FP64 path) are checked against
                 its output in "accuracy_vs_oracle"
  serving        the serving leg: Poisson arrivals at each rate of --serve-rates, mixing the batch's
                 tubes, 2D cavities 32^2-128^2 x 16^2-48^2 and 1D1F3V 24^3 requests (p50 / p99)

    python bench.py [--steps K] [--warmup W] [--instances N] [--no-cpu] [--no-fp64] [--serve N]
"""
from __future__ import annotations

import argparse
import re
import copy
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "kitbase.jl_amd"))

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"
TOL = 1e-4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def macro(prim):
    p = np.asarray(prim)
    return {"rho": p[:, 0], "U": p[:, 1], "T": 1.0 / (2.0 * p[:, 2])}


def errors(pa, pb):
    """plain relative L2 and L-inf errors of a against b in rho, U, T"""
    out = {}
    ma, mb = macro(pa), macro(pb)
    for k in ma:
        d = ma[k] - mb[k]
        out[k + "_l2"] = float(np.linalg.norm(d) / max(np.linalg.norm(mb[k]), 1e-300))
        out[k + "_linf"] = float(np.max(np.abs(d)) / max(np.max(np.abs(mb[k])), 1e-300))
    return out


def cpu_info():
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def build_instances(kb, n):
    inst = kb.configs.batch_instances(n)
    out = []
    for i in inst:
        ks = kb.SolverSet(kb.configs.riemann_1d1f1v(i))
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        out.append((ks, ctr, face, dt))
    return out


def t_end_steps(inst):
    """KitBase's solve!: nt = floor(maxTime / dt) + 1 (solver_main.jl:28)"""
    return np.array([int(np.floor(ks.set.maxTime / dt)) + 1 for ks, _, _, dt in inst])


def profile_traffic(points, schedule):
    """HBM bytes per step of the sweep kernels from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE per step, scaled by the copy-kernel calibration), when it was taken on
    this same workload and schedule; returns (bytes, schedule it was taken under) or (None, None)."""
    p = PROFILES / "sweep1d_fp32_batch512_pmc.txt"
    if not p.exists():
        return None, None
    vals = {}
    for line in p.read_text().splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        if float(vals.get("points_per_step", "nan").split()[0]) != float(points):
            return None, None
        traffic = float(vals["hbm_bytes_per_step"].split()[0])
    except (KeyError, ValueError):
        return None, None
    sched = vals.get("schedule")
    return traffic, sched.split()[0] if sched else None


def traffic_key(schedule):
    """The parts of a schedule string that set the sweep's HBM traffic: the segment length, the
    steps per launch and the velocities per wavefront (the stream count and graph replay only
    change how the same launches are issued)."""
    if not schedule:
        return None
    return re.sub(r"_streams\d+|_graphs", "", schedule)


def run_batch_to(kb, ctx, inst, nts, precision="fp32"):
    """Fresh uploads of every instance; all advance together, each leaving at its own step count."""
    devs = [kb.DeviceSolve(ks, ctr, face, precision=precision, ctx=ctx) for ks, ctr, face, _ in inst]
    done, live = 0, list(np.argsort(nts, kind="stable"))
    t0 = time.perf_counter()
    while live:
        m = int(min(nts[i] for i in live))
        kb.batch_step([devs[i] for i in live], [inst[i][3] for i in live], m - done, ctx=ctx)
        done = m
        live = [i for i in live if nts[i] > done]
    ctx.sync()
    wall = time.perf_counter() - t0
    out = []
    for d, (ks, ctr, face, _) in zip(devs, inst):
        c = copy.deepcopy(ctr)
        d.download(c)
        d.close()
        out.append(c.prim.copy())
    return out, wall


def cavity_leg(kb, ctx, steps):
    """Config 3 geometry (64^2 cells x 32^2 Maxwell-quadrature velocities, UGKS + BGK, Maxwell walls)
    stepped by the fused 2D step in FP32 and FP64; device time from the library's step events."""
    cavity = {"config": "2D2F2V cavity 64^2 x 32^2, UGKS + BGK, Kn 0.075, Maxwell walls", "steps": steps}
    for prec in ("fp32", "fp64"):
        ks = kb.SolverSet(kb.configs.cavity_2d2f2v())
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        d = kb.DeviceSolve(ks, ctr, face, precision=prec, ctx=ctx)
        d.step(dt, 2)
        ctx.sync()
        t1 = time.perf_counter()
        d.step(dt, steps)
        wall_c = time.perf_counter() - t1
        dev_ms = ctx.last_steps_ms() / steps
        d.close()
        pts = ks.ps.nx * ks.ps.ny * ks.vs.nv
        esz = 4 if prec == "fp32" else 8
        alg = 2 * 2 * esz * pts + 2 * 4 * 8 * ks.ps.nx * ks.ps.ny  # f (h, b) read + write, w read + write
        cavity[prec] = {"updates_per_s": pts * steps / wall_c, "ms_per_step": wall_c * 1e3 / steps,
                        "device_ms_per_step": dev_ms, "algorithmic_bytes_per_step": alg,
                        "hbm_frac": alg / (dev_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}
    log(f"[bench] cavity: fp32 {cavity['fp32']['device_ms_per_step']:.3f} ms/step, "
        f"fp64 {cavity['fp64']['device_ms_per_step']:.3f} ms/step (device)")
    return cavity


def unstructured_leg(kb, ctx, steps, nx=64, nu=28):
    """Lid-driven cavity on a triangulated unit square (nx^2 rectangles cut into 2 nx^2 triangles,
    nu^2 velocities, 2D2F2V KFVS + BGK, Maxwell walls) stepped by the fused unstructured step
    (k_ustep) in FP32 and FP64, interpOrder 1; device time from the library's step events."""
    ks = kb.SolverSet(kb.configs.cavity_2d2f2v(nu=nu, nv=nu, knudsen=0.075, interpOrder=1))
    ps = kb.mesh.tri_rectangle(0.0, 1.0, nx, 0.0, 1.0, nx)
    out = {"config": f"2D2F2V cavity on {ps.ncell} triangles ({ps.nface} faces) x {nu}^2 velocities, KFVS + BGK, "
                     f"Kn 0.075, Maxwell walls", "cells": ps.ncell, "faces": ps.nface, "steps": steps}
    for prec in ("fp32", "fp64"):
        st = kb.mesh.init_fvm_us(ks, ps)
        dt = kb.mesh.timestep_us(ks, ps, st)
        d = kb.mesh.DeviceSolveUS(ks, ps, st, precision=prec, ctx=ctx)
        d.step(dt, 2)
        ctx.sync()
        t1 = time.perf_counter()
        d.step(dt, steps)
        wall_c = time.perf_counter() - t1
        dev_ms = ctx.last_steps_ms() / steps
        d.close()
        pts = ps.ncell * ks.vs.nv
        esz = 4 if prec == "fp32" else 8
        alg = 2 * 2 * esz * pts + 2 * 4 * 8 * ps.ncell  # f (h, b) read + write, w read + write
        out[prec] = {"updates_per_s": pts * steps / wall_c, "ms_per_step": wall_c * 1e3 / steps,
                     "device_ms_per_step": dev_ms, "algorithmic_bytes_per_step": alg,
                     "hbm_frac": alg / (dev_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}
    log(f"[bench] unstructured: fp32 {out['fp32']['device_ms_per_step']:.3f} ms/step, "
        f"fp64 {out['fp64']['device_ms_per_step']:.3f} ms/step (device)")
    return out


def fsm_leg(kb, ctx, steps):
    """1D1F3V shock structure with KitBase's fast spectral collision (50 cells, 24^3 velocities,
    M = 5: 20 direction pairs) stepped on the device; device time from the library's step events."""
    ks = kb.SolverSet(kb.configs.shock_1d1f3v(collision="fsm"))
    out = {"config": f"1D1F3V shock, {ks.ps.nx} cells x 24^3 velocities, KFVS + fast spectral collision (M = 5)",
           "steps": steps, "transforms_per_step": (ks.ps.nx - 2) * (2 * 20 + 2) + (ks.ps.nx - 2)}
    for prec in ("fp64", "fp32"):
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        d = kb.DeviceSolve(ks, ctr, face, precision=prec, ctx=ctx)
        d.step(dt, 2)
        ctx.sync()
        t1 = time.perf_counter()
        d.step(dt, steps)
        wall_c = time.perf_counter() - t1
        dev_ms = ctx.last_steps_ms() / steps
        # the 3D transforms of a step (backward spectra + forward mode products) at 5 n log2 n flops each
        n3 = ks.vs.nv
        gflop = out["transforms_per_step"] * 5.0 * n3 * math.log2(n3) * 1e-9
        out[prec] = {"ms_per_step": wall_c * 1e3 / steps, "device_ms_per_step": dev_ms,
                     "transforms_per_s": out["transforms_per_step"] / (dev_ms * 1e-3),
                     "fft_gflop_per_s": gflop / (dev_ms * 1e-3)}
        d.close()
    log(f"[bench] fsm: fp64 {out['fp64']['device_ms_per_step']:.3f} ms/step, fp32 {out['fp32']['device_ms_per_step']:.3f} ms/step")
    return out


def serving_leg(kb, ctx, n, rates, chunk, overlap):
    out = {"mix": "75% 1D tubes of the batch (to t = 0.2), 15% lid-driven cavities {32,48,64,96,128}^2 cells x "
                  "{16,24,32,48}^2 velocities (100 steps), 10% 1D1F3V 24 cells x 24^3 velocities with the fast "
                  "spectral collision (40 steps); FP32",
           "requests_per_rate": n, "budget_bytes": 32 << 30, "chunk_steps": chunk, "overlap": overlap, "rates": {}}
    # server warm-up, once per process: the fast spectral kernel modes of the 24^3 grid (a property
    # of the grid, like a served model's weights; kitbase.jl_amd/fsm.py memoises them) and rocFFT's
    # run-time compiled kernels for its transforms (one step of a small solve on that grid)
    t0 = time.perf_counter()
    ks3 = kb.SolverSet(kb.configs.shock_1d1f3v(nx=4, nu=24, nv=24, nw=24, collision="fsm"))
    c3, f3 = kb.init_fvm(ks3)
    d3 = kb.DeviceSolve(ks3, c3, f3, precision="fp32", ctx=ctx)
    d3.step(kb.timestep(ks3, c3), 1)
    d3.close()
    out["warmup_s"] = time.perf_counter() - t0
    # uploads and downloads through a second context while the chunks run on ctx
    xctx = kb.Context(0) if overlap else None
    xfer = kb.serving.DeviceBackend(kb, ctx=xctx) if overlap else None
    out["transfer_context"] = bool(overlap)
    for rate in rates:
        reqs = [kb.serving.Request(rid=k, cfg=cfg, nsteps=ns, arrival=t)
                for k, (kind, cfg, ns, t) in enumerate(kb.configs.serving_requests(n, rate))]
        srv = kb.serving.Server(kb.serving.DeviceBackend(kb, ctx=ctx), budget_bytes=32 << 30, chunk_steps=chunk,
                                overlap=overlap, transfer=xfer)
        st = srv.run(reqs).summary()
        out["rates"][str(rate)] = st
        log(f"[bench] serving {rate:g}/s: {st['requests']} requests, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, "
            f"{st['updates_per_s']:.3e} updates/s, max live {st['max_live']}")
    ctx.set_sweep_vw(0)
    if xctx is not None:
        xctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--seg", type=int, default=0, help="cells per work item (0: chosen by the autotuner)")
    ap.add_argument("--no-tune", action="store_true", help="skip the schedule autotuner (library defaults)")
    ap.add_argument("--sweep-steps", type=int, default=2, choices=(1, 2),
                    help="time steps per fused sweep launch (2: two per pass over f where the solves allow it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the accuracy leg (FP32 and FP64 runs to t = 0.2)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU work of the baseline sample")
    ap.add_argument("--cavity-steps", type=int, default=20, help="timed steps of the 2D cavity leg (0 skips it)")
    ap.add_argument("--cavity-only", action="store_true", help="run only the 2D cavity leg (profiling)")
    ap.add_argument("--us-steps", type=int, default=50, help="timed steps of the unstructured cavity leg (0 skips it)")
    ap.add_argument("--us-only", action="store_true", help="run only the unstructured cavity leg (profiling)")
    ap.add_argument("--fsm-steps", type=int, default=10, help="timed steps of the fast spectral 1D1F3V leg (0 skips it)")
    ap.add_argument("--fsm-only", action="store_true", help="run only the fast spectral leg (profiling)")
    ap.add_argument("--serve", type=int, default=48, metavar="N",
                    help="serving leg: N requests of the mixed stream per arrival rate (0 skips it)")
    ap.add_argument("--serve-rates", default="10,20,50,100,200,500", help="Poisson arrival rates of the serving leg (1/s)")
    ap.add_argument("--serve-chunk", type=int, default=400, help="steps per continuous-batching chunk of the serving leg")
    ap.add_argument("--serve-serial", action="store_true", help="serving leg: prepare arrivals between chunks only")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.copy_probe:
        # device-to-device hipMemcpy (the runtime's blit kernel) of a known size: reads and writes
        # go through L2 like the sweep's, which calibrates the FETCH_SIZE / WRITE_SIZE counters
        import ctypes

        import torch

        n = a.copy_probe * (1 << 20)
        x = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        for _ in range(5):
            rc = hip.hipMemcpy(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), n, 3)  # 3: device to device
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed: {rc}")
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "copies": 5}))
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: the batched solves are independent, so N GPUs would each "
              f"run their own share with no communication; --gpus {a.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)

    from loader import load_package

    kb = load_package()
    if a.cavity_only:
        print(json.dumps({"cavity_2d": cavity_leg(kb, kb.Context(0), max(1, a.cavity_steps))}), flush=True)
        return
    if a.us_only:
        print(json.dumps({"unstructured_2d": unstructured_leg(kb, kb.Context(0), max(1, a.us_steps))}), flush=True)
        return
    if a.fsm_only:
        print(json.dumps({"fsm_1d1f3v": fsm_leg(kb, kb.Context(0), max(1, a.fsm_steps))}), flush=True)
        return
    K, W = a.steps, a.warmup
    t0 = time.time()
    inst = build_instances(kb, a.instances)
    points = sum(ks.ps.nx * ks.vs.nv for ks, *_ in inst)
    cells = sum(ks.ps.nx for ks, *_ in inst)
    alg_bytes_step = 8 * points + 2 * 3 * 8 * cells  # FP32 f read + write, FP64 w read + write
    log(f"[bench] {len(inst)} instances, {cells} cells, {points} cell x velocity points per step "
        f"(host setup {time.time() - t0:.1f}s)")

    ctx = kb.Context(0)
    ctx.set_sweep_steps(a.sweep_steps)
    dts = np.array([d for *_, d in inst])

    # ---- FP32 batched path: the timed region
    t0 = time.time()
    devs = [kb.DeviceSolve(ks, ctr, face, precision="fp32", ctx=ctx) for ks, ctr, face, _ in inst]
    log(f"[bench] fp32 upload {time.time() - t0:.1f}s")
    tuned, taken = None, 0
    vw_now, graphs_now = 0, False
    if a.seg:
        ctx.set_schedule(a.seg, 0)
    elif not a.no_tune:
        best, trials, taken = kb.autotune.tune(ctx, devs, dts)
        vw_now, graphs_now = best[2], best[3]
        tuned = {"best": {"seg_cells": best[0], "streams": best[1], "sweep_vw": best[2], "graphs": best[3]},
                 "tuner_steps": taken,
                 "ms_per_step": {kb.autotune.key_str(k): round(v, 4) for k, v in trials.items()}}
        log(f"[bench] autotuner: {kb.autotune.key_str(best)} ({taken} trial steps): "
            + ", ".join(f"{kb.autotune.key_str(k)}: {v:.3f} ms" for k, v in trials.items()))
    if W:
        kb.batch_step(devs, dts, W, ctx=ctx)
    ctx.sync()
    t0 = time.perf_counter()
    kb.batch_step(devs, dts, K, ctx=ctx)
    ctx.sync()
    wall = time.perf_counter() - t0
    dev_ms = ctx.last_steps_ms()
    value = points * K / wall
    log(f"[bench] fp32 batched: {K} steps in {wall * 1e3:.2f} ms wall ({dev_ms:.2f} ms device), "
        f"{value:.4e} updates/s")
    for d in devs:
        d.close()
    devs = None
    schedule_now = ctx.schedule()
    ctx.set_graphs(False)  # the other legs run eager, full-width (their results do not depend on it)
    ctx.set_sweep_vw(0)

    # ---- accuracy at t = 0.2: FP32 batch against the FP64 single-solve path, every instance
    nts = t_end_steps(inst)
    res32 = res64 = None
    fp64 = acc = None
    if not a.no_fp64:
        res32, wall32 = run_batch_to(kb, ctx, inst, nts, "fp32")
        res64, tot_wall, tot_dev, upd = [], 0.0, 0.0, 0.0
        for (ks, ctr, face, dt), nt in zip(inst, nts):
            d = kb.DeviceSolve(ks, ctr, face, precision="fp64", ctx=ctx)
            ctx.sync()
            t1 = time.perf_counter()
            d.step(dt, int(nt))
            tot_wall += time.perf_counter() - t1
            tot_dev += ctx.last_steps_ms()
            upd += float(ks.ps.nx * ks.vs.nv) * nt
            c = copy.deepcopy(ctr)
            d.download(c)
            d.close()
            res64.append(c.prim.copy())
        fp64 = {"value": upd / tot_wall, "unit": "cell-velocity updates/s", "device_value": upd / (tot_dev * 1e-3),
                "path": "kbg_step per solve, FP64 storage, one solve at a time, each instance to t = 0.2"}
        rows = []
        for k, ((ks, *_), r32, r64, nt) in enumerate(zip(inst, res32, res64, nts)):
            e = errors(r32, r64)
            rows.append({"instance": k, "nx": ks.ps.nx, "nv": ks.vs.nv, "Kn": ks.gas.Kn, "collision": ks.set.collision,
                         "steps": int(nt), "rms_U": float(np.sqrt(np.mean(r64[:, 1] ** 2))), **e})
        keys = [k for k in rows[0] if k.endswith(("_l2", "_linf"))]
        worst = {k: max((r[k] if np.isfinite(r[k]) else np.inf) for r in rows) for k in keys}
        fails = [r for r in rows if not all(r[k] <= TOL for k in keys if k.endswith("_l2"))]  # NaN fails
        acc = {"t_end": 0.2, "steps": [int(nts.min()), int(nts.max())], "tol_l2": TOL,
               "reference": "FP64 single-solve GPU path (held to the oracle at 1e-12..1e-9 by the GPU tests)",
               "metric": "per instance ||x32 - x64|| / ||x64|| (L2) and max|x32 - x64| / max|x64| (Linf), x = rho, U, T; "
                         "the largest over the instances",
               **{k: v for k, v in worst.items()},
               "instances_over_tol_l2": [{k: r[k] for k in ("instance", "nx", "nv", "Kn", "rms_U", "U_l2")} for r in fails],
               "instances_over_1e-4_linf": int(sum(1 for r in rows if not all(r[k] <= TOL for k in keys if k.endswith("_linf")))),
               "pass": not fails}
        log(f"[bench] accuracy at t = 0.2 (FP32 batch vs FP64): " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items())
            + f"; {len(fails)} instances over {TOL:g} in L2")
        log(f"[bench] fp64 plain: {fp64['value']:.4e} updates/s; ratio {value / fp64['value']:.1f}x")

    # ---- 2D lid-driven cavity (64^2 cells x 32^2 velocities, UGKS, Maxwell walls), fused step
    cavity = cavity_leg(kb, ctx, a.cavity_steps) if a.cavity_steps else None
    # ---- the same cavity on a triangulated mesh (8192 triangles x 28^2 velocities), fused unstructured step
    unstructured = unstructured_leg(kb, ctx, a.us_steps) if a.us_steps else None
    # ---- 1D1F3V with the fast spectral Boltzmann collision (50 cells x 24^3 velocities)
    fsm = fsm_leg(kb, ctx, a.fsm_steps) if a.fsm_steps else None

    # ---- serving leg: Poisson arrivals of 1D tubes, 2D cavities and 1D1F3V requests, served together
    serve = None
    if a.serve:
        rates = [float(x) for x in a.serve_rates.split(",") if x]
        serve = serving_leg(kb, ctx, a.serve, rates, a.serve_chunk, not a.serve_serial)

    # ---- CPU baseline: FP64 oracle with OpenMP on a bounded sample of the instances, run to t = 0.2
    cpu, acc_oracle = None, None
    if not a.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        nthreads = oracle.num_threads()
        # rate estimate from 3 steps of the first instances, then as many instances (in order) as
        # fit the time budget at their full step counts
        probe = [(ks, copy.deepcopy(ctr), copy.deepcopy(face), dt) for ks, ctr, face, dt in inst[:8]]
        t1 = time.perf_counter()
        for ks, c, f, dt in probe:
            oracle.OracleSolve(ks, c, f).run(dt, 3)
        rate = 3 * sum(ks.ps.nx * ks.vs.nv for ks, *_ in probe) / (time.perf_counter() - t1)
        budget, m = a.cpu_seconds * rate, 0
        while m < len(inst) and sum(inst[j][0].ps.nx * inst[j][0].vs.nv * nts[j] for j in range(m + 1)) <= budget:
            m += 1
        m = max(m, 1)
        sample = [(ks, copy.deepcopy(ctr), copy.deepcopy(face), dt) for ks, ctr, face, dt in inst[:m]]
        supd = float(sum(ks.ps.nx * ks.vs.nv * nts[j] for j, (ks, *_) in enumerate(sample)))
        t1 = time.perf_counter()
        outs = []
        for j, (ks, c, f, dt) in enumerate(sample):
            oracle.OracleSolve(ks, c, f).run(dt, int(nts[j]))
            outs.append(c.prim)
        cpu_wall = time.perf_counter() - t1
        cpu = {"value": supd / cpu_wall, "unit": "cell-velocity updates/s", "cores": nthreads, "kind": "port",
               **cpu_info(),
               "sample": f"FP64 oracle (C restatement of KitBase's 1D1F1V step, OpenMP over cells, {nthreads} threads) "
                         f"on the first {m} of the {len(inst)} instances, each to t = 0.2 ({int(supd)} updates), "
                         f"{cpu_wall:.1f} s"}
        log(f"[bench] cpu oracle: {cpu['value']:.4e} updates/s on {nthreads} threads ({cpu_wall:.1f}s, {m} instances)")
        if res32 is not None:
            e32 = [errors(res32[j], outs[j]) for j in range(m)]
            e64 = [errors(res64[j], outs[j]) for j in range(m)]
            acc_oracle = {"instances": m, "t_end": 0.2,
                          "fp32_batch": {k: max(e[k] for e in e32) for k in e32[0]},
                          "fp64_path": {k: max(e[k] for e in e64) for k in e64[0]}}
            acc_oracle["pass"] = all(v <= TOL for k, v in acc_oracle["fp32_batch"].items() if k.endswith("_l2"))
            log(f"[bench] against the oracle ({m} instances, t = 0.2): fp32 max L2 "
                f"{max(v for k, v in acc_oracle['fp32_batch'].items() if k.endswith('_l2')):.2e}, fp64 max L2 "
                f"{max(v for k, v in acc_oracle['fp64_path'].items() if k.endswith('_l2')):.2e}")

    ms_step = dev_ms / K
    achieved = alg_bytes_step / (ms_step * 1e-3) / 1e9
    sched = (f"seg{schedule_now[0]}_streams{schedule_now[1] or 4}_steps{a.sweep_steps}" + (f"_vw{vw_now}" if vw_now else "")
             + ("_graphs" if graphs_now else ""))
    traffic, tsched = profile_traffic(points, sched)
    out = {
        "metric": "cell×velocity updates/s, FP32-batched vs FP64 plain @1e-4 L2; p99 solve latency",
        "value": value,
        "unit": "cell-velocity updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched 1D shock tubes: 512 instances, nx U{64..2048}, nv U{32..256}, "
                               "Kn 10^U(-4,1), BGK/Shakhov, 1D1F1V KFVS, FP32 f (departure from the local Maxwellian) / FP64 w",
                   "instances": len(inst), "seed": kb.configs.BATCH_SEED, "cells": cells,
                   "points_per_step": points, "parallelism": "single GPU"},
        "schedule": tuned,
        "cavity_2d": cavity,
        "unstructured_2d": unstructured,
        "fsm_1d1f3v": fsm,
        "fp64_plain": fp64,
        "speedup_vs_fp64_plain": (value / fp64["value"]) if fp64 else None,
        "l2_vs_fp64": acc,
        "accuracy_vs_oracle": acc_oracle,
        "p99_solve_latency_ms": serve["rates"].get("100.0", {}).get("p99_ms") if serve else None,
        "serving": serve,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic if traffic_key(tsched) == traffic_key(sched) else None,
                     "traffic_schedule": tsched,
                     "kernel": f"k_sweep1d<float,1F,KFVS>: {a.sweep_steps} time step(s) per launch, one launch "
                               f"per stream ({schedule_now[1] or 4} concurrent streams, {schedule_now[0]} cells "
                               f"per segment, {vw_now or 'full-width'} velocities per wavefront, "
                               f"{'HIP graph replay' if graphs_now else 'eager launches'}); "
                               "achieved = per-update algorithmic bytes x updates / device time",
                     "limiter": "VALU issue (the per-velocity arithmetic of two time steps per pass over f), "
                                "not HBM bandwidth: see profiles/sweep1d_fp32_batch512_pmc.txt",
                     "schedule": sched, "steps_per_launch": a.sweep_steps,
                     "algorithmic_bytes_per_step": alg_bytes_step, "device_ms_per_step": ms_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
