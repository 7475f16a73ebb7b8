"""Throughput benchmark of the batched 1D discrete-velocity time step on one MI355X.

Workload: 512 independent 1D1F1V shock tubes (KFVS flux, BGK or Shakhov by coin flip, van Leer
slopes, fixed boundaries) drawn from a fixed seed: nx ~ U{64..2048}, nv ~ U{32..256} on a
[-5, 5] rectangle grid, Kn = 10^U(-4, 1), left/right density and temperature ~ U(0.1, 1).
All 512 solves are resident in HBM in FP32 (conservative variables in FP64) and advanced together
by kbg_batch_step.  One "step" advances every instance by one time step (its own CFL dt).

Printed (one JSON line on stdout):
  value          cell x velocity updates/s of the FP32 batched path (wall clock around the K steps)
  fp64_plain     the same instances stepped one solve at a time in FP64 (kbg_step per solve)
  l2_vs_fp64     max over instances of the relative L2 error in rho, U, T of the FP32 batch
                 against the FP64 single-solve results after the same steps
  roofline       HBM roofline of the sweep kernels: algorithmic bytes per step (8 B per
                 cell x velocity + 48 B per cell) / device time per step (HIP events on the
                 library's stream around the step loop); the sweep advances two steps per pass
                 over f, so its measured HBM traffic ("traffic", per step) is below that figure
  cpu_baseline   the FP64 CPU oracle (OpenMP) on a bounded sample of the same instances

    python bench.py [--steps K] [--warmup W] [--instances N] [--no-cpu] [--no-fp64]
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "kitbase.jl_amd"))

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def macro(prim):
    p = np.asarray(prim)
    return p[:, 0], p[:, 1], 1.0 / (2.0 * p[:, 2])


def rel_l2(a, b, scale=None):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    return float(d / (scale if scale is not None else max(np.linalg.norm(b), 1e-300)))


def build_instances(kb, n):
    inst = kb.configs.batch_instances(n)
    out = []
    for i in inst:
        ks = kb.SolverSet(kb.configs.riemann_1d1f1v(i))
        ctr, face = kb.init_fvm(ks)
        dt = kb.timestep(ks, ctr)
        out.append((ks, ctr, face, dt))
    return out


def traffic_from_profiles(points):
    """HBM bytes per step of the sweep kernels from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE per step, scaled by the copy-kernel calibration), when it was taken on
    this same workload (same number of cell x velocity points)."""
    p = PROFILES / "sweep1d_fp32_batch512_pmc.txt"
    if not p.exists():
        return None
    vals = {}
    for line in p.read_text().splitlines():
        if ":" in line and line.split(":", 1)[0].strip() in ("points_per_step", "hbm_bytes_per_step"):
            k, v = line.split(":", 1)
            vals[k.strip()] = float(v.split()[0])
    if vals.get("points_per_step") != float(points) or "hbm_bytes_per_step" not in vals:
        return None
    return vals["hbm_bytes_per_step"]


def cavity_leg(kb, ctx, steps):
    """Config 3 geometry (64^2 cells x 32^2 Maxwell-quadrature velocities, UGKS + BGK, Maxwell walls)
    stepped by the fused 2D step in FP32 and FP64; device time from the library's step events."""
    cavity = {"config": "2D2F2V cavity 64^2 x 32^2, UGKS + BGK, Kn 0.075, Maxwell walls", "steps": steps,
              "kernels": "k_face2d (FP64 face chain) + k_flux2d (face fluxes, inline slopes) + k_cell2d (update)",
              "bound": "valu: the UGKS face flux is about 100 FP32 operations per face and velocity"}
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
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 single-solve comparison")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU work of the baseline sample")
    ap.add_argument("--cavity-steps", type=int, default=20, help="timed steps of the 2D cavity leg (0 skips it)")
    ap.add_argument("--cavity-only", action="store_true", help="run only the 2D cavity leg (profiling)")
    ap.add_argument("--serve", type=int, default=48, metavar="N",
                    help="serving leg: N requests of the mixed stream (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="Poisson arrival rate of the serving leg (1/s)")
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

    # ---- FP32 batched path
    t0 = time.time()
    devs = [kb.DeviceSolve(ks, ctr, face, precision="fp32", ctx=ctx) for ks, ctr, face, _ in inst]
    log(f"[bench] fp32 upload {time.time() - t0:.1f}s")
    tuned, taken = None, 0
    if a.seg:
        ctx.set_schedule(a.seg, 0)
    elif not a.no_tune:
        # the tuner's trial steps are warm-up steps: they count toward the steps compared with FP64
        best, trials, taken = kb.autotune.tune(ctx, devs, dts)
        tuned = {"best": {"seg_cells": best[0], "streams": best[1]},
                 "ms_per_step": {f"seg{k[0]}_streams{k[1]}": round(v, 4) for k, v in trials.items()}}
        log(f"[bench] autotuner: seg_cells {best[0]}, streams {best[1]} ({taken} trial steps): "
            + ", ".join(f"{k}: {v:.3f} ms" for k, v in trials.items()))
    if W:
        kb.batch_step(devs, dts, W, ctx=ctx)
    W += taken
    log(f"[bench] warmup {W} steps in total")
    ctx.sync()
    t0 = time.perf_counter()
    kb.batch_step(devs, dts, K, ctx=ctx)
    ctx.sync()
    wall = time.perf_counter() - t0
    dev_ms = ctx.last_steps_ms()
    value = points * K / wall
    log(f"[bench] fp32 batched: {K} steps in {wall * 1e3:.2f} ms wall ({dev_ms:.2f} ms device), "
        f"{value:.4e} updates/s")
    res32 = []
    for d, (ks, ctr, face, _) in zip(devs, inst):
        c = copy.deepcopy(ctr)
        d.download(c)
        res32.append(c.prim.copy())
        d.close()
    devs = None

    # ---- FP64 plain single-solve path (same instances, same steps)
    fp64 = None
    l2 = None
    if not a.no_fp64:
        tot_wall, tot_dev = 0.0, 0.0
        worst = [0.0, 0.0, 0.0]
        worst_case = None
        worst_u_strict = 0.0
        for idx, (ks, ctr, face, dt) in enumerate(inst):
            d = kb.DeviceSolve(ks, ctr, face, precision="fp64", ctx=ctx)
            if W:
                d.step(dt, W)
            ctx.sync()
            t1 = time.perf_counter()
            d.step(dt, K)
            tot_wall += time.perf_counter() - t1
            tot_dev += ctx.last_steps_ms()
            c = copy.deepcopy(ctr)
            d.download(c)
            d.close()
            r64, u64, T64 = macro(c.prim)
            r32, u32, T32 = macro(res32[idx])
            # velocity errors relative to the characteristic speed |U| + c (the CFL speed of
            # solver_main.jl:240-288); ||dU|| / ||U|| alone is ill-conditioned where U ~ 0
            speed = np.abs(u64) + np.sqrt(0.5 * ks.gas.gamma / c.prim[:, 2])
            e = (rel_l2(r32, r64), rel_l2(u32, u64, np.linalg.norm(speed)), rel_l2(T32, T64))
            u_strict = rel_l2(u32, u64, max(np.linalg.norm(u64), 1e-300))
            if max(e) > max(worst):
                worst_case = (idx, e, ks.ps.nx, ks.vs.nv, ks.set.collision, ks.gas.Kn,
                              float(np.max(np.abs(u64))), float(np.max(np.abs(u32 - u64))))
            worst = [max(x, y) for x, y in zip(worst, e)]
            worst_u_strict = max(worst_u_strict, u_strict)
        fp64 = {"value": points * K / tot_wall, "unit": "cell-velocity updates/s",
                "device_value": points * K / (tot_dev * 1e-3),
                "path": "kbg_step per solve, FP64 storage, one solve at a time"}
        l2 = {"rho": worst[0], "U": worst[1], "T": worst[2], "tol": 1e-4,
              "U_scale": "L2 norm of |U| + sound speed",
              "U_over_norm_U": worst_u_strict, "pass": max(worst) <= 1e-4}
        log(f"[bench] fp64 plain: {fp64['value']:.4e} updates/s; ratio {value / fp64['value']:.1f}x; "
            f"max rel L2 rho {worst[0]:.2e} U {worst[1]:.2e} T {worst[2]:.2e}")
        log(f"[bench] worst instance {worst_case[0]}: errors {worst_case[1]}, nx {worst_case[2]}, nv {worst_case[3]}, "
            f"{worst_case[4]}, Kn {worst_case[5]:.3e}, max|U| {worst_case[6]:.3e}, max|dU| {worst_case[7]:.3e}")

    # ---- 2D lid-driven cavity (64^2 cells x 32^2 velocities, UGKS, Maxwell walls), fused step
    cavity = cavity_leg(kb, ctx, a.cavity_steps) if a.cavity_steps else None

    # ---- serving leg: Poisson arrivals of 1D tubes, 2D cavities and 1D1F3V requests, served together
    serve = None
    if a.serve:
        reqs = [kb.serving.Request(rid=k, cfg=cfg, nsteps=ns, arrival=t)
                for k, (kind, cfg, ns, t) in enumerate(kb.configs.serving_requests(a.serve, a.serve_rate))]
        srv = kb.serving.Server(kb.serving.DeviceBackend(kb, ctx=ctx), budget_bytes=32 << 30, chunk_steps=a.serve_chunk,
                                overlap=not a.serve_serial)
        st = srv.run(reqs)
        serve = dict(st.summary(), rate_per_s=a.serve_rate, budget_bytes=32 << 30, chunk_steps=a.serve_chunk,
                     overlap=not a.serve_serial,
                     mix="75% 1D tubes (to t=0.2), 15% 2D cavities 32^2-48^2 x 16^2-24^2 (100 steps), "
                         "10% 1D1F3V 24 cells x 24^3 (40 steps); FP32")
        log(f"[bench] serving: {serve['requests']} requests at {a.serve_rate}/s, p50 {serve['p50_ms']:.1f} ms, "
            f"p99 {serve['p99_ms']:.1f} ms, {serve['updates_per_s']:.3e} updates/s, max live {serve['max_live']}")

    # ---- CPU baseline: FP64 oracle with OpenMP on a bounded sample
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        nthreads = oracle.num_threads()
        sample = inst[: min(64, len(inst))]
        spts = sum(ks.ps.nx * ks.vs.nv for ks, *_ in sample)
        runs = [(ks, copy.deepcopy(ctr), copy.deepcopy(face), dt) for ks, ctr, face, dt in sample]
        solvers = [oracle.OracleSolve(ks, c, f) for ks, c, f, _ in runs]
        t1 = time.perf_counter()
        for o, (_, _, _, dt) in zip(solvers, runs):
            o.run(dt, 1)
        per_step = time.perf_counter() - t1
        nst = int(max(1, min(1000, a.cpu_seconds / max(per_step, 1e-6))))
        t1 = time.perf_counter()
        for o, (_, _, _, dt) in zip(solvers, runs):
            o.run(dt, nst)
        cpu_wall = time.perf_counter() - t1
        cpu = {"value": spts * nst / cpu_wall, "unit": "cell-velocity updates/s", "cores": nthreads,
               "kind": "port",
               "sample": f"FP64 oracle (C restatement of KitBase's 1D1F1V step, OpenMP over cells) on the "
                         f"first {len(sample)} of the {len(inst)} instances ({spts} points), {nst} steps each, "
                         f"{cpu_wall:.1f} s"}
        log(f"[bench] cpu oracle: {cpu['value']:.4e} updates/s on {nthreads} threads ({cpu_wall:.1f}s)")

    ms_step = dev_ms / K
    achieved = alg_bytes_step / (ms_step * 1e-3) / 1e9
    traffic = traffic_from_profiles(points)
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
                               "Kn 10^U(-4,1), BGK/Shakhov, 1D1F1V KFVS, FP32 f / FP64 w",
                   "instances": len(inst), "seed": kb.configs.BATCH_SEED, "cells": cells,
                   "points_per_step": points, "parallelism": "single GPU"},
        "schedule": tuned,
        "cavity_2d": cavity,
        "fp64_plain": fp64,
        "speedup_vs_fp64_plain": (value / fp64["value"]) if fp64 else None,
        "l2_vs_fp64": l2,
        "p99_solve_latency_ms": serve["p99_ms"] if serve else None,
        "serving": serve,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_sweep1d<float,1F,KFVS>: {a.sweep_steps} time step(s) per launch, one launch "
                               f"per stream ({ctx.schedule()[1] or 4} concurrent streams, {ctx.schedule()[0]} cells "
                               "per segment); achieved = per-update algorithmic bytes x updates / device time",
                     "steps_per_launch": a.sweep_steps,
                     "algorithmic_bytes_per_step": alg_bytes_step, "device_ms_per_step": ms_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
