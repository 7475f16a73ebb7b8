"""Schedule autotuner for the fused sweep.

The knobs, searched one after another on the caller's own batch:
  1. the context's schedule (kbg_ctx_set_schedule): cells per sweep workgroup (segment length,
     1..60) and the number of concurrent streams the batch is spread over (1..4);
  2. the most velocities per sweep wavefront (kbg_ctx_set_sweep_vw: 0 = the kernel's full width,
     64 FP32 / 32 FP64 for one distribution; 32 or 16 put more wavefronts on each work item);
  3. HIP graph replay of the step rounds against launching them one by one (kbg_ctx_set_graphs).
Each candidate is timed with the library's step-loop events (kbg_batch_step records
KBG_EV_STEPS_BEGIN / END around its steps); the fastest is kept and left set on the context.
Trials advance the solves like any other steps; tune() returns how many steps it took, so a
caller can count them as warm-up.  Results do not depend on the knobs (a solve's arithmetic and
sums keep their order in any schedule).

Results are cached per batch signature (solve count, cells, velocity counts, precision and flux
mix) in a JSON file when `cache` is given, so a serving process tunes a shape once.
"""
from __future__ import annotations

import json
from pathlib import Path

DEFAULT_CANDIDATES = [(60, 4), (60, 3), (60, 2), (60, 1), (44, 4), (30, 4)]
VW_CANDIDATES = (32, 16)


def signature(devs) -> str:
    import hashlib

    parts = sorted((d.ks.dim, d.ks.nf, d.ks.set.flux, d.precision, d.ks.vs.nv,
                    d.ks.ps.nx * (d.ks.ps.ny if d.ks.dim == 2 else 1)) for d in devs)
    return hashlib.sha1(repr(parts).encode()).hexdigest()[:16]


def _apply(ctx, cand):
    seg, streams, vw, graphs = cand
    ctx.set_schedule(seg, streams)
    ctx.set_sweep_vw(vw)
    ctx.set_graphs(bool(graphs))


def key_str(cand) -> str:
    seg, streams, vw, graphs = cand
    return f"seg{seg}_streams{streams}_vw{vw}" + ("_graphs" if graphs else "")


def tune(ctx, devs, dts, candidates=None, steps: int = 20, repeat: int = 2, cache: str | Path | None = None,
         vws=VW_CANDIDATES, graphs: bool = True):
    """Returns (best (seg_cells, streams, sweep_vw, graphs), {candidate: ms per step}, steps taken)."""
    from . import solver

    key = signature(devs)
    if cache is not None and Path(cache).exists():
        known = json.loads(Path(cache).read_text())
        if key in known:
            best = tuple(known[key]["best"])
            _apply(ctx, best)
            return best, {tuple(k): v for k, v in known[key]["ms"]}, 0
    taken = 0
    results = {}

    def trial(cand):
        nonlocal taken
        _apply(ctx, cand)
        solver.batch_step(devs, dts, 1, ctx=ctx)  # first call with this schedule: descriptors, items, graphs
        taken += 1
        best_ms = float("inf")
        for _ in range(repeat):
            solver.batch_step(devs, dts, steps, ctx=ctx)
            taken += steps
            best_ms = min(best_ms, ctx.last_steps_ms() / steps)
        results[cand] = best_ms

    for seg, streams in list(candidates or DEFAULT_CANDIDATES):
        trial((seg, streams, 0, False))
    best = min(results, key=results.get)
    for vw in vws:
        trial((best[0], best[1], vw, False))
    best = min(results, key=results.get)
    if graphs:
        trial((best[0], best[1], best[2], True))
        best = min(results, key=results.get)
    _apply(ctx, best)
    if cache is not None:
        p = Path(cache)
        known = json.loads(p.read_text()) if p.exists() else {}
        known[key] = {"best": list(best), "ms": [[list(k), v] for k, v in results.items()]}
        p.write_text(json.dumps(known, indent=1))
    return best, results, taken
