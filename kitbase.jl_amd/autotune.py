"""Schedule autotuner for the fused sweep.

The knobs are the context's schedule (kbg_ctx_set_schedule): cells per sweep workgroup
(segment length, 1..60) and the number of concurrent streams a batch is spread over (1..4).  The
tuner times each candidate on the caller's own batch with the library's step-loop events
(kbg_batch_step records KBG_EV_STEPS_BEGIN / END around its steps), keeps the fastest, and leaves it
set on the context.  Trials advance the solves like any other steps; tune() returns how many steps
it took, so a caller can count them as warm-up.

Results are cached per batch signature (solve count, cells, velocity counts, precision and flux
mix) in a JSON file when `cache` is given, so a serving process tunes a shape once.
"""
from __future__ import annotations

import json
from pathlib import Path

DEFAULT_CANDIDATES = [(60, 4), (60, 3), (60, 2), (60, 1), (44, 4), (30, 4)]


def signature(devs) -> str:
    import hashlib

    parts = sorted((d.ks.dim, d.ks.nf, d.ks.set.flux, d.precision, d.ks.vs.nv,
                    d.ks.ps.nx * (d.ks.ps.ny if d.ks.dim == 2 else 1)) for d in devs)
    return hashlib.sha1(repr(parts).encode()).hexdigest()[:16]


def tune(ctx, devs, dts, candidates=None, steps: int = 20, repeat: int = 2, cache: str | Path | None = None):
    """Returns (best (seg_cells, streams), {candidate: ms per step}, steps taken)."""
    from .solver import batch_step

    cands = list(candidates or DEFAULT_CANDIDATES)
    key = signature(devs)
    if cache is not None and Path(cache).exists():
        known = json.loads(Path(cache).read_text())
        if key in known:
            best = tuple(known[key]["best"])
            ctx.set_schedule(*best)
            return best, {tuple(map(int, k.split(","))): v for k, v in known[key]["ms"].items()}, 0
    taken = 0
    results = {}
    for c in cands:
        ctx.set_schedule(*c)
        batch_step(devs, dts, 1, ctx=ctx)  # first call with this schedule: descriptors, items
        taken += 1
        best_ms = float("inf")
        for _ in range(repeat):
            batch_step(devs, dts, steps, ctx=ctx)
            taken += steps
            best_ms = min(best_ms, ctx.last_steps_ms() / steps)
        results[c] = best_ms
    best = min(results, key=results.get)
    ctx.set_schedule(*best)
    if cache is not None:
        p = Path(cache)
        known = json.loads(p.read_text()) if p.exists() else {}
        known[key] = {"best": list(best), "ms": {f"{a},{b}": v for (a, b), v in results.items()}}
        p.write_text(json.dumps(known, indent=1))
    return best, results, taken
