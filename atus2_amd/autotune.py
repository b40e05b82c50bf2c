"""Autotuner for the FFT kernels: picks, per kernel family, the number of
values each thread holds (8 or 16), and for 3D grids whether the y/z
passes run slab by slab (atus_set_slab), by timing profiled blocks of the
actual workload on the actual context (HIP events on the context's
stream).

The blocks it runs advance the context's state like any other block, so it
belongs in the warm-up of a run.  Choices are cached per (grid shape,
components, step kind) for the life of the process.
"""
from __future__ import annotations

import time

from . import hip

FAMILIES = ["pass_contig", "pass_strided", "kin", "vfuse"]
_cache: dict = {}


def autotune(ctx: "hip.Context", kind: int, dt: float, t: float, las=None, amp=None, Nk: int = 10,
             reps: int = 2):
    """Returns (t after the tuning blocks, {family: values_per_thread})."""
    key = (ctx.shape, ctx.ncomp, kind)
    if key in _cache:
        _apply(ctx, _cache[key])
        return t, dict(_cache[key])
    best = {f: (float("inf"), 16) for f in FAMILIES}
    ctx.set_slab(0)
    for rm in (16, 8):
        for f in FAMILIES:
            ctx.tune(f, rm)
        for _ in range(reps):
            t, prof = ctx.run_block_profiled(kind, dt, Nk, t, las, amp)
            for f in FAMILIES:
                ms, n, grid = prof[f]
                if n:
                    avg = ms / grid     # per whole-grid pass
                    if avg < best[f][0]:
                        best[f] = (avg, rm)
    choice = {f: best[f][1] for f in FAMILIES}
    _apply(ctx, choice)
    if len(ctx.shape) == 3:
        # slab-chained passes on two streams against whole-grid passes: wall
        # time of ordinary blocks (profiled blocks run on one stream)
        total = {}
        for slab in (0, -1):
            ctx.set_slab(slab)
            t = ctx.run_block(kind, dt, Nk, t, las, amp)
            ctx.synchronize()
            total[slab] = float("inf")
            for _ in range(reps):
                tic = time.perf_counter()
                t = ctx.run_block(kind, dt, Nk, t, las, amp)
                ctx.synchronize()
                total[slab] = min(total[slab], time.perf_counter() - tic)
        choice["slab"] = min(total, key=total.get)
    _apply(ctx, choice)
    _cache[key] = dict(choice)
    return t, choice


def _apply(ctx, choice):
    for f, v in choice.items():
        if f == "slab":
            ctx.set_slab(v)
        else:
            ctx.tune(f, v)
