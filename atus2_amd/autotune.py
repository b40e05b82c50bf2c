"""Autotuner for the FFT kernels: picks, per kernel family, the number of
values each thread holds (8 or 16: the radix of the transform stages) and
the threads per workgroup (the pencils, or tile of columns, a workgroup
takes), and for 3D grids whether the y/z passes run slab by slab
(atus_set_slab), by timing profiled blocks of the actual workload on the
actual context (HIP events on the context's stream).

The blocks it runs advance the context's state like any other block, so it
belongs in the warm-up of a run.  Choices are cached per (grid shape,
components, step kind) for the life of the process.
"""
from __future__ import annotations

import time

from . import hip

FAMILIES = ["pass_contig", "pass_strided", "kin", "vfuse"]
_cache: dict = {}


THREADS = (0, 128, 64)   # 0: as many as the kernel allows (256, or 512 for long columns)


def autotune(ctx: "hip.Context", kind: int, dt: float, t: float, las=None, amp=None, Nk: int = 10,
             reps: int = 2):
    """Returns (t after the tuning blocks, choice): choice[family] = values
    per thread, choice["threads"][family] = threads per workgroup, and for
    3D single-run grids choice["slab"].  A BatchContext is tuned on blocks
    of all its runs with unit amplitudes."""
    key = (ctx.shape, ctx.ncomp, kind, type(ctx).__name__)
    if key in _cache:
        _apply(ctx, _cache[key])
        return t, _copy(_cache[key])
    best = {f: (float("inf"), 16, 0) for f in FAMILIES}
    ctx.set_slab(0)
    for rm in (16, 8):
        for th in THREADS:
            for f in FAMILIES:
                ctx.tune(f, rm)
                ctx.tune_threads(f, th)
            for _ in range(reps):
                t, prof = _profiled(ctx, kind, dt, Nk, t, las, amp)
                for f in FAMILIES:
                    ms, n, grid = prof[f]
                    if n:
                        avg = ms / grid     # per whole-grid pass
                        if avg < best[f][0]:
                            best[f] = (avg, rm, th)
    choice = {f: best[f][1] for f in FAMILIES}
    choice["threads"] = {f: best[f][2] for f in FAMILIES}
    _apply(ctx, choice)
    if len(ctx.shape) == 3 and not isinstance(ctx, hip.BatchContext):
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
    _cache[key] = _copy(choice)
    return t, choice


def _profiled(ctx, kind, dt, Nk, t, las, amp):
    """One profiled block: per family (ms, launches, grid share)."""
    if isinstance(ctx, hip.BatchContext):
        return ctx.run_batch_block_profiled(kind, dt, Nk, t)
    return ctx.run_block_profiled(kind, dt, Nk, t, las, amp)


def _copy(choice):
    out = dict(choice)
    if "threads" in out:
        out["threads"] = dict(out["threads"])
    return out


def _apply(ctx, choice):
    for f, v in choice.items():
        if f == "slab":
            ctx.set_slab(v)
        elif f == "threads":
            for g, th in v.items():
                ctx.tune_threads(g, th)
        else:
            ctx.tune(f, v)
