"""Autotuner for the FFT kernels: picks, per kernel family, the number of
values each thread holds (8 or 16: the radix of the transform stages) and
the threads per workgroup (the pencils, or tile of columns, a workgroup
takes), and for 3D grids whether the y/z passes run slab by slab
(atus_set_slab), by timing profiled blocks of the actual workload on the
actual context (HIP events on the context's stream).  On the slab schedule
the families of the slab chains are then re-timed by the wall time of
ordinary blocks, where their launches share the device with the other
stream's.

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
        def wall(t):
            t = ctx.run_block(kind, dt, Nk, t, las, amp)
            ctx.synchronize()
            best_s = float("inf")
            for _ in range(reps):
                tic = time.perf_counter()
                t = ctx.run_block(kind, dt, Nk, t, las, amp)
                ctx.synchronize()
                best_s = min(best_s, time.perf_counter() - tic)
            return t, best_s

        total = {}
        for slab in (0, -1):
            ctx.set_slab(slab)
            t, total[slab] = wall(t)
        choice["slab"] = min(total, key=total.get)
        if choice["slab"] != 0:
            # on the slab schedule a pass shares the device with the other
            # stream's launches, so the fastest pass alone need not give the
            # fastest step: each family of the slab chains in turn takes the
            # (values per thread, threads) that minimises the block's wall
            # time with the others fixed
            ctx.set_slab(choice["slab"])
            cur = total[choice["slab"]]
            for f in ("pass_strided", "vfuse"):
                for rm in (16, 8):
                    for th in THREADS:
                        if (rm, th) == (choice[f], choice["threads"][f]):
                            continue
                        ctx.tune(f, rm)
                        ctx.tune_threads(f, th)
                        t, w = wall(t)
                        if w < cur:
                            cur = w
                            choice[f] = rm
                            choice["threads"][f] = th
                ctx.tune(f, choice[f])
                ctx.tune_threads(f, choice["threads"][f])
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
