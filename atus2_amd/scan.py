"""Parameter scans of the two-component Bragg pulse: many independent runs
(different grids, Rabi amplitudes, pulse widths, detunings and couplings)
advanced together on one GPU.

The reference scans by launching one bragg process per case
(tammy/src/batch_case_generation_4_tammy.cpp:45-291); each case runs the
bragg_ad sequence of bragg.cpp with the Rabi populations of component 0
after every block (CRT_Base_IF.h:208-314).  Here the runs of one grid shape
share a batch context (atus_batch_*), so that every launch advances all of
them, and the amplitudes of a block are evaluated for all runs at once.
The blocks of all shapes, with their Rabi populations, are captured as one
HIP graph (atus_scan_graph_*): each shape's launches on its own stream, so
the device runs the small 1D groups beside the large 2D ones, and a block
of the whole scan is one graph launch.  Each shape's kernels take the
autotuner's values per thread and threads per workgroup.

`python -m atus2_amd.scan` times the 512-run scan three ways (the graph,
the shape groups one after another with eager launches, and every run
one after another through the single-run path) and prints one JSON line;
with --family-times it times each kernel family of each shape group
instead, one group at a time.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import hip
from .workloads import axis_coords, grid_spacings, particle_number


@dataclass
class ScanRun:
    """One case: grid shape on [-L/2, L/2) per axis, Omega_peak (amp0),
    pulse width sigma of AMP_T = exp(-((t - t0)/sigma)^2 / 2), detuning
    laser_domh and couplings g11 = g22 = g, g12 = g21 = 0.95 g."""
    shape: tuple
    amp0: float
    sigma: float
    laser_domh: float
    g: float
    L: float = 64.0
    t0: float = 0.3
    laser_k: float = 4.0

    @property
    def gs(self):
        return [self.g, 0.95 * self.g, 0.95 * self.g, self.g]

    def amp(self, t):
        return math.exp(-0.5 * ((t - self.t0) / self.sigma) ** 2)

    def laser(self):
        return hip.laser(amp0=self.amp0, laser_k=self.laser_k, laser_domh=self.laser_domh, deltaL=(0.0, 0.0))

    def rabi_k0(self):
        k = 2 * self.laser_k
        return [(0.0,) * len(self.shape), (k,) + (0.0,) * (len(self.shape) - 1),
                (-k,) + (0.0,) * (len(self.shape) - 1)]

    def initial_state(self):
        """psi1: Gaussian of widths (6, 4) normalised to 1; psi2 = 0."""
        widths = (6.0, 4.0, 4.0)
        psi = np.ones(self.shape)
        for a, n in enumerate(self.shape):
            x = axis_coords(n, self.L)
            shp = [1] * len(self.shape)
            shp[a] = n
            psi = psi * np.exp(-x ** 2 / (2 * widths[a] ** 2)).reshape(shp)
        psi = psi.astype(np.complex128)
        dx, _ = grid_spacings(self.shape, self.L)
        psi /= math.sqrt(particle_number(psi, dx))
        return psi, np.zeros_like(psi)


def scan_runs(seed=4, n1d=384, n2d=128):
    """The 512-case scan: 1D n in {1024 .. 16384}, 2D n^2 in {256^2 .. 1024^2},
    Omega_peak ~ U[1,10], sigma ~ U[0.2,1], detuning ~ 8 + U[-2,2],
    g ~ U[1,200]."""
    rng = np.random.default_rng(seed)
    runs = []
    for i in range(n1d + n2d):
        if i < n1d:
            shape = (int(rng.choice([1024, 2048, 4096, 8192, 16384])),)
        else:
            n = int(rng.choice([256, 512, 1024]))
            shape = (n, n)
        runs.append(ScanRun(shape=shape, amp0=rng.uniform(1.0, 10.0), sigma=rng.uniform(0.2, 1.0),
                            laser_domh=8.0 + rng.uniform(-2.0, 2.0), g=rng.uniform(1.0, 200.0)))
    return runs


RABI_THRESHOLD = 2.0


def run_batched(runs, dt=2e-4, Nk=100, steps=3000, max_batch=4096):
    """All runs, grouped by grid shape; returns ({run index: Rabi populations
    per block, shape (blocks, 3)}, seconds of the block loops)."""
    groups = {}
    for i, r in enumerate(runs):
        groups.setdefault(r.shape, []).append(i)
    out = {}
    elapsed = 0.0
    for shape, idx in sorted(groups.items()):
        for s0 in range(0, len(idx), max_batch):
            part = idx[s0:s0 + max_batch]
            rs = [runs[i] for i in part]
            dx, dk = grid_spacings(shape, rs[0].L)
            with hip.BatchContext(shape, dx, dk, len(rs)) as B:
                for j, r in enumerate(rs):
                    p1, p2 = r.initial_state()
                    B.upload(2 * j, p1)
                    B.upload(2 * j + 1, p2)
                B.set_alpha([0.5] * len(shape))
                B.set_runs([r.gs for r in rs], [r.laser() for r in rs])
                k0 = rs[0].rabi_k0()
                sig = np.array([r.sigma for r in rs])
                t0s = np.array([r.t0 for r in rs])
                series = []
                B.synchronize()
                tic = time.perf_counter()
                t = 0.0
                for _ in range(steps // Nk):
                    t1 = np.array(B.step_times(t, dt, Nk))[:, None]
                    F = np.exp(-0.5 * ((t1 - t0s[None, :]) / sig[None, :]) ** 2)
                    t = B.run_batch_block(hip.STEP_BRAGG_AD, dt, Nk, t, F)
                    series.append(B.batch_rabi_integrals(0, k0, RABI_THRESHOLD))
                B.synchronize()
                elapsed += time.perf_counter() - tic
            series = np.stack(series, axis=1)  # runs x blocks x 3
            for j, i in enumerate(part):
                out[i] = series[j]
    return out, elapsed


def make_contexts(runs, dt=2e-4, Nk=100, tune=True, log=None):
    """One batch context per grid shape, loaded with its runs' initial
    states; with tune=True each takes the autotuner's choice (values per
    thread and threads per workgroup per kernel family), measured on a
    scratch context of the same runs.  Returns [(shape, run indices, ctx)]."""
    from .autotune import _apply, autotune
    groups = {}
    for i, r in enumerate(runs):
        groups.setdefault(r.shape, []).append(i)
    out = []
    for shape, idx in sorted(groups.items()):
        rs = [runs[i] for i in idx]
        dx, dk = grid_spacings(shape, rs[0].L)

        def load(B):
            for j, r in enumerate(rs):
                p1, p2 = r.initial_state()
                B.upload(2 * j, p1)
                B.upload(2 * j + 1, p2)
            B.set_alpha([0.5] * len(shape))
            B.set_runs([r.gs for r in rs], [r.laser() for r in rs])

        B = hip.BatchContext(shape, dx, dk, len(rs))
        load(B)
        if tune:
            with hip.BatchContext(shape, dx, dk, len(rs)) as T:
                load(T)
                _, choice = autotune(T, hip.STEP_BRAGG_AD, dt, 0.0, Nk=10, reps=2)
            _apply(B, choice)
            if log is not None:
                log(f"{shape} x {len(rs)}: {choice}")
        out.append((shape, idx, B))
    return out


def run_graph(runs, dt=2e-4, Nk=100, steps=3000, tune=True, log=None):
    """All runs at once: one batch context per grid shape, and a block of
    every context (with its Rabi populations) captured as one HIP graph
    whose branches the device runs side by side (atus_scan_graph_*).
    Returns ({run index: Rabi populations per block}, seconds of the block
    loop, seconds of device time inside the graph launches)."""
    groups = make_contexts(runs, dt, Nk, tune, log)
    ctxs = [B for _, _, B in groups]
    series = [[] for _ in groups]
    try:
        k0 = runs[0].rabi_k0()
        sig = [np.array([runs[i].sigma for i in idx]) for _, idx, _ in groups]
        t0s = [np.array([runs[i].t0 for i in idx]) for _, idx, _ in groups]
        with hip.ScanGraph(ctxs, hip.STEP_BRAGG_AD, dt, Nk, k0, RABI_THRESHOLD) as G:
            for B in ctxs:
                B.synchronize()
            gpu = 0.0
            tic = time.perf_counter()
            t = 0.0
            for _ in range(steps // Nk):
                t1 = np.array(hip.BatchContext.step_times(t, dt, Nk))[:, None]
                F = [np.exp(-0.5 * ((t1 - t0s[g][None, :]) / sig[g][None, :]) ** 2) for g in range(len(groups))]
                t, rab, ms = G.launch(t, F)
                gpu += ms * 1e-3
                for g in range(len(groups)):
                    series[g].append(rab[g])
            elapsed = time.perf_counter() - tic
    finally:
        for B in ctxs:
            B.close()
    out = {}
    for g, (_, idx, _) in enumerate(groups):
        s_ = np.stack(series[g], axis=1)  # runs x blocks x 3
        for j, i in enumerate(idx):
            out[i] = s_[j]
    return out, elapsed, gpu


def run_single(runs, dt=2e-4, Nk=100, steps=3000, log=None):
    """The same runs one after another through the single-run path (the
    amplitude by callback per step, as CRT_Base_IF evaluates AMP_T)."""
    out = {}
    elapsed = 0.0
    for i, r in enumerate(runs):
        dx, dk = grid_spacings(r.shape, r.L)
        with hip.Context(r.shape, dx, dk, 2) as S:
            p1, p2 = r.initial_state()
            S.upload(0, p1)
            S.upload(1, p2)
            S.set_alpha([0.5] * len(r.shape))
            S.set_g(r.gs)
            las = r.laser()
            k0 = r.rabi_k0()
            series = []
            S.synchronize()
            tic = time.perf_counter()
            t = 0.0
            for _ in range(steps // Nk):
                t = S.run_block(hip.STEP_BRAGG_AD, dt, Nk, t, las, r.amp)
                series.append(S.rabi_integrals(0, k0, RABI_THRESHOLD))
            S.synchronize()
            elapsed += time.perf_counter() - tic
        out[i] = np.array(series)
        if log is not None and (i + 1) % 64 == 0:
            log(f"{i + 1} runs, {elapsed:.2f} s")
    return out, elapsed


def family_times(runs, dt=2e-4, Nk=20, reps=3, tune=True, log=None):
    """Per shape group, alone on the device: device time of each kernel
    family per whole-grid pass (the best of `reps` profiled blocks, HIP
    events around every launch) and its rate in algorithmic bytes (a pass
    reads and writes both components of every run: 64 B per point).
    Returns a list of dicts, one per shape."""
    out = []
    for shape, idx, B in make_contexts(runs, dt, 100, tune, log):
        try:
            rs = [runs[i] for i in idx]
            sig = np.array([r.sigma for r in rs])
            t0s = np.array([r.t0 for r in rs])
            t = 0.0
            best = {}
            for rep in range(reps + 1):
                t1 = np.array(B.step_times(t, dt, Nk))[:, None]
                F = np.exp(-0.5 * ((t1 - t0s[None, :]) / sig[None, :]) ** 2)
                t, prof = B.run_batch_block_profiled(hip.STEP_BRAGG_AD, dt, Nk, t, F)
                if rep == 0:
                    continue  # first block: kernel loading and LDS set-up
                for f, (ms, n, grid) in prof.items():
                    if n:
                        best[f] = min(best.get(f, float("inf")), ms / grid)
            gb = 64.0 * len(idx) * math.prod(shape) / 1e9
            out.append({"shape": list(shape), "runs": len(idx),
                        "families": {f: {"us_per_grid": 1e3 * ms, "GBps": gb / (ms * 1e-3)}
                                     for f, ms in sorted(best.items())}})
        finally:
            B.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--n1d", type=int, default=384)
    ap.add_argument("--n2d", type=int, default=128)
    ap.add_argument("--no-single", action="store_true", help="skip the one-after-another baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the eager group-by-group batched run")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of the autotuner's choices")
    ap.add_argument("--family-times", action="store_true",
                    help="only time each kernel family per shape group, one group at a time (JSON line per shape)")
    a = ap.parse_args(argv)
    runs = scan_runs(n1d=a.n1d, n2d=a.n2d)
    pts = sum(math.prod(r.shape) for r in runs) * a.steps

    def log(m):
        print(f"[scan] {m}", file=sys.stderr, flush=True)

    if a.family_times:
        for row in family_times(runs, tune=not a.no_tune, log=log):
            print(json.dumps(row), flush=True)
        return

    log(f"{len(runs)} runs, {a.steps} steps each")
    rg, tg, gpu = run_graph(runs, steps=a.steps, tune=not a.no_tune, log=log)
    log(f"one graph per block: {tg:.2f} s, device busy {gpu / tg:.3f}")
    res = {"workload": f"{len(runs)} Bragg runs ({a.n1d} 1D n=1024..16384, {a.n2d} 2D 256^2..1024^2), "
                       f"{a.steps} steps of dt=2e-4, Rabi populations every 100 steps",
           "graph_s": tg, "graph_Gpt_steps_per_s": pts / tg / 1e9, "gpu_busy": gpu / tg}
    if not a.no_batched:
        rb, tb = run_batched(runs, steps=a.steps)
        log(f"groups one after another, eager: {tb:.2f} s")
        res.update({"batched_s": tb, "batched_Gpt_steps_per_s": pts / tb / 1e9,
                    "graph_vs_batched_rabi_difference": max(float(np.max(np.abs(rg[i] - rb[i])))
                                                            for i in range(len(runs)))})
    if not a.no_single:
        rs, ts = run_single(runs, steps=a.steps, log=lambda m: log(f"one after another: {m}"))
        log(f"one after another: {ts:.2f} s")
        diff = max(float(np.max(np.abs(rg[i] - rs[i]))) for i in range(len(runs)))
        res.update({"single_s": ts, "single_Gpt_steps_per_s": pts / ts / 1e9, "speedup": ts / tg,
                    "max_rabi_difference": diff})
    print(json.dumps(res))


if __name__ == "__main__":
    main()
