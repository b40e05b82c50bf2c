#!/usr/bin/env python3
"""Benchmark: 3D two-component Bragg pulse, 256^3, complex128, on one MI355X.

A step is one full Strang step of both components (one Bragg 2x2 step plus
one kinetic step); the block structure is the reference's hot loop (Nk = 100
steps per block, Rabi populations of component 0 after every block).
Prints one JSON line with the throughput in Gpt-steps/s (10^9 grid points
advanced by one step per second), the roofline of the dominant kernel
(HIP-event timing on the context's stream) and a CPU baseline (the oracle,
on a bounded sample of the same workload).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# HBM bytes per step of the 256^3 workload, per schedule, from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel of known size
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_bragg256.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_run(args, wl):
    from atus2_amd import hip
    shape = wl.shape
    dx, dk = wl.spacings
    psi1, psi2 = wl.initial_state()
    las = hip.laser(amp0=wl.amp0, laser_k=wl.laser_k, laser_domh=wl.laser_domh, deltaL=(0.0, 0.0))
    ctx = hip.Context(shape, dx, dk, 2)
    ctx.upload(0, psi1)
    ctx.upload(1, psi2)
    ctx.set_g(wl.gs)
    ctx.set_alpha(wl.alpha)
    npts = float(math.prod(shape))
    N0 = ctx.particle_number(0) + ctx.particle_number(1)
    t = 0.0
    choice = {}
    if args.tuning:
        # fixed choices, e.g. "pass_strided=8,vfuse=8,slab=-1,threads.vfuse=64"
        for item in args.tuning.split(","):
            fam, v = (w.strip() for w in item.split("="))
            if fam == "slab":
                ctx.set_slab(int(v))
                choice[fam] = int(v)
            elif fam.startswith("threads."):
                ctx.tune_threads(fam.split(".", 1)[1], int(v))
                choice.setdefault("threads", {})[fam.split(".", 1)[1]] = int(v)
            else:
                ctx.tune(fam, int(v))
                choice[fam] = int(v)
        log(f"[bench] fixed choices: {choice}")
    elif not args.no_autotune:
        from atus2_amd.autotune import autotune
        t, choice = autotune(ctx, hip.STEP_BRAGG_AD, wl.dt, t, las, wl.amp, Nk=10, reps=2)
        log(f"[bench] autotuned values per thread: {choice}")
    log(f"[bench] warmup {args.warmup} steps")
    k0 = wl.rabi_k0()
    left = args.warmup
    while left > 0:
        # the timed loop's work: a block, then the Rabi populations (whose
        # first call allocates the transform's buffer)
        nk = min(wl.Nk, left)
        t = ctx.run_block(hip.STEP_BRAGG_AD, wl.dt, nk, t, las, wl.amp)
        ctx.rabi_integrals(0, k0, wl.rabi_threshold)
        left -= nk
    ctx.synchronize()
    log(f"[bench] timing {args.steps} steps")
    rabi = None
    t0 = time.perf_counter()
    left = args.steps
    while left > 0:
        nk = min(wl.Nk, left)
        t = ctx.run_block(hip.STEP_BRAGG_AD, wl.dt, nk, t, las, wl.amp)
        rabi = ctx.rabi_integrals(0, k0, wl.rabi_threshold)
        left -= nk
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    N1 = ctx.particle_number(0) + ctx.particle_number(1)
    nsteps_done = args.warmup + args.steps
    log(f"[bench] {elapsed:.3f} s for {args.steps} steps; norm drift {abs(N1 - N0):.2e} over {nsteps_done} steps")

    # per-kernel device times of one block (HIP events on the context stream)
    t, prof = ctx.run_block_profiled(hip.STEP_BRAGG_AD, wl.dt, wl.Nk, t, las, wl.amp)
    fams = {k: v for k, v in prof.items() if v[1] > 0}
    dom = max(fams, key=lambda k: fams[k][0])
    ms_tot, nl, grid = fams[dom]
    avg_ms = ms_tot / nl
    # every pass family reads and writes each point of both components once
    # per whole-grid pass; a launch over a slab of x-planes moves its share
    grid_bytes = 2 * npts * 32.0
    bytes_per_launch = grid_bytes * grid / nl
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"avg_us": 1e3 * v[0] / v[1], "launches": v[1], "grid_share_per_launch": v[2] / v[1],
                   "us_per_grid": 1e3 * v[0] / v[2], "GBps": grid_bytes * v[2] / (v[0] * 1e-3) / 1e9}
               for k, v in fams.items()}
    copy_gbps = ctx.copy_bandwidth(20)
    log(f"[bench] copy probe {copy_gbps:.0f} GB/s; dominant kernel {dom} {achieved:.0f} GB/s")
    ctx.close()
    return {
        "elapsed": elapsed,
        "norm_drift_per_step": abs(N1 - N0) / nsteps_done,
        "rabi": [float(v) for v in rabi],
        "dominant": dom,
        "dom_avg_ms": avg_ms,
        "dom_share": grid / nl,
        "achieved": achieved,
        "bytes_per_launch": bytes_per_launch,
        "kernels": kernels,
        "copy_gbps": copy_gbps,
        "tuning": choice,
    }


def rt_solver_run(wl, steps, energy, tune=True):
    """One-component rt_solver freeprop (V + g|psi|^2, kinetic step) on the
    GPU: `steps` steps in blocks of Nk after the autotuner's blocks and one
    untimed block, the norm (and the energy) after every block."""
    from atus2_amd import hip
    dx, dk = wl.spacings
    npts = float(math.prod(wl.shape))
    ctx = hip.Context(wl.shape, dx, dk, 1)
    try:
        ctx.upload(0, wl.initial_state())
        ctx.set_g([wl.g])
        ctx.set_alpha([0.5] * len(wl.shape))
        ctx.set_potential(0, wl.potential())
        if getattr(wl, "momentum", None):
            ctx.set_momentum(0, wl.momentum)
        N0 = ctx.particle_number(0)
        t, choice = 0.0, {}
        nsteps = steps + wl.Nk
        if tune:
            from atus2_amd.autotune import autotune
            t, choice = autotune(ctx, hip.STEP_FREEPROP, wl.dt, t, Nk=10, reps=2)
            nsteps += 2 * 2 * 10 + (2 * (1 + 2) * 10 if len(wl.shape) == 3 else 0)
        t = ctx.run_block(hip.STEP_FREEPROP, wl.dt, wl.Nk, t)
        ctx.synchronize()
        t0 = time.perf_counter()
        left = steps
        while left > 0:
            nk = min(wl.Nk, left)
            t = ctx.run_block(hip.STEP_FREEPROP, wl.dt, nk, t)
            ctx.particle_number(0)
            if energy:
                ctx.energy()
            left -= nk
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        drift = abs(ctx.particle_number(0) - N0) / nsteps
        t, prof = ctx.run_block_profiled(hip.STEP_FREEPROP, wl.dt, wl.Nk, t)
    finally:
        ctx.close()
    kernels = {k: {"avg_us": 1e3 * v[0] / v[1], "launches": v[1]} for k, v in prof.items() if v[1] > 0}
    return npts * steps / elapsed / 1e9, 1e3 * elapsed / steps, drift, kernels, choice


def other_configs_run(tune=True):
    """The reference's 1D (4096, V = x^2/2, g = 10) and 2D (2048^2 double
    slit) rt_solver workloads, with the bytes per point and step of the
    reference's unfused split step (one nonlinear sweep and two sweeps per
    axis, plus the potential) and of this design: in 2D one pass per axis,
    the nonlinear step and the kinetic phase fused into them, plus the
    potential; in 1D the block runs in one launch with the state in
    registers, so the state moves once per block of Nk steps and the
    potential is read every step."""
    from atus2_amd.workloads import DoubleSlitWorkload, HarmonicLineWorkload
    out = []
    for wl, steps, energy, name, bound in (
            (HarmonicLineWorkload(), 1000, True,
             "1D rt_solver: 4096 on [-40,40), V = x^2/2, g = 10, dt = 1e-3, norm and energy per 100 steps",
             "latency: one workgroup holds the whole run"),
            (DoubleSlitWorkload(), 500, False,
             "2D rt_solver double slit: 2048^2 on [-64,64)^2, g = 1, dt = 5e-4, norm per 100 steps",
             "transform latency of 2048-point pencils; the state and potential (96 MiB) stay in the "
             "256 MiB last-level cache")):
        d = len(wl.shape)
        ref_bytes = (1 + 2 * d) * 32 + 8
        own_bytes = 32.0 / wl.Nk + 8 if d == 1 else d * 32 + 8
        value, ms, drift, kernels, choice = rt_solver_run(wl, steps, energy, tune)
        npts = float(math.prod(wl.shape))
        gbps = own_bytes * npts / (ms * 1e-3) / 1e9
        log(f"[bench] {name}: {value:.3f} Gpt-steps/s, {ms * 1e3:.1f} us per step")
        out.append({"workload": name, "grid": list(wl.shape), "steps": steps, "value": value,
                    "unit": "Gpt-steps/s", "ms_per_step": ms, "norm_drift_per_step": drift,
                    "bytes_per_point_step": {"reference_model": ref_bytes, "this_design": own_bytes},
                    "achieved_GBps": gbps, "frac": gbps / HBM_PEAK_GBS, "bound": bound,
                    "state_MiB": npts * 24 / 2 ** 20, "kernels": kernels, "tuning": choice})
    return out


def scan_run(steps):
    """The 512-run Bragg scan (seed 4: 384 1D and 128 2D runs) for `steps`
    steps per run: all shape groups in one captured graph per block
    (autotuned), and the same runs one after another."""
    import numpy as np
    from atus2_amd import scan
    runs = scan.scan_runs()
    pts = sum(math.prod(r.shape) for r in runs) * steps
    rg, tg, gpu = scan.run_graph(runs, steps=steps)
    rs, ts = scan.run_single(runs, steps=steps)
    diff = max(float(np.max(np.abs(rg[i] - rs[i]))) for i in range(len(runs)))
    log(f"[bench] scan of {len(runs)} runs x {steps} steps: one graph per block {tg:.2f} s "
        f"(device busy {gpu / tg:.3f}), one after another {ts:.2f} s")
    return {"runs": len(runs), "steps_per_run": steps, "graph_s": tg, "gpu_busy": gpu / tg, "single_s": ts,
            "speedup": ts / tg, "graph_Gpt_steps_per_s": pts / tg / 1e9, "max_rabi_difference": diff,
            "sample": f"{steps} of the configuration's 3000 steps per run (python -m atus2_amd.scan runs all); "
                      f"autotuning of each shape group not timed"}


def measured_step_traffic(n, tuning):
    """HBM bytes per Strang step measured by counters (FETCH_SIZE +
    WRITE_SIZE over every launch of the step's kernels, calibrated by a copy
    kernel) for the schedule the bench ran, or None when the table holds no
    measurement of that schedule."""
    if n != 256 or not os.path.exists(TRAFFIC_TABLE):
        return None
    with open(TRAFFIC_TABLE) as f:
        tab = json.load(f)
    def norm(t):
        # threads per workgroup default to 0 (automatic) for every family
        out = {k: int(v) for k, v in t.items() if k != "threads"}
        out["threads"] = {f: int(v) for f, v in t.get("threads", {}).items() if int(v) != 0}
        return out
    for entry in tab.get("steps", []):
        if norm(entry["tuning"]) == norm(tuning):
            return entry["bytes_per_step"]
    return None


def ground_state_run(ncases):
    """Sobolev ground states of held-out traps at 128^3 (res <= 1e-10):
    GPU wall time from the cold start and from the learned warm start."""
    from atus2_amd import learned as Lr
    net = Lr.load_net()
    rows = []
    try:
        for c in Lr.held_out(0, ncases):
            ic, tc, _, _ = Lr.solve_large(c, "cold")
            iw, tw, _, _ = Lr.solve_large(c, "warm", net)
            rows.append((ic, tc, iw, tw))
            log(f"[bench] ground state w={tuple(round(v, 2) for v in c.w)} g={c.g:.0f}: "
                f"cold {ic} its {tc:.2f} s, warm {iw} its {tw:.2f} s")
    finally:
        net.close()
    cold = sum(r[1] for r in rows)
    warm = sum(r[3] for r in rows)
    return {"grid": [128, 128, 128], "eps": 1e-10, "cases": ncases, "cold_s": cold, "warm_s": warm,
            "warm_over_cold": warm / cold, "cold_iters": [r[0] for r in rows], "warm_iters": [r[2] for r in rows],
            "sample": "held-out (w, g) of seed 55; warm time includes the network and the interpolation"}


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or 'unknown'."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, steps, reps=3):
    """The oracle on the same 256^3 Bragg workload: `reps` blocks of `steps`
    steps, the median block time."""
    from oracle import oracle as O
    g = O.make_grid(wl.shape, L=wl.L)
    psi1, psi2 = wl.initial_state()
    S = O.RTSolver(g, [psi1, psi2], list(wl.gs), list(wl.alpha), dt=wl.dt, amp=wl.amp,
                   beta0=0.0, deltaL=[0.0, 0.0], amp0=wl.amp0, laser_k=wl.laser_k,
                   laser_domh=wl.laser_domh)
    del psi1, psi2
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        S.block(O.BRAGG_AD, steps)
        times.append(time.perf_counter() - t0)
    el = sorted(times)[len(times) // 2]
    npts = float(math.prod(wl.shape))
    return {"value": npts * steps / el / 1e9, "unit": "Gpt-steps/s", "cores": O.threads(), "kind": "port",
            "sample": f"median of {reps} blocks of {steps} Strang steps of the 256^3 two-component Bragg workload "
                      f"({', '.join(f'{x:.2f}' for x in times)} s), radix-2 OpenMP FFT oracle on {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="",
                    help="fixed choices, e.g. pass_strided=8,vfuse=8,slab=-1,threads.vfuse=64")
    ap.add_argument("--gs-cases", type=int, default=4, help="ground states timed cold and warm (0: none)")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the 1D 4096 and 2D 2048^2 rt_solver workloads")
    ap.add_argument("--scan-steps", type=int, default=300,
                    help="steps per run of the 512-run scan, batched and one after another (0: none)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the propagator has no multi-GPU path "
              "(it would shard by x-slabs with one all-to-all per transform); rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from atus2_amd.workloads import BraggWorkload
    wl = BraggWorkload(n=args.n)
    r = gpu_run(args, wl)
    npts = float(math.prod(wl.shape))
    value = npts * args.steps / r["elapsed"] / 1e9
    # a step is four passes, each reading and writing both components once
    # (256 B per point and step); its HBM traffic is measured by counters
    # for the schedule that ran (None if that schedule was not measured)
    step_bytes = 4 * 2 * 32.0 * math.prod(wl.shape)
    step_gbps = step_bytes / (r["elapsed"] / args.steps) / 1e9
    step_traffic = measured_step_traffic(wl.n, r["tuning"])
    gs = ground_state_run(args.gs_cases) if args.gs_cases > 0 else None
    others = None if args.no_other_configs else other_configs_run(not (args.no_autotune or args.tuning))
    scan = scan_run(args.scan_steps) if args.scan_steps > 0 else None
    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: oracle, {args.cpu_steps} steps")
        cpu = cpu_baseline(wl, args.cpu_steps)
        log(f"[bench] CPU baseline {cpu['value']:.4f} Gpt-steps/s on {cpu['cores']} threads")
    out = {
        "metric": "Gpt-steps/s + %HBM roof, 256³ 2-comp Bragg FP64; ground-state time warm vs cold",
        "value": value,
        "unit": "Gpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D two-component Bragg pulse, {wl.n}^3 on [-32,32]^3, complex128, "
                               f"dt=2e-4, Nk=100, Rabi populations per block",
                   "grid": list(wl.shape), "components": 2, "Nk": wl.Nk},
        # the step's kernels overlap (slab chains on two streams), so the
        # roofline is that of the whole step: its algorithmic bytes over its
        # wall time; the dominant kernel's own figure follows it
        "roofline": {"bound": "hbm", "achieved": step_gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": step_gbps / HBM_PEAK_GBS, "traffic": step_traffic,
                     "kernel": "3D step: k_vfuse + k_pass (y) x 2 + k_kin, y/z passes slab-chained on two streams",
                     "avg_us": 1e6 * r["elapsed"] / args.steps, "algorithmic_bytes_per_launch": step_bytes,
                     "measured_copy_GBps": r["copy_gbps"], "copy_frac": step_gbps / r["copy_gbps"],
                     # the reference's unfused split step moves C (1 + 2d) 32 B
                     # per point (448 B at d = 3, C = 2); the same step rate
                     # expressed in those bytes
                     "reference_model": {"bytes_per_point_step": 448,
                                         "effective_GBps": 448 * npts / (r["elapsed"] / args.steps) / 1e9},
                     "dominant_kernel": {"kernel": r["dominant"], "avg_us": 1e3 * r["dom_avg_ms"],
                                         "algorithmic_bytes_per_launch": r["bytes_per_launch"],
                                         "achieved": r["achieved"], "frac": r["achieved"] / HBM_PEAK_GBS}},
        "cpu_baseline": cpu,
        "ground_state": gs,
        "other_configs": others,
        "scan": scan,
        "kernels": r["kernels"],
        "tuning": r["tuning"],
        "checks": {"norm_drift_per_step": r["norm_drift_per_step"], "rabi_last_block": r["rabi"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
