#!/usr/bin/env python3
"""Long parity runs of the 3D two-component Bragg workload, HIP path against
the oracle (minutes of oracle time, so not part of the pytest suite):

  --n 128              the 128^3 grid of the same spacing over the full
                       bragg sequence (bragg_ad 1.0, freeprop 0.4; 70 blocks)
  --n 256 --blocks 5   the 256^3 grid itself over its first 500 steps

Same gates as tests/test_gpu_fullrun.py; prints one line per block and a
final summary line, exits non-zero when a gate fails.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--blocks", type=int, default=None, help="stop after this many blocks")
    args = ap.parse_args()
    from atus2_amd import hip
    from atus2_amd.workloads import BraggWorkload
    from oracle import oracle as orc
    from test_gpu_fullrun import _run
    hip.load()
    wl = BraggWorkload() if args.n == 256 else BraggWorkload.proxy(args.n)
    print(f"oracle threads {orc.threads()}; grid {wl.shape}, L={wl.L}, gs={wl.gs}", flush=True)
    err, t, done = _run(hip, orc, wl, max_blocks=args.blocks, log=lambda m: print(m, flush=True))
    ok = err <= 1e-10
    print(f"SUMMARY {args.n}^3 blocks={done} t={t:.4f} rel_l2={err:.3e} gate=1e-10 {'PASS' if ok else 'FAIL'}",
          flush=True)
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
