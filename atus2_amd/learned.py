"""Learned warm start of the Sobolev ground-state minimiser (sobmin).

Ground states of the anisotropic harmonic trap V = (wx^2 x^2 + wy^2 y^2 +
wz^2 z^2)/2 with contact interaction g (H = -Delta/2 + V + g|psi|^2, N = 1,
CSOB_Base conventions with Alpha = 1/2) on 128^3 points of [-16, 16)^3.

* Training data: the plain minimiser (atus_sob_run) solves each (w, g) on a
  32^3 grid whose box [-s_a, s_a) is fitted to the cloud (half widths from
  the Thomas-Fermi radius and the oscillator length), from the cold start.
  The targets are those 32^3 states divided by their peak.
* Network: (log w, log g) scaled to about [-1, 1] -> 256 -> 256 -> 64 ->
  32^3 (tanh hidden layers, linear output; 2.2M parameters), trained with
  the library's BF16 MFMA kernels (atus_net_*), 60000 Adam steps.
* Warm start: the network's 32^3 output, read as periodic samples on the
  fitted box and carried onto the 128^3 grid by the trigonometric
  interpolant of each axis (atus_net_warm_start); the
  minimiser then polishes it to the same residual as a cold start.  The
  cold start is the Gaussian or the Thomas-Fermi profile, whichever has
  the lower energy.

`python -m atus2_amd.learned train` builds the training set and trains;
`python -m atus2_amd.learned eval` times cold and warm starts on held-out
cases and prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import hip

N_GRID, L_BOX = 128, 32.0
N_SMALL = 32
DIMS = (4, 256, 256, 64, N_SMALL ** 3)
ALPHA = 0.5
STEP_128 = 1e-3        # stable for every case of the range on [-16, 16)^3
EPS = 1e-10
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_net.npz")


@dataclass
class Case:
    w: tuple
    g: float

    @property
    def mu_tf(self):
        """Thomas-Fermi chemical potential of N = 1 atoms."""
        return 0.5 * (15.0 * self.g * self.w[0] * self.w[1] * self.w[2] / (4 * math.pi)) ** 0.4

    def half_widths(self):
        """Box [-s_a, s_a) of the 32^3 grid: 1.3 Thomas-Fermi radii or 4.5
        oscillator lengths, whichever is larger."""
        r = math.sqrt(2 * self.mu_tf)
        return tuple(max(1.3 * r / w, 4.5 / math.sqrt(w)) for w in self.w)

    def features(self):
        lw = [math.log(w) / math.log(2.0) for w in self.w]       # [-1, 1]
        lg = (math.log(self.g) - 0.5 * math.log(10 * 5000)) / (0.5 * math.log(500))
        return np.array(lw + [lg], dtype=np.float32)


def cases(n, seed):
    """w_a ~ U[0.5, 2], g ~ log-U[10, 5000] (the n values of g are drawn
    after all the w, so the first cases of a longer draw differ in g)."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 2.0, (n, 3))
    g = np.exp(rng.uniform(math.log(10.0), math.log(5000.0), n))
    return [Case(tuple(float(v) for v in w[i]), float(g[i])) for i in range(n)]


HELD_OUT = 256


def held_out(first=0, count=HELD_OUT):
    """Cases first .. first+count-1 of the held-out set: cases(256, 55)
    (never trained on; the training set is cases(4096, 5))."""
    if first < 0 or first + count > HELD_OUT:
        raise ValueError(f"the held-out set has {HELD_OUT} cases")
    return cases(HELD_OUT, 55)[first:first + count]


def _grid(n, half):
    dx = [2.0 * h / n for h in half]
    dk = [2 * math.pi / (2.0 * h) for h in half]
    xs = [((np.arange(n) - n // 2) * dx[a]).reshape([n if b == a else 1 for b in range(3)]) for a in range(3)]
    return dx, dk, xs


def _potential(case, xs):
    return 0.5 * sum((case.w[a] * xs[a]) ** 2 for a in range(3)) * np.ones([x.size for x in xs])


def _setup(ctx, case, V, step):
    ctx.set_alpha([ALPHA] * 3)
    ctx.set_potential(0, V)
    ctx.set_g([case.g])
    ctx.sob_setup([ALPHA], step)


def cold_start(ctx, case, xs, V):
    """Uploads the Gaussian or the Thomas-Fermi start, whichever has the
    lower energy; returns its name."""
    gauss = np.exp(-0.5 * sum(case.w[a] * xs[a] ** 2 for a in range(3))).astype(np.complex128)
    tf = np.sqrt(np.maximum(case.mu_tf - V, 0.0) / case.g).astype(np.complex128)
    best = None
    for name, psi in (("gaussian", gauss), ("thomas-fermi", tf)):
        ctx.upload(0, psi)
        ctx.sob_renormalize()
        e = ctx.energy()
        if best is None or e < best[0]:
            best = (e, name, psi)
    ctx.upload(0, best[2])
    return best[1]


def solve_small(case, eps=1e-9, maxit=200000):
    """The plain minimiser on the fitted 32^3 box; returns (state / peak as
    float32, iterations)."""
    half = case.half_widths()
    dx, dk, xs = _grid(N_SMALL, half)
    V = _potential(case, xs)
    step = 1.0 / (float(V.max()) + case.mu_tf + 1.0)
    with hip.Context((N_SMALL,) * 3, dx, dk, 1) as ctx:
        _setup(ctx, case, V, step)
        cold_start(ctx, case, xs, V)
        ctx.sob_renormalize()
        its, res = ctx.sob_run(eps, maxit)
        psi = np.abs(ctx.download(0))
    return (psi / psi.max()).astype(np.float32), its


def make_dataset(n, seed, log=None):
    X = np.empty((n, DIMS[0]), np.float32)
    T = np.empty((n, DIMS[4]), np.float32)
    total = 0
    for i, c in enumerate(cases(n, seed)):
        X[i] = c.features()
        t, its = solve_small(c)
        T[i] = t.ravel()
        total += its
        if log and (i + 1) % 256 == 0:
            log(f"{i + 1}/{n} ground states, {total / (i + 1):.0f} iterations on average")
    return X, T


def load_net(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as f:
        p = f["params"]
        dims = tuple(int(v) for v in f["dims"])
    net = hip.Net(dims, max_batch=1)
    net.set_params(p)
    return net


def solve_large(case, start, net=None, eps=EPS, maxit=0, state=False):
    """The minimiser on 128^3 points of [-16, 16)^3 from the cold or the
    warm start: (iterations, seconds from the start state to res <= eps,
    mu, res); with state=True also (psi, mu, energy) after renormalising
    the result to N = 1 exactly (the minimiser leaves |N - 1| <= 1e-5
    alone, CSOB_Base.h:470-479, so two runs end at slightly different N).
    Two runs stopped at res <= eps agree in psi and mu to first order in
    the norm of the remaining gradient (sqrt(eps) = 1e-5: mu carries the
    first variation of the interaction term), in the energy to second."""
    half = (L_BOX / 2,) * 3
    dx, dk, xs = _grid(N_GRID, half)
    V = _potential(case, xs)
    with hip.Context((N_GRID,) * 3, dx, dk, 1) as ctx:
        _setup(ctx, case, V, STEP_128)
        # the clock runs from the start state on the device (cold) or from
        # the network's input (warm: inference and interpolation included)
        if start == "cold":
            cold_start(ctx, case, xs, V)
            ctx.synchronize()
            tic = time.perf_counter()
        else:
            ctx.synchronize()
            tic = time.perf_counter()
            net.warm_start(case.features(), case.half_widths(), ctx, 0)
        ctx.sob_renormalize()
        its, res = ctx.sob_run(eps, maxit)
        ctx.synchronize()
        el = time.perf_counter() - tic
        mu = float(ctx.sob_mu()[0])
        if not state:
            return its, el, mu, res
        ctx.sob_renormalize()
        return its, el, mu, res, ctx.download(0), float(ctx.sob_mu()[0]), float(ctx.energy())


def compare_states(a, b):
    """Relative L2 difference of two ground states of N = 1 (real up to a
    global phase, which is removed first)."""
    ph = np.vdot(b, a)
    a = a * (abs(ph) / ph) if ph != 0 else a
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main(argv=None):
    ap = argparse.ArgumentParser(description="learned warm start of sobmin")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--cases", type=int, default=4096)
    tr.add_argument("--steps", type=int, default=60000)
    tr.add_argument("--batch", type=int, default=256)
    tr.add_argument("--lr", type=float, default=2e-3)
    tr.add_argument("--out", default=WEIGHTS)
    ev = sub.add_parser("eval")
    ev.add_argument("--cases", type=int, default=HELD_OUT)
    ev.add_argument("--first", type=int, default=0, help="start at this held-out case")
    ev.add_argument("--weights", default=WEIGHTS)
    ev.add_argument("--rows", default=None, help="write one CSV row per case here")
    a = ap.parse_args(argv)

    def log(m):
        print(f"[learned] {m}", file=sys.stderr, flush=True)

    if a.cmd == "train":
        tic = time.perf_counter()
        X, T = make_dataset(a.cases, 5, log)
        tdata = time.perf_counter() - tic
        log(f"training set: {a.cases} ground states in {tdata:.1f} s")
        with hip.Net(DIMS, max_batch=a.batch, seed=5) as net:
            net.set_data(X, T)
            tic = time.perf_counter()
            losses = net.train(a.steps, a.batch, a.lr, a.lr * 0.02, seed=5)
            ttrain = time.perf_counter() - tic
            p = net.get_params()
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.savez(a.out, params=p, dims=np.array(DIMS))
        log(f"trained {a.steps} steps in {ttrain:.1f} s; loss {losses[0]:.3e} -> {losses[-20:].mean():.3e}")
        print(json.dumps({"cases": a.cases, "data_s": tdata, "train_s": ttrain, "steps": a.steps,
                          "loss_first": float(losses[0]), "loss_last": float(losses[-20:].mean())}))
        return
    net = load_net(a.weights)
    rows = []
    fout = open(a.rows, "w") if a.rows else None
    if fout:
        fout.write("case,wx,wy,wz,g,cold_iters,cold_s,warm_iters,warm_s,mu_cold_N1,mu_warm_N1,E_cold_N1,E_warm_N1,"
                   "psi_rel_l2_N1\n")
    for i, c in enumerate(held_out(a.first, a.cases), start=a.first):
        ic, tc, _, _, pc, mc, ec = solve_large(c, "cold", state=True)
        iw, tw, _, _, pw, mw, ew = solve_large(c, "warm", net, state=True)
        d = compare_states(pw, pc)
        rows.append((ic, tc, iw, tw, abs(mc - mw) / abs(mc), d, abs(ec - ew) / abs(ec)))
        if fout:
            fout.write(f"{i},{c.w[0]:.6f},{c.w[1]:.6f},{c.w[2]:.6f},{c.g:.4f},{ic},{tc:.4f},{iw},{tw:.4f},"
                       f"{mc:.15f},{mw:.15f},{ec:.15f},{ew:.15f},{d:.3e}\n")
            fout.flush()
        if (i + 1 - a.first) % 8 == 0:
            r = np.array(rows)
            log(f"{i + 1 - a.first} cases: cold {r[:, 1].sum():.1f} s, warm {r[:, 3].sum():.1f} s")
    r = np.array(rows)
    print(json.dumps({"cases": a.cases, "first": a.first, "grid": [N_GRID] * 3, "eps": EPS, "step": STEP_128,
                      "cold_s": float(r[:, 1].sum()), "warm_s": float(r[:, 3].sum()),
                      "ratio": float(r[:, 3].sum() / r[:, 1].sum()),
                      "cold_iters_mean": float(r[:, 0].mean()), "warm_iters_mean": float(r[:, 2].mean()),
                      "worst_case_ratio": float(np.max(r[:, 3] / r[:, 1])),
                      "max_rel_mu_difference_N1": float(r[:, 4].max()),
                      "max_psi_rel_l2_N1": float(r[:, 5].max()),
                      "max_rel_energy_difference_N1": float(r[:, 6].max())}))


if __name__ == "__main__":
    main()
