"""Parity over the full run length at the workloads' own sizes, HIP path
(through the C-ABI) against the oracle:

* the 3D two-component Bragg sequence (bragg_ad 1.0 with the Rabi
  populations after every block, then freeprop 0.4; 7000 steps) on the
  128^3 grid of the same spacing, and the 256^3 grid itself over its first
  500 steps (the oracle needs about 0.1 s per 256^3 step on the host);
* the 2D double slit at 2048^2 over all 2000 steps.

Gates: relative L2 error of psi <= 1e-10 (checked after every range of
blocks and at the end), total particle number drift <= 1e-12 per step on
the device, Rabi populations of every block within 1e-10 of the oracle's.
Each test covers a range of blocks, so that no single test runs long."""
import pytest

from atus2_amd.workloads import BraggWorkload, DoubleSlitWorkload
from fullrun import BraggSequence, SlitSequence, chunks

pytestmark = pytest.mark.gpu

_RUNS = {}


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _get(key, make):
    if key not in _RUNS:
        _RUNS[key] = make()
    return _RUNS[key]


def _finish(key):
    run = _RUNS.pop(key, None)
    if run is not None:
        run.close()


C3_128 = chunks(70, 10)


@pytest.mark.parametrize("start,stop", C3_128)
def test_c3_128_full_sequence(hip, orc, margin, start, stop):
    run = _get("c3_128", lambda: BraggSequence(hip, orc, BraggWorkload.proxy(128)))
    try:
        err, worst = run.run(start, stop)
        print(f"128^3 blocks {start}..{stop - 1}, t={run.t:.4f}: rel L2 {err:.2e}, worst Rabi difference {worst:.2e}")
        margin(f"128^3 blocks {start}-{stop - 1}", err, 1e-10)
        if worst:
            margin(f"128^3 blocks {start}-{stop - 1} Rabi", worst, 1e-10)
        assert err <= 1e-10
        if stop == C3_128[-1][1]:
            assert run.done == 70 and abs(run.t - 1.4) < 1e-9
            margin("full run: Bragg 128^3, bragg_ad 1.0 + freeprop 0.4 (7000 steps), psi rel L2", err, 1e-10)
    finally:
        if stop == C3_128[-1][1]:
            _finish("c3_128")


@pytest.mark.parametrize("start,stop", chunks(5, 1))
def test_c3_256_first_500_steps(hip, orc, margin, start, stop):
    run = _get("c3_256", lambda: BraggSequence(hip, orc, BraggWorkload()))
    try:
        err, worst = run.run(start, stop)
        print(f"256^3 block {start}, t={run.t:.4f}: rel L2 {err:.2e}, Rabi difference {worst:.2e}")
        margin(f"256^3 block {start}", err, 1e-10)
        margin(f"256^3 block {start} Rabi", worst, 1e-10)
        assert err <= 1e-10
        if stop == 5:
            margin("full run: Bragg 256^3, first 500 steps of bragg_ad, psi rel L2", err, 1e-10)
    finally:
        if stop == 5:
            _finish("c3_256")


C2 = chunks(20, 5)


@pytest.mark.parametrize("start,stop", C2)
def test_c2_2048_full_run(hip, orc, margin, start, stop):
    run = _get("c2", lambda: SlitSequence(hip, orc, DoubleSlitWorkload()))
    try:
        err = run.run(start, stop)
        print(f"2048^2 blocks {start}..{stop - 1}, t={run.t:.4f}: rel L2 {err:.2e}")
        margin(f"2048^2 blocks {start}-{stop - 1}", err, 1e-10)
        assert err <= 1e-10
        if stop == C2[-1][1]:
            assert run.done == 20 and abs(run.t - 1.0) < 1e-9
            margin("full run: double slit 2048^2, 2000 steps, psi rel L2", err, 1e-10)
    finally:
        if stop == C2[-1][1]:
            _finish("c2")


@pytest.mark.parametrize("shape", [(16384,), (1024, 1024)])
def test_c4_scan_runs_match_oracle(hip, orc, margin, shape):
    """Two runs of the 512-case parameter scan (seed 4) with the largest
    grids of each dimension, batched as the scan runs them (atus_batch_*,
    the amplitude table of a block evaluated on the host), over 300 steps
    against the oracle run by run.  The batched Rabi integrals transform
    component 0 in place and back as the reference does
    (CRT_Base_IF.h:208-281), so the oracle's component 0 takes the same
    round trip after every block."""
    import numpy as np
    from atus2_amd.scan import RABI_THRESHOLD, scan_runs
    from atus2_amd.workloads import grid_spacings
    from fullrun import rel_l2
    runs = scan_runs()
    idx = [i for i, r in enumerate(runs) if r.shape == shape][:2]
    rs = [runs[i] for i in idx]
    dt, Nk, nblocks = 2e-4, 100, 3
    g = orc.make_grid(shape, L=rs[0].L)
    dx, dk = grid_spacings(shape, rs[0].L)
    orcs = []
    with hip.BatchContext(shape, dx, dk, len(rs)) as B:
        for j, r in enumerate(rs):
            p1, p2 = r.initial_state()
            B.upload(2 * j, p1)
            B.upload(2 * j + 1, p2)
            orcs.append(orc.RTSolver(g, [p1, p2], r.gs, [0.5] * len(shape), dt=dt, amp=r.amp, beta0=0.0,
                                     deltaL=[0.0, 0.0], amp0=r.amp0, laser_k=r.laser_k, laser_domh=r.laser_domh))
        B.set_alpha([0.5] * len(shape))
        B.set_runs([r.gs for r in rs], [r.laser() for r in rs])
        k0 = rs[0].rabi_k0()
        t = 0.0
        worst = 0.0
        for b in range(nblocks):
            # AMP_T of each run at the t1 the block runner uses, with the
            # same libm exp the oracle's callback calls
            t1 = B.step_times(t, dt, Nk)
            F = np.array([[r.amp(v) for r in rs] for v in t1])
            t = B.run_batch_block(hip.STEP_BRAGG_AD, dt, Nk, t, F)
            rab = B.batch_rabi_integrals(0, k0, RABI_THRESHOLD)
            for j, S in enumerate(orcs):
                S.block(orc.BRAGG_AD, Nk)
                assert t == S.t
                r_orc, psi0 = orc.rabi_integrals(g, S.psis[0], k0, RABI_THRESHOLD)
                S.psis[0][...] = psi0
                worst = max(worst, float(np.max(np.abs(rab[j] - r_orc))))
                assert np.all(np.abs(rab[j] - r_orc) <= 1e-10), (b, j, rab[j], r_orc)
        for j, (i, S) in enumerate(zip(idx, orcs)):
            err = rel_l2([B.download(2 * j), B.download(2 * j + 1)], S.psis)
            print(f"scan run {i} {shape}: g={rs[j].g:.1f}, Omega={rs[j].amp0:.2f}: rel L2 {err:.2e}")
            margin(f"full run: scan run {i} {shape}, 300 steps batched, psi rel L2", err, 1e-10)
            assert err <= 1e-10
        margin(f"scan runs {shape} Rabi", worst, 1e-10)


class SobolevRun:
    """The ground state of held-out trap 238 (w = (1.25, 1.14, 1.26),
    g = 4482; about 3000 iterations from the warm start) on 128^3 points of
    [-16, 16)^3, from the learned warm start: the HIP minimiser and the
    oracle's (CSOB_Base.h:427-505) iterate side by side from the same start
    state."""

    CASE = 238

    def __init__(self, hip, orc):
        import numpy as np
        from atus2_amd import learned as Lr
        self.np = np
        self.orc = orc
        c = Lr.held_out(self.CASE, 1)[0]
        half = (Lr.L_BOX / 2,) * 3
        dx, dk, xs = Lr._grid(Lr.N_GRID, half)
        V = Lr._potential(c, xs)
        self.ctx = hip.Context((Lr.N_GRID,) * 3, dx, dk, 1)
        Lr._setup(self.ctx, c, V, Lr.STEP_128)
        net = Lr.load_net()
        try:
            net.warm_start(c.features(), c.half_widths(), self.ctx, 0)
        finally:
            net.close()
        start = self.ctx.download(0)
        # the iteration count of the whole polish, from the same start
        self.ctx.sob_renormalize()
        self.its, res = self.ctx.sob_run(Lr.EPS, 100000)
        assert res <= Lr.EPS
        self.ctx.upload(0, start)
        self.ctx.sob_renormalize()
        g = orc.make_grid((Lr.N_GRID,) * 3, L=Lr.L_BOX)
        self.S = orc.SobSolver(g, [start], [V], [c.g], [Lr.ALPHA], stepsize=Lr.STEP_128)
        self.S.renormalize()
        self.done = 0
        self.first_below = None

    def advance(self, n):
        res = []
        for _ in range(n):
            r = self.S.iterate()
            res.append(r)
            self.done += 1
            if self.first_below is None and r <= 1e-10:
                self.first_below = self.done
        got, r_gpu, _ = self.ctx.sob_iterate(n)
        assert got == n
        return res[-1], r_gpu

    def compare(self):
        np = self.np
        psi = self.ctx.download(0)
        err = float(np.linalg.norm(psi - self.S.psi[0]) / np.linalg.norm(self.S.psi[0]))
        mu_gpu = float(self.ctx.sob_mu()[0])
        mu_orc = float(self.S.mu()[0])
        return err, mu_gpu, mu_orc

    def close(self):
        self.ctx.close()
        self.S = None


C5 = [(0, 200), (200, 1100), (1100, 2000), (2000, 2900), (2900, None)]


@pytest.mark.parametrize("start,stop", C5)
def test_c5_128_warm_start_polish_matches_oracle(hip, orc, margin, start, stop):
    run = _get("c5", lambda: SobolevRun(hip, orc))
    try:
        assert 2900 < run.its <= 2900 + 1024, run.its
        end = run.its if stop is None else stop
        if run.done < start:
            run.advance(start - run.done)
        r_orc, r_gpu = run.advance(end - run.done)
        assert abs(r_gpu / r_orc - 1) < 1e-8, (r_gpu, r_orc)
        err, mu_gpu, mu_orc = run.compare()
        print(f"128^3 Sobolev, {run.done} iterations of {run.its}: res {r_gpu:.3e} / {r_orc:.3e}, "
              f"psi rel L2 {err:.2e}, mu {mu_gpu:.15f} / {mu_orc:.15f}")
        margin(f"128^3 Sobolev after {run.done} iterations: psi rel L2", err, 1e-10)
        margin(f"128^3 Sobolev after {run.done} iterations: mu", abs(mu_gpu - mu_orc), 1e-9)
        assert err <= 1e-10
        assert abs(mu_gpu - mu_orc) <= 1e-9
        if start == 0:
            margin("full run: ground state 128^3, first 200 Sobolev iterations from the warm start, mu",
                   abs(mu_gpu - mu_orc), 1e-9)
        if stop is None:
            # the oracle's do-while stops at the same iteration
            assert run.first_below == run.its, (run.first_below, run.its)
            margin(f"full run: ground state 128^3, warm start polished to res 1e-10 ({run.its} iterations), mu",
                   abs(mu_gpu - mu_orc), 1e-9)
            margin(f"full run: ground state 128^3, warm start polished to res 1e-10 ({run.its} iterations), "
                   "psi rel L2", err, 1e-10)
    finally:
        if stop is None:
            _finish("c5")
