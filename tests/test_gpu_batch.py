"""Batches of independent Bragg runs (atus_batch_*): every run of a batch
must evolve exactly as the same run through the single-run path, and as
the oracle.  Runs differ in couplings, Rabi amplitude, detuning, pulse
width and the detuning of the second level, as in a parameter scan."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _runs(nb, seed):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nb):
        g = rng.uniform(1.0, 200.0)
        out.append(dict(gs=[g, 0.95 * g, 0.95 * g, g], amp0=rng.uniform(1.0, 10.0),
                        laser_domh=8.0 + rng.uniform(-2.0, 2.0), sigma=rng.uniform(0.02, 0.1),
                        deltaL=[0.0, rng.uniform(-0.5, 0.5)], beta0=rng.uniform(-0.02, 0.02)))
    return out


def _amp(run, t0):
    s = run["sigma"]
    return lambda t: float(np.exp(-0.5 * ((t - t0) / s) ** 2))


def _initial(orc, g, r):
    xs = orc.coords_x(g)
    base = np.exp(-0.5 * sum((x / (2.0 + 0.1 * r + a)) ** 2 for a, x in enumerate(xs))).astype(np.complex128)
    base /= np.sqrt(orc.particle_number(g, base))
    return base, np.zeros_like(base)


@pytest.mark.parametrize("shape,rm", [((1024,), 16), ((256,), 8), ((16384,), 16), ((8192,), 8), ((64, 128), 16),
                                      ((32, 64), 8), ((16, 8, 32), 16)])
@pytest.mark.parametrize("kind", ["bragg_ad", "freeprop_if"])
def test_batch_matches_single_runs_and_oracle(hip, orc, shape, rm, kind):
    dim = len(shape)
    L = [40.0, 30.0, 36.0][:dim]
    g = orc.make_grid(shape, L=L)
    nb, dt, Nk, t0 = 5, 1e-3, 20, 0.05
    runs = _runs(nb, 11 + dim)
    laser_k = 3 * (2 * np.pi / L[0])
    hkind = hip.STEP_BRAGG_AD if kind == "bragg_ad" else hip.STEP_FREEPROP_IF
    okind = orc.BRAGG_AD if kind == "bragg_ad" else orc.FREEPROP_IF
    lasers = [hip.laser(amp0=r["amp0"], laser_k=laser_k, laser_domh=r["laser_domh"], beta0=r["beta0"],
                        deltaL=tuple(r["deltaL"])) for r in runs]
    alpha = [0.5] * dim
    dx = [g.dx[a] for a in range(dim)]
    dk = [g.dk[a] for a in range(dim)]
    with hip.BatchContext(shape, dx, dk, nb) as B:
        for f in hip.KFAM:
            B.tune(f, rm)
        states = [_initial(orc, g, r) for r in range(nb)]
        for r in range(nb):
            B.upload(2 * r, states[r][0])
            B.upload(2 * r + 1, states[r][1])
        B.set_alpha(alpha)
        B.set_runs([r["gs"] for r in runs], lasers)
        t = t0
        for blk in range(2):
            F = np.array([[_amp(runs[r], t0 + 0.02)(t1) for r in range(nb)] for t1 in B.step_times(t, dt, Nk)])
            if kind != "bragg_ad":
                F = None
            t = B.run_batch_block(hkind, dt, Nk, t, F)
        out = [(B.download(2 * r), B.download(2 * r + 1)) for r in range(nb)]
        Ns = B.particle_numbers()
        k0 = [(0.0,) * dim, (2 * laser_k,) + (0.0,) * (dim - 1)]
        rabi = B.batch_rabi_integrals(0, k0, 0.5 * laser_k)
        after_rabi = [(B.download(2 * r), B.download(2 * r + 1)) for r in range(nb)]
    assert abs(t - (t0 + 2 * Nk * dt)) < 1e-12
    for r in range(nb):
        # the in-place transform and back of the Rabi integrals leaves psi
        # unchanged to rounding
        assert rel_l2(after_rabi[r], out[r]) < 1e-14
        # the single-run path with the same parameters
        with hip.Context(shape, dx, dk, 2) as S:
            S.upload(0, states[r][0])
            S.upload(1, states[r][1])
            S.set_alpha(alpha)
            S.set_g(runs[r]["gs"])
            for f in hip.KFAM:
                S.tune(f, rm)
            ts = t0
            for blk in range(2):
                ts = S.run_block(hkind, dt, Nk, ts, lasers[r], _amp(runs[r], t0 + 0.02) if kind == "bragg_ad" else None)
            single = (S.download(0), S.download(1))
            assert rel_l2(out[r], single) < 1e-13, r
            for c in range(2):
                assert abs(Ns[r, c] - S.particle_number(c)) < 1e-14
            rs = S.rabi_integrals(0, k0, 0.5 * laser_k)
            assert np.allclose(rabi[r], rs, rtol=0, atol=1e-13), (r, rabi[r], rs)
    # the oracle, for two of the runs
    for r in (0, nb - 1):
        run = runs[r]
        ifp = dict(beta0=run["beta0"], deltaL=list(run["deltaL"]), amp0=run["amp0"], laser_k=laser_k,
                   laser_domh=run["laser_domh"])
        S = orc.RTSolver(g, list(states[r]), run["gs"], alpha, dt=dt, t=t0,
                         amp=_amp(run, t0 + 0.02) if kind == "bragg_ad" else None, **ifp)
        S.block(okind, Nk)
        S.block(okind, Nk)
        err = rel_l2(out[r], S.psis)
        assert err < 1e-11, (r, err)


def test_batch_rejects_single_run_calls(hip):
    with hip.BatchContext((64,), [0.5], [2 * np.pi / 32], 3) as B:
        with pytest.raises(hip.AtusError, match="batch"):
            B.set_g([1.0, 0.0, 0.0, 1.0])
        with pytest.raises(hip.AtusError, match="batch"):
            B.run_block(hip.STEP_BRAGG_AD, 1e-3, 2, 0.0, hip.laser())
        with pytest.raises(hip.AtusError, match="run parameters not set"):
            B.run_batch_block(hip.STEP_BRAGG_AD, 1e-3, 2, 0.0)
        B.set_runs([[1.0, 0.0, 0.0, 1.0]] * 3, [hip.laser()] * 3)
        with pytest.raises(hip.AtusError, match="bragg_ad or freeprop"):
            B.run_batch_block(hip.STEP_FREEPROP, 1e-3, 2, 0.0)
        with pytest.raises(ValueError):
            B.run_batch_block(hip.STEP_BRAGG_AD, 1e-3, 2, 0.0, np.ones((3, 3)))


def test_scan_batched_matches_one_after_another(hip):
    from atus2_amd.scan import run_batched, run_single, scan_runs
    runs = scan_runs(seed=7, n1d=6, n2d=2)
    rb, _ = run_batched(runs, steps=300)
    rs, _ = run_single(runs, steps=300)
    for i in range(len(runs)):
        assert rb[i].shape == (3, 3)
        # populations agree; the amplitudes come from numpy's and libm's exp
        # (a last-bit difference, amplified near the nodes of Omega)
        assert np.max(np.abs(rb[i] - rs[i])) < 1e-11, (i, runs[i].shape)


def test_scan_graph_equals_eager_batches(hip, orc):
    """Batch contexts of three grids (a resident 1D group, a four-step 1D
    group and a 2D group) advanced by one captured scan graph per block:
    states and Rabi populations bit for bit those of the eager batch blocks
    and batched Rabi integrals, block after block."""
    shapes = [(1024,), (8192,), (64, 128)]
    nb, dt, Nk, blocks = 3, 1e-3, 20, 3
    k0 = [(0.0,), (8.0,), (-8.0,)]
    made = []
    try:
        for s_i, shape in enumerate(shapes):
            dim = len(shape)
            L = [40.0, 30.0][:dim]
            g = orc.make_grid(shape, L=L)
            runs = _runs(nb, 30 + s_i)
            lasers = [hip.laser(amp0=r["amp0"], laser_k=4.0, laser_domh=r["laser_domh"], beta0=r["beta0"],
                                deltaL=tuple(r["deltaL"])) for r in runs]
            pair = []
            # one set of initial states for both contexts: the oracle's
            # OpenMP norm (inside _initial) may round differently per call
            init = [_initial(orc, g, r) for r in range(nb)]
            for _ in range(2):
                B = hip.BatchContext(shape, [g.dx[a] for a in range(dim)], [g.dk[a] for a in range(dim)], nb)
                for r, (p1, p2) in enumerate(init):
                    B.upload(2 * r, p1)
                    B.upload(2 * r + 1, p2 + 0.1 * p1)
                B.set_alpha([0.5] * dim)
                B.set_runs([r["gs"] for r in runs], lasers)
                pair.append(B)
            made.append((runs, pair))
        eager = [p[0] for _, p in made]
        graphed = [p[1] for _, p in made]
        G = hip.ScanGraph(graphed, hip.STEP_BRAGG_AD, dt, Nk, k0, 2.0)
        t = 0.02
        for blk in range(blocks):
            t1 = np.array(hip.BatchContext.step_times(t, dt, Nk))
            F = [np.array([[_amp(r, 0.04)(v) for r in runs] for v in t1]) for runs, _ in made]
            te = None
            rab_e = []
            for B, f in zip(eager, F):
                te = B.run_batch_block(hip.STEP_BRAGG_AD, dt, Nk, t, f)
                rab_e.append(B.batch_rabi_integrals(0, k0, 2.0))
            tg, rab_g, ms = G.launch(t, F)
            assert tg == te and ms > 0
            for s_i, (Be, Bg) in enumerate(zip(eager, graphed)):
                assert np.array_equal(rab_g[s_i], rab_e[s_i]), (blk, shapes[s_i])
                for c in range(2 * nb):
                    assert np.array_equal(Bg.download(c), Be.download(c)), (blk, shapes[s_i], c)
            t = te
        # the graph keeps its contexts: the C-ABI refuses to destroy one,
        # and the Python context closes the graph before itself
        rc = hip._lib.atus_ctx_destroy(graphed[0]._h)
        assert rc == -5
        graphed[0].close()
        with pytest.raises(hip.AtusError, match="closed"):
            G.launch(t, F)
    finally:
        for _, pair in made:
            for B in pair:
                B.close()


def test_scan_family_times_cover_every_group(hip):
    """`python -m atus2_amd.scan --family-times`: one row per shape group
    with a positive time and rate for each kernel family it launched."""
    from atus2_amd.scan import family_times, scan_runs
    runs = scan_runs(seed=7, n1d=3, n2d=1)
    rows = family_times(runs, Nk=4, reps=1, tune=False)
    assert sorted(tuple(r["shape"]) for r in rows) == sorted({r.shape for r in runs})
    for r in rows:
        assert r["runs"] == sum(1 for x in runs if list(x.shape) == r["shape"])
        assert r["families"], r
        for f, v in r["families"].items():
            assert v["us_per_grid"] > 0 and v["GBps"] > 0, (r["shape"], f, v)
