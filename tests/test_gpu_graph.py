"""Blocks captured as HIP graphs and replayed (atus_graph_*) against the same
blocks run eagerly: every replay must leave exactly the eager state, with
the program's own eager work (reductions, eager blocks, host copies)
between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _state(orc, g, ncomp, seed):
    xs = orc.coords_x(g)
    rng = np.random.default_rng(seed)
    base = np.exp(-0.5 * sum((x / (2.0 + a)) ** 2 for a, x in enumerate(xs)))
    out = []
    for c in range(ncomp):
        psi = (base * (1.0 if c == 0 else 0.4) + 1e-3 * (rng.standard_normal(base.shape)
                                                          + 1j * rng.standard_normal(base.shape)))
        out.append(psi.astype(np.complex128))
    return out


def _amp(t):
    return float(np.exp(-0.5 * ((t - 0.05) / 0.03) ** 2))


@pytest.mark.parametrize("shape,kind", [((4096,), "freeprop"), ((8192,), "bragg_ad"), ((64, 128), "bragg_ad"),
                                        ((16, 32, 64), "freeprop_if"), ((32, 16, 64), "bragg_ad")])
def test_graph_replays_equal_eager_blocks(hip, orc, shape, kind):
    dim = len(shape)
    g = orc.make_grid(shape, L=[40.0, 30.0, 36.0][:dim])
    ncomp = 1 if kind == "freeprop" else 2
    psis = _state(orc, g, ncomp, 5)
    hk = {"freeprop": hip.STEP_FREEPROP, "freeprop_if": hip.STEP_FREEPROP_IF, "bragg_ad": hip.STEP_BRAGG_AD}[kind]
    las = hip.laser(amp0=5.0, laser_k=3 * 2 * np.pi / 40.0, laser_domh=8.0, beta0=0.01, deltaL=(0.0, 0.3))
    dt, Nk = 1e-3, 20
    ctxs = []
    for _ in range(2):
        ctx = hip.Context.from_grid(g, ncomp)
        for c in range(ncomp):
            ctx.upload(c, psis[c])
        ctx.set_alpha([0.5] * dim)
        ctx.set_g([10.0] if ncomp == 1 else [100.0, 95.0, 95.0, 100.0])
        if kind == "freeprop":
            ctx.set_potential(0, 0.5 * sum(x ** 2 for x in orc.coords_x(g)) * np.ones(shape))
        ctxs.append(ctx)
    eager, rep = ctxs
    amp = _amp if kind == "bragg_ad" else None
    try:
        with hip.Graph(rep, hk, dt, Nk, las) as G:
            te = tr = 0.0
            for blk in range(4):
                te = eager.run_block(hk, dt, Nk, te, las, amp)
                F = None if amp is None else np.array([amp(t1) for t1 in hip.BatchContext.step_times(tr, dt, Nk)])
                tr = G.launch(tr, F)
                assert tr == te
                for c in range(ncomp):
                    assert np.array_equal(rep.download(c), eager.download(c)), (blk, c)
                # the program's own eager work between replays
                assert rep.particle_number(0) == eager.particle_number(0)
                if blk == 1:
                    te = eager.run_block(hk, dt, 3, te, las, amp)
                    tr = rep.run_block(hk, dt, 3, tr, las, amp)
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("shape", [(1024,), (16384,), (32, 64)])
def test_graph_replays_equal_eager_batch_blocks(hip, orc, shape):
    dim = len(shape)
    g = orc.make_grid(shape, L=[40.0, 30.0][:dim])
    nb, dt, Nk = 4, 1e-3, 10
    dx = [g.dx[a] for a in range(dim)]
    dk = [g.dk[a] for a in range(dim)]
    rng = np.random.default_rng(2)
    gsl = [[gg, 0.95 * gg, 0.95 * gg, gg] for gg in rng.uniform(1, 200, nb)]
    lasers = [hip.laser(amp0=a, laser_k=3 * 2 * np.pi / 40.0, laser_domh=d)
              for a, d in zip(rng.uniform(1, 10, nb), rng.uniform(6, 10, nb))]
    sig = rng.uniform(0.02, 0.05, nb)
    bs = []
    for _ in range(2):
        B = hip.BatchContext(shape, dx, dk, nb)
        for r in range(nb):
            p = _state(orc, g, 2, 10 + r)
            B.upload(2 * r, p[0])
            B.upload(2 * r + 1, p[1])
        B.set_alpha([0.5] * dim)
        B.set_runs(gsl, lasers)
        bs.append(B)
    eager, rep = bs
    try:
        with hip.Graph(rep, hip.STEP_BRAGG_AD, dt, Nk) as G:
            t = 0.02
            for blk in range(3):
                t1 = np.array(eager.step_times(t, dt, Nk))[:, None]
                F = np.exp(-0.5 * ((t1 - 0.04) / sig[None, :]) ** 2)
                te = eager.run_batch_block(hip.STEP_BRAGG_AD, dt, Nk, t, F)
                tr = G.launch(t, F)
                assert tr == te
                for c in range(2 * nb):
                    assert np.array_equal(rep.download(c), eager.download(c)), (blk, c)
                assert np.array_equal(rep.particle_numbers(), eager.particle_numbers())
                t = te
    finally:
        for B in bs:
            B.close()


def test_graph_of_a_slab_chained_3d_block(hip, orc):
    # a block whose y/z passes run per slab on two streams (fork / join by
    # events) captured as a graph: replays equal eager whole-grid blocks bit
    # for bit
    shape = (32, 16, 64)
    g = orc.make_grid(shape, L=[40.0, 30.0, 36.0])
    psis = _state(orc, g, 2, 9)
    las = hip.laser(amp0=5.0, laser_k=3 * 2 * np.pi / 40.0, laser_domh=8.0, deltaL=(0.0, 0.3))
    dt, Nk = 1e-3, 12
    ctxs = []
    for slab in (0, 5):
        ctx = hip.Context.from_grid(g, 2)
        for c in range(2):
            ctx.upload(c, psis[c])
        ctx.set_alpha([0.5] * 3)
        ctx.set_g([100.0, 95.0, 95.0, 100.0])
        ctx.set_slab(slab)
        ctxs.append(ctx)
    eager, rep = ctxs
    try:
        with hip.Graph(rep, hip.STEP_BRAGG_AD, dt, Nk, las) as G:
            te = tr = 0.0
            for blk in range(3):
                te = eager.run_block(hip.STEP_BRAGG_AD, dt, Nk, te, las, _amp)
                F = np.array([_amp(t1) for t1 in hip.BatchContext.step_times(tr, dt, Nk)])
                tr = G.launch(tr, F)
                assert tr == te
                for c in range(2):
                    assert np.array_equal(rep.download(c), eager.download(c)), (blk, c)
                assert rep.particle_number(1) == eager.particle_number(1)
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("shape", [(4096,), (32, 64), (16, 32, 64)])
def test_graph_rebuilds_its_kinetic_tables_after_a_block_of_another_dt(hip, orc, shape):
    # run_sequence gives every sequence item its own dt (Set_dt,
    # CRT_Base.h:757-758): an eager block at dt2 between two replays of a
    # graph captured at dt1 must not change what the replay computes
    dim = len(shape)
    g = orc.make_grid(shape, L=[40.0, 30.0, 36.0][:dim])
    psis = _state(orc, g, 2, 21)
    las = hip.laser(amp0=4.0, laser_k=3 * 2 * np.pi / 40.0, laser_domh=8.0, deltaL=(0.0, 0.3))
    dt1, dt2, Nk = 1e-3, 3e-3, 10
    ctxs = []
    for _ in range(2):
        ctx = hip.Context.from_grid(g, 2)
        for c in range(2):
            ctx.upload(c, psis[c])
        ctx.set_alpha([0.5] * dim)
        ctx.set_g([100.0, 95.0, 95.0, 100.0])
        ctxs.append(ctx)
    eager, rep = ctxs
    try:
        with hip.Graph(rep, hip.STEP_BRAGG_AD, dt1, Nk, las) as G:
            te = tr = 0.0
            for blk in range(3):
                # the dt2 block on both, then the dt1 block: eager vs replay
                te = eager.run_block(hip.STEP_FREEPROP_IF, dt2, 4, te, las)
                tr = rep.run_block(hip.STEP_FREEPROP_IF, dt2, 4, tr, las)
                te = eager.run_block(hip.STEP_BRAGG_AD, dt1, Nk, te, las, _amp)
                F = np.array([_amp(t1) for t1 in hip.BatchContext.step_times(tr, dt1, Nk)])
                tr = G.launch(tr, F)
                assert tr == te
                for c in range(2):
                    assert np.array_equal(rep.download(c), eager.download(c)), (blk, c)
    finally:
        for c in ctxs:
            c.close()


def test_graph_refuses_a_changed_context_and_outlives_nothing(hip, orc):
    g = orc.make_grid((32, 64), L=[40.0, 30.0])
    psis = _state(orc, g, 1, 3)
    V = 0.5 * sum(x ** 2 for x in orc.coords_x(g)) * np.ones((32, 64))
    ctx = hip.Context.from_grid(g, 1)
    ctx.upload(0, psis[0])
    ctx.set_alpha([0.5, 0.5])
    ctx.set_g([10.0])
    ctx.set_potential(0, V)
    G = hip.Graph(ctx, hip.STEP_FREEPROP, 1e-3, 5)
    t = G.launch(0.0)
    # amplitudes of the wrong length are refused before the C call
    with pytest.raises(ValueError):
        G.launch(t, np.ones(4))
    # the potential buffer the graph reads cannot be freed under it
    with pytest.raises(hip.AtusError, match="captured graphs"):
        ctx.set_potential(0, None)
    # new values in the same buffer are fine; new couplings are refused
    ctx.set_potential(0, 2.0 * V)
    t = G.launch(t)
    ctx.set_g([11.0])
    with pytest.raises(hip.AtusError, match="couplings changed"):
        G.launch(t)
    ctx.set_g([10.0])
    t = G.launch(t)
    # the C-ABI refuses to destroy a context with a live graph ...
    rc = hip._lib.atus_ctx_destroy(ctx._h)
    assert rc == -5 and "graph" in hip._lib.atus_last_error().decode()
    # ... and the Python context closes its graphs before itself
    ctx.close()
    with pytest.raises(hip.AtusError, match="closed"):
        G.launch(t)


def test_copy_components_refuses_overlapping_ranges(hip, orc):
    g = orc.make_grid((64,), L=20.0)
    with hip.Context.from_grid(g, 3) as ctx:
        rng = np.random.default_rng(4)
        ps = [rng.standard_normal(64) + 1j * rng.standard_normal(64) for _ in range(3)]
        for c in range(3):
            ctx.upload(c, ps[c])
        with pytest.raises(hip.AtusError, match="overlapping"):
            ctx.copy_components(1, ctx, 0, 2)
        ctx.copy_components(0, ctx, 0, 3)  # identical ranges: nothing to do
        ctx.copy_components(2, ctx, 0, 1)  # disjoint ranges
        assert np.array_equal(ctx.download(2), ps[0])
        assert np.array_equal(ctx.download(1), ps[1])
