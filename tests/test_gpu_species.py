"""Two species on the GPU (CRT_Base_2 kinetic tables, CRT_Base_IF_2 free
propagation, the bragg_ds two-species Bragg step) against the oracle on
the same seeded inputs."""
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


def _states(orc, g, ncomp, seed):
    xs = orc.coords_x(g)
    rng = np.random.default_rng(seed)
    base = np.exp(-0.5 * sum((x / (2.0 + a)) ** 2 for a, x in enumerate(xs)))
    out = []
    for c in range(ncomp):
        psi = base * (1.0 if c % 2 == 0 else 0.3) + 1e-3 * (rng.standard_normal(base.shape)
                                                            + 1j * rng.standard_normal(base.shape))
        psi = psi.astype(np.complex128)
        psi /= np.sqrt(orc.particle_number(g, psi))
        out.append(psi)
    return out


@pytest.mark.parametrize("shape", [(256,), (32, 64), (16, 8, 32)])
@pytest.mark.parametrize("frac", [0.5, 1.0])
def test_kinetic_uses_the_species_tables(hip, orc, shape, frac):
    dim = len(shape)
    g = orc.make_grid(shape, L=[20.0, 24.0, 28.0][:dim])
    a1, a2 = [0.5, 0.3, 0.7][:dim], [0.21, 0.45, 0.12][:dim]
    psis = _states(orc, g, 4, 3)
    dt = 3e-3
    ref = orc.ft_step2(g, psis, orc.kinetic_table(g, a1, dt, frac), orc.kinetic_table(g, a2, dt, frac))
    with hip.Context.from_grid(g, 4) as ctx:
        for c in range(4):
            ctx.upload(c, psis[c])
        ctx.set_alpha(a1)
        ctx.set_alpha2(a2)
        ctx.kinetic(frac, dt)
        out = [ctx.download(c) for c in range(4)]
    assert rel_l2(out, ref) < 1e-13


def _laser_pair(hip, L0):
    l1 = dict(beta0=0.01, deltaL=[0.0, 0.4, 0.1, -0.3], amp0=5.0, laser_k=3 * 2 * np.pi / L0, laser_domh=8.0,
              laser_dk=0.05, chirp_rate0=0.2, phase0=0.3)
    l2 = dict(beta2=-0.02, amp2=3.5, laser_k2=2 * 2 * np.pi / L0, laser_domh2=6.0, laser_dk2=0.03, chirp_rate2=-0.1,
              phase2=0.7)
    las = hip.laser(beta0=l1["beta0"], deltaL=tuple(l1["deltaL"]), amp0=l1["amp0"], laser_k=l1["laser_k"],
                    laser_domh=l1["laser_domh"], laser_dk=l1["laser_dk"], chirp_rate0=l1["chirp_rate0"],
                    phase0=l1["phase0"])
    las2 = hip.laser(beta0=l2["beta2"], amp0=l2["amp2"], laser_k=l2["laser_k2"], laser_domh=l2["laser_domh2"],
                     laser_dk=l2["laser_dk2"], chirp_rate0=l2["chirp_rate2"], phase0=l2["phase2"])
    return l1, l2, las, las2


@pytest.mark.parametrize("shape", [(512,), (64, 32), (32, 16, 16)])
@pytest.mark.parametrize("kind,ncomp", [("freeprop_if", 2), ("freeprop_if", 4), ("bragg_ds", 4)])
def test_two_species_blocks_match_oracle(hip, orc, shape, kind, ncomp):
    dim = len(shape)
    L = [40.0, 30.0, 36.0][:dim]
    g = orc.make_grid(shape, L=L)
    psis = _states(orc, g, ncomp, 17)
    rng = np.random.default_rng(2)
    gs = list(rng.uniform(5.0, 60.0, ncomp * ncomp))
    a1, a2 = [0.5] * dim, [0.35] * dim
    l1, l2, las, las2 = _laser_pair(hip, L[0])
    dt, Nk, t0 = 2e-3, 25, 0.1
    # the second pulse switches off inside the second block
    seq_time, duration = 0.05, (1.0, 0.05 + 0.1 + 1.5 * Nk * dt - 0.05)
    okind = orc.FREEPROP_IF2 if kind == "freeprop_if" else orc.BRAGG_DS
    sp2 = dict(l2, seq_time=seq_time, duration=duration)
    l1o = dict(l1)
    l1o["deltaL"] = l1["deltaL"][:ncomp]
    S = orc.RTSolver(g, psis, gs, a1, dt=dt, t=t0, alpha2=a2, species2=sp2, **l1o)
    Sp = None
    if kind == "bragg_ds":
        # the noise floor of the reference formula (see the single-species tests)
        pert = [p * (1 + 2e-16 * np.random.default_rng(99).standard_normal(p.shape)) for p in psis]
        Sp = orc.RTSolver(g, pert, gs, a1, dt=dt, t=t0, alpha2=a2, species2=sp2, **l1o)
    hk = hip.STEP_FREEPROP_IF if kind == "freeprop_if" else hip.STEP_BRAGG_AD
    with hip.Context.from_grid(g, ncomp) as ctx:
        for c in range(ncomp):
            ctx.upload(c, psis[c])
        ctx.set_g(gs)
        ctx.set_alpha(a1)
        ctx.set_alpha2(a2)
        t = t0
        n_start = sum(ctx.particle_number(c) for c in range(ncomp))
        for b in range(3):
            S.block(okind, Nk)
            t = ctx.run_block_ds(hk, dt, Nk, t, las, las2, seq_time, duration)
            assert t == S.t
            out = [ctx.download(c) for c in range(ncomp)]
            err = rel_l2(out, S.psis)
            tol = 1e-12
            if Sp is not None:
                Sp.block(okind, Nk)
                tol = max(tol, 20 * rel_l2(Sp.psis, S.psis))
            assert err < tol, (b, err, tol)
            n_now = sum(ctx.particle_number(c) for c in range(ncomp))
            assert abs(n_now - n_start) <= 1e-12 * Nk * (b + 1)


def test_two_species_rejections(hip):
    with hip.Context((64,), [0.5], [2 * np.pi / 32], 3) as ctx:
        with pytest.raises(hip.AtusError, match="even number"):
            ctx.set_alpha2([0.3])
    with hip.Context((64,), [0.5], [2 * np.pi / 32], 2) as ctx:
        with pytest.raises(hip.AtusError, match="four components"):
            ctx.run_block_ds(hip.STEP_BRAGG_AD, 1e-3, 2, 0.0, hip.laser(), hip.laser())
    with hip.Context((8192,), [0.01], [2 * np.pi / 81.92], 2) as ctx:
        with pytest.raises(hip.AtusError, match="not covered"):
            ctx.set_alpha2([0.3])


@pytest.mark.parametrize("shape", [(512,), (32, 16, 16)])
@pytest.mark.parametrize("kind", ["freeprop_if", "bragg_ds"])
def test_single_v_steps_compose_to_the_block(hip, orc, shape, kind):
    # half, (V, full) x (Nk-1), V, half from single device steps, as the
    # host templates call them when the kinetic steps are replaced
    dim = len(shape)
    L = [40.0, 30.0, 36.0][:dim]
    g = orc.make_grid(shape, L=L)
    psis = _states(orc, g, 4, 23)
    gs = list(np.random.default_rng(4).uniform(5.0, 60.0, 16))
    a1, a2 = [0.5] * dim, [0.35] * dim
    _, _, las, las2 = _laser_pair(hip, L[0])
    dt, Nk, t0 = 2e-3, 6, 0.1
    seq_time, duration = 0.05, (1.0, 0.05 + 3.5 * dt)
    hk = hip.STEP_FREEPROP_IF if kind == "freeprop_if" else hip.STEP_BRAGG_AD
    outs = []
    for fused in (True, False):
        with hip.Context.from_grid(g, 4) as ctx:
            for c in range(4):
                ctx.upload(c, psis[c])
            ctx.set_g(gs)
            ctx.set_alpha(a1)
            ctx.set_alpha2(a2)
            if fused:
                t = ctx.run_block_ds(hk, dt, Nk, t0, las, las2, seq_time, duration)
            else:
                t = t0
                ctx.kinetic(0.5, dt)
                t += 0.5 * dt
                for m in range(1, Nk + 1):
                    ctx.v_step_ds(hk, dt, t, las, las2, seq_time, duration)
                    ctx.kinetic(1.0 if m < Nk else 0.5, dt)
                    t += dt if m < Nk else 0.5 * dt
            outs.append(([ctx.download(c) for c in range(4)], t))
    assert outs[0][1] == outs[1][1]
    assert rel_l2(outs[1][0], outs[0][0]) <= 1e-15


def test_single_v_step_rejections(hip):
    with hip.Context((64,), [0.5], [2 * np.pi / 32], 2) as ctx:
        with pytest.raises(hip.AtusError, match="four components"):
            ctx.v_step_ds(hip.STEP_BRAGG_AD, 1e-3, 0.0, hip.laser(), hip.laser())
        with pytest.raises(hip.AtusError, match="freeprop"):
            ctx.v_step_ds(hip.STEP_FREEPROP, 1e-3, 0.0, hip.laser(), hip.laser())
