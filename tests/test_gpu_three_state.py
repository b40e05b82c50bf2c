"""Three internal states on the GPU: the double Bragg step (closed form of
double_bragg.cpp), the numerical Bragg and Raman steps (Hermitian matrix
exponentials of CRT_Base_IF) against the oracle, as single V steps and as
blocks with kinetic steps."""
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


def _states(orc, g, n, seed):
    xs = orc.coords_x(g)
    rng = np.random.default_rng(seed)
    base = np.exp(-0.5 * sum((x / (2.0 + a)) ** 2 for a, x in enumerate(xs)))
    out = []
    for c in range(n):
        p = base * (1.0, 0.5, 0.3, 0.2)[c] + 1e-3 * (rng.standard_normal(base.shape)
                                                      + 1j * rng.standard_normal(base.shape))
        p = p.astype(np.complex128)
        p /= np.sqrt(orc.particle_number(g, p)) * (1.0 + c)
        out.append(p)
    return out


KINDS = {
    # name: (hip kind, oracle kind, ncomp, laser parameters)
    "double_bragg": ("STEP_DOUBLE_BRAGG", "DOUBLE_BRAGG", 3,
                     dict(amp0=6.0, laser_k=1.1, laser_domh=0.6, laser_dk=0.3, chirp_rate0=0.2, phase0=0.3,
                          deltaL=[0.0, -0.7, 0.0])),
    "num_bragg2": ("STEP_NUM_BRAGG", "NUM_BRAGG", 2,
                   dict(amp0=5.0, amp1=3.0, laser_k=1.1, laser_domh=0.6, chirp_rate0=0.2, phase0=0.3,
                        deltaL=[0.1, -0.7])),
    "num_bragg3": ("STEP_NUM_BRAGG", "NUM_BRAGG", 3,
                   dict(amp0=5.0, amp1=3.0, laser_k=1.1, laser_domh=0.6, chirp_rate0=0.2, phase0=0.3,
                        deltaL=[0.1, -0.7, 0.4])),
    "raman3": ("STEP_RAMAN", "RAMAN", 3,
               dict(amp0=5.0, amp1=3.0, laser_k=1.1, laser_domh=4.0, deltaL=[0.0, -2.0, 1.5])),
    "raman4": ("STEP_RAMAN", "RAMAN", 4,
               dict(amp0=5.0, amp1=3.0, laser_k=1.1, laser_domh=4.0, deltaL=[0.0, -2.0, 1.5, 0.3])),
}


def _setup(hip, orc, shape, name, seed, amp=None, dt=2e-3, t0=0.1):
    hk, ok, nc, lp = KINDS[name]
    dim = len(shape)
    g = orc.make_grid(shape, L=[20.0, 16.0, 18.0][:dim])
    psis = _states(orc, g, nc, seed)
    gs = list(np.random.default_rng(seed + 1).uniform(2.0, 30.0, nc * nc))
    beta = [0.03, -0.02, 0.01][:dim]
    lp = dict(lp, beta=beta)
    S = orc.RTSolver(g, psis, gs, [0.5] * dim, dt=dt, t=t0, amp=amp, **lp)
    # the noise floor of the formulas: the same run from inputs perturbed
    # in the last bits
    pert = [p * (1 + 2e-16 * np.random.default_rng(99).standard_normal(p.shape)) for p in psis]
    Sp = orc.RTSolver(g, pert, gs, [0.5] * dim, dt=dt, t=t0, amp=amp, **lp)
    las = hip.laser(**dict(lp, beta=beta + [0.0] * (3 - dim)))
    return g, psis, gs, S, Sp, las, getattr(hip, hk), getattr(orc, ok), nc


@pytest.mark.parametrize("shape", [(256,), (32, 16), (16, 8, 8)])
@pytest.mark.parametrize("name", list(KINDS))
def test_single_v_step_matches_oracle(hip, orc, shape, name):
    g, psis, gs, S, Sp, las, hk, ok, nc = _setup(hip, orc, shape, name, 3)
    t, F = 0.37, 0.8
    ref = orc.three_state_step(ok, g, psis, S, t, F)
    pert = [p * (1 + 2e-16 * np.random.default_rng(98).standard_normal(p.shape)) for p in psis]
    refp = orc.three_state_step(ok, g, pert, S, t, F)
    with hip.Context.from_grid(g, nc) as ctx:
        for c in range(nc):
            ctx.upload(c, psis[c])
        ctx.set_g(gs)
        ctx.v_step(hk, S.s.dt, t, las, F)
        out = [ctx.download(c) for c in range(nc)]
    tol = max(2e-14, 20 * rel_l2(refp, ref))
    assert rel_l2(out, ref) < tol, (rel_l2(out, ref), tol)


@pytest.mark.parametrize("shape", [(512,), (32, 64), (16, 16, 16)])
@pytest.mark.parametrize("name", list(KINDS))
def test_blocks_match_oracle(hip, orc, shape, name):
    amp = (lambda t: float(np.exp(-0.5 * ((t - 0.12) / 0.05) ** 2))) if name == "double_bragg" else None
    g, psis, gs, S, Sp, las, hk, ok, nc = _setup(hip, orc, shape, name, 5, amp=amp)
    Nk = 15
    with hip.Context.from_grid(g, nc) as ctx:
        for c in range(nc):
            ctx.upload(c, psis[c])
        ctx.set_g(gs)
        ctx.set_alpha([0.5] * len(shape))
        t = S.t
        N0 = sum(ctx.particle_number(c) for c in range(nc))
        for b in range(2):
            S.block(ok, Nk)
            Sp.block(ok, Nk)
            t = ctx.run_block(hk, S.s.dt, Nk, t, las, amp)
            assert t == S.t
            out = [ctx.download(c) for c in range(nc)]
            tol = max(1e-12, 20 * rel_l2(Sp.psis, S.psis))
            assert rel_l2(out, S.psis) < tol, (b, rel_l2(out, S.psis), tol)
        N = sum(ctx.particle_number(c) for c in range(nc))
        assert abs(N - sum(orc.particle_number(g, q) for q in S.psis)) < 1e-12 * N0
        # the numerical steps are unitary; the double Bragg closed form is
        # not when laser_dk != 0 (its O23 / O32 phases), in the reference too
        if name != "double_bragg":
            assert abs(N - N0) < 1e-12 * N0


def test_single_steps_compose_to_the_block(hip, orc):
    g, psis, gs, S, Sp, las, hk, ok, nc = _setup(hip, orc, (32, 16), "raman3", 7)
    dt, Nk, t0 = S.s.dt, 5, 0.1
    outs = []
    for fused in (True, False):
        with hip.Context.from_grid(g, nc) as ctx:
            for c in range(nc):
                ctx.upload(c, psis[c])
            ctx.set_g(gs)
            ctx.set_alpha([0.5, 0.5])
            if fused:
                t = ctx.run_block(hk, dt, Nk, t0, las)
            else:
                t = t0
                ctx.kinetic(0.5, dt)
                t += 0.5 * dt
                for m in range(1, Nk + 1):
                    ctx.v_step(hk, dt, t, las)
                    ctx.kinetic(1.0 if m < Nk else 0.5, dt)
                    t += dt if m < Nk else 0.5 * dt
            outs.append(([ctx.download(c) for c in range(nc)], t))
    assert outs[0][1] == outs[1][1]
    assert rel_l2(outs[1][0], outs[0][0]) <= 1e-15


def test_three_state_rejections(hip):
    with hip.Context((64,), [0.5], [2 * np.pi / 32], 2) as ctx:
        with pytest.raises(hip.AtusError, match="three components"):
            ctx.v_step(hip.STEP_DOUBLE_BRAGG, 1e-3, 0.0, hip.laser())
        with pytest.raises(hip.AtusError, match="three or four"):
            ctx.run_block(hip.STEP_RAMAN, 1e-3, 2, 0.0, hip.laser())
    with hip.Context((64,), [0.5], [2 * np.pi / 32], 4) as ctx:
        with pytest.raises(hip.AtusError, match="two or three"):
            ctx.v_step(hip.STEP_NUM_BRAGG, 1e-3, 0.0, hip.laser())
    with hip.Context((64,), [0.5], [2 * np.pi / 32], 3) as ctx:
        with pytest.raises(hip.AtusError, match="graphs replay"):
            hip.Graph(ctx, hip.STEP_RAMAN, 1e-3, 2, hip.laser())
