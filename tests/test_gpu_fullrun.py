"""Full-sequence parity of the 3D two-component Bragg workload between the
HIP path (through the C-ABI) and the oracle.

The sequence is the bragg driver's (bragg.cpp:80-120, CRT_Base_IF.h:639-946):
bragg_ad for 1.0 with the Rabi populations of component 0 after every
block, then freeprop for 0.4; dt = 2e-4, Nk = 100, so 50 + 20 blocks.
This file runs the whole sequence on the 64^3 grid of the same spacing
(dx = 0.25, couplings scaled so that g|psi|^2 is unchanged) and the 2D
double slit on a 512^2 grid; the workloads' own sizes follow in
test_gpu_fullrun_sizes.py.

Gates: relative L2 error of psi <= 1e-10 after the run, total particle
number drift <= 1e-12 per step on the device, Rabi populations of every
block within 1e-10 of the oracle's.
"""
import numpy as np
import pytest

from atus2_amd.workloads import BraggWorkload

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


def _solvers(hip, orc, wl):
    g = orc.make_grid(wl.shape, L=wl.L)
    p1, p2 = wl.initial_state()
    S = orc.RTSolver(g, [p1, p2], list(wl.gs), list(wl.alpha), dt=wl.dt, amp=wl.amp, beta0=0.0,
                     deltaL=[0.0, 0.0], amp0=wl.amp0, laser_k=wl.laser_k, laser_domh=wl.laser_domh)
    dx, dk = wl.spacings
    ctx = hip.Context(wl.shape, dx, dk, 2)
    ctx.upload(0, p1)
    ctx.upload(1, p2)
    ctx.set_g(wl.gs)
    ctx.set_alpha(wl.alpha)
    las = hip.laser(amp0=wl.amp0, laser_k=wl.laser_k, laser_domh=wl.laser_domh, deltaL=(0.0, 0.0))
    return g, S, ctx, las


def _run(hip, orc, wl, max_blocks=None, log=None):
    g, S, ctx, las = _solvers(hip, orc, wl)
    kinds = {"bragg_ad": (orc.BRAGG_AD, hip.STEP_BRAGG_AD), "freeprop": (orc.FREEPROP_IF, hip.STEP_FREEPROP_IF)}
    t = 0.0
    n_prev = ctx.particle_number(0) + ctx.particle_number(1)
    done = 0
    worst_rabi = 0.0
    with ctx:
        for name, duration in wl.SEQUENCE:
            okind, hkind = kinds[name]
            for b in range(wl.blocks(duration)):
                if max_blocks is not None and done == max_blocks:
                    break
                S.block(okind, wl.Nk)
                t = ctx.run_block(hkind, wl.dt, wl.Nk, t, las, wl.amp)
                assert t == S.t
                if name == "bragg_ad":
                    r_gpu = np.asarray(ctx.rabi_integrals(0, wl.rabi_k0(), wl.rabi_threshold))
                    r_orc, _ = orc.rabi_integrals(g, S.psis[0], wl.rabi_k0(), wl.rabi_threshold)
                    worst_rabi = max(worst_rabi, float(np.max(np.abs(r_gpu - r_orc))))
                    assert np.all(np.abs(r_gpu - r_orc) <= 1e-10), (name, b, r_gpu, r_orc)
                n_now = ctx.particle_number(0) + ctx.particle_number(1)
                assert abs(n_now - n_prev) <= 1e-12 * wl.Nk, (name, b, n_now - n_prev)
                n_prev = n_now
                done += 1
                if log is not None:
                    log(f"{name} block {b}: t={t:.4f} N={n_now:.15f}")
        out = [ctx.download(0), ctx.download(1)]
    err = rel_l2(out, S.psis)
    print(f"{wl.n}^3, {done} blocks, t={t:.4f}: relative L2 {err:.2e}, worst Rabi difference {worst_rabi:.2e}")
    return err, t, done


def test_c3_proxy_full_sequence(hip, orc):
    wl = BraggWorkload.proxy(64)
    err, t, done = _run(hip, orc, wl)
    assert done == 70
    assert abs(t - 1.4) < 1e-9
    assert err <= 1e-10


def _rect(x, xl, xr):
    # rect = H(x-xl) - H(x-xr) with H(+0) = 1, i.e. 1 on [xl, xr) (misc.cpp:40-53)
    return ((x >= xl) & (x < xr)).astype(np.float64)


def _c2_run(hip, orc, n, L, nblocks):
    """rt_solver 2D double-slit workload: V = 1e3 rect(x,-0.5,0.5)(1 - rect(y,-5,-3)
    - rect(y,3,5)), psi0 = exp(-((x+20)^2 + y^2)/(2*4^2)) normalised, then
    set_momentum 3,0,0 (e^{+i 3 x}) and freeprop at dt = 5e-4, Nk = 100, g = 1,
    alpha = 1/2 (rt_solver.cpp:74-131, CRT_Base.h:397-473, 685-841)."""
    g = orc.make_grid((n, n), L=L)
    x, y = orc.coords_x(g)
    X, Y = np.broadcast_arrays(x, y)
    V = 1e3 * _rect(X, -0.5, 0.5) * (1.0 - _rect(Y, -5.0, -3.0) - _rect(Y, 3.0, 5.0))
    psi = np.exp(-((X + 20.0) ** 2 + Y ** 2) / (2 * 4.0 ** 2)).astype(np.complex128)
    psi /= np.sqrt(orc.particle_number(g, psi))
    dt, Nk = 5e-4, 100
    with hip.Context.from_grid(g, 1) as ctx:
        ctx.upload(0, psi)
        ctx.set_g([1.0])
        ctx.set_alpha([0.5, 0.5])
        ctx.set_potential(0, V)
        ctx.set_momentum(0, (3.0, 0.0))
        S = orc.RTSolver(g, [psi * np.exp(1j * 3.0 * X)], [1.0], [0.5, 0.5], dt=dt, pots=[V])
        assert rel_l2([ctx.download(0)], S.psis) < 1e-15
        t = 0.0
        n_prev = ctx.particle_number(0)
        for b in range(nblocks):
            S.block(orc.FREEPROP, Nk)
            t = ctx.run_block(hip.STEP_FREEPROP, dt, Nk, t)
            assert t == S.t
            n_now = ctx.particle_number(0)
            assert abs(n_now - n_prev) <= 1e-12 * Nk, (b, n_now - n_prev)
            n_prev = n_now
        out = ctx.download(0)
    err = rel_l2([out], S.psis)
    print(f"{n}^2, {nblocks} blocks, t={t:.4f}: relative L2 {err:.2e}")
    return err, t


def test_c2_proxy_full_run(hip, orc):
    # 512^2 on [-32,32)^2 (dx = 1/8): the packet (velocity 3) stays left of
    # the wall for the whole run, as at 2048^2 on [-64,64)^2
    err, t = _c2_run(hip, orc, 512, 64.0, int(1.0 / 5e-4) // 100)
    assert abs(t - 1.0) < 1e-9
    assert err <= 1e-10
