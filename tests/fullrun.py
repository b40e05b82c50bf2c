"""Long parity runs shared by the full-size tests: a reference sequence
advanced block by block on the HIP path (through the C-ABI) and on the
oracle side by side, split over several tests so that no single test runs
for minutes.  Each test advances the run over its own range of blocks (and
first catches up to that range's start when run on its own)."""
import numpy as np


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def chunks(total, size):
    return [(s, min(total, s + size)) for s in range(0, total, size)]


class BraggSequence:
    """The bragg driver's sequence on the 3D two-component workload
    (bragg.cpp:80-120, CRT_Base_IF.h:639-946): bragg_ad for 1.0 with the
    Rabi populations of component 0 after every block, then freeprop for
    0.4; dt = 2e-4, Nk = 100, so 50 + 20 blocks."""

    def __init__(self, hip, orc, wl):
        self.hip, self.orc, self.wl = hip, orc, wl
        self.g = orc.make_grid(wl.shape, L=wl.L)
        p1, p2 = wl.initial_state()
        self.S = orc.RTSolver(self.g, [p1, p2], list(wl.gs), list(wl.alpha), dt=wl.dt, amp=wl.amp, beta0=0.0,
                              deltaL=[0.0, 0.0], amp0=wl.amp0, laser_k=wl.laser_k, laser_domh=wl.laser_domh)
        dx, dk = wl.spacings
        self.ctx = hip.Context(wl.shape, dx, dk, 2)
        self.ctx.upload(0, p1)
        self.ctx.upload(1, p2)
        self.ctx.set_g(wl.gs)
        self.ctx.set_alpha(wl.alpha)
        self.las = hip.laser(amp0=wl.amp0, laser_k=wl.laser_k, laser_domh=wl.laser_domh, deltaL=(0.0, 0.0))
        kinds = {"bragg_ad": (orc.BRAGG_AD, hip.STEP_BRAGG_AD), "freeprop": (orc.FREEPROP_IF, hip.STEP_FREEPROP_IF)}
        self.plan = [(name,) + kinds[name] for name, duration in wl.SEQUENCE for _ in range(wl.blocks(duration))]
        self.t = 0.0
        self.done = 0
        self.n_prev = self.ctx.particle_number(0) + self.ctx.particle_number(1)

    def block(self, check=True):
        """One block on both sides; returns the worst Rabi difference."""
        name, okind, hkind = self.plan[self.done]
        self.S.block(okind, self.wl.Nk)
        self.t = self.ctx.run_block(hkind, self.wl.dt, self.wl.Nk, self.t, self.las, self.wl.amp)
        assert self.t == self.S.t
        worst = 0.0
        if name == "bragg_ad":
            k0 = self.wl.rabi_k0()
            r_gpu = np.asarray(self.ctx.rabi_integrals(0, k0, self.wl.rabi_threshold))
            r_orc, _ = self.orc.rabi_integrals(self.g, self.S.psis[0], k0, self.wl.rabi_threshold)
            worst = float(np.max(np.abs(r_gpu - r_orc)))
            if check:
                assert worst <= 1e-10, (self.done, r_gpu, r_orc)
        n_now = self.ctx.particle_number(0) + self.ctx.particle_number(1)
        if check:
            # total particle number drift <= 1e-12 per step on the device
            assert abs(n_now - self.n_prev) <= 1e-12 * self.wl.Nk, (self.done, n_now - self.n_prev)
        self.n_prev = n_now
        self.done += 1
        return worst

    def run(self, start, stop):
        while self.done < start:
            self.block(check=False)
        worst = 0.0
        while self.done < stop:
            worst = max(worst, self.block())
        err = rel_l2([self.ctx.download(0), self.ctx.download(1)], self.S.psis)
        return err, worst

    def close(self):
        self.ctx.close()
        self.S = None


class SlitSequence:
    """The rt_solver double-slit run (rt_solver.cpp:74-131,
    CRT_Base.h:397-473, 685-841): set_momentum 3,0 then freeprop at
    dt = 5e-4 in blocks of Nk = 100, V read from the potential array."""

    def __init__(self, hip, orc, wl):
        self.hip, self.orc, self.wl = hip, orc, wl
        self.g = orc.make_grid(wl.shape, L=wl.L)
        psi = wl.initial_state()
        V = wl.potential()
        dx, dk = wl.spacings
        self.ctx = hip.Context(wl.shape, dx, dk, 1)
        self.ctx.upload(0, psi)
        self.ctx.set_g([wl.g])
        self.ctx.set_alpha([0.5, 0.5])
        self.ctx.set_potential(0, V)
        self.ctx.set_momentum(0, wl.momentum)
        X = orc.coords_x(self.g)[0] * np.ones(wl.shape)
        self.S = orc.RTSolver(self.g, [psi * np.exp(1j * wl.momentum[0] * X)], [wl.g], [0.5, 0.5], dt=wl.dt,
                              pots=[V])
        assert rel_l2([self.ctx.download(0)], self.S.psis) < 1e-15
        self.nblocks = int(1.0 / wl.dt) // wl.Nk
        self.t = 0.0
        self.done = 0
        self.n_prev = self.ctx.particle_number(0)

    def block(self, check=True):
        self.S.block(self.orc.FREEPROP, self.wl.Nk)
        self.t = self.ctx.run_block(self.hip.STEP_FREEPROP, self.wl.dt, self.wl.Nk, self.t)
        assert self.t == self.S.t
        n_now = self.ctx.particle_number(0)
        if check:
            assert abs(n_now - self.n_prev) <= 1e-12 * self.wl.Nk, (self.done, n_now - self.n_prev)
        self.n_prev = n_now
        self.done += 1

    def run(self, start, stop):
        while self.done < start:
            self.block(check=False)
        while self.done < stop:
            self.block()
        return rel_l2([self.ctx.download(0)], self.S.psis)

    def close(self):
        self.ctx.close()
        self.S = None
