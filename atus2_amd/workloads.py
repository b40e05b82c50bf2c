"""Synthetic workloads of the benchmark configurations (initial states,
potentials and step parameters), built on the host with numpy.

Grids follow the reference's header conventions: dx = L/n, dk = 2 pi/L,
x_i = (i - n/2) dx (gen_psi_0.cpp:60-88, 199-220).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def axis_coords(n: int, L: float):
    dx = L / n
    return (np.arange(n) - n // 2) * dx


def grid_spacings(shape, L):
    Ls = [L] * len(shape) if np.isscalar(L) else list(L)
    dx = [Ls[a] / shape[a] for a in range(len(shape))]
    dk = [2 * math.pi / Ls[a] for a in range(len(shape))]
    return dx, dk


def particle_number(psi, dx):
    return float(np.prod(dx) * np.sum(psi.real ** 2 + psi.imag ** 2))


@dataclass
class BraggWorkload:
    """3D two-component Bragg pulse: 256^3 on [-32,32]^3, psi1 a Gaussian of
    widths (6,4,4) normalised to 1, psi2 = 0; g11=g22=100, g12=95;
    Omega(t) = 5 exp(-((t-0.5)/0.5)^2/2) cos(4 x - 8 t) (the coupling runs
    along the first axis, as in Do_Bragg_ad); dt = 2e-4, Nk = 100; the Rabi
    populations of component 0 at k = 0, +-8 are taken after every block."""
    n: int = 256
    L: float = 64.0
    widths: tuple = (6.0, 4.0, 4.0)
    gs: tuple = (100.0, 95.0, 95.0, 100.0)
    alpha: tuple = (0.5, 0.5, 0.5)
    amp0: float = 5.0
    laser_k: float = 4.0
    laser_domh: float = 8.0
    pulse_t0: float = 0.5
    pulse_sigma: float = 0.5
    dt: float = 2e-4
    Nk: int = 100
    rabi_threshold: float = 2.0
    shape: tuple = field(init=False)

    def __post_init__(self):
        self.shape = (self.n, self.n, self.n)

    @property
    def spacings(self):
        return grid_spacings(self.shape, self.L)

    def initial_state(self):
        x = axis_coords(self.n, self.L)
        wx, wy, wz = self.widths
        gx = np.exp(-x ** 2 / (2 * wx ** 2))
        gy = np.exp(-x ** 2 / (2 * wy ** 2))
        gz = np.exp(-x ** 2 / (2 * wz ** 2))
        psi1 = (gx[:, None, None] * gy[None, :, None] * gz[None, None, :]).astype(np.complex128)
        dx, _ = self.spacings
        psi1 /= math.sqrt(particle_number(psi1, dx))
        psi2 = np.zeros_like(psi1)
        return psi1, psi2

    def amp(self, t):
        return math.exp(-0.5 * ((t - self.pulse_t0) / self.pulse_sigma) ** 2)

    def rabi_k0(self):
        k = 2 * self.laser_k
        return [(0.0, 0.0, 0.0), (k, 0.0, 0.0), (-k, 0.0, 0.0)]

    # the reference's bragg sequence for this workload: bragg_ad for 1.0
    # (Rabi populations after every block), then freeprop for 0.4, both at
    # dt = 2e-4 and Nk = 100 (int(1.0/2e-4) = 5000, int(0.4/2e-4) = 2000)
    SEQUENCE = (("bragg_ad", 1.0), ("freeprop", 0.4))

    def blocks(self, duration):
        """Number of Nk-step blocks of a sequence item (CRT_Base.h:742-755)."""
        return int(duration / self.dt) // self.Nk

    @classmethod
    def proxy(cls, n):
        """The same physics on an n^3 grid of the same spacing (dx = 0.25):
        box and widths scaled by s = n/256, couplings by s^3, so that the
        peak g*|psi|^2 (the nonlinear phase rate; the density scales as
        1/s^3) is that of the 256^3 case."""
        s = n / 256.0
        return cls(n=n, L=64.0 * s, widths=tuple(w * s for w in cls.widths),
                   gs=tuple(g * s ** 3 for g in cls.gs))


@dataclass
class HarmonicLineWorkload:
    """1D one-component rt_solver run: N = 4096 on [-40,40), V = x^2/2,
    g = 10, alpha = 1/2, psi0 = pi^-1/4 exp(-x^2/2) plus seeded 1e-3 complex
    noise, normalised to 1; freeprop at dt = 1e-3 in blocks of Nk = 100,
    norm and energy after every block (rt_solver.cpp:74-131,
    CRT_Base.h:397-441, 685-841)."""
    n: int = 4096
    L: float = 80.0
    g: float = 10.0
    dt: float = 1e-3
    Nk: int = 100
    seed: int = 1
    shape: tuple = field(init=False)

    def __post_init__(self):
        self.shape = (self.n,)

    @property
    def spacings(self):
        return grid_spacings(self.shape, self.L)

    def potential(self):
        x = axis_coords(self.n, self.L)
        return 0.5 * x ** 2

    def initial_state(self):
        x = axis_coords(self.n, self.L)
        rng = np.random.default_rng(self.seed)
        psi = np.pi ** -0.25 * np.exp(-0.5 * x ** 2) + 1e-3 * (rng.standard_normal(self.n)
                                                             + 1j * rng.standard_normal(self.n))
        psi = psi.astype(np.complex128)
        dx, _ = self.spacings
        psi /= math.sqrt(particle_number(psi, dx))
        return psi


@dataclass
class DoubleSlitWorkload:
    """2D one-component rt_solver run: 2048^2 on [-64,64)^2, a wall
    V = 1e3 rect(x,-0.5,0.5) (1 - rect(y,-5,-3) - rect(y,3,5)) (rect is 1 on
    [xl,xr), misc.cpp:40-53), psi0 = exp(-((x+20)^2 + y^2)/(2*4^2))
    normalised, then the momentum (3, 0) (CRT_Base.h:449-473); g = 1,
    alpha = 1/2, freeprop at dt = 5e-4 in blocks of Nk = 100."""
    n: int = 2048
    L: float = 128.0
    g: float = 1.0
    dt: float = 5e-4
    Nk: int = 100
    momentum: tuple = (3.0, 0.0)
    shape: tuple = field(init=False)

    def __post_init__(self):
        self.shape = (self.n, self.n)

    @property
    def spacings(self):
        return grid_spacings(self.shape, self.L)

    def _xy(self):
        x = axis_coords(self.n, self.L)
        return np.broadcast_arrays(x[:, None], x[None, :])

    def potential(self):
        X, Y = self._xy()

        def rect(v, lo, hi):
            return ((v >= lo) & (v < hi)).astype(np.float64)
        return 1e3 * rect(X, -0.5, 0.5) * (1.0 - rect(Y, -5.0, -3.0) - rect(Y, 3.0, 5.0))

    def initial_state(self):
        X, Y = self._xy()
        psi = np.exp(-((X + 20.0) ** 2 + Y ** 2) / (2 * 4.0 ** 2)).astype(np.complex128)
        dx, _ = self.spacings
        psi /= math.sqrt(particle_number(psi, dx))
        return psi
