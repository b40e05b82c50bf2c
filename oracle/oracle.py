"""ctypes binding of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, always as the checker.  The product path
(atus2_amd) never imports this module.

Fields are numpy complex128 arrays shaped (nx,), (nx, ny) or (nx, ny, nz),
C-contiguous -- the fftw_complex layout of the reference (cft_base.h:209).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_longlong * 3),
                ("dx", C.c_double * 3), ("dk", C.c_double * 3)]


class BraggParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "dt", "t", "F", "amp0", "laser_k", "laser_domh", "laser_dk", "chirp",
        "chirp_rate0", "phase0", "beta0", "deltaL1")] + [("gs", C.c_double * 4)]


AMP_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class RT(C.Structure):
    _fields_ = [("g", Grid), ("ncomp", C.c_int),
                ("psi", C.c_void_p * 4), ("pot", C.c_void_p * 4),
                ("gs", C.c_double * 16), ("alpha", C.c_double * 3),
                ("dt", C.c_double), ("t", C.c_double),
                ("half_tab", C.c_void_p), ("full_tab", C.c_void_p),
                ("tab_dt", C.c_double),
                ("beta0", C.c_double), ("deltaL", C.c_double * 4),
                ("amp0", C.c_double), ("laser_k", C.c_double),
                ("laser_domh", C.c_double), ("laser_dk", C.c_double),
                ("chirp", C.c_double), ("chirp_rate0", C.c_double),
                ("phase0", C.c_double),
                ("amp", AMP_FN), ("amp_user", C.c_void_p),
                ("two_species", C.c_int), ("alpha2", C.c_double * 3),
                ("half_tab2", C.c_void_p), ("full_tab2", C.c_void_p),
                ("beta2", C.c_double), ("amp2", C.c_double), ("laser_k2", C.c_double),
                ("laser_domh2", C.c_double), ("laser_dk2", C.c_double), ("chirp_rate2", C.c_double),
                ("phase2", C.c_double), ("seq_time", C.c_double), ("duration", C.c_double * 2),
                ("amp1", C.c_double), ("beta", C.c_double * 3),
                ("jitter", C.c_double), ("jitter_seed", C.c_ulonglong), ("jitter_count", C.c_longlong)]


class SOB(C.Structure):
    _fields_ = [("g", Grid), ("nwf", C.c_int),
                ("psi", C.c_void_p * 2), ("lap", C.c_void_p * 2),
                ("sob", C.c_void_p * 2), ("grad", C.c_void_p * 2),
                ("op", C.c_void_p * 2), ("lapop", C.c_void_p * 2),
                ("pot", C.c_void_p * 2),
                ("gs", C.c_double * 4), ("alpha", C.c_double * 2),
                ("stepsize", C.c_double), ("res", C.c_double * 2),
                ("res_tot", C.c_double), ("N", C.c_double * 2),
                ("renormalised", C.c_int)]


FREEPROP, FREEPROP_IF, BRAGG_AD, FREEPROP_IF2, BRAGG_DS = 0, 1, 2, 3, 4
DOUBLE_BRAGG, NUM_BRAGG, RAMAN = 5, 6, 7

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P, D, I, LL = C.c_void_p, C.c_double, C.c_int, C.c_longlong
        G = C.POINTER(Grid)
        sig = {
            "orc_npts": (LL, [G]), "orc_threads": (I, []), "orc_ar": (D, [G]), "orc_ar_k": (D, [G]),
            "orc_get_x": (None, [G, LL, P]), "orc_get_k": (None, [G, LL, P]),
            "orc_fft": (None, [G, P, I]), "orc_ft": (None, [G, P, I]),
            "orc_kinetic_table": (None, [G, P, D, D, P]),
            "orc_ft_step": (None, [G, I, P, P]),
            "orc_nl_step": (None, [G, I, P, P, D, P]),
            "orc_nl_if_step": (None, [G, I, P, P, D, D, P]),
            "orc_bragg_ad": (None, [G, P, P, C.POINTER(BraggParams)]),
            "orc_ft_step2": (None, [G, I, P, P, P]),
            "orc_nl_if2_step": (None, [G, I, P, P, D, D, D, P]),
            "orc_bragg_ds": (None, [G, P, C.POINTER(RT), D, I, I]),
            "orc_herm_expm": (None, [I, P, D, P]),
            "orc_double_bragg": (None, [G, P, C.POINTER(RT), D, D]),
            "orc_num_bragg": (None, [G, I, P, C.POINTER(RT), D]),
            "orc_raman": (None, [G, I, P, C.POINTER(RT)]),
            "orc_particle_number": (D, [G, P]),
            "orc_expval_x": (None, [G, P, P]), "orc_expval_k": (None, [G, P, P]),
            "orc_rabi_integrals": (None, [G, P, P, I, D, P]),
            "orc_energy": (D, [G, I, P, P, P, P]),
            "orc_rt_block": (None, [C.POINTER(RT), I, I]),
            "orc_rt_set_dt": (None, [C.POINTER(RT), D]),
            "orc_sob_init_ops": (None, [C.POINTER(SOB)]),
            "orc_sob_renormalize": (None, [C.POINTER(SOB)]),
            "orc_sob_iter": (None, [C.POINTER(SOB)]),
            "orc_sob_mu": (None, [C.POINTER(SOB), P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def make_grid(shape, L=None, dx=None, dk=None) -> Grid:
    """Grid of a symmetric box; `L` (per axis) gives dx = L/n, dk = 2 pi/L as
    gen_psi_0.cpp:60-88 writes them into the header."""
    shape = tuple(int(s) for s in shape)
    dim = len(shape)
    n = list(shape) + [1] * (3 - dim)
    if L is not None:
        Ls = [float(L)] * dim if np.isscalar(L) else [float(v) for v in L]
        dx = [Ls[a] / n[a] for a in range(dim)]
        dk = [2 * np.pi / Ls[a] for a in range(dim)]
    dx = list(dx) + [0.0] * (3 - dim)
    dk = list(dk) + [0.0] * (3 - dim)
    g = Grid()
    g.dim = dim
    for a in range(3):
        g.n[a] = n[a]
        g.dx[a] = dx[a]
        g.dk[a] = dk[a]
    return g


def grid_shape(g: Grid):
    return tuple(int(g.n[a]) for a in range(g.dim))


def threads():
    return lib().orc_threads()


def ar(g):
    return lib().orc_ar(C.byref(g))


def ar_k(g):
    return lib().orc_ar_k(C.byref(g))


def coords_x(g):
    """x arrays per axis (x_i = (i - n/2) dx), broadcastable to the grid."""
    out = []
    for a in range(g.dim):
        n = int(g.n[a])
        v = (np.arange(n) - n // 2) * g.dx[a]
        shp = [1] * g.dim
        shp[a] = n
        out.append(v.reshape(shp))
    return out


def coords_k(g):
    """k arrays per axis in native FFT order."""
    out = []
    for a in range(g.dim):
        n = int(g.n[a])
        t = (np.arange(n) + n // 2) % n
        v = g.dk[a] * (t - n // 2).astype(np.float64)
        shp = [1] * g.dim
        shp[a] = n
        out.append(v.reshape(shp))
    return out


def fft(g, psi, sign):
    a = _c128(psi).copy()
    lib().orc_fft(C.byref(g), _ptr(a), int(sign))
    return a


def ft(g, psi, isign):
    a = _c128(psi).copy()
    lib().orc_ft(C.byref(g), _ptr(a), int(isign))
    return a


def kinetic_table(g, alpha, dt, s):
    tab = np.empty(grid_shape(g), dtype=np.complex128)
    al = (C.c_double * 3)(*(list(alpha) + [0.0] * (3 - len(alpha))))
    lib().orc_kinetic_table(C.byref(g), al, float(dt), float(s), _ptr(tab))
    return tab


def _ptr_array(arrs, n=4):
    arr = (C.c_void_p * n)()
    for i, a in enumerate(arrs):
        arr[i] = a.ctypes.data if a is not None else None
    return arr


def ft_step(g, psis, tab):
    psis = [_c128(p).copy() for p in psis]
    tab = _c128(tab)
    lib().orc_ft_step(C.byref(g), len(psis), _ptr_array(psis), _ptr(tab))
    return psis


def nl_step(g, psis, gs, dt, pots=None):
    psis = [_c128(p).copy() for p in psis]
    gs = np.ascontiguousarray(gs, dtype=np.float64).ravel()
    pp = None
    if pots is not None:
        pots = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
                for v in pots]
        pp = _ptr_array(pots)
    lib().orc_nl_step(C.byref(g), len(psis), _ptr_array(psis), _ptr(gs),
                      float(dt), pp)
    return psis


def nl_if_step(g, psis, gs, dt, beta0, deltaL):
    psis = [_c128(p).copy() for p in psis]
    gs = np.ascontiguousarray(gs, dtype=np.float64).ravel()
    dl = np.ascontiguousarray(deltaL, dtype=np.float64)
    lib().orc_nl_if_step(C.byref(g), len(psis), _ptr_array(psis), _ptr(gs),
                         float(dt), float(beta0), _ptr(dl))
    return psis


def bragg_params(**kw) -> BraggParams:
    p = BraggParams()
    defaults = dict(dt=1e-3, t=0.0, F=1.0, amp0=0.0, laser_k=0.0, laser_domh=0.0,
                    laser_dk=0.0, chirp=0.0, chirp_rate0=0.0, phase0=0.0,
                    beta0=0.0, deltaL1=0.0)
    gs = kw.pop("gs", (0.0, 0.0, 0.0, 0.0))
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, float(v))
    for i in range(4):
        p.gs[i] = float(gs[i])
    return p


def bragg_ad(g, psi1, psi2, params: BraggParams):
    a = _c128(psi1).copy()
    b = _c128(psi2).copy()
    lib().orc_bragg_ad(C.byref(g), _ptr(a), _ptr(b), C.byref(params))
    return a, b


def particle_number(g, psi):
    return lib().orc_particle_number(C.byref(g), _ptr(_c128(psi)))


def expval_x(g, psi):
    out = (C.c_double * 3)()
    lib().orc_expval_x(C.byref(g), _ptr(_c128(psi)), out)
    return np.array(out[:g.dim])


def expval_k(g, psi):
    a = _c128(psi).copy()
    out = (C.c_double * 3)()
    lib().orc_expval_k(C.byref(g), _ptr(a), out)
    return np.array(out[:g.dim]), a


def rabi_integrals(g, psi, k0list, threshold):
    a = _c128(psi).copy()
    k0 = np.zeros((len(k0list), 3))
    for i, k in enumerate(k0list):
        k0[i, :len(k)] = k
    k0 = np.ascontiguousarray(k0)
    out = np.zeros(len(k0list))
    lib().orc_rabi_integrals(C.byref(g), _ptr(a), _ptr(k0), len(k0list),
                             float(threshold), _ptr(out))
    return out, a


def energy(g, psis, gs, alpha, pots=None):
    psis = [_c128(p) for p in psis]
    gs = np.ascontiguousarray(gs, dtype=np.float64).ravel()
    al = (C.c_double * 3)(*(list(alpha) + [0.0] * (3 - len(alpha))))
    pp = None
    if pots is not None:
        pots = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
                for v in pots]
        pp = _ptr_array(pots)
    return lib().orc_energy(C.byref(g), len(psis), _ptr_array(psis), _ptr(gs),
                            al, pp)


class RTSolver:
    """Owns the arrays of an orc_rt and runs reference blocks on them."""

    def __init__(self, g, psis, gs, alpha, dt, t=0.0, pots=None, amp=None, alpha2=None, species2=None,
                 jitter=0.0, jitter_seed=1, **ifp):
        """alpha2: kinetic coefficients of species 2 (components [n/2, n),
        CRT_Base_2); species2: dict of beta2, amp2, laser_k2, laser_domh2,
        laser_dk2, chirp_rate2, phase2, seq_time, duration (CRT_Base_IF_2,
        bragg_ds); jitter: relative size of the rounding model applied after
        every kinetic and V application of block() (0: none)."""
        self.g = g
        self.psis = [_c128(p).copy() for p in psis]
        self.pots = None
        if pots is not None:
            self.pots = [np.ascontiguousarray(v, dtype=np.float64) if v is not None
                         else None for v in pots]
        shape = grid_shape(g)
        self.half = np.empty(shape, np.complex128)
        self.full = np.empty(shape, np.complex128)
        s = RT()
        s.g = g
        s.ncomp = len(psis)
        for i, p in enumerate(self.psis):
            s.psi[i] = p.ctypes.data
        if self.pots is not None:
            for i, v in enumerate(self.pots):
                s.pot[i] = v.ctypes.data if v is not None else None
        gsa = np.ravel(gs)
        for i in range(len(gsa)):
            s.gs[i] = float(gsa[i])
        for a in range(len(alpha)):
            s.alpha[a] = float(alpha[a])
        s.t = float(t)
        s.jitter = float(jitter)
        s.jitter_seed = int(jitter_seed)
        s.half_tab = self.half.ctypes.data
        s.full_tab = self.full.ctypes.data
        dl = ifp.pop("deltaL", (0.0, 0.0, 0.0, 0.0))
        for i in range(len(dl)):
            s.deltaL[i] = float(dl[i])
        for k in ("beta0", "amp0", "laser_k", "laser_domh", "laser_dk", "chirp",
                  "chirp_rate0", "phase0", "amp1"):
            setattr(s, k, float(ifp.pop(k, 0.0)))
        beta = ifp.pop("beta", (0.0, 0.0, 0.0))
        for a in range(len(beta)):
            s.beta[a] = float(beta[a])
        if ifp:
            raise TypeError(f"unknown parameters {sorted(ifp)}")
        self._amp_cb = AMP_FN(lambda t, u: float(amp(t))) if amp else AMP_FN()
        s.amp = self._amp_cb
        if alpha2 is not None:
            s.two_species = 1
            for a in range(len(alpha2)):
                s.alpha2[a] = float(alpha2[a])
            self.half2 = np.empty(shape, np.complex128)
            self.full2 = np.empty(shape, np.complex128)
            s.half_tab2 = self.half2.ctypes.data
            s.full_tab2 = self.full2.ctypes.data
        sp2 = dict(species2 or {})
        dur = sp2.pop("duration", (0.0, 0.0))
        s.duration[0], s.duration[1] = float(dur[0]), float(dur[1])
        for k in ("beta2", "amp2", "laser_k2", "laser_domh2", "laser_dk2", "chirp_rate2", "phase2", "seq_time"):
            setattr(s, k, float(sp2.pop(k, 0.0)))
        if sp2:
            raise TypeError(f"unknown species-2 parameters {sorted(sp2)}")
        self.s = s
        lib().orc_rt_set_dt(C.byref(s), float(dt))

    @property
    def t(self):
        return self.s.t

    def set_dt(self, dt):
        lib().orc_rt_set_dt(C.byref(self.s), float(dt))

    def block(self, kind, Nk):
        lib().orc_rt_block(C.byref(self.s), int(kind), int(Nk))


class SobSolver:
    """Owns the arrays of an orc_sob (CSOB_Base state)."""

    def __init__(self, g, psis, pots, gs_scaled, alpha, stepsize):
        self.g = g
        shape = grid_shape(g)
        nw = len(psis)
        self.psi = [_c128(p).copy() for p in psis]
        self.lap = [np.zeros(shape, np.complex128) for _ in range(nw)]
        self.sob = [np.zeros(shape, np.complex128) for _ in range(nw)]
        self.grad = [np.zeros(shape, np.complex128) for _ in range(nw)]
        self.op = [np.zeros(shape) for _ in range(nw)]
        self.lapop = [np.zeros(shape) for _ in range(nw)]
        self.pot = [np.ascontiguousarray(v, dtype=np.float64) for v in pots]
        s = SOB()
        s.g = g
        s.nwf = nw
        for i in range(nw):
            s.psi[i] = self.psi[i].ctypes.data
            s.lap[i] = self.lap[i].ctypes.data
            s.sob[i] = self.sob[i].ctypes.data
            s.grad[i] = self.grad[i].ctypes.data
            s.op[i] = self.op[i].ctypes.data
            s.lapop[i] = self.lapop[i].ctypes.data
            s.pot[i] = self.pot[i].ctypes.data
            s.alpha[i] = float(alpha[i])
        gsa = np.ravel(gs_scaled)
        for i in range(len(gsa)):
            s.gs[i] = float(gsa[i])
        s.stepsize = float(stepsize)
        self.s = s
        lib().orc_sob_init_ops(C.byref(s))

    def renormalize(self):
        lib().orc_sob_renormalize(C.byref(self.s))

    def iterate(self):
        lib().orc_sob_iter(C.byref(self.s))
        return self.s.res_tot

    def mu(self):
        out = np.zeros(self.s.nwf)
        lib().orc_sob_mu(C.byref(self.s), _ptr(out))
        return out


def ft_step2(g, psis, tab1, tab2):
    """CRT_Base_2 kinetic step: tab1 on components [0, n/2), tab2 on the rest."""
    out = [_c128(p).copy() for p in psis]
    lib().orc_ft_step2(C.byref(g), len(out), _ptr_array(out), _ptr(_c128(tab1)), _ptr(_c128(tab2)))
    return out


def nl_if2_step(g, psis, gs, dt, beta, beta2, deltaL):
    out = [_c128(p).copy() for p in psis]
    gsa = np.ascontiguousarray(np.ravel(gs), dtype=np.float64)
    dl = np.ascontiguousarray(deltaL, dtype=np.float64)
    lib().orc_nl_if2_step(C.byref(g), len(out), _ptr_array(out), _ptr(gsa), float(dt), float(beta), float(beta2),
                          _ptr(dl))
    return out


def bragg_ds(g, psis, solver, t1, mode1, mode2):
    """One two-species Bragg step (bragg_ds) with the parameters of an
    RTSolver built with species2=...; returns new arrays."""
    out = [_c128(p).copy() for p in psis]
    lib().orc_bragg_ds(C.byref(g), _ptr_array(out), C.byref(solver.s), float(t1), int(mode1), int(mode2))
    return out


def herm_expm(H, dt):
    """exp(-i dt H) of a Hermitian n x n matrix (n <= 4), by the oracle's
    long-double Jacobi eigen-decomposition."""
    H = np.ascontiguousarray(H, dtype=np.complex128)
    n = H.shape[0]
    U = np.zeros((n, n), np.complex128)
    lib().orc_herm_expm(n, _ptr(H), float(dt), _ptr(U))
    return U


def three_state_step(kind, g, psis, solver, t, F=1.0):
    """One three-state V step (DOUBLE_BRAGG at t1 = t + dt/2 with amplitude
    F, NUM_BRAGG at Get_t() = t, RAMAN) with the parameters of an RTSolver;
    returns new arrays."""
    out = [_c128(p).copy() for p in psis]
    if kind == DOUBLE_BRAGG:
        lib().orc_double_bragg(C.byref(g), _ptr_array(out), C.byref(solver.s), float(t + 0.5 * solver.s.dt), float(F))
    elif kind == NUM_BRAGG:
        lib().orc_num_bragg(C.byref(g), len(out), _ptr_array(out), C.byref(solver.s), float(t))
    elif kind == RAMAN:
        lib().orc_raman(C.byref(g), len(out), _ptr_array(out), C.byref(solver.s))
    else:
        raise ValueError(kind)
    return out
