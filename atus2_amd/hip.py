"""ctypes binding of libatus_hip.so (include/atus_hip.h).

This is the Python face of the C-ABI: it loads the in-tree library and
raises if it is missing -- there is no CPU fallback.  Fields are numpy
complex128 arrays in the reference's fftw_complex layout.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libatus_hip.so")

STEP_FREEPROP, STEP_FREEPROP_IF, STEP_BRAGG_AD = 0, 1, 2
STEP_DOUBLE_BRAGG, STEP_NUM_BRAGG, STEP_RAMAN = 3, 4, 5

EXPORTED = [
    "atus_last_error", "atus_version", "atus_device_info", "atus_ctx_create", "atus_ctx_destroy",
    "atus_synchronize", "atus_upload", "atus_download", "atus_set_potential", "atus_set_g",
    "atus_set_alpha", "atus_ft", "atus_kinetic", "atus_nl_step", "atus_nl_if_step", "atus_bragg_ad",
    "atus_set_momentum", "atus_run_block", "atus_particle_number", "atus_expval_x", "atus_expval_k",
    "atus_rabi_integrals", "atus_energy", "atus_sob_setup", "atus_sob_renormalize", "atus_sob_iter",
    "atus_sob_mu", "atus_sob_iterate", "atus_set_timing", "atus_last_block_ms", "atus_run_block_profiled", "atus_copy_bandwidth", "atus_tune", "atus_tune_threads",
    "atus_batch_create", "atus_batch_set_runs", "atus_batch_run_block", "atus_batch_particle_numbers",
    "atus_batch_run_block_profiled",
    "atus_batch_particle_numbers_noalloc",
    "atus_batch_rabi_integrals", "atus_sob_run", "atus_graph_create", "atus_graph_launch", "atus_graph_destroy",
    "atus_scan_graph_create", "atus_scan_graph_launch", "atus_scan_graph_destroy",
    "atus_set_alpha2", "atus_run_block_ds", "atus_v_step_ds",
    "atus_snapshot_save", "atus_snapshot_restore", "atus_copy_components", "atus_v_step",
    "atus_get_moments", "atus_inflate", "atus_set_slab", "atus_set_resident",
    "atus_net_create", "atus_net_destroy", "atus_net_set_precision", "atus_net_param_count", "atus_net_get_params", "atus_net_set_params",
    "atus_net_forward", "atus_net_gradients", "atus_net_set_data", "atus_net_train", "atus_net_warm_start",
]

KFAM = ["pass_contig", "pass_strided", "kin", "vfuse", "point"]


class AtusError(RuntimeError):
    pass


class Laser(C.Structure):
    _fields_ = [("amp0", C.c_double), ("laser_k", C.c_double), ("laser_domh", C.c_double),
                ("laser_dk", C.c_double), ("chirp", C.c_double), ("chirp_rate0", C.c_double),
                ("phase0", C.c_double), ("beta0", C.c_double), ("deltaL", C.c_double * 4),
                ("amp1", C.c_double), ("beta", C.c_double * 3)]


AMP_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class DsParams(C.Structure):
    _fields_ = [("laser2", Laser), ("seq_time", C.c_double), ("duration", C.c_double * 2)]


class BatchRun(C.Structure):
    _fields_ = [("gs", C.c_double * 4), ("laser", Laser)]


class Moments(C.Structure):
    _fields_ = [("N", C.c_double), ("x", C.c_double * 3), ("x2", C.c_double * 3), ("k", C.c_double * 3),
                ("k2", C.c_double * 3), ("phase", C.c_double)]


class StepParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("dt", C.c_double), ("laser", Laser), ("amp", AMP_FN),
                ("amp_user", C.c_void_p)]


_lib = None


def load():
    """Load libatus_hip.so; raises AtusError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AtusError(f"{LIB_PATH} is missing: build it with `make -C atus2_amd/csrc` "
                        "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, D, I = C.c_void_p, C.c_double, C.c_int
    pctx = C.POINTER(C.c_void_p)
    sig = {
        "atus_last_error": (C.c_char_p, []),
        "atus_version": (C.c_char_p, []),
        "atus_device_info": (I, [C.c_char_p, I, C.POINTER(I), C.POINTER(D)]),
        "atus_ctx_create": (I, [pctx, I, C.POINTER(C.c_longlong), C.POINTER(D), C.POINTER(D), I]),
        "atus_ctx_destroy": (I, [P]),
        "atus_synchronize": (I, [P]),
        "atus_upload": (I, [P, I, P]),
        "atus_download": (I, [P, I, P]),
        "atus_set_potential": (I, [P, I, P]),
        "atus_set_g": (I, [P, P]),
        "atus_set_alpha": (I, [P, C.POINTER(D)]),
        "atus_ft": (I, [P, I, I]),
        "atus_kinetic": (I, [P, D, D]),
        "atus_nl_step": (I, [P, D]),
        "atus_nl_if_step": (I, [P, D, D, P]),
        "atus_bragg_ad": (I, [P, D, D, D, C.POINTER(Laser)]),
        "atus_set_momentum": (I, [P, I, C.POINTER(D)]),
        "atus_run_block": (I, [P, C.POINTER(StepParams), I, C.POINTER(D)]),
        "atus_particle_number": (I, [P, I, C.POINTER(D)]),
        "atus_expval_x": (I, [P, I, C.POINTER(D)]),
        "atus_expval_k": (I, [P, I, C.POINTER(D)]),
        "atus_rabi_integrals": (I, [P, I, P, I, D, P]),
        "atus_energy": (I, [P, C.POINTER(D)]),
        "atus_sob_setup": (I, [P, P, D]),
        "atus_sob_renormalize": (I, [P]),
        "atus_sob_iter": (I, [P, C.POINTER(D), C.POINTER(I), C.POINTER(D)]),
        "atus_sob_iterate": (I, [P, I, D, C.POINTER(I), C.POINTER(D), P]),
        "atus_sob_mu": (I, [P, P]),
        "atus_set_timing": (I, [P, I]),
        "atus_last_block_ms": (I, [P, C.POINTER(D)]),
        "atus_run_block_profiled": (I, [P, C.POINTER(StepParams), I, C.POINTER(D), C.POINTER(D),
                                        C.POINTER(I), C.POINTER(D)]),
        "atus_copy_bandwidth": (I, [P, I, C.POINTER(D)]),
        "atus_tune": (I, [P, I, I]),
        "atus_tune_threads": (I, [P, I, I]),
        "atus_batch_create": (I, [pctx, I, C.POINTER(C.c_longlong), C.POINTER(D), C.POINTER(D), I, I]),
        "atus_batch_set_runs": (I, [P, C.POINTER(BatchRun), I]),
        "atus_batch_run_block": (I, [P, I, D, I, C.POINTER(D), P]),
        "atus_batch_run_block_profiled": (I, [P, I, D, I, C.POINTER(D), P, C.POINTER(D), C.POINTER(I),
                                              C.POINTER(D)]),
        "atus_batch_particle_numbers": (I, [P, P]),
        "atus_batch_particle_numbers_noalloc": (I, [P, P]),
        "atus_batch_rabi_integrals": (I, [P, I, P, I, D, P]),
        "atus_sob_run": (I, [P, D, C.c_longlong, C.POINTER(C.c_longlong), C.POINTER(D)]),
        "atus_graph_create": (I, [P, I, D, C.POINTER(Laser), I, pctx]),
        "atus_graph_launch": (I, [P, C.POINTER(D), P]),
        "atus_graph_destroy": (I, [P]),
        "atus_scan_graph_create": (I, [P, I, I, D, I, P, I, D, pctx]),
        "atus_scan_graph_launch": (I, [P, C.POINTER(D), P, P, C.POINTER(D)]),
        "atus_scan_graph_destroy": (I, [P]),
        "atus_set_alpha2": (I, [P, C.POINTER(D)]),
        "atus_run_block_ds": (I, [P, C.POINTER(StepParams), C.POINTER(DsParams), I, C.POINTER(D)]),
        "atus_v_step_ds": (I, [P, I, D, D, C.POINTER(Laser), C.POINTER(DsParams)]),
        "atus_snapshot_save": (I, [P, I]),
        "atus_v_step": (I, [P, I, D, D, D, C.POINTER(Laser)]),
        "atus_get_moments": (I, [P, I, C.POINTER(Moments)]),
        "atus_inflate": (I, [P, P, I, I]),
        "atus_set_slab": (I, [P, I]),
        "atus_set_resident": (I, [P, I]),
        "atus_snapshot_restore": (I, [P, I]),
        "atus_copy_components": (I, [P, I, P, I, I]),
        "atus_net_create": (I, [pctx, I, I, I, I, I, I, C.c_ulonglong]),
        "atus_net_destroy": (I, [P]),
        "atus_net_set_precision": (I, [P, I]),
        "atus_net_param_count": (I, [P, C.POINTER(C.c_longlong)]),
        "atus_net_get_params": (I, [P, P]),
        "atus_net_set_params": (I, [P, P]),
        "atus_net_forward": (I, [P, P, I, P]),
        "atus_net_gradients": (I, [P, P, P, I, C.POINTER(C.c_float), P]),
        "atus_net_set_data": (I, [P, P, P, I]),
        "atus_net_train": (I, [P, I, I, C.c_float, C.c_float, C.c_ulonglong, P]),
        "atus_net_warm_start": (I, [P, P, C.POINTER(D), P, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = _lib.atus_last_error().decode(errors="replace")
        raise AtusError(f"atus call failed ({rc}): {msg}")


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def laser(**kw) -> Laser:
    L = Laser()
    dl = kw.pop("deltaL", (0.0,) * 4)
    beta = kw.pop("beta", (0.0,) * 3)
    for k, v in kw.items():
        setattr(L, k, float(v))
    for i, v in enumerate(dl):
        L.deltaL[i] = float(v)
    for i, v in enumerate(beta):
        L.beta[i] = float(v)
    return L


class Context:
    """One grid with `ncomp` device-resident complex128 components."""

    def __init__(self, shape, dx, dk, ncomp=1):
        L = load()
        shape = tuple(int(s) for s in shape)
        self.shape = shape
        self.dim = len(shape)
        self.ncomp = int(ncomp)
        n = (C.c_longlong * 3)(*(list(shape) + [1] * (3 - self.dim)))
        dxa = (C.c_double * 3)(*(list(dx)[:self.dim] + [0.0] * (3 - self.dim)))
        dka = (C.c_double * 3)(*(list(dk)[:self.dim] + [0.0] * (3 - self.dim)))
        h = C.c_void_p()
        _check(L.atus_ctx_create(C.byref(h), self.dim, n, dxa, dka, self.ncomp))
        self._h = h
        self._amp_keep = None
        self._graphs = weakref.WeakSet()

    @classmethod
    def from_grid(cls, g, ncomp=1):
        """From an oracle-style Grid structure (dim, n, dx, dk)."""
        shape = tuple(int(g.n[a]) for a in range(g.dim))
        return cls(shape, [g.dx[a] for a in range(g.dim)], [g.dk[a] for a in range(g.dim)], ncomp)

    def close(self):
        """Destroys the context, after the graphs captured on it (they hold
        its device buffers; atus_ctx_destroy refuses while any is alive)."""
        if getattr(self, "_h", None):
            for g in list(getattr(self, "_graphs", ())):
                g.close()
            _check(_lib.atus_ctx_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state
    def upload(self, comp, psi):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != grid {self.shape}")
        _check(_lib.atus_upload(self._h, comp, _dp(a)))

    def download(self, comp):
        a = np.empty(self.shape, np.complex128)
        _check(_lib.atus_download(self._h, comp, _dp(a)))
        return a

    def v_step(self, kind, dt, t, las: Laser, F=1.0):
        """One V step of any block kind at Get_t() = t, no kinetic step."""
        _check(_lib.atus_v_step(self._h, int(kind), float(dt), float(t), float(F), C.byref(las)))

    def moments(self, comp=0):
        """The moments ana_expval prints: dict of N, x, x2, k, k2 (dim
        entries each) and phase."""
        m = Moments()
        _check(_lib.atus_get_moments(self._h, int(comp), C.byref(m)))
        d = self.dim
        return {"N": m.N, "x": list(m.x[:d]), "x2": list(m.x2[:d]), "k": list(m.k[:d]), "k2": list(m.k2[:d]),
                "phase": m.phase}

    def inflate_from(self, src, fak, fix=False):
        """This context := src inflated by fak (inflate_domain)."""
        _check(_lib.atus_inflate(self._h, src._h, int(fak), int(bool(fix))))

    def snapshot_save(self, slot=0):
        _check(_lib.atus_snapshot_save(self._h, int(slot)))

    def snapshot_restore(self, slot=0):
        _check(_lib.atus_snapshot_restore(self._h, int(slot)))

    def copy_components(self, dst_comp, src, src_comp, n):
        """Device copy of n components of context `src` into this one."""
        _check(_lib.atus_copy_components(self._h, int(dst_comp), src._h, int(src_comp), int(n)))

    def set_potential(self, comp, V):
        if V is None:
            _check(_lib.atus_set_potential(self._h, 0, None))
            return
        a = np.ascontiguousarray(V, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("potential shape mismatch")
        _check(_lib.atus_set_potential(self._h, comp, _dp(a)))

    def set_g(self, gs):
        a = np.ascontiguousarray(np.ravel(gs), dtype=np.float64)
        _check(_lib.atus_set_g(self._h, _dp(a)))

    def set_alpha(self, alpha):
        a = (C.c_double * 3)(*(list(alpha) + [0.0] * (3 - len(alpha))))
        _check(_lib.atus_set_alpha(self._h, a))

    # steps
    def ft(self, comp, isign):
        _check(_lib.atus_ft(self._h, comp, isign))

    def kinetic(self, frac, dt):
        _check(_lib.atus_kinetic(self._h, frac, dt))

    def nl_step(self, dt):
        _check(_lib.atus_nl_step(self._h, dt))

    def nl_if_step(self, dt, beta0, deltaL):
        a = np.ascontiguousarray(deltaL, dtype=np.float64)
        _check(_lib.atus_nl_if_step(self._h, dt, beta0, _dp(a)))

    def bragg_ad(self, dt, t, F, las: Laser):
        _check(_lib.atus_bragg_ad(self._h, dt, t, F, C.byref(las)))

    def set_momentum(self, comp, p):
        a = (C.c_double * 3)(*(list(p) + [0.0] * (3 - len(p))))
        _check(_lib.atus_set_momentum(self._h, comp, a))

    def run_block(self, kind, dt, Nk, t, las: Laser | None = None, amp=None):
        sp = StepParams()
        sp.kind = kind
        sp.dt = dt
        if las is not None:
            sp.laser = las
        if amp is not None:
            cb = AMP_FN(lambda tt, u: float(amp(tt)))
            self._amp_keep = cb
            sp.amp = cb
        tt = C.c_double(t)
        _check(_lib.atus_run_block(self._h, C.byref(sp), int(Nk), C.byref(tt)))
        return tt.value

    def run_block_profiled(self, kind, dt, Nk, t, las: Laser | None = None, amp=None):
        """Runs one block like run_block; returns (t, {family: (ms_total,
        launches, whole-grid equivalents)})."""
        sp = StepParams()
        sp.kind = kind
        sp.dt = dt
        if las is not None:
            sp.laser = las
        if amp is not None:
            cb = AMP_FN(lambda tt, u: float(amp(tt)))
            self._amp_keep = cb
            sp.amp = cb
        tt = C.c_double(t)
        ms = (C.c_double * len(KFAM))()
        cnt = (C.c_int * len(KFAM))()
        grid = (C.c_double * len(KFAM))()
        _check(_lib.atus_run_block_profiled(self._h, C.byref(sp), int(Nk), C.byref(tt), ms, cnt, grid))
        return tt.value, {KFAM[i]: (ms[i], cnt[i], grid[i]) for i in range(len(KFAM))}

    # reductions
    def particle_number(self, comp=0):
        out = C.c_double()
        _check(_lib.atus_particle_number(self._h, comp, C.byref(out)))
        return out.value

    def expval_x(self, comp=0):
        out = (C.c_double * 3)()
        _check(_lib.atus_expval_x(self._h, comp, out))
        return np.array(out[:self.dim])

    def expval_k(self, comp=0):
        out = (C.c_double * 3)()
        _check(_lib.atus_expval_k(self._h, comp, out))
        return np.array(out[:self.dim])

    def rabi_integrals(self, comp, k0list, threshold):
        k0 = np.zeros((len(k0list), 3))
        for i, k in enumerate(k0list):
            k0[i, :len(k)] = k
        k0 = np.ascontiguousarray(k0)
        out = np.zeros(len(k0list))
        _check(_lib.atus_rabi_integrals(self._h, comp, _dp(k0), len(k0list), threshold, _dp(out)))
        return out

    def energy(self):
        out = C.c_double()
        _check(_lib.atus_energy(self._h, C.byref(out)))
        return out.value

    # Sobolev
    def sob_setup(self, alpha_iso, stepsize):
        a = np.ascontiguousarray(alpha_iso, dtype=np.float64)
        _check(_lib.atus_sob_setup(self._h, _dp(a), stepsize))

    def sob_renormalize(self):
        _check(_lib.atus_sob_renormalize(self._h))

    def sob_iter(self):
        """One Sobolev iteration: (res_tot, renormalised, particle numbers)."""
        r = C.c_double()
        rn = C.c_int()
        N = (C.c_double * 4)()
        _check(_lib.atus_sob_iter(self._h, C.byref(r), C.byref(rn), N))
        return r.value, bool(rn.value), list(N[:self.ncomp])

    def set_alpha2(self, alpha2):
        """Kinetic coefficients of species 2 (components [ncomp/2, ncomp))."""
        a = (C.c_double * 3)(*(list(alpha2) + [0.0] * (3 - len(alpha2))))
        _check(_lib.atus_set_alpha2(self._h, a))

    def run_block_ds(self, kind, dt, Nk, t, las: Laser, las2: Laser, seq_time=0.0, duration=(1e300, 1e300)):
        """One two-species block (CRT_Base_IF_2 / bragg_ds); returns the new t."""
        sp = StepParams()
        sp.kind = kind
        sp.dt = dt
        sp.laser = las
        ds = DsParams()
        ds.laser2 = las2
        ds.seq_time = float(seq_time)
        ds.duration[0], ds.duration[1] = float(duration[0]), float(duration[1])
        tt = C.c_double(t)
        _check(_lib.atus_run_block_ds(self._h, C.byref(sp), C.byref(ds), int(Nk), C.byref(tt)))
        return tt.value

    def v_step_ds(self, kind, dt, t, las: Laser, las2: Laser, seq_time=0.0, duration=(1e300, 1e300)):
        """One two-species V step at t1 = t + dt/2, no kinetic step."""
        ds = DsParams()
        ds.laser2 = las2
        ds.seq_time = float(seq_time)
        ds.duration[0], ds.duration[1] = float(duration[0]), float(duration[1])
        _check(_lib.atus_v_step_ds(self._h, int(kind), float(dt), float(t), C.byref(las), C.byref(ds)))

    def sob_iterate(self, count, eps=-1.0):
        """Up to `count` loop bodies on the device (atus_sob_iterate):
        (performed, res_tot, log rows (N0, N1, renormalised, res_tot))."""
        done = C.c_int()
        r = C.c_double()
        log = np.zeros((int(count), 4))
        _check(_lib.atus_sob_iterate(self._h, int(count), float(eps), C.byref(done), C.byref(r), _dp(log)))
        return done.value, r.value, log[:done.value]

    def sob_run(self, eps, maxit=0):
        """CSOB_Base::run: iterations until res_tot <= eps (maxit > 0 caps
        them); returns (iterations, res_tot)."""
        it = C.c_longlong()
        r = C.c_double()
        _check(_lib.atus_sob_run(self._h, float(eps), int(maxit), C.byref(it), C.byref(r)))
        return it.value, r.value

    def sob_mu(self):
        out = np.zeros(self.ncomp)
        _check(_lib.atus_sob_mu(self._h, _dp(out)))
        return out

    # timing
    def set_timing(self, on=True):
        _check(_lib.atus_set_timing(self._h, 1 if on else 0))

    def last_block_ms(self):
        out = C.c_double()
        _check(_lib.atus_last_block_ms(self._h, C.byref(out)))
        return out.value

    def synchronize(self):
        _check(_lib.atus_synchronize(self._h))

    def tune(self, family, values_per_thread):
        """family: index or name in KFAM; values_per_thread: 8 or 16."""
        f = KFAM.index(family) if isinstance(family, str) else int(family)
        _check(_lib.atus_tune(self._h, f, int(values_per_thread)))

    def tune_threads(self, family, threads):
        """Threads per workgroup at most for a kernel family (0: automatic)."""
        f = KFAM.index(family) if isinstance(family, str) else int(family)
        _check(_lib.atus_tune_threads(self._h, f, int(threads)))

    def set_slab(self, planes):
        """3D blocks: y/z passes per slab of `planes` x-planes (-1 automatic,
        0 whole grid)."""
        _check(_lib.atus_set_slab(self._h, int(planes)))

    def set_resident(self, on):
        """1D blocks in one launch with the pencils resident (True, the
        default) or a launch per step (False)."""
        _check(_lib.atus_set_resident(self._h, 1 if on else 0))

    def copy_bandwidth(self, reps=20):
        """GB/s of a plain streaming copy of all components (HIP events)."""
        out = C.c_double()
        _check(_lib.atus_copy_bandwidth(self._h, int(reps), C.byref(out)))
        return out.value


def device_info():
    L = load()
    name = C.create_string_buffer(256)
    cu = C.c_int()
    bw = C.c_double()
    _check(L.atus_device_info(name, 256, C.byref(cu), C.byref(bw)))
    return name.value.decode(), cu.value, bw.value


class BatchContext(Context):
    """`nbatch` independent two-component runs of one grid (atus_batch_*).
    Component c of run r is component r*2 + c of the underlying context."""

    def __init__(self, shape, dx, dk, nbatch):
        L = load()
        shape = tuple(int(v) for v in shape)
        self.shape = shape
        self.dim = len(shape)
        self.ncomp = 2 * int(nbatch)
        self.nbatch = int(nbatch)
        n = (C.c_longlong * 3)(*(list(shape) + [1] * (3 - self.dim)))
        dxa = (C.c_double * 3)(*(list(dx)[:self.dim] + [0.0] * (3 - self.dim)))
        dka = (C.c_double * 3)(*(list(dk)[:self.dim] + [0.0] * (3 - self.dim)))
        h = C.c_void_p()
        _check(L.atus_batch_create(C.byref(h), self.dim, n, dxa, dka, 2, self.nbatch))
        self._h = h
        self._amp_keep = None
        self._graphs = weakref.WeakSet()

    def set_runs(self, gs, lasers):
        """gs: nbatch x 4 couplings; lasers: nbatch Laser structures."""
        arr = (BatchRun * self.nbatch)()
        g = np.asarray(gs, dtype=np.float64).reshape(self.nbatch, 4)
        for r in range(self.nbatch):
            for i in range(4):
                arr[r].gs[i] = float(g[r, i])
            arr[r].laser = lasers[r]
        _check(_lib.atus_batch_set_runs(self._h, arr, self.nbatch))

    @staticmethod
    def step_times(t, dt, Nk):
        """The t1 of V steps 1..Nk of a block starting at t, accumulated as
        the block runner does (t += dt/2, then t1 = t + dt/2, t += dt), so
        that amplitudes evaluated here equal the single-run path's."""
        out = []
        tt = t + 0.5 * dt
        for m in range(1, Nk + 1):
            out.append(tt + 0.5 * dt)
            if m < Nk:
                tt += dt
        return out

    def run_batch_block(self, kind, dt, Nk, t, F=None):
        """One block of every run; F: (Nk, nbatch) amplitudes of V steps
        m = 1..Nk (t1 = t + m dt) or None (all 1).  Returns the new t."""
        tt = C.c_double(t)
        if F is None:
            _check(_lib.atus_batch_run_block(self._h, kind, dt, Nk, C.byref(tt), None))
        else:
            a = np.ascontiguousarray(F, dtype=np.float64)
            if a.shape != (Nk, self.nbatch):
                raise ValueError(f"amplitude table shape {a.shape} != {(Nk, self.nbatch)}")
            _check(_lib.atus_batch_run_block(self._h, kind, dt, Nk, C.byref(tt), _dp(a)))
        return tt.value

    def run_batch_block_profiled(self, kind, dt, Nk, t, F=None):
        """One block with per-family device times: (t, {family: (ms,
        launches, grid share)})."""
        tt = C.c_double(t)
        ms = (C.c_double * len(KFAM))()
        cnt = (C.c_int * len(KFAM))()
        grid = (C.c_double * len(KFAM))()
        a = None if F is None else np.ascontiguousarray(F, dtype=np.float64)
        if a is not None and a.shape != (Nk, self.nbatch):
            raise ValueError(f"amplitude table shape {a.shape} != {(Nk, self.nbatch)}")
        _check(_lib.atus_batch_run_block_profiled(self._h, int(kind), float(dt), int(Nk), C.byref(tt),
                                                  None if a is None else _dp(a), ms, cnt, grid))
        return tt.value, {KFAM[i]: (ms[i], cnt[i], grid[i]) for i in range(len(KFAM))}

    def particle_numbers(self):
        out = np.empty((self.nbatch, 2))
        _check(_lib.atus_batch_particle_numbers(self._h, _dp(out)))
        return out

    def batch_rabi_integrals(self, comp, k0list, threshold):
        k0 = np.zeros((len(k0list), 3))
        for i, k in enumerate(k0list):
            k0[i, :len(k)] = k
        out = np.empty((self.nbatch, len(k0list)))
        _check(_lib.atus_batch_rabi_integrals(self._h, comp, _dp(np.ascontiguousarray(k0)), len(k0list),
                                              float(threshold), _dp(out)))
        return out


class Net:
    """The warm-start perceptron (atus_net_*): dims = (n_in, h1, h2, h3, n_out)."""

    def __init__(self, dims, max_batch=256, seed=1):
        L = load()
        h = C.c_void_p()
        _check(L.atus_net_create(C.byref(h), *[int(d) for d in dims], int(max_batch), int(seed)))
        self._h = h
        self.dims = tuple(int(d) for d in dims)
        n = C.c_longlong()
        _check(L.atus_net_param_count(h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.atus_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_precision(self, bf16_terms):
        """1: bf16 operands; 3: hi/lo bf16 pairs (three products)."""
        _check(_lib.atus_net_set_precision(self._h, int(bf16_terms)))

    def get_params(self):
        out = np.empty(self.nparams, np.float32)
        _check(_lib.atus_net_get_params(self._h, _dp(out)))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.nparams:
            raise ValueError(f"{a.size} parameters, expected {self.nparams}")
        _check(_lib.atus_net_set_params(self._h, _dp(a)))

    def forward(self, x):
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        y = np.empty((a.shape[0], self.dims[4]), np.float32)
        _check(_lib.atus_net_forward(self._h, _dp(a), a.shape[0], _dp(y)))
        return y

    def gradients(self, x, t):
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        b = np.ascontiguousarray(t, dtype=np.float32).reshape(a.shape[0], self.dims[4])
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        _check(_lib.atus_net_gradients(self._h, _dp(a), _dp(b), a.shape[0], C.byref(loss), _dp(g)))
        return loss.value, g

    def set_data(self, x, t):
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        b = np.ascontiguousarray(t, dtype=np.float32).reshape(a.shape[0], self.dims[4])
        _check(_lib.atus_net_set_data(self._h, _dp(a), _dp(b), a.shape[0]))

    def train(self, steps, batch, lr, lr_final=0.0, seed=1):
        losses = np.empty(steps, np.float32)
        _check(_lib.atus_net_train(self._h, int(steps), int(batch), float(lr), float(lr_final), int(seed),
                                   _dp(losses)))
        return losses

    def warm_start(self, x, half_width, ctx, comp=0):
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(self.dims[0])
        hw = (C.c_double * 3)(*[float(v) for v in half_width])
        _check(_lib.atus_net_warm_start(self._h, _dp(a), hw, ctx._h, int(comp)))


class Graph:
    """One block of `ctx` (Context or BatchContext) captured as a HIP graph
    (atus_graph_*); launch() replays it with the block's amplitudes."""

    def __init__(self, ctx, kind, dt, Nk, las: Laser | None = None):
        L = load()
        h = C.c_void_p()
        _check(L.atus_graph_create(ctx._h, int(kind), float(dt), C.byref(las) if las is not None else None, int(Nk),
                                   C.byref(h)))
        self._h = h
        # the context outlives its graphs, and closes them before itself
        self._ctx = ctx
        ctx._graphs.add(self)
        self.Nk = int(Nk)
        self.dt = float(dt)
        self.nbatch = getattr(ctx, "nbatch", 1)

    def launch(self, t, F=None):
        """F: amplitudes of V steps 1..Nk ((Nk,) or (Nk, nbatch)) or None."""
        if not getattr(self, "_h", None):
            raise AtusError("graph is closed")
        tt = C.c_double(t)
        if F is None:
            _check(_lib.atus_graph_launch(self._h, C.byref(tt), None))
        else:
            a = np.ascontiguousarray(F, dtype=np.float64)
            if a.size != self.Nk * self.nbatch:
                raise ValueError(f"amplitudes: expected {self.Nk * self.nbatch} values (Nk x nbatch), got shape {a.shape}")
            _check(_lib.atus_graph_launch(self._h, C.byref(tt), _dp(a)))
        return tt.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.atus_graph_destroy(self._h)
            self._h = None
            ctx = getattr(self, "_ctx", None)
            if ctx is not None:
                ctx._graphs.discard(self)
                self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class ScanGraph:
    """The blocks and Rabi integrals of several BatchContexts (runs of
    different grids) captured as one HIP graph (atus_scan_graph_*)."""

    def __init__(self, ctxs, kind, dt, Nk, k0list=(), threshold=0.0):
        L = load()
        self.ctxs = list(ctxs)
        arr = (C.c_void_p * len(self.ctxs))(*[c._h.value for c in self.ctxs])
        k0 = np.zeros((max(1, len(k0list)), 3))
        for i, k in enumerate(k0list):
            k0[i, :len(k)] = k
        k0 = np.ascontiguousarray(k0)
        h = C.c_void_p()
        _check(L.atus_scan_graph_create(C.cast(arr, C.c_void_p), len(self.ctxs), int(kind), float(dt), int(Nk),
                                        _dp(k0), len(k0list), float(threshold), C.byref(h)))
        self._h = h
        self.Nk = int(Nk)
        self.nk0 = len(k0list)
        for c in self.ctxs:
            c._graphs.add(self)

    def launch(self, t, F=None):
        """F: a list of (Nk, nbatch) amplitude tables, one per context, or
        None.  Returns (t, [Rabi populations (nbatch, nk0) per context],
        device ms of the graph)."""
        if not getattr(self, "_h", None):
            raise AtusError("scan graph is closed")
        tt = C.c_double(t)
        ms = C.c_double()
        fa = None
        if F is not None:
            parts = []
            for c, f in zip(self.ctxs, F):
                a = np.asarray(f, dtype=np.float64)
                if a.shape != (self.Nk, c.nbatch):
                    raise ValueError(f"amplitude table shape {a.shape} != {(self.Nk, c.nbatch)}")
                parts.append(a.ravel())
            if len(parts) != len(self.ctxs):
                raise ValueError("one amplitude table per context")
            fa = np.ascontiguousarray(np.concatenate(parts))
        nr = sum(c.nbatch for c in self.ctxs) * self.nk0
        rabi = np.zeros(max(nr, 1))
        _check(_lib.atus_scan_graph_launch(self._h, C.byref(tt), None if fa is None else _dp(fa), _dp(rabi),
                                           C.byref(ms)))
        out, o = [], 0
        for c in self.ctxs:
            n = c.nbatch * self.nk0
            out.append(rabi[o:o + n].reshape(c.nbatch, self.nk0))
            o += n
        return tt.value, out, ms.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.atus_scan_graph_destroy(self._h)
            self._h = None
            for c in self.ctxs:
                c._graphs.discard(self)
            self.ctxs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
