"""Pins the CPU oracle (oracle/atus_oracle.c) before it is trusted.

The reference ships no golden vectors with expected values (its cft_*_test
programs only write files for visual inspection), so every check here is one
of: a library routine (numpy/pocketfft, scipy.linalg.expm), a brute-force
DFT, a closed form of the mathematics, or the known-answer setups of the
reference's own test programs (cft_1d_test.cpp:47-53,81-92 and
cft_3d_test.cpp:54-60,106).  Each is chosen so that a plausible slip -- a
wrong FFT sign, a k-order shift, a missing scale factor, a transposed 2x2
element, a step taken at the wrong time -- fails at least one assertion.
"""
import numpy as np
import pytest
import scipy.linalg


# ---------------------------------------------------------------- FFT (a3)

@pytest.mark.parametrize("shape", [(1,), (2,), (64,), (4096,), (16, 32),
                                   (8, 16, 32), (12,), (6, 10), (5, 4, 3)])
def test_fft_matches_pocketfft(orc, shape):
    rng = np.random.default_rng(1)
    g = orc.make_grid(shape, L=7.0)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fwd = orc.fft(g, x, -1)
    bwd = orc.fft(g, x, +1)
    ref_f = np.fft.fftn(x)
    ref_b = np.fft.ifftn(x) * x.size
    assert np.abs(fwd - ref_f).max() <= 1e-14 * np.abs(ref_f).max()
    assert np.abs(bwd - ref_b).max() <= 1e-14 * np.abs(ref_b).max()


def test_fft_brute_force_dft(orc):
    for n in (2, 4, 8, 16, 3, 7):
        g = orc.make_grid((n,), L=1.0)
        x = np.random.default_rng(n).standard_normal(n) * (1 + 0.5j)
        j = np.arange(n)
        W = np.exp(-2j * np.pi * (np.outer(j, j) % n) / n)
        tol = 1e-14 * np.abs(x).sum()
        assert np.abs(orc.fft(g, x, -1) - W @ x).max() < tol
        assert np.abs(orc.fft(g, x, +1) - W.conj() @ x).max() < tol


@pytest.mark.parametrize("shape", [(256,), (32, 64), (16, 8, 32)])
def test_ft_round_trip_and_parseval(orc, shape):
    rng = np.random.default_rng(2)
    g = orc.make_grid(shape, L=[10.0, 12.0, 14.0][:len(shape)])
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    k = orc.ft(g, x, -1)
    back = orc.ft(g, k, +1)
    assert np.abs(back - x).max() < 1e-14 * np.abs(x).max() * 10
    n_x = orc.ar(g) * np.sum(np.abs(x) ** 2)
    n_k = orc.ar_k(g) * np.sum(np.abs(k) ** 2)
    assert abs(n_x - n_k) < 1e-13 * n_x


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_ft_unitary_gaussian_pair(orc, dim):
    # e^{-x^2/2} <-> e^{-k^2/2} under the 1/sqrt(2 pi) per-axis convention,
    # in native order: checks scale factor, sign and the k index map together
    # box and spacing both wide enough that e^{-x^2/2} and e^{-k^2/2} fall
    # below 1e-16 at the edges (no aliasing)
    n = {1: 256, 2: 64, 3: 64}[dim]
    L = {1: 40.0, 2: 20.0, 3: 20.0}[dim]
    g = orc.make_grid((n,) * dim, L=L)
    xs = orc.coords_x(g)
    ks = orc.coords_k(g)
    psi = np.exp(-0.5 * sum(x ** 2 for x in xs)) * np.ones((n,) * dim)
    # shift by x0 along axis 0: transform picks up e^{-i k x0}
    x0 = 2 * g.dx[0]
    psi_s = np.exp(-0.5 * ((xs[0] - x0) ** 2 + sum(x ** 2 for x in xs[1:]))) * np.ones((n,) * dim)
    # x is centred on index n/2 (cft_3d.cpp:134-136) and scale() does not
    # undo it (only fix() would, cft_3d.cpp:580-615), so the transform
    # carries (-1)^(i+j+k) against the continuous one
    idx = np.indices((n,) * dim).sum(axis=0)
    expect = (-1.0) ** idx * np.exp(-0.5 * sum(k ** 2 for k in ks))
    assert np.abs(orc.ft(g, psi, -1) - expect).max() < 1e-12
    assert np.abs(orc.ft(g, psi_s, -1) - expect * np.exp(-1j * ks[0] * x0)).max() < 1e-12
    assert np.abs(orc.ft(g, expect, +1) - psi).max() < 1e-12


def test_reference_known_answer_cft_1d(orc):
    # cft_1d_test.cpp:47-53,81-92: N=512 on [-10,10], psi = e^{-x^2},
    # ft(-1), multiply by -k^2, ft(+1) gives (4x^2-2) e^{-x^2}
    g = orc.make_grid((512,), L=20.0)
    (x,) = orc.coords_x(g)
    (k,) = orc.coords_k(g)
    psi = np.exp(-x ** 2) + 0j
    out = orc.ft(g, orc.ft(g, psi, -1) * (-k ** 2), +1)
    assert np.abs(out - (4 * x ** 2 - 2) * np.exp(-x ** 2)).max() < 1e-12


def test_reference_known_answer_cft_3d(orc):
    # cft_3d_test.cpp:54-60,106: d/dz e^{-r^2} = -2 z e^{-r^2} (grid reduced
    # from 256^3 to 64^3 on the same [-10,10]^3 box)
    g = orc.make_grid((64, 64, 64), L=20.0)
    x, y, z = orc.coords_x(g)
    kx, ky, kz = orc.coords_k(g)
    psi = np.exp(-(x ** 2 + y ** 2 + z ** 2)) + 0j
    out = orc.ft(g, orc.ft(g, psi, -1) * (1j * kz), +1)
    assert np.abs(out - (-2 * z * psi)).max() < 1e-10


def test_get_x_get_k_maps(orc):
    import ctypes as C
    g = orc.make_grid((4, 6, 8), L=[4.0, 6.0, 8.0])
    L = orc.lib()
    xb = (C.c_double * 3)()
    kb = (C.c_double * 3)()
    for l in (0, 1, 7, 8, 47, 48, 191):
        i, j, k = np.unravel_index(l, (4, 6, 8))
        L.orc_get_x(C.byref(g), l, xb)
        L.orc_get_k(C.byref(g), l, kb)
        assert list(xb) == [(i - 2) * 1.0, (j - 3) * 1.0, (k - 4) * 1.0]
        exp_k = [2 * np.pi / 4 * (((i + 2) % 4) - 2), 2 * np.pi / 6 * (((j + 3) % 6) - 3),
                 2 * np.pi / 8 * (((k + 4) % 8) - 4)]
        assert np.allclose(list(kb), exp_k, rtol=0, atol=1e-15)


# ------------------------------------------------------- kinetic step (a4, a5)

@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kinetic_plane_wave_exact_phase(orc, dim):
    n = {1: 64, 2: 32, 3: 16}[dim]
    g = orc.make_grid((n,) * dim, L=[8.0, 9.0, 10.0][:dim])
    alpha = [0.5, 0.7, 0.3][:dim]
    dt = 0.013
    xs = orc.coords_x(g)
    m = [3, -5, 2][:dim]
    kvec = [m[a] * g.dk[a] for a in range(dim)]
    psi = np.exp(1j * sum(kvec[a] * xs[a] for a in range(dim))) * np.ones((n,) * dim)
    for s in (0.5, 1.0):
        tab = orc.kinetic_table(g, alpha, dt, s)
        (out,) = orc.ft_step(g, [psi], tab)
        ph = -s * dt * sum(alpha[a] * kvec[a] ** 2 for a in range(dim))
        assert np.abs(out - psi * np.exp(1j * ph)).max() < 1e-12


def test_free_gaussian_spreading_closed_form(orc):
    # i dpsi/dt = -alpha psi'' (alpha = 1/2): a Gaussian of width s0 evolves to
    # (s0^2/(s0^2 + i t))^{1/2} exp(-x^2 / (2 (s0^2 + i t)))
    n, L, s0, dt, steps = 1024, 80.0, 1.5, 0.01, 50
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    psi = np.exp(-x ** 2 / (2 * s0 ** 2)) + 0j
    tab = orc.kinetic_table(g, [0.5], dt, 1.0)
    for _ in range(steps):
        (psi,) = orc.ft_step(g, [psi], tab)
    t = dt * steps
    c = s0 ** 2 + 1j * t
    expect = np.sqrt(s0 ** 2 / c) * np.exp(-x ** 2 / (2 * c))
    assert np.abs(psi - expect).max() < 1e-12


# ------------------------------------------------------- nonlinear step (a6, a7)

def test_nl_step_phase_and_simultaneous_update(orc):
    rng = np.random.default_rng(3)
    shape = (8, 4, 6)
    g = orc.make_grid(shape, L=3.0)
    p1 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    V1 = rng.standard_normal(shape)
    V2 = rng.standard_normal(shape)
    gs = np.array([[1.3, 0.7], [0.6, 2.1]])
    dt = 0.05
    o1, o2 = orc.nl_step(g, [p1, p2], gs, dt, [V1, V2])
    n1, n2 = np.abs(p1) ** 2, np.abs(p2) ** 2
    e1 = p1 * np.exp(-1j * dt * (V1 + gs[0, 0] * n1 + gs[0, 1] * n2))
    e2 = p2 * np.exp(-1j * dt * (V2 + gs[1, 0] * n1 + gs[1, 1] * n2))
    assert np.abs(o1 - e1).max() < 1e-13 and np.abs(o2 - e2).max() < 1e-13
    assert np.abs(np.abs(o1) - np.abs(p1)).max() < 1e-14
    # without a potential (Init_Potential never called) the phase is g-only
    (o,) = orc.nl_step(g, [p1], [2.0], dt, None)
    assert np.abs(o - p1 * np.exp(-1j * dt * 2.0 * n1)).max() < 1e-13


def test_nl_if_step_gravity_detuning(orc):
    rng = np.random.default_rng(4)
    shape = (16, 4)
    g = orc.make_grid(shape, L=[8.0, 2.0])
    p1 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    gs = np.array([1.0, 0.5, 0.4, 1.5])
    dt, beta0, dL = 0.02, 0.3, [0.1, -2.0]
    o1, o2 = orc.nl_if_step(g, [p1, p2], gs, dt, beta0, dL)
    x = orc.coords_x(g)[0]
    n1, n2 = np.abs(p1) ** 2, np.abs(p2) ** 2
    e1 = p1 * np.exp(-1j * dt * (gs[0] * n1 + gs[1] * n2 + beta0 * x - dL[0]))
    e2 = p2 * np.exp(-1j * dt * (gs[2] * n1 + gs[3] * n2 + beta0 * x - dL[1]))
    assert np.abs(o1 - e1).max() < 1e-13 and np.abs(o2 - e2).max() < 1e-13


# ----------------------------------------------------------- Bragg 2x2 (a8)

def _bragg_expected(x, p1, p2, P):
    t1 = P.t + 0.5 * P.dt
    n1, n2 = np.abs(p1) ** 2, np.abs(p2) ** 2
    V11 = P.gs[0] * n1 + P.gs[1] * n2 + P.beta0 * x
    V22 = P.gs[2] * n1 + P.gs[3] * n2 - P.deltaL1 + P.beta0 * x
    Om = P.F * P.amp0 * np.cos(P.laser_k * x - (P.laser_domh + P.chirp * t1 +
                                                 P.chirp_rate0 * t1) * t1 + P.phase0 / 2)
    eta = np.exp(-0.5j * P.laser_dk * x)
    o1 = np.empty_like(p1)
    o2 = np.empty_like(p2)
    for idx in np.ndindex(p1.shape):
        H = np.array([[V11[idx], Om[idx] * eta[idx]],
                      [Om[idx] * np.conj(eta[idx]), V22[idx]]])
        U = scipy.linalg.expm(-1j * P.dt * H)
        v = U @ np.array([p1[idx], p2[idx]])
        o1[idx], o2[idx] = v
    return o1, o2


def test_bragg_ad_matches_expm_and_is_unitary(orc):
    rng = np.random.default_rng(5)
    shape = (24, 3)
    g = orc.make_grid(shape, L=[12.0, 3.0])
    p1 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    P = orc.bragg_params(dt=0.07, t=0.31, F=0.8, amp0=2.5, laser_k=1.7,
                         laser_domh=0.9, laser_dk=0.6, chirp=0.2, chirp_rate0=0.1,
                         phase0=0.4, beta0=0.05, deltaL1=-1.2, gs=(1.0, 0.3, 0.25, 0.8))
    o1, o2 = orc.bragg_ad(g, p1, p2, P)
    x = orc.coords_x(g)[0] * np.ones(shape)
    e1, e2 = _bragg_expected(x, p1, p2, P)
    assert np.abs(o1 - e1).max() < 1e-12 and np.abs(o2 - e2).max() < 1e-12
    tot0 = np.abs(p1) ** 2 + np.abs(p2) ** 2
    tot1 = np.abs(o1) ** 2 + np.abs(o2) ** 2
    assert np.abs(tot1 - tot0).max() < 1e-13 * tot0.max()


def test_bragg_ad_omega_zero_branch(orc):
    rng = np.random.default_rng(6)
    shape = (16,)
    g = orc.make_grid(shape, L=4.0)
    p1 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    P = orc.bragg_params(dt=0.1, amp0=0.0, gs=(1.0, 0.5, 0.5, 2.0), deltaL1=0.3, beta0=0.2)
    o1, o2 = orc.bragg_ad(g, p1, p2, P)
    x = orc.coords_x(g)[0]
    n1, n2 = np.abs(p1) ** 2, np.abs(p2) ** 2
    assert np.abs(o1 - p1 * np.exp(-1j * 0.1 * (n1 + 0.5 * n2 + 0.2 * x))).max() < 1e-14
    assert np.abs(o2 - p2 * np.exp(-1j * 0.1 * (0.5 * n1 + 2.0 * n2 - 0.3 + 0.2 * x))).max() < 1e-14


def test_bragg_uniform_rabi_closed_form(orc):
    # uniform Omega (laser_k = 0), g = 0, beta = 0, deltaL = 0: resonant Rabi
    # oscillation, |psi2|^2 = sin^2(Omega t) after n steps of dt (exact, the
    # per-point propagator is exact for constant H)
    g = orc.make_grid((8,), L=1.0)
    p1 = np.ones(8, np.complex128)
    p2 = np.zeros(8, np.complex128)
    Om, dt, n = 1.3, 0.01, 40
    for i in range(n):
        P = orc.bragg_params(dt=dt, t=i * dt, amp0=Om)
        p1, p2 = orc.bragg_ad(g, p1, p2, P)
    T = n * dt
    assert np.abs(np.abs(p2) ** 2 - np.sin(Om * T) ** 2).max() < 1e-13
    assert np.abs(np.abs(p1) ** 2 - np.cos(Om * T) ** 2).max() < 1e-13


# ------------------------------------------------------------ reductions (a10)

def test_norm_expval_of_boosted_gaussian(orc):
    g = orc.make_grid((64, 64), L=24.0)
    x, y = orc.coords_x(g)
    x0, y0, px, py, s = 1.25, -0.75, 1.5, -0.5, 1.3
    psi = np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * s ** 2)) * np.exp(1j * (px * x + py * y))
    N = orc.particle_number(g, psi)
    assert abs(N - np.pi * s ** 2) < 1e-12
    ex = orc.expval_x(g, psi) / N
    assert np.allclose(ex, [x0, y0], atol=1e-12)
    ek, back = orc.expval_k(g, psi)
    assert np.allclose(ek / N, [px, py], atol=1e-10)
    assert np.abs(back - psi).max() < 1e-14


def test_rabi_bins_plane_waves(orc):
    # superposition of three plane waves at k = 0, +-K along x; each bin
    # holds exactly its wave's weight (orthogonal modes, unitary transform)
    n, L = 128, 32.0
    g = orc.make_grid((n, 8), L=[L, 4.0])
    x, _ = orc.coords_x(g)
    K = 8 * g.dk[0]
    c = [0.6, 0.3 + 0.2j, -0.5j]
    psi = (c[0] + c[1] * np.exp(1j * K * x) + c[2] * np.exp(-1j * K * x)) * np.ones((n, 8))
    out, _ = orc.rabi_integrals(g, psi, [(0, 0), (K, 0), (-K, 0)], 0.5 * K)
    area = n * g.dx[0] * 8 * g.dx[1]
    assert np.allclose(out, [abs(v) ** 2 * area for v in c], rtol=1e-12)


def test_rabi_bins_overlap_slot_quirk(orc):
    # CRT_Base_IF.h:244-255: a hit does not advance the slot counter, so with
    # overlapping spheres the second hit lands in the first slot
    g = orc.make_grid((32,), L=8.0)
    (x,) = orc.coords_x(g)
    psi = np.ones(32, np.complex128)          # all weight at k = 0
    out, _ = orc.rabi_integrals(g, psi, [(0.0,), (0.1,), (5.0,)], 1.0)
    N = orc.particle_number(g, psi)
    assert np.allclose(out, [2 * N, 0.0, 0.0], rtol=1e-12)


# -------------------------------------------------------- Strang scheme (a9)

def test_ho_ground_state_stationary(orc):
    # alpha = 1/2, V = x^2/2, g = 0: pi^{-1/4} e^{-x^2/2} is an eigenstate;
    # Strang splitting keeps it stationary to O(dt^2) up to e^{-i t/2}
    n, L = 512, 32.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    psi0 = np.pi ** -0.25 * np.exp(-x ** 2 / 2) + 0j
    S = orc.RTSolver(g, [psi0], [0.0], [0.5], dt=1e-3, pots=[0.5 * x ** 2])
    for _ in range(10):
        S.block(orc.FREEPROP, 100)
    assert abs(S.t - 1.0) < 1e-12
    psi = S.psis[0]
    ov = np.sum(np.conj(psi0) * psi) * g.dx[0]
    assert abs(abs(ov) - 1) < 1e-9
    assert abs(np.angle(ov) - (-0.5)) < 1e-6          # E = 1/2, t = 1
    assert abs(orc.particle_number(g, psi) - 1) < 1e-12


def test_coherent_state_oscillates(orc):
    n, L, x0 = 512, 32.0, 2.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    psi0 = np.pi ** -0.25 * np.exp(-(x - x0) ** 2 / 2) + 0j
    S = orc.RTSolver(g, [psi0], [0.0], [0.5], dt=2e-3, pots=[0.5 * x ** 2])
    for _ in range(5):
        S.block(orc.FREEPROP, 100)
    t = S.t
    assert abs(orc.expval_x(g, S.psis[0])[0] - x0 * np.cos(t)) < 1e-5


def test_block_time_bookkeeping_matches_reference_rule(orc):
    # V-step m of a block sees t1 = t_block + m dt (CRT_Base.h:776-786 with
    # Get_t()+0.5dt in Amplitude_at_time, CRT_Base_IF.h:183)
    seen = []
    g = orc.make_grid((8,), L=1.0)
    S = orc.RTSolver(g, [np.ones(8), np.zeros(8)], [0, 0, 0, 0], [0.5], dt=0.25,
                     t=1.0, amp=lambda t: seen.append(t) or 0.0)
    S.block(orc.BRAGG_AD, 4)
    assert np.allclose(seen, [1.25, 1.5, 1.75, 2.0], rtol=0, atol=1e-15)
    assert S.t == 2.0


def test_energy_ho_closed_form_and_conservation(orc):
    g = orc.make_grid((32, 32), L=16.0)
    x, y = orc.coords_x(g)
    gs0 = np.pi ** -0.5 * np.exp(-(x ** 2 + y ** 2) / 2) + 0j
    V = 0.5 * (x ** 2 + y ** 2)
    E = orc.energy(g, [gs0], [0.0], [0.5, 0.5], [V])
    assert abs(E - 1.0) < 1e-12                     # 1/2 per dimension
    # nonlinear, displaced state: E conserved to O(dt^2) over 200 steps
    psi = np.pi ** -0.5 * np.exp(-((x - 1) ** 2 + y ** 2) / 2) + 0j
    E0 = orc.energy(g, [psi], [5.0], [0.5, 0.5], [V])
    S = orc.RTSolver(g, [psi], [5.0], [0.5, 0.5], dt=1e-3, pots=[V])
    S.block(orc.FREEPROP, 200)
    E1 = orc.energy(g, S.psis, [5.0], [0.5, 0.5], [V])
    assert abs(E1 - E0) < 1e-5 * abs(E0)


# ------------------------------------------------------- Sobolev (a12-a14)

def test_sobolev_ho_ground_state_mu(orc):
    # H = -alpha Lap + V (no 1/2), V = w^2 x^2 / 2 per axis, g = 0:
    # mu = sum_a w_a sqrt(alpha/2); ground state exp(-w x^2 / (2 sqrt(2 alpha))).
    # sobmin only renormalises when |N-1| > 1e-5 (CSOB_Base.h:470-481), so the
    # converged state is compared after dividing out its norm.
    g = orc.make_grid((32, 32), L=12.0)
    x, y = orc.coords_x(g)
    alpha, wx, wy = 0.5, 1.0, 1.5
    V = 0.5 * (wx ** 2 * x ** 2 + wy ** 2 * y ** 2)
    guess = np.exp(-(x ** 2 + y ** 2) / 4.0) + 0j
    S = orc.SobSolver(g, [guess], [V], [0.0], [alpha], stepsize=0.02)
    S.renormalize()
    res = []
    for _ in range(4000):
        res.append(S.iterate())
        if res[-1] < 1e-22:
            break
    assert res[-1] < 1e-22
    assert all(b <= a * 1.0001 for a, b in zip(res[10:], res[11:]))  # monotone tail
    N = orc.particle_number(g, S.psi[0])
    mu = S.mu()[0] / N
    assert abs(mu - (wx + wy) * np.sqrt(alpha / 2)) < 1e-9
    gs_exact = np.exp(-(wx * x ** 2 + wy * y ** 2) / (2 * np.sqrt(2 * alpha))) + 0j
    gs_exact = gs_exact / np.sqrt(orc.particle_number(g, gs_exact))
    psi = S.psi[0] / np.sqrt(N)
    psi = psi * np.exp(-1j * np.angle(np.sum(psi)))
    assert np.abs(psi - gs_exact).max() < 1e-8


def test_sobolev_projected_gradient_orthogonal(orc):
    rng = np.random.default_rng(7)
    g = orc.make_grid((16, 16), L=10.0)
    x, y = orc.coords_x(g)
    V = 0.5 * (x ** 2 + y ** 2)
    guess = np.exp(-(x ** 2 + y ** 2) / 3.0) * (1 + 0.1 * rng.standard_normal((16, 16))) + 0j
    S = orc.SobSolver(g, [guess], [V], [3.0], [0.5], stepsize=0.3)
    S.renormalize()
    psi_before = S.psi[0].copy()
    S.iterate()
    # projection makes Re<psi, grad> vanish against the pre-update psi
    # (CSOB_Base.h:340-367): <psi, g - (<psi,g>/<psi,psi_sob>) psi_sob> = 0
    v = np.sum(psi_before.real * S.grad[0].real + psi_before.imag * S.grad[0].imag)
    assert abs(v) < 1e-12 * np.sum(np.abs(S.grad[0]) ** 2) ** 0.5 * np.sum(np.abs(psi_before) ** 2) ** 0.5 + 1e-14


# ------------------------------------------------ two species (CRT_Base_2, bragg_ds)

def test_ft_step2_applies_the_species_tables(orc):
    # plane waves e^{i k x}: species 1 picks up exp(-i dt a1 k^2), species 2 exp(-i dt a2 k^2)
    n, L = 64, 20.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    k = 5 * 2 * np.pi / L
    psis = [np.exp(1j * k * x) * (1 + 0.1 * c) for c in range(4)]
    dt, a1, a2 = 0.013, 0.5, 0.21
    out = orc.ft_step2(g, psis, orc.kinetic_table(g, [a1], dt, 1.0), orc.kinetic_table(g, [a2], dt, 1.0))
    for c in range(4):
        a = a1 if c < 2 else a2
        assert np.abs(out[c] - psis[c] * np.exp(-1j * dt * a * k * k)).max() < 1e-13


def test_nl_if2_step_per_species_gravity(orc):
    rng = np.random.default_rng(7)
    g = orc.make_grid((16, 4), L=[8.0, 2.0])
    psis = [rng.standard_normal((16, 4)) + 1j * rng.standard_normal((16, 4)) for _ in range(4)]
    gs = rng.uniform(0.1, 2.0, 16)
    dt, b1, b2, dL = 0.02, 0.3, -0.7, [0.1, -2.0, 0.5, 1.5]
    out = orc.nl_if2_step(g, psis, gs, dt, b1, b2, dL)
    x = orc.coords_x(g)[0]
    n = [np.abs(p) ** 2 for p in psis]
    for i in range(4):
        phi = sum(gs[j + 4 * i] * n[j] for j in range(4)) + (b1 if i < 2 else b2) * x - dL[i]
        assert np.abs(out[i] - psis[i] * np.exp(-1j * dt * phi)).max() < 1e-13


def _ds_solver(orc, g, psis, gs, sp2=None, **sp1):
    return orc.RTSolver(g, psis, gs, [0.5] * g.dim, dt=0.05, alpha2=[0.3] * g.dim, species2=sp2, **sp1)


def _ds_expected(orc, g, psis, gs, S, t1, m1, m2):
    """Per point: phi_i from all four densities; each species the
    exponential expm(-i dt [[V11, Om eta], [Om conj(eta), V22]]) with
    eta = exp(-i (laser_dk x + phase)/2), or diagonal phases when its Omega
    is 0."""
    s = S.s
    x = orc.coords_x(g)[0] * np.ones(psis[0].shape)
    n = [np.abs(p) ** 2 for p in psis]
    phi = [sum(gs[j + 4 * i] * n[j] for j in range(4)) + (s.beta0 if i < 2 else s.beta2) * x - s.deltaL[i]
           for i in range(4)]
    Om = m1 * s.amp0 * np.cos(s.laser_k * x - (s.laser_domh + s.chirp_rate0 * t1) * t1 + s.phase0 / 2)
    Om2 = m2 * s.amp2 * np.cos(s.laser_k2 * x - (s.laser_domh2 + s.chirp_rate2 * t1) * t1 + s.phase2 / 2)
    eta = np.exp(-0.5j * (s.laser_dk * x + s.phase0))
    eta2 = np.exp(-0.5j * (s.laser_dk2 * x + s.phase2))
    out = [p.copy() for p in psis]
    for idx in np.ndindex(x.shape):
        for i0, om, et in ((0, Om[idx], eta[idx]), (2, Om2[idx], eta2[idx])):
            if om == 0:
                for i in (i0, i0 + 1):
                    out[i][idx] = psis[i][idx] * np.exp(-1j * s.dt * phi[i][idx])
                continue
            H = np.array([[phi[i0][idx], om * et], [om * np.conj(et), phi[i0 + 1][idx]]])
            v = scipy.linalg.expm(-1j * s.dt * H) @ np.array([psis[i0][idx], psis[i0 + 1][idx]])
            out[i0][idx], out[i0 + 1][idx] = v
    return out


@pytest.mark.parametrize("modes", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_bragg_ds_matches_expm_per_species(orc, modes):
    rng = np.random.default_rng(11)
    shape = (20, 3)
    g = orc.make_grid(shape, L=[10.0, 3.0])
    psis = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(4)]
    gs = rng.uniform(0.1, 1.5, 16)
    S = _ds_solver(orc, g, psis, gs,
                   sp2=dict(beta2=-0.04, amp2=1.9, laser_k2=2.3, laser_domh2=0.7, laser_dk2=0.35, chirp_rate2=0.15,
                            phase2=0.9),
                   beta0=0.05, deltaL=[0.2, -1.2, 0.4, -0.8], amp0=2.5, laser_k=1.7, laser_domh=0.9, laser_dk=0.6,
                   chirp_rate0=0.1, phase0=0.4)
    t1 = 0.37
    out = orc.bragg_ds(g, psis, S, t1, *modes)
    exp = _ds_expected(orc, g, psis, gs, S, t1, *modes)
    for c in range(4):
        assert np.abs(out[c] - exp[c]).max() < 1e-12, (modes, c)
    # each species keeps its population point by point
    for i0 in (0, 2):
        n_in = np.abs(psis[i0]) ** 2 + np.abs(psis[i0 + 1]) ** 2
        n_out = np.abs(out[i0]) ** 2 + np.abs(out[i0 + 1]) ** 2
        assert np.abs(n_out - n_in).max() < 1e-12 * n_in.max()


def test_bragg_ds_block_with_an_empty_second_species_is_single_bragg(orc):
    # species 2 empty and uncoupled, alpha2 = alpha: species 1 follows the
    # single-species bragg_ad block (no chirp, AMP_T = 1, phase 0, DeltaL_0 = 0)
    n, L = 64, 24.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    p1 = np.exp(-x ** 2 / 8).astype(np.complex128)
    p1 /= np.sqrt(orc.particle_number(g, p1))
    z = np.zeros_like(p1)
    g2 = [3.0, 2.5, 2.5, 3.1]
    g4 = np.zeros(16)
    g4[0], g4[1], g4[4], g4[5] = g2
    lp = dict(beta0=0.02, amp0=4.0, laser_k=2 * 2 * np.pi / L * 3, laser_domh=1.5, laser_dk=0.2, chirp_rate0=0.05)
    ds = orc.RTSolver(g, [p1, z, z, z], g4, [0.5], dt=2e-3, alpha2=[0.5], deltaL=[0.0, -0.3, 0.0, 0.0],
                      species2=dict(amp2=3.0, laser_k2=1.0, duration=(10.0, 10.0)), **lp)
    single = orc.RTSolver(g, [p1, z], g2, [0.5], dt=2e-3, deltaL=[0.0, -0.3], **lp)
    for _ in range(3):
        ds.block(orc.BRAGG_DS, 20)
        single.block(orc.BRAGG_AD, 20)
    assert ds.t == single.t
    err = np.linalg.norm(np.concatenate([ds.psis[0] - single.psis[0], ds.psis[1] - single.psis[1]]))
    assert err / np.linalg.norm(np.concatenate(single.psis)) < 1e-12
    assert not np.any(ds.psis[2]) and not np.any(ds.psis[3])


# ------------------------------------------------------------ three states (f3)

def _rand_states(shape, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(n)]


@pytest.mark.parametrize("n", [2, 3, 4])
def test_herm_expm_matches_scipy_also_when_degenerate(orc, n):
    rng = np.random.default_rng(40 + n)
    for trial in range(60):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        H = (A + A.conj().T) * rng.uniform(0.1, 80.0)
        if trial % 3 == 0:   # a repeated eigenvalue
            w, V = np.linalg.eigh(H)
            w[-1] = w[0]
            H = (V * w) @ V.conj().T
            H = (H + H.conj().T) / 2
        if trial % 7 == 0:   # already diagonal
            H = np.diag(np.diag(H).real).astype(np.complex128)
        dt = rng.uniform(1e-3, 1.5)
        U = orc.herm_expm(H, dt)
        R = scipy.linalg.expm(-1j * dt * H)
        scale = max(1.0, np.abs(H).max() * dt)
        assert np.abs(U - R).max() < 1e-14 * scale, (trial, np.abs(U - R).max())
        assert np.abs(U @ U.conj().T - np.eye(n)).max() < 1e-14


def _three_state_solver(orc, g, psis, gs, **kw):
    return orc.RTSolver(g, psis, gs, [0.5] * g.dim, dt=kw.pop("dt", 0.05), **kw)


def test_num_bragg_two_states_is_the_analytic_bragg_step(orc):
    # the numerical Bragg step with Amp_1 = (A, A), no chirp and no phase
    # couples the two states by A cos(k x - domh t): the Hamiltonian of
    # Do_Bragg_ad with laser_dk = 0, whose closed form is independent code
    rng = np.random.default_rng(7)
    shape = (48,)
    g = orc.make_grid(shape, L=12.0)
    p1, p2 = _rand_states(shape, 2, 8)
    gs = (1.0, 0.3, 0.25, 0.8)
    t, dt = 0.31, 0.07
    P = orc.bragg_params(dt=dt, t=t, F=1.0, amp0=2.5, laser_k=1.7, laser_domh=0.9, beta0=0.05, deltaL1=-1.2, gs=gs)
    e1, e2 = orc.bragg_ad(g, p1, p2, P)
    S = _three_state_solver(orc, g, [p1, p2], list(gs), dt=dt, amp0=2.5, amp1=2.5, laser_k=1.7, laser_domh=0.9,
                            beta=[0.05], deltaL=[0.0, -1.2])
    o1, o2 = orc.three_state_step(orc.NUM_BRAGG, g, [p1, p2], S, t + 0.5 * dt)
    scale = np.abs(np.concatenate([p1, p2])).max()
    assert np.abs(o1 - e1).max() < 1e-13 * scale and np.abs(o2 - e2).max() < 1e-13 * scale


def test_num_bragg_three_states_matches_expm_and_is_unitary(orc):
    shape = (10, 3)
    g = orc.make_grid(shape, L=[6.0, 2.0])
    psis = _rand_states(shape, 3, 9)
    gs = list(np.random.default_rng(10).uniform(0.1, 1.2, 9))
    kw = dict(amp0=2.0, amp1=0.7, laser_k=1.3, laser_domh=0.8, chirp_rate0=0.3, phase0=0.5,
              beta=[0.04, -0.02], deltaL=[0.1, -0.9, 0.6])
    S = _three_state_solver(orc, g, psis, gs, dt=0.09, **kw)
    t = 0.42
    out = orc.three_state_step(orc.NUM_BRAGG, g, psis, S, t)
    x0, x1 = [np.broadcast_to(c, shape) for c in orc.coords_x(g)]
    lk, cr = [1.3, -1.3], [0.3, -0.3]
    for idx in np.ndindex(shape):
        v = np.array([p[idx] for p in psis])
        n = np.abs(v) ** 2
        H = np.zeros((3, 3), complex)
        for i in range(3):
            H[i, i] = sum(gs[j + 3 * i] * n[j] for j in range(3)) + 0.04 * x0[idx] - 0.02 * x1[idx] - kw["deltaL"][i]
        for i in range(2):
            th = (-0.8 + cr[i] * t) * t + lk[i] * x0[idx] - 0.25
            H[i + 1, 0] = (2.0 + 0.7) / 2 * np.cos(th) + 1j * (2.0 - 0.7) / 2 * np.sin(th)
            H[0, i + 1] = np.conj(H[i + 1, 0])
        ref = scipy.linalg.expm(-1j * 0.09 * H) @ v
        got = np.array([o[idx] for o in out])
        assert np.abs(got - ref).max() < 1e-13 * np.abs(v).max()
        assert abs(np.sum(np.abs(got) ** 2) - np.sum(n)) < 1e-13 * np.sum(n)


def test_raman_matches_expm_and_reduces_to_phases(orc):
    shape = (16,)
    g = orc.make_grid(shape, L=8.0)
    psis = _rand_states(shape, 3, 12)
    gs = list(np.random.default_rng(13).uniform(0.1, 1.0, 9))
    kw = dict(amp0=3.0, amp1=1.1, laser_k=2.2, laser_domh=4.0, beta=[0.06], deltaL=[0.0, -2.0, 1.5])
    S = _three_state_solver(orc, g, psis, gs, dt=0.05, **kw)
    out = orc.three_state_step(orc.RAMAN, g, psis, S, 0.0)
    (x,) = orc.coords_x(g)
    for l in range(shape[0]):
        v = np.array([p[l] for p in psis])
        n = np.abs(v) ** 2
        H = np.diag([sum(gs[j + 3 * i] * n[j] for j in range(3)) + 0.06 * x[l] - kw["deltaL"][i]
                     for i in range(3)]).astype(complex)
        H[2, 2] += 4.0
        e = np.exp(1j * 2.2 * x[l])
        H[0, 2], H[2, 0] = 1.5 * e, 1.5 * np.conj(e)
        H[1, 2], H[2, 1] = 0.55 * np.conj(e), 0.55 * e
        ref = scipy.linalg.expm(-1j * 0.05 * H) @ v
        assert np.abs(np.array([o[l] for o in out]) - ref).max() < 1e-13 * np.abs(v).max()
    # no light: each state only picks up its phase
    S0 = _three_state_solver(orc, g, psis, gs, dt=0.05, laser_domh=4.0, beta=[0.06], deltaL=[0.0, -2.0, 1.5])
    out0 = orc.three_state_step(orc.RAMAN, g, psis, S0, 0.0)
    n = [np.abs(p) ** 2 for p in psis]
    for i in range(3):
        V = sum(gs[j + 3 * i] * n[j] for j in range(3)) + 0.06 * x - kw["deltaL"][i] + (4.0 if i == 2 else 0.0)
        assert np.abs(out0[i] - psis[i] * np.exp(-1j * 0.05 * V)).max() < 1e-14 * np.abs(psis[i]).max()


def _dbragg_columns(orc, g, S, t, F):
    """The 3x3 matrix of the double Bragg step at every point (g = 0, so the
    step is linear): column j is the step applied to the unit state j."""
    shape = tuple(int(v) for v in orc.grid_shape(g))
    cols = []
    for j in range(3):
        e = [np.full(shape, 1.0 + 0j) if c == j else np.zeros(shape, complex) for c in range(3)]
        cols.append(orc.three_state_step(orc.DOUBLE_BRAGG, g, e, S, t, F))
    return cols   # cols[j][i] = O_ij


def test_double_bragg_closed_form_is_the_exponential_without_dk(orc):
    # laser_dk = 0: the reference's O is exp(-i dt H) for
    # H = [[V11, Op, Om], [Op, V22, 0], [Om, 0, V22]]
    shape = (24,)
    g = orc.make_grid(shape, L=9.0)
    kw = dict(amp0=2.4, laser_k=1.1, laser_domh=0.6, chirp_rate0=0.2, phase0=0.3, beta=[0.07],
              deltaL=[0.0, -1.4, 0.0])
    S = _three_state_solver(orc, g, [np.zeros(shape, complex)] * 3, [0.0] * 9, dt=0.08, **kw)
    t, F = 0.55, 0.9
    cols = _dbragg_columns(orc, g, S, t, F)
    (x,) = orc.coords_x(g)
    t1 = t + 0.04
    for l in range(shape[0]):
        V11 = 0.07 * x[l]
        V22 = 1.4 + 0.07 * x[l]          # - DeltaL[1] with DeltaL[1] = -1.4
        Op = F * 2.4 * np.cos(1.1 * x[l] - (0.6 + 0.2 * t1) * t1 - 0.15)
        Om = F * 2.4 * np.cos(-1.1 * x[l] - (0.6 - 0.2 * t1) * t1 - 0.15)
        H = np.array([[V11, Op, Om], [Op, V22, 0], [Om, 0, V22]], complex)
        U = scipy.linalg.expm(-1j * 0.08 * H)
        O = np.array([[cols[j][i][l] for j in range(3)] for i in range(3)])
        assert np.abs(O - U).max() < 1e-12, (l, np.abs(O - U).max())


def test_double_bragg_dk_phases_and_clamped_amplitude(orc):
    # laser_dk != 0 multiplies O12, O31 by eta and O21, O13 by conj(eta),
    # eta = exp(-i dk x / 2); the diagonal is unchanged.  O23 and O32 take
    # the reference's own factors (double_bragg.cpp:216-222), restated here
    shape = (20,)
    g = orc.make_grid(shape, L=7.0)
    base = dict(amp0=1.9, laser_k=0.9, laser_domh=0.4, chirp_rate0=0.1, phase0=0.2, beta=[0.03],
                deltaL=[0.0, -0.8, 0.0])
    z = [np.zeros(shape, complex)] * 3
    S0 = _three_state_solver(orc, g, z, [0.0] * 9, dt=0.06, laser_dk=0.0, **base)
    S1 = _three_state_solver(orc, g, z, [0.0] * 9, dt=0.06, laser_dk=0.7, **base)
    t, F = 0.2, 1.3
    c0, c1 = _dbragg_columns(orc, g, S0, t, F), _dbragg_columns(orc, g, S1, t, F)
    (x,) = orc.coords_x(g)
    eta = np.exp(-0.5j * 0.7 * x)
    O0 = lambda i, j: c0[j][i]
    O1 = lambda i, j: c1[j][i]
    for i in range(3):
        assert np.abs(O1(i, i) - O0(i, i)).max() < 1e-14
    assert np.abs(O1(0, 1) - eta * O0(0, 1)).max() < 1e-14
    assert np.abs(O1(2, 0) - eta * O0(2, 0)).max() < 1e-14
    assert np.abs(O1(1, 0) - np.conj(eta) * O0(1, 0)).max() < 1e-14
    assert np.abs(O1(0, 2) - np.conj(eta) * O0(0, 2)).max() < 1e-14
    e0, e1 = eta.real, eta.imag
    z23 = O0(1, 2)
    o23 = (e0 ** 2 + e1 ** 2) * z23.real + 2 * e0 * e1 * z23.imag + 1j * ((e0 ** 2 + e1 ** 2) * z23.imag
                                                                         - 2 * e0 * e1 * z23.real)
    o32 = (e0 ** 2 - e1 ** 2) * z23.real - 2 * e0 * e1 * z23.imag + 1j * ((e0 ** 2 - e1 ** 2) * o23.imag
                                                                         + 2 * e0 * e1 * z23.real)
    assert np.abs(O1(1, 2) - o23).max() < 1e-14 and np.abs(O1(2, 1) - o32).max() < 1e-14
    # F < 0 is clamped to 0: no light, pure phases (V22 on states 2 and 3)
    cn = _dbragg_columns(orc, g, S1, t, -0.5)
    assert np.abs(cn[0][0] - np.exp(-1j * 0.06 * 0.03 * x)).max() < 1e-15
    assert np.abs(cn[2][2] - np.exp(-1j * 0.06 * (0.8 + 0.03 * x))).max() < 1e-15
    assert not np.any(cn[0][1]) and not np.any(cn[1][2])


def test_three_state_blocks_conserve_the_norm(orc):
    n, L = 64, 16.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    p = np.exp(-x ** 2 / 4).astype(np.complex128)
    p /= np.sqrt(orc.particle_number(g, p))
    psis = [p, 0.3 * p, 0.2j * p]
    gs = [1.0, 0.9, 0.9, 0.9, 1.1, 0.8, 0.9, 0.8, 1.2]
    N0 = sum(orc.particle_number(g, q) for q in psis)
    for kind, kw in ((orc.DOUBLE_BRAGG, dict(amp0=3.0, laser_k=1.2, laser_domh=0.5, deltaL=[0, -0.4, 0])),
                     (orc.NUM_BRAGG, dict(amp0=3.0, amp1=2.0, laser_k=1.2, laser_domh=0.5, deltaL=[0, -0.4, 0.2])),
                     (orc.RAMAN, dict(amp0=3.0, amp1=2.0, laser_k=1.2, laser_domh=2.5, deltaL=[0, -0.4, 0.2]))):
        S = orc.RTSolver(g, psis, gs, [0.5], dt=2e-3, amp=lambda t: 1.0, **kw)
        for _ in range(3):
            S.block(kind, 10)
        N = sum(orc.particle_number(g, q) for q in S.psis)
        assert abs(N - N0) < 1e-12, (kind, N - N0)
