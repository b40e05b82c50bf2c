#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from an independent numpy
restatement of the reference's sequence block (CRT_Base.h:776-786):

    half kinetic, (V, full kinetic) x (Nk-1), V, half kinetic

with numpy's pocketfft for the transforms (scale prod(dx)/(2pi)^{d/2}
forward, prod(dk)/(2pi)^{d/2} backward, cft_3d.cpp:627-637), the kinetic
phase exp(-i s dt sum alpha_a k_a^2) in native order, the V steps of
CRT_Base::Do_NL_Step, CRT_Base_IF::Do_NL_Step and, for Bragg_single::
Do_Bragg_ad, the 2x2 exponential written in the rotation form

    U = exp(-i dt (V11+V22)/2) [cos(th) 1 - i sin(th)/th * dt (H - (V11+V22)/2)],
    th = dt sqrt(((V11-V22)/2)^2 + Omega^2)

(mathematically equal to the reference's eigenvector form, bragg.cpp:242-289,
and well conditioned everywhere).  V-step m of a block sees the amplitude at
t1 = t_block + m dt (accumulated as the reference does).

The fixtures are inputs and expected outputs only; run this script to
regenerate them (deterministic).
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def grid(shape, L):
    dx = [L[a] / shape[a] for a in range(len(shape))]
    dk = [2 * math.pi / L[a] for a in range(len(shape))]
    xs, ks = [], []
    for a, n in enumerate(shape):
        sh = [1] * len(shape)
        sh[a] = n
        xs.append(((np.arange(n) - n // 2) * dx[a]).reshape(sh))
        ks.append((dk[a] * (((np.arange(n) + n // 2) % n) - n // 2)).reshape(sh))
    return dx, dk, xs, ks


def scale(d, s):
    if d == 1:
        return s[0] / math.sqrt(2 * math.pi)
    if d == 2:
        return 0.5 * s[0] * s[1] / math.pi
    return s[0] * s[1] * s[2] / (2 * math.pi) ** 1.5


def kinetic(psis, frac, dt, alpha, dx, dk, ks):
    d = len(dx)
    ph = np.exp(1j * frac * (-dt) * sum(alpha[a] * ks[a] ** 2 for a in range(d)))
    out = []
    for p in psis:
        f = np.fft.fftn(p) * scale(d, dx)
        out.append(np.fft.ifftn(f * ph) * p.size * scale(d, dk))
    return out


def v_nl(psis, gs, dt, pots):
    n = [np.abs(p) ** 2 for p in psis]
    C = len(psis)
    out = []
    for i in range(C):
        phi = (pots[i] if pots is not None else 0.0) + sum(gs[j + C * i] * n[j] for j in range(C))
        out.append(psis[i] * np.exp(-1j * dt * phi))
    return out


def v_nl_if(psis, gs, dt, beta0, deltaL, x0):
    n = [np.abs(p) ** 2 for p in psis]
    C = len(psis)
    return [psis[i] * np.exp(-1j * dt * (sum(gs[j + C * i] * n[j] for j in range(C)) + beta0 * x0 - deltaL[i]))
            for i in range(C)]


def v_bragg(psis, p, dt, t1, F, x0):
    a, b = psis
    n1, n2 = np.abs(a) ** 2, np.abs(b) ** 2
    V11 = p["gs"][0] * n1 + p["gs"][1] * n2 + p["beta0"] * x0
    V22 = p["gs"][2] * n1 + p["gs"][3] * n2 - p["deltaL"][1] + p["beta0"] * x0
    Om = F * p["amp0"] * np.cos(p["laser_k"] * x0 - (p["laser_domh"] + p["chirp"] * t1) * t1)
    eta = np.exp(-0.5j * p["laser_dk"] * x0)
    m = 0.5 * (V11 + V22)
    h = 0.5 * (V11 - V22)
    w = np.sqrt(h * h + Om * Om)
    th = dt * w
    sinc = np.where(th > 0, np.sin(th) / np.where(th > 0, th, 1.0), 1.0)
    ph = np.exp(-1j * dt * m)
    c = np.cos(th)
    U11 = ph * (c - 1j * sinc * dt * h)
    U22 = ph * (c + 1j * sinc * dt * h)
    U12 = ph * (-1j * sinc * dt * Om) * eta
    U21 = ph * (-1j * sinc * dt * Om) * np.conj(eta)
    return [U11 * a + U12 * b, U21 * a + U22 * b]


def run(kind, shape, L, psis, dt, Nk, nblocks, t0, alpha, gs, pots=None, ifp=None, amp=None):
    dx, dk, xs, ks = grid(shape, L)
    x0 = xs[0] * np.ones(shape)
    t = t0
    out = [p.copy() for p in psis]
    norms = []
    for _ in range(nblocks):
        out = kinetic(out, 0.5, dt, alpha, dx, dk, ks)
        t += 0.5 * dt
        for m in range(1, Nk + 1):
            if kind == "freeprop":
                out = v_nl(out, gs, dt, pots)
            elif kind == "freeprop_if":
                out = v_nl_if(out, gs, dt, ifp["beta0"], ifp["deltaL"], x0)
            else:
                t1 = t + 0.5 * dt
                out = v_bragg(out, dict(ifp, gs=gs), dt, t1, amp(t1), x0)
            frac = 1.0 if m < Nk else 0.5
            out = kinetic(out, frac, dt, alpha, dx, dk, ks)
            t += dt if m < Nk else 0.5 * dt
        norms.append([float(np.prod(dx) * np.sum(np.abs(p) ** 2)) for p in out])
    return out, t, np.array(norms)


def main():
    cases = {}
    # 1D rt_solver: harmonic trap, g = 10, seeded noise
    shape, L = (256,), [32.0]
    _, _, xs, _ = grid(shape, L)
    rng = np.random.default_rng(11)
    psi = np.pi ** -0.25 * np.exp(-xs[0] ** 2 / 2) + 1e-3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    psi /= np.sqrt(L[0] / shape[0] * np.sum(np.abs(psi) ** 2))
    V = 0.5 * xs[0] ** 2
    out, t, norms = run("freeprop", shape, L, [psi], 1e-3, 50, 2, 0.0, [0.5], [10.0], pots=[V])
    cases["rt1d"] = dict(kind="freeprop", shape=shape, L=L, dt=1e-3, Nk=50, nblocks=2, t0=0.0, alpha=[0.5],
                         gs=[10.0], psi0=[psi], pots=[V], out=out, t=t, norms=norms)
    # 2D freeprop with gravity and detuning, two components
    shape, L = (32, 16), [16.0, 12.0]
    _, _, xs, _ = grid(shape, L)
    a = np.exp(-(xs[0] ** 2 + xs[1] ** 2) / 4) * np.exp(1j * 0.5 * xs[0])
    b = 0.5 * np.exp(-((xs[0] - 1) ** 2 + xs[1] ** 2) / 3)
    ps = []
    for p in (a, b):
        p = p.astype(np.complex128)
        p /= np.sqrt(np.prod([L[i] / shape[i] for i in range(2)]) * np.sum(np.abs(p) ** 2))
        ps.append(p)
    ifp = dict(beta0=0.05, deltaL=[0.0, 0.3])
    out, t, norms = run("freeprop_if", shape, L, ps, 2e-3, 20, 2, 0.5, [0.5, 0.5], [2.0, 1.0, 1.0, 3.0], ifp=ifp)
    cases["if2d"] = dict(kind="freeprop_if", shape=shape, L=L, dt=2e-3, Nk=20, nblocks=2, t0=0.5, alpha=[0.5, 0.5],
                         gs=[2.0, 1.0, 1.0, 3.0], psi0=ps, ifp=ifp, out=out, t=t, norms=norms)
    # 3D Bragg pulse, small grid, two components
    shape, L = (16, 8, 8), [32.0, 16.0, 16.0]
    _, _, xs, _ = grid(shape, L)
    a = np.exp(-xs[0] ** 2 / 18 - xs[1] ** 2 / 8 - xs[2] ** 2 / 8) * np.ones(shape)
    a = a.astype(np.complex128)
    a /= np.sqrt(np.prod([L[i] / shape[i] for i in range(3)]) * np.sum(np.abs(a) ** 2))
    b = np.zeros_like(a)
    ifp = dict(beta0=0.0, deltaL=[0.0, 0.0], amp0=5.0, laser_k=2 * math.pi / 32.0 * 4, laser_domh=1.5,
               laser_dk=0.0, chirp=0.0)
    amp = lambda t: math.exp(-0.5 * ((t - 0.05) / 0.05) ** 2)
    out, t, norms = run("bragg_ad", shape, L, [a, b], 1e-3, 25, 2, 0.0, [0.5] * 3, [10.0, 9.5, 9.5, 10.0],
                        ifp=ifp, amp=amp)
    cases["bragg3d"] = dict(kind="bragg_ad", shape=shape, L=L, dt=1e-3, Nk=25, nblocks=2, t0=0.0, alpha=[0.5] * 3,
                            gs=[10.0, 9.5, 9.5, 10.0], psi0=[a, b], ifp=ifp, amp_t0=0.05, amp_sigma=0.05,
                            out=out, t=t, norms=norms)
    for name, c in cases.items():
        arrays = {"psi0": np.stack(c.pop("psi0")), "out": np.stack(c.pop("out")), "norms": c.pop("norms")}
        if c.get("pots") is not None:
            arrays["pots"] = np.stack(c.pop("pots"))
        meta = {k: (np.array(v) if not isinstance(v, dict) else np.array(repr(v))) for k, v in c.items()}
        np.savez(os.path.join(HERE, f"{name}.npz"), **arrays, **meta)
        print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
