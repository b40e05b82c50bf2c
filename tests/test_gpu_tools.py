"""The .bin ecosystem around the path on the GPU: the moments of ana_expval,
inflate_domain, the two-component ground state of sobmin_2, and the
pipeline the reference's batch script runs (sobmin_2, inflate_domain
--fak, bragg_ds)."""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_drivers import rel_l2, write_bin
from test_host import read_bin, tool

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _axes(shape, L):
    return [((np.arange(n) - n // 2) * (L[a] / n)).reshape([n if b == a else 1 for b in range(len(shape))])
            for a, n in enumerate(shape)]


def _numpy_moments(psi, L):
    """ana_expval's two-pass moments (ana_expval.cpp:50-185) in numpy."""
    shape = psi.shape
    dim = psi.ndim
    dx = [L[a] / shape[a] for a in range(dim)]
    dk = [2 * np.pi / L[a] for a in range(dim)]
    ar, ark = np.prod(dx), np.prod(dk)
    d = np.abs(psi) ** 2
    N = ar * d.sum()
    xs = _axes(shape, L)
    x = [ar * (xs[a] * d).sum() / N for a in range(dim)]
    x2 = [ar * ((xs[a] - x[a]) ** 2 * d).sum() / N for a in range(dim)]
    pk = np.fft.fftn(psi) * ar / (2 * np.pi) ** (dim / 2)
    dkk = np.abs(pk) ** 2
    ks = [(np.fft.fftfreq(n, 1.0 / n) * dk[a]).reshape([n if b == a else 1 for b in range(dim)])
          for a, n in enumerate(shape)]
    k = [ark * (ks[a] * dkk).sum() / N for a in range(dim)]
    k2 = [ark * ((ks[a] - k[a]) ** 2 * dkk).sum() / N for a in range(dim)]
    phase = (np.arctan2(psi.imag, psi.real) * d).sum() / d.sum()
    return dict(N=N, x=x, x2=x2, k=k, k2=k2, phase=phase)


@pytest.mark.parametrize("shape", [(1024,), (128, 64), (64, 64, 64)])
def test_moments_of_a_moving_gaussian(hip, shape):
    # grids on which the Gaussian and its spectrum both decay to below
    # 1e-16 of their peaks, so the sums equal the integrals
    dim = len(shape)
    L = [40.0, 32.0, 36.0][:dim]
    c, s, p = [1.5, -2.0, 0.7][:dim], [2.0, 1.6, 1.8][:dim], [1.2, -0.8, 0.5][:dim]
    xs = _axes(shape, L)
    psi = np.ones(shape, np.complex128)
    for a in range(dim):
        psi = psi * np.exp(-(xs[a] - c[a]) ** 2 / (2 * s[a] ** 2) + 1j * p[a] * xs[a])
    with hip.Context(shape, [L[a] / shape[a] for a in range(dim)], [2 * np.pi / L[a] for a in range(dim)], 1) as ctx:
        ctx.upload(0, psi)
        m = ctx.moments(0)
    ref = _numpy_moments(psi, L)
    N_exact = np.prod([np.sqrt(np.pi) * s[a] for a in range(dim)])
    assert abs(m["N"] - N_exact) < 1e-12 * N_exact
    for a in range(dim):
        # closed forms of the Gaussian: <x> = c, <(x-c)^2> = s^2/2,
        # <k> = p, <(k-p)^2> = 1/(2 s^2)
        assert abs(m["x"][a] - c[a]) < 1e-12 and abs(m["x2"][a] - s[a] ** 2 / 2) < 1e-12
        assert abs(m["k"][a] - p[a]) < 1e-12 and abs(m["k2"][a] - 1 / (2 * s[a] ** 2)) < 1e-12
        for key in ("x", "x2", "k", "k2"):
            assert abs(m[key][a] - ref[key][a]) < 1e-12 * max(1.0, abs(ref[key][a])), key
    assert abs(m["phase"] - ref["phase"]) < 1e-12


def test_moments_phase_of_a_constant_phase(hip):
    n, L, th = 512, 20.0, 0.83
    (x,) = _axes((n,), [L])
    psi = np.exp(-x ** 2 / 4 + 1j * th)
    with hip.Context((n,), [L / n], [2 * np.pi / L], 1) as ctx:
        ctx.upload(0, psi)
        assert abs(ctx.moments(0)["phase"] - th) < 1e-14


def _inflate_numpy(psi, fak, fix):
    """inflate_{1,2,3}d[_fix] index maps (inflate_domain.cpp:41-223)."""
    shape = psi.shape
    if fix:
        out = np.zeros(shape, np.complex128)
        off = [n * (fak - 1) // 2 // fak for n in shape]
        sub = psi[tuple(slice(0, n, fak) for n in shape)]
        out[tuple(slice(o, o + m) for o, m in zip(off, sub.shape))] = sub
        return out
    out = np.zeros(tuple(n * fak for n in shape), np.complex128)
    off = [(n * fak - n) // 2 for n in shape]
    out[tuple(slice(o, o + n) for o, n in zip(off, shape))] = psi
    return out


@pytest.mark.parametrize("shape,fak,fix", [((256,), 4, False), ((64, 32), 2, False), ((16, 8, 32), 2, False),
                                           ((64, 32), 4, True), ((16, 16, 32), 2, True)])
def test_inflate_matches_the_index_map(hip, shape, fak, fix):
    dim = len(shape)
    rng = np.random.default_rng(3)
    psis = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(2)]
    dst_shape = shape if fix else tuple(n * fak for n in shape)
    with hip.Context(shape, [0.5] * dim, [1.0] * dim, 2) as S, hip.Context(dst_shape, [0.5] * dim, [1.0] * dim, 2) as D:
        for c in range(2):
            S.upload(c, psis[c])
            D.upload(c, np.full(dst_shape, 7.0 + 0j))   # everything must be overwritten
        D.inflate_from(S, fak, fix)
        for c in range(2):
            assert np.array_equal(D.download(c), _inflate_numpy(psis[c], fak, fix))
        with pytest.raises(hip.AtusError, match="inflation needs"):
            D.inflate_from(S, fak if fix else 2 * fak, False)


def test_inflate_domain_and_ana_expval_tools(tmp_path):
    n, L = 256, 20.0
    (x,) = _axes((n,), [L])
    psi = np.exp(-(x - 1.0) ** 2 / 4 + 0.6j * x)
    write_bin(tmp_path / "a.bin", psi, [L])
    r = subprocess.run([tool("inflate_domain"), "--fak", "4", "a.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "# new header" in r.stdout and "# nDimX    == 1024" in r.stdout
    h, data = read_bin(tmp_path / "inf_a.bin")
    assert h["nDimX"] == 1024 and h["dx"] == L / n and h["dkx"] == 2 * np.pi / L / 4 and h["dt"] == 0.001
    assert h["xMin"] == -L / 2 * 4 and h["xMax"] == L / 2 * 4
    assert np.array_equal(data, _inflate_numpy(psi, 4, False))
    # the moments of the inflated state are those of the original
    out = []
    for f in ("a.bin", "inf_a.bin"):
        r = subprocess.run([tool("ana_expval"), f], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        out.append([float(v) for v in r.stdout.split()])
    ref = _numpy_moments(psi, [L])
    want = [ref["phase"], ref["N"], ref["x"][0], ref["x2"][0], ref["k"][0], ref["k2"][0]]
    assert np.allclose(out[0], want, rtol=1e-5, atol=1e-9)    # 6 significant digits
    assert np.allclose(out[1][1:], want[1:], rtol=1e-5, atol=1e-9)
    # --fix keeps the number of points
    write_bin(tmp_path / "b.bin", psi.reshape(16, 16), [L, L])
    r = subprocess.run([tool("inflate_domain"), "--fix=2", "b.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    h, data = read_bin(tmp_path / "inf_b.bin")
    assert h["nDimX"] == 16 and h["dx"] == 2 * L / 16
    assert np.array_equal(data.reshape(16, 16), _inflate_numpy(psi.reshape(16, 16), 2, True))


def _sobmin2_xml(n, L, eps, g, N):
    return f"""<SIMULATION><DIM>1</DIM><FILENAME>K.bin</FILENAME><FILENAME_3>Rb.bin</FILENAME_3>
      <GUESS_1D>exp(-0.2*x^2)</GUESS_1D><GUESS_1D_2>exp(-0.1*x^2)</GUESS_1D_2>
      <POTENTIAL_1D>0.5*x^2</POTENTIAL_1D><POTENTIAL_1D_2>0.5*0.64*x^2</POTENTIAL_1D_2>
      <CONSTANTS><epsilon>{eps!r}</epsilon><stepsize>0.01</stepsize></CONSTANTS>
      <VCONSTANTS><Alpha_1>0.5</Alpha_1><Alpha_2>0.3</Alpha_2>
        <GS_1>{g[0]},{g[1]}</GS_1><GS_2>{g[2]},{g[3]}</GS_2><N>{N[0]},{N[1]}</N></VCONSTANTS>
      <ALGORITHM><NX>{n}</NX><XMIN>{-L / 2!r}</XMIN><XMAX>{L / 2!r}</XMAX></ALGORITHM></SIMULATION>"""


def test_sobmin_2_ground_states(orc, tmp_path):
    n, L = 256, 24.0
    g, N = [3.0, 1.5, 1.5, 2.0], [1.0, 0.5]
    (tmp_path / "s.xml").write_text(_sobmin2_xml(n, L, 1e-16, g, N))
    r = subprocess.run([tool("sobmin_2"), "s.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ATUS_SOB_MAXITER="2000"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    iters = r.stdout.count("--- ")
    outs = [read_bin(tmp_path / f)[1] for f in ("K.bin", "Rb.bin")]
    assert (tmp_path / "zero.bin").exists()
    grid = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(grid)
    V = [0.5 * x ** 2, 0.5 * 0.64 * x ** 2]
    guess = [np.exp(-0.2 * x ** 2).astype(np.complex128), np.exp(-0.1 * x ** 2).astype(np.complex128)]
    gsc = [g[0] * N[0], g[1] * np.sqrt(N[0] * N[1]), g[2] * np.sqrt(N[0] * N[1]), g[3] * N[1]]
    S = orc.SobSolver(grid, guess, V, gsc, [0.5, 0.3], stepsize=0.01)
    S.renormalize()
    for _ in range(iters):
        S.iterate()
    # Save(filenames, false): the states are saved as the minimiser holds
    # them, normalised to 1 up to its renormalisation threshold
    for c in range(2):
        assert rel_l2(outs[c], S.psi[c]) < 1e-9, c
        Nc = orc.particle_number(grid, outs[c])
        assert abs(Nc - orc.particle_number(grid, S.psi[c])) < 1e-9 and abs(Nc - 1.0) < 1e-4


def test_sobmin_2_inflate_bragg_ds_pipeline(orc, tmp_path):
    # the per-case pipeline of the reference's batch script
    # (batch_case_generation_4_tammy.cpp:271-283): ground states of two
    # species, domains inflated, then the two-species Bragg pulses
    n, L, fak = 128, 24.0, 4
    (tmp_path / "s.xml").write_text(_sobmin2_xml(n, L, 1e-14, [3.0, 1.5, 1.5, 2.0], [1.0, 0.5]))
    env = dict(os.environ, ATUS_SOB_MAXITER="3000")
    r = subprocess.run([tool("sobmin_2"), "s.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    for f in ("K.bin", "Rb.bin", "zero.bin"):
        r = subprocess.run([tool("inflate_domain"), "--fak", str(fak), f], cwd=tmp_path, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
    states = [read_bin(tmp_path / f)[1] for f in ("inf_K.bin", "inf_zero.bin", "inf_Rb.bin", "inf_zero.bin")]
    small = [read_bin(tmp_path / f)[1] for f in ("K.bin", "Rb.bin")]
    assert np.array_equal(states[0], _inflate_numpy(small[0], fak, False))
    assert np.array_equal(states[2], _inflate_numpy(small[1], fak, False))
    gs = [2.0, 1.9, 0.5, 0.4, 1.9, 2.0, 0.4, 0.5, 0.5, 0.4, 3.0, 2.8, 0.4, 0.5, 2.8, 3.0]
    gsx = "".join(f"<GS_{i + 1}>{','.join(str(v) for v in gs[4 * i:4 * i + 4])}</GS_{i + 1}>" for i in range(4))
    (tmp_path / "p.xml").write_text(f"""<SIMULATION><DIM>1</DIM>
      <FILENAME>inf_K.bin</FILENAME><FILENAME_2>inf_zero.bin</FILENAME_2>
      <FILENAME_3>inf_Rb.bin</FILENAME_3><FILENAME_4>inf_zero.bin</FILENAME_4>
      <CONSTANTS><laser_k>1.0</laser_k><laser_k_2>0.75</laser_k_2><laser_domh>0.5</laser_domh>
        <laser_domh_2>0.28</laser_domh_2><laser_dk>0</laser_dk><laser_dk_2>0</laser_dk_2>
        <rabi_threshold>0.3</rabi_threshold></CONSTANTS>
      <VCONSTANTS>{gsx}<Alpha_1>0.5</Alpha_1><Alpha_2>0.3</Alpha_2>
        <Beta>0</Beta><Beta_2>0</Beta_2><Delta_L>0,0.1,0,-0.1</Delta_L>
        <Amp_1>4,4</Amp_1><Amp_2>3,3</Amp_2></VCONSTANTS>
      <ALGORITHM><DT>1e-3</DT></ALGORITHM>
      <SEQUENCE><bragg_ad dt="1e-3" Nk="50" rabi_output_freq="each" output_freq="last">0.1,0.05</bragg_ad>
      </SEQUENCE></SIMULATION>""")
    r = subprocess.run([tool("bragg_ds"), "p.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    grid = orc.make_grid((n * fak,), L=L * fak)
    sp2 = dict(amp2=3.0, laser_k2=0.75, laser_domh2=0.28, seq_time=0.0, duration=(0.1, 0.05))
    S = orc.RTSolver(grid, states, gs, [0.5], dt=1e-3, alpha2=[0.3], species2=sp2, deltaL=[0, 0.1, 0, -0.1],
                     amp0=4.0, laser_k=1.0, laser_domh=0.5)
    pert = [p * (1 + 2e-16 * np.random.default_rng(99).standard_normal(p.shape)) for p in states]
    Sp = orc.RTSolver(grid, pert, gs, [0.5], dt=1e-3, alpha2=[0.3], species2=sp2, deltaL=[0, 0.1, 0, -0.1],
                      amp0=4.0, laser_k=1.0, laser_domh=0.5)
    for _ in range(2):
        S.block(orc.BRAGG_DS, 50)
        Sp.block(orc.BRAGG_DS, 50)
    out = np.concatenate([read_bin(tmp_path / f"0.100_{c}.bin")[1] for c in range(1, 5)])
    tol = max(1e-11, 20 * rel_l2(np.concatenate(Sp.psis), np.concatenate(S.psis)))
    assert rel_l2(out, np.concatenate(S.psis)) < tol
    assert (tmp_path / "Spezie_1_Rabi_1_0.txt").exists() and (tmp_path / "Spezie_2_Rabi_1_0.txt").exists()
