"""The rebuilt drivers end to end on the GPU: an XML file and .bin inputs in,
.bin / Rabi_*.txt out, compared with the oracle run from the same files."""
import os
import subprocess

import numpy as np
import pytest

from test_host import read_bin, tool

pytestmark = pytest.mark.gpu


def write_bin(path, psi, L, dt=0.001, t=0.0):
    """A .bin file as gen_psi_0 writes it (my_structs.h:40-72)."""
    import struct
    shape = psi.shape
    dim = len(shape)
    n = list(shape) + [1] * (3 - dim)
    dx = [L[a] / shape[a] for a in range(dim)] + [0.0] * (3 - dim)
    dk = [2 * np.pi / L[a] for a in range(dim)] + [0.0] * (3 - dim)
    mins = [-L[a] / 2 for a in range(dim)] + [-10.0] * (3 - dim)
    maxs = [L[a] / 2 for a in range(dim)] + [10.0] * (3 - dim)
    head = struct.pack("<7q2i14d2i", 1380, 16, 0, dim, *n, 1, 1, t, mins[0], maxs[0], mins[1], maxs[1],
                       mins[2], maxs[2], *dx, *dk, dt, 0, 0)
    head += b"\0" * (1380 - len(head))
    with open(path, "wb") as f:
        f.write(head)
        f.write(np.ascontiguousarray(psi, dtype=np.complex128).tobytes())


def rel_l2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def test_rt_solver_1d_full_run(orc, tmp_path):
    # the 1D rt_solver workload (N=4096 on [-40,40], g=10, V=x^2/2, 1000 steps)
    n, L = 4096, 80.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    rng = np.random.default_rng(1)
    psi = np.pi ** -0.25 * np.exp(-x ** 2 / 2) + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    psi /= np.sqrt(orc.particle_number(g, psi))
    write_bin(tmp_path / "psi0.bin", psi, [L])
    (tmp_path / "rt.xml").write_text("""<SIMULATION><DIM>1</DIM><FILENAME>psi0.bin</FILENAME>
      <POTENTIAL_1D>0.5*x^2</POTENTIAL_1D>
      <VCONSTANTS><GS_1>10</GS_1><Alpha_1>0.5</Alpha_1></VCONSTANTS>
      <ALGORITHM><DT>0.001</DT></ALGORITHM>
      <SEQUENCE><freeprop dt="0.001" Nk="100" pn_freq="each" output_freq="last">1.0</freeprop></SEQUENCE>
      </SIMULATION>""")
    r = subprocess.run([tool("rt_solver"), "rt.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # pn_freq="each" prints N after every block and nothing more at the end
    # (CRT_Base.h:807-832)
    assert "FYI: Na          : 10" in r.stdout and r.stdout.count("N[0] = ") == 10
    h, out = read_bin(tmp_path / "1.000_1.bin")
    assert abs(h["t"] - 1.0) < 1e-12 and h["dt"] == 0.001
    S = orc.RTSolver(g, [psi], [10.0], [0.5], dt=1e-3, pots=[0.5 * x ** 2])
    for _ in range(10):
        S.block(orc.FREEPROP, 100)
    assert rel_l2(out, S.psis[0]) <= 1e-10


def test_bragg_3d_pulse_with_rabi_output(orc, tmp_path):
    # the 3D two-component Bragg workload on a 64^3 proxy of the 256^3 grid
    n, L = 64, 64.0
    g = orc.make_grid((n, n, n), L=L)
    x, y, z = orc.coords_x(g)
    psi1 = (np.exp(-x ** 2 / 72 - y ** 2 / 32 - z ** 2 / 32) * np.ones((n, n, n))).astype(np.complex128)
    psi1 /= np.sqrt(orc.particle_number(g, psi1))
    psi2 = np.zeros_like(psi1)
    write_bin(tmp_path / "a.bin", psi1, [L] * 3)
    write_bin(tmp_path / "b.bin", psi2, [L] * 3)
    (tmp_path / "br.xml").write_text("""<SIMULATION><DIM>3</DIM>
      <FILENAME>a.bin</FILENAME><FILENAME_2>b.bin</FILENAME_2>
      <AMP_T>exp(-0.5*((t-0.5)/0.5)^2)</AMP_T>
      <CONSTANTS><laser_k>4</laser_k><laser_domh>8</laser_domh><laser_dk>0</laser_dk>
        <rabi_threshold>2</rabi_threshold><chirp>0</chirp></CONSTANTS>
      <VCONSTANTS><GS_1>100,95</GS_1><GS_2>95,100</GS_2><Alpha_1>0.5,0.5,0.5</Alpha_1>
        <Beta>0,0,0</Beta><Delta_L>0,0</Delta_L><Amp_1>5,5</Amp_1></VCONSTANTS>
      <ALGORITHM><DT>2e-4</DT></ALGORITHM>
      <SEQUENCE><bragg_ad dt="2e-4" Nk="100" rabi_output_freq="each" output_freq="last">0.04</bragg_ad></SEQUENCE>
      </SIMULATION>""")
    r = subprocess.run([tool("bragg"), "br.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    amp = lambda t: float(np.exp(-0.5 * ((t - 0.5) / 0.5) ** 2))
    S = orc.RTSolver(g, [psi1, psi2], [100., 95., 95., 100.], [0.5] * 3, dt=2e-4, amp=amp, deltaL=[0., 0.],
                     amp0=5.0, laser_k=4.0, laser_domh=8.0)
    rabi_ref = []
    for _ in range(2):
        S.block(orc.BRAGG_AD, 100)
        rb, _ = orc.rabi_integrals(g, S.psis[0], [(0, 0, 0), (8, 0, 0), (-8, 0, 0)], 2.0)
        rabi_ref.append(rb)
    outs = [read_bin(tmp_path / f"0.040_{c}.bin")[1].reshape(n, n, n) for c in (1, 2)]
    err = rel_l2(np.concatenate([o.ravel() for o in outs]), np.concatenate([s_.ravel() for s_ in S.psis]))
    assert err < 1e-11
    lines = (tmp_path / "Rabi_1_0.txt").read_text().splitlines()
    assert lines[0].startswith("# time \t0,0,0\t8,0,0\t-8,0,0")
    rows = [[float(v) for v in ln.split("\t") if v] for ln in lines[1:]]
    assert [r_[0] for r_ in rows] == pytest.approx([0.0, 0.02])
    for r_, ref in zip(rows, rabi_ref):
        assert np.allclose(r_[1:], ref, rtol=1e-5, atol=1e-9)   # 6 significant digits in the file


def test_sobmin_3d_ground_state(orc, tmp_path):
    n = 32
    (tmp_path / "gs.xml").write_text(f"""<SIMULATION><DIM>3</DIM><FILENAME>gs.bin</FILENAME>
      <GUESS_3D>exp(-0.25*(x^2+y^2+z^2))</GUESS_3D>
      <POTENTIAL_3D>0.5*(x^2+1.5^2*y^2+0.75^2*z^2)</POTENTIAL_3D>
      <CONSTANTS><epsilon>1e-18</epsilon><stepsize>0.02</stepsize></CONSTANTS>
      <VCONSTANTS><Alpha_1>0.5</Alpha_1><GS_1>5</GS_1><N>2</N></VCONSTANTS>
      <ALGORITHM><NX>{n}</NX><NY>{n}</NY><NZ>{n}</NZ><XMIN>-8</XMIN><XMAX>8</XMAX>
        <YMIN>-8</YMIN><YMAX>8</YMAX><ZMIN>-8</ZMIN><ZMAX>8</ZMAX></ALGORITHM></SIMULATION>""")
    env = dict(os.environ, ATUS_SOB_MAXITER="3000")
    r = subprocess.run([tool("sobmin"), "gs.xml"], cwd=tmp_path, capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    iters = r.stdout.count("--- ")
    h, out = read_bin(tmp_path / "gs.bin")
    assert h["nDimX"] == n and (tmp_path / "zero.bin").exists()
    # the oracle from the same guess, same number of iterations
    g = orc.make_grid((n, n, n), L=16.0)
    x, y, z = orc.coords_x(g)
    V = 0.5 * (x ** 2 + 2.25 * y ** 2 + 0.5625 * z ** 2) * np.ones((n, n, n))
    guess = (np.exp(-0.25 * (x ** 2 + y ** 2 + z ** 2)) * np.ones((n, n, n))).astype(np.complex128)
    S = orc.SobSolver(g, [guess], [V], [5.0 * 2.0], [0.5], stepsize=0.02)
    S.renormalize()
    for _ in range(iters):
        res = S.iterate()
    ref = S.psi[0] * np.sqrt(2.0)             # Save(..., rescale = true) multiplies by sqrt(N)
    assert rel_l2(out.reshape(n, n, n), ref) < 1e-9
    assert res < 1e-18 or iters == 3000


def test_bragg_ds_two_species_pulses_then_free_flight(orc, tmp_path):
    # bragg_ds: four components (two species of two states), pulses of
    # different lengths, then the two-species free step; .bin outputs and
    # both species' Rabi files against the oracle from the same inputs
    n, L, dt = 2048, 80.0, 1e-3
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    rng = np.random.default_rng(8)
    psis = []
    for c, (w, amp) in enumerate([(6.0, 1.0), (6.0, 0.0), (4.0, 1.0), (4.0, 0.0)]):
        p = (amp * np.exp(-x ** 2 / (2 * w ** 2)) + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)))
        p = p.astype(np.complex128)
        p /= np.sqrt(orc.particle_number(g, p)) if amp else 10.0
        psis.append(p)
        write_bin(tmp_path / f"p{c + 1}.bin", p, [L], dt=dt)
    gs = [2.0, 1.9, 0.5, 0.4, 1.9, 2.0, 0.4, 0.5, 0.5, 0.4, 3.0, 2.8, 0.4, 0.5, 2.8, 3.0]
    gsx = "".join(f"<GS_{i + 1}>{','.join(str(v) for v in gs[4 * i:4 * i + 4])}</GS_{i + 1}>" for i in range(4))
    (tmp_path / "ds.xml").write_text(f"""<SIMULATION><DIM>1</DIM>
      <FILENAME>p1.bin</FILENAME><FILENAME_2>p2.bin</FILENAME_2>
      <FILENAME_3>p3.bin</FILENAME_3><FILENAME_4>p4.bin</FILENAME_4>
      <CONSTANTS><laser_k>1.0</laser_k><laser_k_2>0.75</laser_k_2><laser_domh>0.5</laser_domh>
        <laser_domh_2>0.28</laser_domh_2><laser_dk>0.01</laser_dk><laser_dk_2>0.02</laser_dk_2>
        <rabi_threshold>0.3</rabi_threshold></CONSTANTS>
      <VCONSTANTS>{gsx}<Alpha_1>0.5</Alpha_1><Alpha_2>0.35</Alpha_2>
        <Beta>0.01</Beta><Beta_2>-0.02</Beta_2><Delta_L>0,0.3,0.1,-0.2</Delta_L>
        <Amp_1>5,5</Amp_1><Amp_2>3.5,3.5</Amp_2></VCONSTANTS>
      <ALGORITHM><DT>1e-3</DT></ALGORITHM>
      <SEQUENCE>
        <bragg_ad dt="1e-3" Nk="50" rabi_output_freq="each" output_freq="last" pn_freq="last">0.2,0.13025</bragg_ad>
        <freeprop dt="1e-3" Nk="25" output_freq="last">0.05</freeprop>
      </SEQUENCE></SIMULATION>""")
    r = subprocess.run([tool("bragg_ds"), "ds.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # pn_freq defaults to "last": four lines after each sequence
    assert r.stdout.count("N[") == 8 and "FYI: Na          : 4" in r.stdout
    sp2 = dict(beta2=-0.02, amp2=3.5, laser_k2=0.75, laser_domh2=0.28, laser_dk2=0.02, seq_time=0.0,
               duration=(0.2, 0.13025))
    l1 = dict(beta0=0.01, deltaL=[0.0, 0.3, 0.1, -0.2], amp0=5.0, laser_k=1.0, laser_domh=0.5, laser_dk=0.01)
    S = orc.RTSolver(g, psis, gs, [0.5], dt=dt, alpha2=[0.35], species2=sp2, **l1)
    # the noise floor of the reference formula near its Omega = 0 branches
    pert = [p * (1 + 2e-16 * np.random.default_rng(99).standard_normal(n)) for p in psis]
    Sp = orc.RTSolver(g, pert, gs, [0.5], dt=dt, alpha2=[0.35], species2=sp2, **l1)
    rabi1, rabi2 = [], []
    for _ in range(4):
        S.block(orc.BRAGG_DS, 50)
        Sp.block(orc.BRAGG_DS, 50)
        rabi1.append(orc.rabi_integrals(g, S.psis[0], [(0.0,), (2.0,), (-2.0,)], 0.3)[0])
        rabi2.append(orc.rabi_integrals(g, S.psis[2], [(0.0,), (1.5,), (-1.5,)], 0.3)[0])
    tol = max(1e-11, 20 * rel_l2(np.concatenate(Sp.psis), np.concatenate(S.psis)))
    out = np.concatenate([read_bin(tmp_path / f"0.200_{c}.bin")[1] for c in range(1, 5)])
    assert read_bin(tmp_path / "0.200_1.bin")[0]["t"] == pytest.approx(0.2, abs=1e-12)
    assert rel_l2(out, np.concatenate(S.psis)) < tol
    for name, ref, ks in (("Spezie_1_Rabi_1_0.txt", rabi1, "0\t2\t-2"), ("Spezie_2_Rabi_1_0.txt", rabi2, "0\t1.5\t-1.5")):
        lines = (tmp_path / name).read_text().splitlines()
        assert lines[0].startswith("# time \t" + ks), lines[0]
        rows = [[float(v) for v in ln.split("\t") if v] for ln in lines[1:]]
        assert [r_[0] for r_ in rows] == pytest.approx([0.0, 0.05, 0.1, 0.15])
        for r_, rf in zip(rows, ref):
            assert np.allclose(r_[1:], rf, rtol=1e-5, atol=1e-9)
    # the free flight of both species from the end of the pulses
    S.s.seq_time = S.t
    for _ in range(2):
        S.block(orc.FREEPROP_IF2, 25)
    out = np.concatenate([read_bin(tmp_path / f"0.250_{c}.bin")[1] for c in range(1, 5)])
    assert rel_l2(out, np.concatenate(S.psis)) < tol + 1e-11


def _chirp_scan_reference(orc, g, psis, t0, Na, Nk, dt, nch, chirp_min, chirp_max, mode, amp, ifp, k0, thr,
                          output_each=False):
    """The chirp loop of CRT_Base_IF::run_sequence (CRT_Base_IF.h:744-921)
    on the oracle, with a dictionary standing in for the file system: every
    Save_Phi stores the state under its "%.3f_%d.bin" name and the restarts
    read it back from there, as the reference reads its files."""
    files = {}

    def save(S):
        for k in range(2):
            files[f"{S.t:.3f}_{k + 1}"] = S.psis[k].copy()

    def load(t):
        return [files[f"{t:.3f}_{k + 1}"].copy() for k in range(2)]

    dw, dphi = [0.0] * max(nch, 2), [0.0] * max(nch, 2)
    state, t = [p.copy() for p in psis], t0
    files.update({f"{t0:.3f}_{k + 1}": psis[k].copy() for k in range(2)})
    dw[0], dw[1] = chirp_min, chirp_max
    rows, rabi_per_chirp, end = [], [], None
    for s in range(nch):
        phase = s * 2 * np.pi / nch if mode == 1 else 0.0
        S = orc.RTSolver(g, state, [100., 95., 95., 100.], [0.5], dt=dt, t=t, amp=amp, chirp_rate0=dw[s], phase0=phase,
                         **ifp)
        per_block = []
        for _ in range(Na):
            S.block(orc.BRAGG_AD, Nk)
            if output_each and s == 0:
                save(S)
            per_block.append(orc.rabi_integrals(g, S.psis[0], k0, thr)[0])
        rabi_per_chirp.append(per_block)
        if s == 0:
            save(S)
            end = (S.t, [p.copy() for p in S.psis])
        last = list(orc.rabi_integrals(g, S.psis[0], k0, thr)[0])
        if mode == 1:
            rows.append([s, phase] + last)
        else:
            dphi[s] = last[0]
            rows.append([s, dw[s]] + last)
        if s < nch - 1:
            state, t = load(t0), t0
        else:
            state, t = [p.copy() for p in S.psis], S.t
        if 0 < s < nch - 1:
            dw[s + 1] = (dw[s] + dw[s - 1]) / 2
            if dphi[s] - dphi[s - 1] < 0:
                dw[s], dphi[s] = dw[s - 1], dphi[s - 1]
    return rows, rabi_per_chirp, end, load(end[0])


@pytest.mark.parametrize("case", ["bisection", "phase_scan", "restart_name_collision"])
def test_bragg_chirp_scan(orc, tmp_path, case):
    # chirp scans of the bragg driver: Rabi_<seq>_<chirp>.txt per chirp,
    # Chirp_<seq>.txt, restarts of every chirp from the sequence's start
    # (device snapshots) and the end state of chirp 0 handed to the next
    # sequence
    n, L = 1024, 40.0
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    psi1 = np.exp(-x ** 2 / 32).astype(np.complex128)
    psi1 /= np.sqrt(orc.particle_number(g, psi1))
    psis = [psi1, np.zeros_like(psi1)]
    k = 3 * 2 * np.pi / L
    if case == "bisection":
        dt, Nk, dur, nch, cmin, cmax, mode, each = 1e-3, 20, "0.1", 5, -1.0, 1.0, 0, False
    elif case == "phase_scan":
        dt, Nk, dur, nch, cmin, cmax, mode, each = 1e-3, 20, "0.1", 4, 0.5, 0.5, 1, False
    else:
        # blocks of 4e-4: the first block's output "0.000_k.bin" overwrites
        # the start files, and every later chirp restarts from it, as in the
        # reference
        dt, Nk, dur, nch, cmin, cmax, mode, each = 1e-4, 4, "0.0016", 3, -1.0, 1.0, 0, True
    Na = int(float(dur) / dt) // Nk
    for c in range(2):
        write_bin(tmp_path / f"s{c}.bin", psis[c], [L], dt=dt)
    (tmp_path / "c.xml").write_text(f"""<SIMULATION><DIM>1</DIM>
      <FILENAME>s0.bin</FILENAME><FILENAME_2>s1.bin</FILENAME_2>
      <AMP_T>exp(-0.5*((t-0.05)/0.03)^2)</AMP_T>
      <CONSTANTS><laser_k>{k!r}</laser_k><laser_domh>0.40</laser_domh><laser_dk>0</laser_dk>
        <rabi_threshold>0.3</rabi_threshold><chirp>0</chirp></CONSTANTS>
      <VCONSTANTS><GS_1>100,95</GS_1><GS_2>95,100</GS_2><Alpha_1>0.5</Alpha_1>
        <Beta>0</Beta><Delta_L>0,0</Delta_L><Amp_1>20,20</Amp_1></VCONSTANTS>
      <ALGORITHM><DT>{dt!r}</DT></ALGORITHM>
      <SEQUENCE>
        <bragg_ad dt="{dt!r}" Nk="{Nk}" rabi_output_freq="each" output_freq="{'each' if each else 'last'}"
          no_of_chirps="{nch}" chirp_min="{cmin!r}" chirp_max="{cmax!r}" chirp_mode="{mode}">{dur}</bragg_ad>
        <freeprop dt="{dt!r}" Nk="{Nk}" output_freq="last">{float(dur) / 2!r}</freeprop>
      </SEQUENCE></SIMULATION>""")
    r = subprocess.run([tool("bragg"), "c.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    amp = lambda t: float(np.exp(-0.5 * ((t - 0.05) / 0.03) ** 2))
    ifp = dict(amp0=20.0, laser_k=k, laser_domh=0.40, deltaL=[0.0, 0.0])
    k0 = [(0.0,), (2 * k,), (-2 * k,)]
    rows, rabi, (t_end, end_state), restored = _chirp_scan_reference(orc, g, psis, 0.0, Na, Nk, dt, nch, cmin, cmax,
                                                                     mode, amp, ifp, k0, 0.3, output_each=each)
    # Chirp_1.txt: step, chirp rate (or phase), the Rabi populations, 8 digits
    lines = (tmp_path / "Chirp_1.txt").read_text().splitlines()
    assert lines[0].startswith("# Step \t Chirp \t ")
    got = [[float(v) for v in ln.split("\t") if v] for ln in lines[1:]]
    assert len(got) == nch
    for gr, rr in zip(got, rows):
        assert gr[0] == rr[0] and gr[1] == pytest.approx(rr[1], rel=1e-7, abs=1e-12)
        assert np.allclose(gr[2:], rr[2:], rtol=1e-7, atol=1e-10), (gr, rr)
    for s in range(nch):
        body = (tmp_path / f"Rabi_1_{s}.txt").read_text().splitlines()[1:]
        vals = np.array([[float(v) for v in ln.split("\t") if v][1:] for ln in body])
        assert vals.shape == (Na, 3)
        assert np.allclose(vals, np.array(rabi[s]), rtol=1e-5, atol=1e-9), s
    # the next sequence starts from the end of chirp 0, read back from its
    # restart files
    S = orc.RTSolver(g, restored, [100., 95., 95., 100.], [0.5], dt=dt, t=t_end)
    for _ in range(int(float(dur) / 2 / dt) // Nk):
        S.block(orc.FREEPROP_IF, Nk)
    out = [read_bin(tmp_path / f"{S.t:.3f}_{c}.bin")[1] for c in (1, 2)]
    assert rel_l2(np.concatenate(out), np.concatenate(S.psis)) < 1e-11
    if case == "restart_name_collision":
        # chirp 1 really restarted from the overwritten file: its first Rabi
        # row differs from chirp 0's
        assert not np.allclose(rabi[1][0], rabi[0][0], rtol=1e-3)
    if case == "phase_scan":
        # the phase scan ran as one batch; the same scan one chirp after
        # another prints the same lines and writes the same files
        seq_dir = tmp_path / "sequential"
        seq_dir.mkdir()
        for f in ("c.xml", "s0.bin", "s1.bin"):
            (seq_dir / f).write_bytes((tmp_path / f).read_bytes())
        env = dict(os.environ, ATUS_CHIRP_BATCH="0")
        r2 = subprocess.run([tool("bragg"), "c.xml"], cwd=seq_dir, capture_output=True, text=True, timeout=120, env=env)
        assert r2.returncode == 0, r2.stdout + r2.stderr
        l1, l2 = r.stdout.splitlines(), r2.stdout.splitlines()
        assert len(l1) == len(l2)
        for a, b in zip(l1, l2):
            if a.startswith("N["):
                assert float(a.split("=")[1]) == pytest.approx(float(b.split("=")[1]), rel=1e-5)
            else:
                assert a == b
        names = sorted(p.name for p in tmp_path.iterdir() if p.is_file())
        assert names == sorted(p.name for p in seq_dir.iterdir() if p.is_file())
        for name in names:
            if name.endswith(".txt"):
                A = [[float(v) for v in ln.split("\t") if v] for ln in (tmp_path / name).read_text().splitlines()[1:]]
                B = [[float(v) for v in ln.split("\t") if v] for ln in (seq_dir / name).read_text().splitlines()[1:]]
                assert np.allclose(A, B, rtol=1e-6, atol=1e-12), name
            elif name.endswith(".bin") and not name.startswith("s"):
                a, b = read_bin(tmp_path / name)[1], read_bin(seq_dir / name)[1]
                assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), name


def _three_state_inputs(orc, tmp_path, n, L, ncomp, dt):
    g = orc.make_grid((n,), L=L)
    (x,) = orc.coords_x(g)
    p = np.exp(-x ** 2 / 18).astype(np.complex128)
    p /= np.sqrt(orc.particle_number(g, p))
    rng = np.random.default_rng(21)
    psis = [p] + [1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)) for _ in range(ncomp - 1)]
    for c in range(ncomp):
        write_bin(tmp_path / f"q{c + 1}.bin", psis[c], [L], dt=dt)
    files = "".join(f"<FILENAME{'' if c == 0 else '_' + str(c + 1)}>q{c + 1}.bin</FILENAME{'' if c == 0 else '_' + str(c + 1)}>"
                    for c in range(ncomp))
    return g, psis, files


@pytest.mark.parametrize("driver", ["double_bragg", "raman", "bragg_numerical"])
def test_three_state_drivers(orc, tmp_path, driver):
    # double_bragg ("bragg_ad": closed-form 3x3 step), raman ("raman") and
    # bragg with the numerical two-state step ("bragg") from XML and .bin
    # inputs, against the oracle
    n, L, dt, Nk = 1024, 40.0, 1e-3, 25
    ncomp = 2 if driver == "bragg_numerical" else 3
    g, psis, files = _three_state_inputs(orc, tmp_path, n, L, ncomp, dt)
    gs = [20.0, 19.0, 18.0, 19.0, 21.0, 18.5, 18.0, 18.5, 20.5][:ncomp * ncomp] if ncomp == 3 else [20.0, 19.0, 19.0, 21.0]
    gsx = "".join(f"<GS_{i + 1}>{','.join(str(v) for v in gs[ncomp * i:ncomp * (i + 1)])}</GS_{i + 1}>"
                  for i in range(ncomp))
    k = 4 * 2 * np.pi / L
    dl = {"double_bragg": [0.0, -0.5, 0.0], "raman": [0.0, -3.0, 2.0], "bragg_numerical": [0.0, -0.5]}[driver]
    domh = {"double_bragg": 0.8, "raman": 5.0, "bragg_numerical": 0.8}[driver]
    seq = {"double_bragg": "bragg_ad", "raman": "raman", "bragg_numerical": "bragg"}[driver]
    amp_t = "<AMP_T>exp(-0.5*((t-0.05)/0.02)^2)</AMP_T>" if driver == "double_bragg" else ""
    (tmp_path / "t.xml").write_text(f"""<SIMULATION><DIM>1</DIM>{files}{amp_t}
      <CONSTANTS><laser_k>{k!r}</laser_k><laser_domh>{domh}</laser_domh><laser_dk>0.05</laser_dk>
        <rabi_threshold>0.3</rabi_threshold><chirp>0</chirp></CONSTANTS>
      <VCONSTANTS>{gsx}<Alpha_1>0.5</Alpha_1><Beta>0.02</Beta>
        <Delta_L>{','.join(str(v) for v in dl)}</Delta_L><Amp_1>8,5</Amp_1></VCONSTANTS>
      <ALGORITHM><DT>{dt!r}</DT></ALGORITHM>
      <SEQUENCE><{seq} dt="{dt!r}" Nk="{Nk}" rabi_output_freq="each" output_freq="last">0.1</{seq}></SEQUENCE>
      </SIMULATION>""")
    exe = "bragg" if driver == "bragg_numerical" else driver
    r = subprocess.run([tool(exe), "t.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    Na = int(0.1 / dt) // Nk
    okind = {"double_bragg": orc.DOUBLE_BRAGG, "raman": orc.RAMAN, "bragg_numerical": orc.NUM_BRAGG}[driver]
    amp = (lambda t: float(np.exp(-0.5 * ((t - 0.05) / 0.02) ** 2))) if driver == "double_bragg" else None
    lp = dict(amp0=8.0, amp1=5.0, laser_k=k, laser_domh=domh, laser_dk=0.05, beta0=0.02, beta=[0.02], deltaL=dl)
    S = orc.RTSolver(g, psis, gs, [0.5], dt=dt, amp=amp, **lp)
    pert = [p * (1 + 2e-16 * np.random.default_rng(99).standard_normal(n)) for p in psis]
    Sp = orc.RTSolver(g, pert, gs, [0.5], dt=dt, amp=amp, **lp)
    rabi = []
    for _ in range(Na):
        S.block(okind, Nk)
        Sp.block(okind, Nk)
        rabi.append(orc.rabi_integrals(g, S.psis[0], [(0.0,), (2 * k,), (-2 * k,)], 0.3)[0])
    tol = max(1e-11, 20 * rel_l2(np.concatenate(Sp.psis), np.concatenate(S.psis)))
    out = np.concatenate([read_bin(tmp_path / f"{S.t:.3f}_{c}.bin")[1] for c in range(1, ncomp + 1)])
    assert rel_l2(out, np.concatenate(S.psis)) < tol
    body = (tmp_path / "Rabi_1_0.txt").read_text().splitlines()[1:]
    vals = np.array([[float(v) for v in ln.split("\t") if v][1:] for ln in body])
    assert np.allclose(vals, np.array(rabi), rtol=1e-5, atol=1e-9)
    if driver == "double_bragg":
        # one amplitude line per V step (6 digits): AMP_T at t1 = Get_t() +
        # dt/2, which is m*dt for the V step m after the half kinetic step
        lines = [ln for ln in r.stdout.splitlines()
                 if ln and not ln.startswith(("FYI", "t =", "N[")) and "Amp" not in ln]
        assert len(lines) == Na * Nk
        t1 = dt * (1 + np.arange(Na * Nk))
        assert np.allclose([float(v) for v in lines], np.exp(-0.5 * ((t1 - 0.05) / 0.02) ** 2), rtol=1e-5, atol=1e-7)
