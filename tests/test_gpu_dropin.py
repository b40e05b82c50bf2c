"""The reference's own drivers, built unchanged against this project's host
headers and libraries (oracle/_ref/: rt_solver.cpp, bragg.cpp and
sob_solver.cpp of the reference compiled by oracle/Makefile's `ref`
target), run on the GPU beside the rebuilt drivers of atus2_amd/bin from the
same XML and .bin files.

* rt_solver and sobmin: the reference's drivers only set up the potential
  and the guess on the host; the steps are this project's device steps, so
  the outputs must be identical bit for bit.
* bragg: the reference's Do_Bragg_ad is host code on Getp2In() (the host
  mirror of each field, downloaded and uploaded around every V step); the
  rebuilt driver runs the device step with the same operation order.  The
  results agree to rounding.

Skipped where oracle/_ref/ was not built (no reference sources there)."""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_drivers import write_bin
from test_host import read_bin, tool

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def ref_tool(name):
    p = os.path.join(REF_BIN, name)
    if not os.path.exists(p):
        pytest.skip(f"{p} not built (oracle/Makefile target `ref` needs the reference sources)")
    return p


def _both(name, xml, files, tmp_path, ref_env=None, tag="ref"):
    """Runs the reference's driver (with ref_env added to its environment)
    and the rebuilt one, each in its own directory holding the same inputs;
    returns the two directories."""
    dirs = []
    for kind, exe, extra in ((tag, ref_tool(name), ref_env or {}), ("ours", tool(name), {})):
        d = tmp_path / kind
        if d.exists():
            dirs.append(d)
            continue
        d.mkdir()
        for fname, writer in files.items():
            writer(d / fname)
        (d / "run.xml").write_text(xml)
        r = subprocess.run([exe, "run.xml"], cwd=d, capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, **extra))
        assert r.returncode == 0, (kind, r.stdout[-3000:], r.stderr[-3000:])
        (d / "stdout.txt").write_text(r.stdout)
        dirs.append(d)
    return dirs


@pytest.mark.parametrize("dim", [1, 2])
def test_reference_rt_solver_runs_unchanged(tmp_path, dim):
    n = 4096 if dim == 1 else 128
    L = 40.0
    x = (np.arange(n) - n // 2) * (L / n)
    psi = np.exp(-x ** 2 / 2) if dim == 1 else np.exp(-(x[:, None] + 3) ** 2 / 2 - x[None, :] ** 2 / 3)
    psi = (psi / np.sqrt(np.sum(psi ** 2) * (L / n) ** dim)).astype(np.complex128)
    pot = "0.5*x^2" if dim == 1 else "0.5*(x^2+2*y^2)+10*rect(x,-0.5,0.5)"
    xml = f"""<SIMULATION><DIM>{dim}</DIM><FILENAME>psi0.bin</FILENAME>
      <POTENTIAL_{dim}D>{pot}</POTENTIAL_{dim}D>
      <VCONSTANTS><GS_1>10</GS_1><Alpha_1>{",".join(["0.5"] * dim)}</Alpha_1></VCONSTANTS>
      <ALGORITHM><DT>0.001</DT></ALGORITHM>
      <SEQUENCE><set_momentum comp="0">{",".join(["1"] * (dim + 1))}</set_momentum>
        <freeprop dt="0.001" Nk="50" pn_freq="each" output_freq="last">0.2</freeprop></SEQUENCE>
      </SIMULATION>"""
    ref, ours = _both("rt_solver", xml, {"psi0.bin": lambda p: write_bin(p, psi, [L] * dim)}, tmp_path)
    a, b = (ref / "0.200_1.bin").read_bytes(), (ours / "0.200_1.bin").read_bytes()
    assert a == b
    na = [ln for ln in (ref / "stdout.txt").read_text().splitlines() if ln.startswith("N[")]
    nb = [ln for ln in (ours / "stdout.txt").read_text().splitlines() if ln.startswith("N[")]
    assert len(na) == 4 and na == nb


def test_reference_sobmin_runs_unchanged(tmp_path):
    n = 64
    xml = f"""<SIMULATION><DIM>2</DIM><FILENAME>gs.bin</FILENAME>
      <GUESS_2D>exp(-0.25*(x^2+y^2))</GUESS_2D>
      <POTENTIAL_2D>0.5*(x^2+1.5^2*y^2)</POTENTIAL_2D>
      <CONSTANTS><epsilon>1e-12</epsilon><stepsize>0.02</stepsize></CONSTANTS>
      <VCONSTANTS><Alpha_1>0.5</Alpha_1><GS_1>5</GS_1><N>2</N></VCONSTANTS>
      <ALGORITHM><NX>{n}</NX><NY>{n}</NY><XMIN>-8</XMIN><XMAX>8</XMAX><YMIN>-8</YMIN><YMAX>8</YMAX></ALGORITHM>
      </SIMULATION>"""
    ref, ours = _both("sobmin", xml, {}, tmp_path)
    assert (ref / "gs.bin").read_bytes() == (ours / "gs.bin").read_bytes()
    assert (ref / "zero.bin").read_bytes() == (ours / "zero.bin").read_bytes()
    ra = [ln for ln in (ref / "stdout.txt").read_text().splitlines() if ln.startswith("res ")]
    rb = [ln for ln in (ours / "stdout.txt").read_text().splitlines() if ln.startswith("res ")]
    assert len(ra) > 10 and ra == rb
    assert float(ra[-1].split()[1]) <= 1e-12


def test_reference_bragg_runs_unchanged_with_its_host_step(tmp_path):
    n, L = 64, 32.0
    x = (np.arange(n) - n // 2) * (L / n)
    psi1 = (np.exp(-x[:, None] ** 2 / 18 - x[None, :] ** 2 / 8)).astype(np.complex128)
    psi1 /= np.sqrt(np.sum(np.abs(psi1) ** 2) * (L / n) ** 2)
    xml = """<SIMULATION><DIM>2</DIM><FILENAME>a.bin</FILENAME><FILENAME_2>b.bin</FILENAME_2>
      <AMP_T>exp(-0.5*((t-0.03)/0.02)^2)</AMP_T>
      <CONSTANTS><laser_k>4</laser_k><laser_domh>8</laser_domh><laser_dk>0</laser_dk>
        <rabi_threshold>2</rabi_threshold><chirp>0</chirp></CONSTANTS>
      <VCONSTANTS><GS_1>100,95</GS_1><GS_2>95,100</GS_2><Alpha_1>0.5,0.5</Alpha_1>
        <Beta>0,0</Beta><Delta_L>0,0</Delta_L><Amp_1>5,5</Amp_1></VCONSTANTS>
      <ALGORITHM><DT>2e-4</DT></ALGORITHM>
      <SEQUENCE><bragg_ad dt="2e-4" Nk="50" rabi_output_freq="each" output_freq="last">0.04</bragg_ad></SEQUENCE>
      </SIMULATION>"""
    files = {"a.bin": lambda p: write_bin(p, psi1, [L, L], dt=2e-4),
             "b.bin": lambda p: write_bin(p, np.zeros_like(psi1), [L, L], dt=2e-4)}
    ref, ours = _both("bragg", xml, files, tmp_path)
    a = np.concatenate([read_bin(ref / f"0.040_{c}.bin")[1] for c in (1, 2)])
    b = np.concatenate([read_bin(ours / f"0.040_{c}.bin")[1] for c in (1, 2)])
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"reference bragg (host step) vs rebuilt bragg (device step): rel L2 {err:.2e}")
    assert err < 1e-12
    assert np.linalg.norm(b[n * n:]) > 1e-3      # the pulse moved population
    ra = (ref / "Rabi_1_0.txt").read_text().splitlines()
    rb = (ours / "Rabi_1_0.txt").read_text().splitlines()
    assert len(ra) == 5 and ra[0] == rb[0]
    for la, lb in zip(ra[1:], rb[1:]):
        assert np.allclose([float(v) for v in la.split()], [float(v) for v in lb.split()], rtol=1e-5, atol=1e-9)
    # ATUS_DEVICE_STEPS=bragg_ad: the same unchanged driver takes the
    # device step, and matches the rebuilt driver bit for bit
    dev, ours = _both("bragg", xml, files, tmp_path, {"ATUS_DEVICE_STEPS": "bragg_ad"}, tag="ref_device")
    assert "bragg_ad runs as the device step" in (dev / "stdout.txt").read_text()
    for c in (1, 2):
        assert (dev / f"0.040_{c}.bin").read_bytes() == (ours / f"0.040_{c}.bin").read_bytes()
    assert (dev / "Rabi_1_0.txt").read_text() == (ours / "Rabi_1_0.txt").read_text()
