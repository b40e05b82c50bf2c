"""Failure detection and resume (the reference's error paths,
CRT_Base.h:207-240, 724-725, 760-768, CRT_shared.h:131-135), through the
C-ABI and through the rebuilt drivers: bad inputs end with the reference's
message and a non-zero exit status, a state with NaN in it is reported as
a numeric error, and a run resumed from a mid-run .bin (whose header
carries t) ends bit-identical to the run without the stop."""
import subprocess

import numpy as np
import pytest

from test_gpu_drivers import write_bin
from test_host import read_bin, tool

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def test_nan_in_psi_is_a_numeric_error(hip, orc):
    g = orc.make_grid((64, 32), L=[20.0, 10.0])
    psi = np.exp(-sum(x ** 2 for x in orc.coords_x(g))) * np.ones((64, 32)) + 0j
    psi[5, 7] = np.nan
    with hip.Context.from_grid(g, 1) as ctx:
        ctx.upload(0, psi)
        with pytest.raises(hip.AtusError, match=r"\(-6\).*not finite"):
            ctx.particle_number(0)
        # the NaN spreads through a block; the next norm check reports it
        ctx.set_g([1.0])
        ctx.set_alpha([0.5, 0.5])
        ctx.run_block(hip.STEP_FREEPROP_IF, 1e-3, 3, 0.0, hip.laser())
        with pytest.raises(hip.AtusError, match="not finite"):
            ctx.particle_number(0)
        # the Sobolev loop stops with the same code
        ctx.set_potential(0, np.zeros((64, 32)))
        ctx.sob_setup([0.5], 1e-3)
        with pytest.raises(hip.AtusError, match=r"\(-6\)"):
            ctx.sob_iterate(4)


def _rt_xml(seq, filename="psi0.bin"):
    return f"""<SIMULATION><DIM>1</DIM><FILENAME>{filename}</FILENAME>
      <POTENTIAL_1D>0.5*x^2</POTENTIAL_1D>
      <VCONSTANTS><GS_1>10</GS_1><Alpha_1>0.5</Alpha_1></VCONSTANTS>
      <ALGORITHM><DT>0.001</DT></ALGORITHM>
      <SEQUENCE>{seq}</SEQUENCE></SIMULATION>"""


def _run(drv, xml, cwd):
    return subprocess.run([tool(drv), xml], cwd=cwd, capture_output=True, text=True, timeout=120)


def test_driver_failures_exit_nonzero_with_the_reference_messages(tmp_path):
    n, L = 256, 20.0
    x = (np.arange(n) - n // 2) * (L / n)
    psi = (np.pi ** -0.25 * np.exp(-x ** 2 / 2)).astype(np.complex128)
    write_bin(tmp_path / "psi0.bin", psi, [L])
    free = '<freeprop dt="0.001" Nk="10" pn_freq="each">0.02</freeprop>'
    # an unknown sequence name (CRT_Base.h:760-768)
    (tmp_path / "a.xml").write_text(_rt_xml('<warp dt="0.001" Nk="10">0.02</warp>'))
    r = _run("rt_solver", "a.xml", tmp_path)
    assert r.returncode != 0 and "Invalid squence name warp" in r.stdout, r.stdout + r.stderr
    # set_momentum needs more than dim comma-separated values (CRT_Base.h:724)
    (tmp_path / "b.xml").write_text(_rt_xml('<set_momentum comp="0">3</set_momentum>' + free))
    r = _run("rt_solver", "b.xml", tmp_path)
    assert r.returncode != 0 and "set_momentum needs more than dim" in r.stdout, r.stdout + r.stderr
    (tmp_path / "b2.xml").write_text(_rt_xml('<set_momentum comp="0">3,0</set_momentum>' + free))
    r = _run("rt_solver", "b2.xml", tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    # a file shorter than its header announces (LoadFiles, CRT_Base.h:207-240)
    raw = (tmp_path / "psi0.bin").read_bytes()
    (tmp_path / "short.bin").write_bytes(raw[:-16])
    (tmp_path / "c.xml").write_text(_rt_xml(free, "short.bin"))
    r = _run("rt_solver", "c.xml", tmp_path)
    assert r.returncode != 0 and "fewer points than its header" in r.stdout, r.stdout + r.stderr
    # NaN in the state: the first norm check ends the run
    bad = psi.copy()
    bad[100] = np.nan
    write_bin(tmp_path / "nan.bin", bad, [L])
    (tmp_path / "d.xml").write_text(_rt_xml(free, "nan.bin"))
    r = _run("rt_solver", "d.xml", tmp_path)
    assert r.returncode != 0 and "not finite" in r.stdout, r.stdout + r.stderr


BRAGG_XML = """<SIMULATION><DIM>2</DIM>
  <FILENAME>{f1}</FILENAME><FILENAME_2>{f2}</FILENAME_2>
  <AMP_T>exp(-0.5*((t-0.03)/0.02)^2)</AMP_T>
  <CONSTANTS><laser_k>4</laser_k><laser_domh>8</laser_domh><laser_dk>0</laser_dk>
    <rabi_threshold>2</rabi_threshold><chirp>0</chirp></CONSTANTS>
  <VCONSTANTS><GS_1>100,95</GS_1><GS_2>95,100</GS_2><Alpha_1>0.5,0.5</Alpha_1>
    <Beta>0,0</Beta><Delta_L>0,0</Delta_L><Amp_1>5,5</Amp_1></VCONSTANTS>
  <ALGORITHM><DT>2e-4</DT></ALGORITHM>
  <SEQUENCE><bragg_ad dt="2e-4" Nk="50" rabi_output_freq="each" output_freq="{of}">{dur}</bragg_ad></SEQUENCE>
  </SIMULATION>"""


def test_bragg_resumed_from_a_mid_run_bin_is_bit_identical(tmp_path):
    n, L = 64, 32.0
    x = (np.arange(n) - n // 2) * (L / n)
    psi1 = (np.exp(-x[:, None] ** 2 / 18 - x[None, :] ** 2 / 8)).astype(np.complex128)
    psi1 /= np.sqrt(np.sum(np.abs(psi1) ** 2) * (L / n) ** 2)
    whole, part = tmp_path / "whole", tmp_path / "part"
    for d in (whole, part):
        d.mkdir()
        write_bin(d / "a.bin", psi1, [L, L], dt=2e-4)
        write_bin(d / "b.bin", np.zeros_like(psi1), [L, L], dt=2e-4)
    # 4 blocks of 50 steps without a stop (t = 0.04)
    (whole / "w.xml").write_text(BRAGG_XML.format(f1="a.bin", f2="b.bin", of="last", dur="0.04"))
    r = _run("bragg", "w.xml", whole)
    assert r.returncode == 0, r.stdout + r.stderr
    # 2 blocks with a .bin after each, then the remaining 2 blocks from the
    # files of t = 0.02 (the run takes t from their header)
    (part / "p1.xml").write_text(BRAGG_XML.format(f1="a.bin", f2="b.bin", of="each", dur="0.02"))
    r = _run("bragg", "p1.xml", part)
    assert r.returncode == 0, r.stdout + r.stderr
    h, _ = read_bin(part / "0.020_1.bin")
    assert abs(h["t"] - 0.02) < 1e-15
    (part / "p2.xml").write_text(BRAGG_XML.format(f1="0.020_1.bin", f2="0.020_2.bin", of="last", dur="0.02"))
    r = _run("bragg", "p2.xml", part)
    assert r.returncode == 0, r.stdout + r.stderr
    for c in (1, 2):
        hw, a = read_bin(whole / f"0.040_{c}.bin")
        hp, b = read_bin(part / f"0.040_{c}.bin")
        assert hw["t"] == hp["t"]
        assert np.array_equal(a, b), c
