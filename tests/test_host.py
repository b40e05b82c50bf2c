"""Host side of the drop-in (CPU only): the XML reader / ParameterHandler
defaults, the formula evaluator, the .bin header and gen_psi_0."""
import json
import math
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "atus2_amd", "bin")


def tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} not built (run __graft_entry__.build())")
    return p


XML = """<?xml version="1.0"?>
<SIMULATION>
  <DIM>3</DIM>
  <FILENAME>a.bin</FILENAME>
  <AMP_T>exp(-0.5*((t-0.5)/0.5)^2)</AMP_T>
  <CONSTANTS>
    <V0>1e3</V0>   <!-- wall height -->
    <laser_k>4</laser_k>
    <V0>7</V0>
    <w>2.5</w>
  </CONSTANTS>
  <VCONSTANTS>
    <GS_1>100, 95</GS_1>
    <Alpha_1>0.5,0.5,0.5</Alpha_1>
  </VCONSTANTS>
  <ALGORITHM><NX>256</NX><XMIN>-32</XMIN><XMAX>32</XMAX><DT>2e-4</DT></ALGORITHM>
  <SEQUENCE>
    <set_momentum comp="1">3,0,0,0</set_momentum>
    <bragg_ad dt="2e-4" Nk="50" rabi_output_freq="each" output_freq="last">1.0</bragg_ad>
    <freeprop output_freq="packed" pn_freq="bogus" no_of_chirps="3" chirp_min="-1" chirp_max="2">0.4,0.6</freeprop>
  </SEQUENCE>
</SIMULATION>
"""


@pytest.fixture(scope="module")
def params_xml(tmp_path_factory):
    p = tmp_path_factory.mktemp("xml") / "p.xml"
    p.write_text(XML)
    return str(p)


def parse(path):
    out = subprocess.run([tool("atus_params"), path], capture_output=True, text=True, check=True).stdout
    return json.loads(out)


def evaluate(path, formula, **v):
    args = [tool("atus_params"), path, "--eval", formula] + [f"{k}={val!r}" for k, val in v.items()]
    r = subprocess.run(args, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout)
    return float(r.stdout.strip())


def test_parameter_sections_and_defaults(params_xml):
    d = parse(params_xml)
    # ParameterHandler.cpp:183-210: std::map::insert keeps the first definition
    assert d["constants"] == {"V0": 1000.0, "laser_k": 4.0, "w": 2.5}
    assert d["vconstants"]["GS_1"] == [100.0, 95.0]
    assert d["vconstants"]["Alpha_1"] == [0.5, 0.5, 0.5]
    assert d["simulation"]["DIM"] == "3" and "CONSTANTS" not in d["simulation"]
    assert d["getters"]["dt"] == 2e-4 and d["getters"]["NX"] == 256
    assert d["getters"]["NY"] == 256 and d["getters"]["epsilon"] == 1e-5  # defaults :359-486
    s0, s1, s2 = d["sequence"]
    assert s0["name"] == "set_momentum" and s0["comp"] == 1 and s0["duration"] == [3, 0, 0, 0]
    # sequence defaults (ParameterHandler.cpp:68-86): dt 0.001, Nk 100, pn_freq last
    assert s0["dt"] == 0.001 and s0["Nk"] == 100 and s0["pn_freq"] == 2
    assert s1["Nk"] == 50 and s1["rabi_output_freq"] == 1 and s1["output_freq"] == 2
    # unknown freq names fall back to none; packed = 3
    assert s2["pn_freq"] == 0 and s2["output_freq"] == 3
    assert s2["duration"] == [0.4, 0.6] and s2["no_of_chirps"] == 3
    assert s2["chirp_min"] == -1 and s2["chirp_max"] == 2


@pytest.mark.parametrize("formula,vars_,expect", [
    ("-x^2", dict(x=3.0), -9.0),
    ("2^-1", {}, 0.5),
    ("2^3^2", {}, 512.0),                       # right associative
    ("-2^2", {}, -4.0),
    ("1+2*3-4/8", {}, 6.5),
    ("(1+2)*3", {}, 9.0),
    ("V0*x + laser_k", dict(x=0.5), 504.0),      # first V0 definition
    ("_pi", {}, math.pi),
    ("_e", {}, math.e),
    ("sqrt(abs(-16)) + exp(0) + log(_e) + ln(1)", {}, 6.0),
    ("sin(0.3)^2 + cos(0.3)^2", {}, 1.0),
    ("min(3, x, 2) + max(1, y)", dict(x=-1.0, y=4.0), 3.0),
    ("x < 1 ? 10 : 20", dict(x=0.5), 10.0),
    ("x >= 1 && y != 0", dict(x=1.0, y=2.0), 1.0),
    ("exp(-0.5*((t-0.5)/0.5)^2)", dict(t=0.75), math.exp(-0.125)),
    # misc.cpp:40-53: sign = copysign(1, x), Heaviside = (1+sign)/2,
    # rect(x, a, b) = H(x-a) - H(x-b): 1 on [a, b)
    ("sign(0)", {}, 1.0),
    ("sign(-0)", {}, -1.0),
    ("Heaviside(0)", {}, 1.0),
    ("rect(x,-0.5,0.5)", dict(x=-0.5), 1.0),
    ("rect(x,-0.5,0.5)", dict(x=0.5), 0.0),
    ("rect(x,-0.5,0.5)", dict(x=0.25), 1.0),
    ("w*(1-rect(y,-5,-3)-rect(y,3,5))", dict(y=4.0), 0.0),
])
def test_formula_semantics(params_xml, formula, vars_, expect):
    assert evaluate(params_xml, formula, **vars_) == pytest.approx(expect, rel=1e-15, abs=1e-15)


def test_formula_errors_are_reported(params_xml):
    with pytest.raises(RuntimeError, match="unknown token"):
        evaluate(params_xml, "nope*2")
    with pytest.raises(RuntimeError, match="position"):
        evaluate(params_xml, "(1+2")


def test_malformed_xml_fails_loudly(tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text("<SIMULATION><DIM>1</DIM><CONSTANTS></SIMULATION>")
    r = subprocess.run([tool("atus_params"), str(p)], capture_output=True, text=True)
    assert r.returncode != 0 and "XML parse error" in r.stdout


def read_bin(path):
    raw = open(path, "rb").read()
    # my_structs.h:40-72, pack(4): 7 int64, 2 int32, 14 double, 2 int32, ...
    head = struct.unpack_from("<7q2i14d2i", raw, 0)
    names = ["nself", "nDatatyp", "nself_and_data", "nDims", "nDimX", "nDimY", "nDimZ", "bAtom", "bComplex",
             "t", "xMin", "xMax", "yMin", "yMax", "zMin", "zMax", "dx", "dy", "dz", "dkx", "dky", "dkz", "dt",
             "ks", "fs"]
    h = dict(zip(names, head))
    data = np.frombuffer(raw[1380:], dtype=np.complex128)
    return h, data


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_gen_psi_0_header_and_normalisation(tmp_path, dim):
    n = {1: 512, 2: 64, 3: 16}[dim]
    guess = {1: "exp(-0.5*x^2)", 2: "exp(-0.5*(x^2+2*y^2))", 3: "exp(-0.5*(x^2+y^2+z^2))*(1+x)"}[dim]
    xml = f"""<SIMULATION><DIM>{dim}</DIM><FILENAME>{tmp_path}/psi.bin</FILENAME>
      <GUESS_{dim}D>{guess}</GUESS_{dim}D><CONSTANTS><N>2.5</N></CONSTANTS>
      <ALGORITHM><NX>{n}</NX><NY>{n}</NY><NZ>{n}</NZ><XMIN>-8</XMIN><XMAX>8</XMAX>
      <YMIN>-6</YMIN><YMAX>6</YMAX><ZMIN>-4</ZMIN><ZMAX>4</ZMAX></ALGORITHM></SIMULATION>"""
    (tmp_path / "g.xml").write_text(xml)
    subprocess.run([tool("gen_psi_0"), str(tmp_path / "g.xml")], check=True, capture_output=True)
    h, data = read_bin(tmp_path / "psi.bin")
    assert h["nself"] == 1380 and h["nDatatyp"] == 16 and h["nDims"] == dim
    assert h["bComplex"] == 1 and h["dt"] == 0.001 and h["t"] == 0.0
    L = [16.0, 12.0, 8.0]
    shape = [n] * dim
    assert [h["nDimX"], h["nDimY"], h["nDimZ"]] == shape + [1] * (3 - dim)
    for a, key in enumerate(["dx", "dy", "dz"][:dim]):
        assert h[key] == L[a] / n and h["dk" + key[1]] == 2 * math.pi / L[a]
    psi = data.reshape(shape)
    xs = [(np.arange(n) - n // 2) * (L[a] / n) for a in range(dim)]
    grids = np.meshgrid(*xs, indexing="ij")
    if dim == 1:
        ref = np.exp(-0.5 * grids[0] ** 2)
    elif dim == 2:
        ref = np.exp(-0.5 * (grids[0] ** 2 + 2 * grids[1] ** 2))
    else:
        ref = np.exp(-0.5 * (grids[0] ** 2 + grids[1] ** 2 + grids[2] ** 2)) * (1 + grids[0])
    ar = np.prod([L[a] / n for a in range(dim)])
    ref = ref * np.sqrt(2.5 / (ar * np.sum(ref ** 2)))
    assert np.abs(psi.real - ref).max() < 1e-14 and np.all(psi.imag == 0)


def test_gen_psi_0_seeded_noise_is_reproducible(tmp_path):
    xml = f"""<SIMULATION><DIM>1</DIM><FILENAME>{tmp_path}/psi.bin</FILENAME>
      <GUESS_1D>exp(-0.5*x^2)</GUESS_1D><CONSTANTS><N>1</N></CONSTANTS>
      <ALGORITHM><NX>256</NX><XMIN>-10</XMIN><XMAX>10</XMAX></ALGORITHM></SIMULATION>"""
    (tmp_path / "g.xml").write_text(xml)
    env = dict(os.environ, GEN_PSI_NOISE="1e-3", GEN_PSI_SEED="1")
    subprocess.run([tool("gen_psi_0"), str(tmp_path / "g.xml")], check=True, env=env, capture_output=True)
    a = read_bin(tmp_path / "psi.bin")[1].copy()
    subprocess.run([tool("gen_psi_0"), str(tmp_path / "g.xml")], check=True, env=env, capture_output=True)
    b = read_bin(tmp_path / "psi.bin")[1]
    assert np.array_equal(a, b) and np.abs(a.imag).max() > 1e-4
    assert abs(np.sum(np.abs(a) ** 2) * 20 / 256 - 1) < 1e-14


def test_driver_errors_exit_nonzero(tmp_path):
    # a missing FILENAME fails with the reference's message (CRT_shared.h:135)
    # and a non-zero status (the reference returns 0)
    xml = """<SIMULATION><DIM>1</DIM><FILENAME>does_not_exist.bin</FILENAME>
      <VCONSTANTS><GS_1>1</GS_1><Alpha_1>0.5</Alpha_1></VCONSTANTS></SIMULATION>"""
    (tmp_path / "e.xml").write_text(xml)
    for drv in ("rt_solver", "bragg", "bragg_ds", "double_bragg", "raman"):
        r = subprocess.run([tool(drv), str(tmp_path / "e.xml")], capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode != 0 and "Could not open file does_not_exist.bin" in r.stdout, drv
        r = subprocess.run([tool(drv)], capture_output=True, text=True)
        assert r.returncode != 0 and "No parameter xml file specified" in r.stdout, drv


def test_bin_tools_argument_errors(tmp_path):
    # inflate_domain: no inflation option ("Hmmm", inflate_domain.cpp:266-270),
    # no file name, a missing file; ana_expval: a missing file prints and
    # exits 0, no argument prints nothing (ana_expval.cpp:236-258)
    r = subprocess.run([tool("inflate_domain"), "a.bin"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "Hmmm" in r.stdout
    r = subprocess.run([tool("inflate_domain"), "--fak", "2"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "missing file name" in r.stdout
    r = subprocess.run([tool("inflate_domain"), "--fak=2", "nope.bin"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "Cannot open file: nope.bin" in r.stdout
    r = subprocess.run([tool("ana_expval"), "nope.bin"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and "Could not open file nope.bin." in r.stdout
    r = subprocess.run([tool("ana_expval")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and r.stdout == ""
