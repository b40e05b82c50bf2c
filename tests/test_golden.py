"""The oracle against the golden fixtures: an independent numpy restatement
of the sequence block (numpy FFT, rotation-form 2x2 exponential) pins the
oracle's block structure, time bookkeeping, kinetic phase and V steps."""
import numpy as np
import pytest

from golden_cases import NAMES, load, rel_l2


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_golden(orc, name):
    c = load(name)
    g = orc.make_grid(c["shape"], L=c["L"])
    kind = {"freeprop": orc.FREEPROP, "freeprop_if": orc.FREEPROP_IF, "bragg_ad": orc.BRAGG_AD}[c["kind"]]
    S = orc.RTSolver(g, c["psi0"], c["gs"], c["alpha"], dt=c["dt"], t=c["t0"], pots=c["pots"], amp=c["amp"],
                     **dict(c["ifp"]))
    for b in range(c["nblocks"]):
        S.block(kind, c["Nk"])
        norms = [orc.particle_number(g, p) for p in S.psis]
        assert np.allclose(norms, c["norms"][b], rtol=0, atol=1e-13)
    assert S.t == pytest.approx(c["t"], abs=1e-15)
    assert rel_l2(S.psis, c["out"]) < 1e-12


def test_golden_fixtures_are_reproducible(tmp_path):
    # the fixtures are what make_golden.py writes (deterministic)
    import importlib.util
    import os
    import shutil
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("mg", os.path.join(here, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    mg.HERE = str(tmp_path)
    mg.main()
    for n in NAMES:
        a = np.load(os.path.join(here, f"{n}.npz"))
        b = np.load(os.path.join(str(tmp_path), f"{n}.npz"))
        assert np.array_equal(a["out"], b["out"]) and np.array_equal(a["psi0"], b["psi0"])
