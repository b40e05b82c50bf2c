"""Loads the golden fixtures of tests/golden/ (written by make_golden.py)."""
import ast
import math
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["rt1d", "if2d", "bragg3d"]


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    c = {k: z[k] for k in z.files}
    out = {
        "kind": str(c["kind"]),
        "shape": tuple(int(v) for v in c["shape"]),
        "L": [float(v) for v in c["L"]],
        "dt": float(c["dt"]),
        "Nk": int(c["Nk"]),
        "nblocks": int(c["nblocks"]),
        "t0": float(c["t0"]),
        "alpha": [float(v) for v in c["alpha"]],
        "gs": [float(v) for v in c["gs"]],
        "psi0": list(c["psi0"]),
        "out": list(c["out"]),
        "t": float(c["t"]),
        "norms": c["norms"],
        "pots": list(c["pots"]) if "pots" in c else None,
        "ifp": ast.literal_eval(str(c["ifp"])) if "ifp" in c else {},
    }
    if "amp_t0" in c:
        t0, s = float(c["amp_t0"]), float(c["amp_sigma"])
        out["amp"] = lambda t: math.exp(-0.5 * ((t - t0) / s) ** 2)
    else:
        out["amp"] = None
    return out


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
