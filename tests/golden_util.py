"""Helpers to load the golden fixtures written by tests/golden/make_golden.py."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

INI_KEYS = ["uavg", "vavg", "rho_ini", "viscosity", "inlet_profile", "collision", "curved",
            "outlet_profile", "iterations", "autosave_every", "autosave_after", "postproc",
            "drag_lift", "crit_v", "crit_rho", "read_former"]
INT_KEYS = {"inlet_profile", "collision", "curved", "outlet_profile", "iterations",
            "autosave_every", "autosave_after", "postproc", "drag_lift", "read_former"}


def ini_dict(values):
    d = {}
    for k, v in zip(INI_KEYS, values):
        d[k] = int(v) if k in INT_KEYS else float(v)
    return d


def small_cases():
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    return [n for n in names if not n.startswith(("coarse_", "medium_"))]


def shipped_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "coarse_*.npz")))


def medium_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "medium_*.npz")))


SETUP_FIELDS = ("fluid", "corner", "stream", "id", "boundary", "boundary_id", "bc", "q", "x", "y")


def sha(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def residual_lines(g):
    return str(g["residuals"]).splitlines()
