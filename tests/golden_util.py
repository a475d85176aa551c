"""Shared helpers for the golden fixtures in tests/golden/*.npz (written by
tests/golden/make_golden.py from runs of the reference solver)."""
import hashlib
import os

import numpy as np

import fields

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("smooth8", "posts8", "posts8x8x70", "laminar64x32x32")
LOG_KEYS = ("K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def params(g):
    keys = ("X", "Y", "Z", "Rebs", "ub", "CFL", "TSTR", "N", "uc", "amp", "seed", "utime")
    p = {k: g[k].item() for k in keys}
    p["slip"] = bool(g["slip"])
    return p


def wall_mask(p):
    if p["slip"]:
        m = fields.posts_mask(p["X"], p["Y"])
        return m, m
    return None, None


def initial_state(p):
    return fields.device_poiseuille(p["X"], p["Y"], p["Z"], p["uc"], p["amp"], p["seed"])


def log_text(key, t, v):
    fmt = "%f %.14f\n" if key == "K" else "%f %.12f\n"
    return "".join(fmt % (a, b) for a, b in zip(t, v)).encode()


def lbf_sha(ft, f):
    return hashlib.sha256(np.ascontiguousarray(ft).tobytes() + np.ascontiguousarray(f).tobytes()).hexdigest()
