"""Helpers to read the golden vectors written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("ot_xe_32", "otc_per_36x32x40", "ot_null_45x30x27")
DT = 5.0e-4 / 7.63822291


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        cfg = json.load(f)
    data = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    return cfg, data


def params_for(cfg, capi, **over):
    kw = dict(nx=cfg["nx"], ny=cfg["ny"], nz=cfg["nz"], xmax=cfg["xmax"], ymax=cfg["ymax"],
              zmax=cfg["zmax"], npd=cfg["npd"], icon=cfg["icon"], core=cfg["core"],
              lsolid=int(cfg["lsolid"]), tzmean=cfg["tzmean"])
    kw.update(over)
    return capi.default_params(**kw)


def planes(f):
    nz, ny, nx = f.shape
    return np.concatenate([f[nz // 2, :, :].ravel(), f[:, ny // 2, :].ravel(), f[:, :, nx // 3].ravel()])


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
