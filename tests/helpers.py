"""Shared fixture loading for the parity tests."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def slim_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "slim_*.npz")))


def aabb_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "aabb_*.npz")))


def constraint_kwargs(c):
    """Constraint arguments of a fixture, in the oracle's keyword names."""
    kw = {}
    if "idsT" in c:
        kw.update(idsT=c["idsT"], normalT=c["normalT"], disT=c["disT"], lam_T=float(c["lamT"]))
    if "idsC" in c:
        kw.update(idsC=c["idsC"], Cpos=c["C"], lam_C=float(c["lamC"]))
    if "idsL" in c:
        kw.update(idsL=c["idsL"], axaL=c["axaL"], originL=c["originL"], lam_L=float(c["lamL"]))
    return kw


def gpu_constraint_kwargs(c):
    kw = {}
    if "idsT" in c:
        kw.update(idsT=c["idsT"], normalT=c["normalT"], disT=c["disT"], lamT=float(c["lamT"]))
    if "idsC" in c:
        kw.update(idsC=c["idsC"], Cpos=c["C"], lamC=float(c["lamC"]))
    if "idsL" in c:
        kw.update(idsL=c["idsL"], axaL=c["axaL"], originL=c["originL"], lamL=float(c["lamL"]))
    return kw


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
