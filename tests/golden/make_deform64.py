"""Writes tests/golden/deform_grid64_superellipsoid.npz: the reference's deformation loop
(grid_hex_meshing.cpp:2945-2959, one SLIM iteration per outer step) on the 64^3 jittered grid
of [-1,1]^3 (262,144 hexes, 274,625 vertices, interior jitter sigma = 0.15 h, seed 0) whose
24,578 boundary vertices are projected each step onto the superellipsoid |x|^4+|y|^4+|z|^4 = 1
(level-5 icosphere mapped radially, 20,480 triangles), lambda_T from 1 with bound 1e8.

The start positions are rounded to multiples of 2^-20 so that they are stored exactly as
integers. The compiled reference (oracle/_ref/slim_ref, mode 1) runs the 50 steps in chunks
(each outer step depends only on the positions and lambda_T it starts from, so a chunked run
equals a continuous one). Teacher-forced steps: from the reference's iterate k (rounded to the
same lattice) and its lambda_T, one more reference step gives the iterate the device must
reproduce from the same start.

Run in the development container, in order (a reference step at this size takes ~3 minutes):
  python tests/golden/make_deform64.py start
  python tests/golden/make_deform64.py chunk K N      # steps K .. K+N-1 (K = 0, 5, ... 45)
  python tests/golden/make_deform64.py teacher K      # one step from the rounded iterate K
  python tests/golden/make_deform64.py assemble
Intermediate states go to _work/deform64/ (not part of the repository).

"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "feature-preserving-octree-hex-meshing_amd"))
import refio  # noqa: E402
from pyhexopt import meshgen as mg  # noqa: E402

N, SEED, Q = 64, 0, 2.0 ** -20
STEPS, TEACHER = 50, (25, 49)
LAM0, LAM_BOUND = 1.0, 1e8
NSAMPLE = 1000
WORK = os.path.join(REPO, "_work", "deform64")


def problem():
    V, H = mg.grid_hex(N)
    bm = mg.grid_boundary_mask(N)
    V = mg.jitter(V, ~bm, 0.15 * 2.0 / N, seed=SEED)
    SV, SF = mg.superellipsoid(5)
    ids = np.nonzero(bm)[0].astype(np.int32)
    return np.round(V / Q).astype(np.int64), H, ids, SV, SF


def step(V, H, ids, SV, SF, lam, n):
    return refio.run_slim(1, V, H, energy=2, n_outer=n, n_inner=1, lamT=lam, lamT_bound=LAM_BOUND, idsT=ids,
                          normalT=np.zeros((ids.size, 3)), disT=np.zeros(ids.size), surfV=SV, surfF=SF)


def lam_next(lam, eq, es):
    """grid_hex_meshing.cpp:2958"""
    return min(LAM0 * max(eq / es * lam, 1.0), LAM_BOUND)


def state_path(k):
    return os.path.join(WORK, "state_%03d.npz" % k)


def main(argv):
    os.makedirs(WORK, exist_ok=True)
    Vq0, H, ids, SV, SF = problem()
    if argv[0] == "start":
        np.savez(state_path(0), V=Vq0 * Q, lam=LAM0)
    elif argv[0] == "chunk":
        # one driver call per step: the driver's own lambda update is relative to the lambda
        # it was started with, so the schedule (relative to lambda_0 = 1) is applied here
        k, n = int(argv[1]), int(argv[2])
        for j in range(k, k + n):
            s = np.load(state_path(j))
            r = step(s["V"], H, ids, SV, SF, float(s["lam"]), 1)
            log = r["log"]
            np.savez(state_path(j + 1), V=r["Vtrace"].reshape(-1, 3), lam=lam_next(log[3], log[1], log[2]), log=log)
            print("step %d done, energy %.12g, lambda %g" % (j, log[0], log[3]), flush=True)
    elif argv[0] == "teacher":
        k = int(argv[1])
        s = np.load(state_path(k))
        Vk = np.round(s["V"] / Q) * Q
        r = step(Vk, H, ids, SV, SF, float(s["lam"]), 1)
        np.savez(os.path.join(WORK, "teacher_%03d.npz" % k), V=r["Vtrace"].reshape(-1, 3), log=r["log"],
                 Vq=np.round(s["V"] / Q).astype(np.int64), lam=float(s["lam"]))
        print("teacher step from %d done" % k, flush=True)
    elif argv[0] == "assemble":
        rng = np.random.default_rng(1)
        sample = np.sort(rng.choice(H.max() + 1, size=NSAMPLE, replace=False)).astype(np.int32)
        w = rng.uniform(-1.0, 1.0, size=(H.max() + 1, 3))
        logs, Vs, sums = [], [], []
        for k in range(STEPS + 1):
            s = np.load(state_path(k))
            if k > 0:
                logs.append(s["log"])
            Vs.append(s["V"][sample])
            sums.append([s["V"].sum(), (s["V"] ** 2).sum(), (w * s["V"]).sum()])
        out = dict(n=N, seed=SEED, q=Q, Vq0=Vq0.astype(np.int32), H_n=N, idsT=ids, SV=SV, SF=SF, lamT0=LAM0,
                   lamT_bound=LAM_BOUND, steps=STEPS, sample=sample, weights_seed=1,
                   ref_log=np.array(logs), ref_Vsample=np.array(Vs), ref_sums=np.array(sums))
        for k in TEACHER:
            t = np.load(os.path.join(WORK, "teacher_%03d.npz" % k))
            out["tf%d_Vq" % k] = t["Vq"].astype(np.int32)
            out["tf%d_lam" % k] = float(t["lam"])
            out["tf%d_ref_log" % k] = t["log"]
            out["tf%d_ref_Vsample" % k] = t["V"][sample]
            out["tf%d_ref_sums" % k] = np.array([t["V"].sum(), (t["V"] ** 2).sum(), (w * t["V"]).sum()])
        path = os.path.join(HERE, "deform_grid64_superellipsoid.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
