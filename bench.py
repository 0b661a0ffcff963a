"""Benchmark: 1M-hex optimiser iterations/s, FP32+adjoint vs FP64 plain, on one MI355X.

Workload: the clipped torus (R=1, r=0.35, planar cuts z=+-0.2) as a uniform-level octree
hex mesh of ~1.0M hexes; its boundary vertices carry tangent-plane constraints to the
analytic surface. One step = one optimiser iteration of the reference's deformation loop
shape: compute_referenceMesh + one SLIM iteration (local step, 20-step CG global step,
flip-avoiding line search). `value` is the FP32+adjoint optimiser's iterations/s; the FP64
plain (reference-faithful) path is timed on the same input and reported beside it, with the
min scaled Jacobian both reach after the same number of iterations.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "feature-preserving-octree-hex-meshing_amd"))

HBM_PEAK_GBS = 8000.0


def build_workload(target):
    from pyhexopt import meshgen as mg
    V, H, bnd = mg.torus_hex_mesh(target)
    C, N = mg.clipped_torus_closest(V[bnd])
    dis = (N * C).sum(axis=1)
    return V, H, bnd, N, dis


def make_ctx(V, H, bnd, N, dis, lamT):
    from pyhexopt import capi
    h = capi.HexOpt(V, H)
    h.reference()
    h.set_constraints(idsT=bnd, normalT=N, disT=dis, lamT=lamT)
    return h


def run(h, mode, steps):
    """`steps` optimiser iterations; returns seconds and the per-iteration logs."""
    t0 = time.perf_counter()
    logs = []
    for _ in range(steps):
        h.rest_from_current()              # ts.V = ts.UV (deformation loop, grid_hex_meshing.cpp:2956)
        h.reference()                      # compute_referenceMesh at the current positions
        if mode == "fp64":
            E, Eq, Es, log, _ = h.slim_solve(1)
        else:
            E, log = h.opt32_solve(1)
        logs.append(log[0])
    h.synchronize()
    return time.perf_counter() - t0, np.array(logs)


def cpu_baseline(target_small):
    """The CPU oracle (serial FP64 restatement) on a bounded sample: one SLIM iteration on a
    smaller mesh of the same workload; iterations/s scaled to 1M hexes."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as O
    V, H, bnd, N, dis = build_workload(target_small)
    G = O.grad_ref(V, H)
    t0 = time.perf_counter()
    O.slim_m_opt(V, H, G, V, 1, lam_T=1.0, idsT=bnd, normalT=N, disT=dis)
    dt = time.perf_counter() - t0
    return {"value": (1.0 / dt) * H.shape[0] / 1e6, "unit": "iterations/s (scaled to 1M hexes)", "cores": 1,
            "kind": "port",
            "sample": "1 FP64 SLIM iteration of the CPU oracle on the %d-hex clipped torus (%.1f s), x %d/1e6"
                      % (H.shape[0], dt, H.shape[0])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--hexes", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see the design notes on "
              "how the mesh would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    V, H, bnd, N, dis = build_workload(args.hexes)
    nV, nH = V.shape[0], H.shape[0]
    print("workload: %d hexes, %d vertices, %d constrained boundary vertices" % (nH, nV, bnd.size),
          file=sys.stderr, flush=True)

    res = {}
    for mode in ("fp64", "fp32"):
        h = make_ctx(V, H, bnd, N, dis, 1.0)
        run(h, mode, args.warmup)
        secs, logs = run(h, mode, args.steps)
        mn, avg = h.scaled_jacobian()
        which = 0 if mode == "fp64" else 1
        ms_apply = h.time_kernel(which, 20)
        res[mode] = dict(secs=secs, logs=logs, minsj=mn, avgsj=avg, ms_apply=ms_apply, layout=h.layout_info())
        print("%s: %.3f ms/iter, min SJ %.6f, apply tile kernel %.1f us, final energy %.9g" %
              (mode, 1e3 * secs / args.steps, mn, 1e3 * ms_apply, logs[-1][0]), file=sys.stderr, flush=True)
        h.close()

    r32, r64 = res["fp32"], res["fp64"]
    rV = nV / nH
    # algorithmic bytes per hex of the FP32 CG-operator tile pass (cube reference operator):
    # connectivity 8 x 4 B, K 8 tets x 6 x 4 B, the direction gathered once and the product
    # written once, 3 x 4 B per vertex each
    bytes_per_hex32 = 32 + 192 + 2 * 12 * rV
    achieved = bytes_per_hex32 * nH / (r32["ms_apply"] * 1e-3) / 1e9
    value = args.steps / r32["secs"]
    out = {
        "metric": "1M-hex optimiser iters/s: FP32+adjoint vs FP64 plain, equal min scaled-Jac",
        "value": round(value, 3),
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["secs"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 step bound, sign test, energy sums, CG dot products and residual)",
        "data": "synthetic",
        "config": {"workload": "clipped torus R=1 r=0.35 z=+-0.2, uniform-level octree hex mesh",
                   "hexes": nH, "vertices": nV, "boundary_vertices": int(bnd.size),
                   "iteration": "compute_referenceMesh + 1 SLIM iteration (20-step CG, flip-avoiding line search)"},
        "fp64_plain": {"iterations_per_s": round(args.steps / r64["secs"], 3),
                       "ms_per_step": round(1e3 * r64["secs"] / args.steps, 4), "min_scaled_jacobian": r64["minsj"],
                       "apply_tile_kernel_us": round(1e3 * r64["ms_apply"], 2)},
        "fp32_adjoint": {"min_scaled_jacobian": r32["minsj"], "avg_scaled_jacobian": r32["avgsj"]},
        "speedup_fp32_vs_fp64": round((args.steps / r32["secs"]) / (args.steps / r64["secs"]), 3),
        "min_sj_abs_diff": abs(r32["minsj"] - r64["minsj"]),
        "element_evals_per_s": round(value * nH, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "tile_scatter_kernel<float,3,ApplyOp<float,float,float,float>> (FP32 CG operator)",
                     "bytes_per_hex": round(bytes_per_hex32, 2)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(max(args.hexes // 16, 20000))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
