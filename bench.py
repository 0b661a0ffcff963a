"""Benchmark: 1M-hex optimiser iterations/s, FP32+adjoint vs FP64 plain, on one MI355X.

Workload: the torus R=1, r=0.35 cut by the planes z=+-0.2, triangulated
with 40,000 triangles (two tube bands and two flat annuli meeting at four sharp rim
circles), and a uniform-level octree hex mesh of ~1.0M hexes of its inside. Boundary
vertices near a rim are LINE vertices on that rim (feature curve); the other boundary
vertices are tangent-plane vertices.

One step = one iteration of the reference's feature-alignment loop
(grid_hex_meshing.cpp:3855-3869) on the device: rest <- current, compute_referenceMesh,
tangent-plane projection onto the triangle surface (closest point + pseudonormal), feature
projection onto the rims, one SLIM iteration (local step, 20-step CG global step,
flip-avoiding line search), then lamda_T = lamda_L = min(max(E_q/E_s lamda, 1), 1e8).
`value` is the FP32+adjoint optimiser's iterations/s; the FP64 plain (reference-faithful)
path runs the same loop from the same start and is reported beside it, with the min scaled
Jacobian and the largest boundary-vertex distance to the surface each reaches.

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
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "cg32_apply_pipe_torus1m_hbm_traffic.md")
R_MAJ, R_MIN, Z_CUT = 1.0, 0.35, 0.2
WEIGHT_OPT = 0.001          # feature_alignment's opt.weight_opt = weight_opt * 0.001
LAMDA0, LAMDA_BOUND = 1.0, 1e8


def measured_traffic(nH):
    """HBM bytes per launch of cg32_apply_pipe_kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a device copy of known size), when that summary
    was taken on a mesh of this many hexes; None otherwise."""
    try:
        kv = dict(l.split(":", 1) for l in open(TRAFFIC_PROFILE) if l.startswith(("traffic_bytes_per_launch:", "hexes:")))
        return int(kv["traffic_bytes_per_launch"]) if int(kv["hexes"]) == nH else None
    except (OSError, KeyError, ValueError):
        return None


def build_workload(target):
    from pyhexopt import meshgen as mg
    V, H, bnd = mg.torus_hex_mesh(target, R_MAJ, R_MIN, Z_CUT)
    SV, SF, rims = mg.clipped_torus_surface(R_MAJ, R_MIN, Z_CUT, nu=200, nband=30, nring=20)
    h = np.linalg.norm(V[H[0, 1]] - V[H[0, 0]])
    w = np.sqrt(R_MIN ** 2 - Z_CUT ** 2)
    rim_geom = [(R_MAJ - w, -Z_CUT), (R_MAJ - w, Z_CUT), (R_MAJ + w, -Z_CUT), (R_MAJ + w, Z_CUT)]
    P = V[bnd]
    rho = np.sqrt(P[:, 0] ** 2 + P[:, 1] ** 2)
    d = np.stack([np.hypot(rho - rr, P[:, 2] - zz) for rr, zz in rim_geom], axis=1)
    k = d.argmin(axis=1)
    line = d[np.arange(k.size), k] < 0.75 * h
    return dict(V=V, H=H, idsT=bnd[~line].astype(np.int32), idsL=bnd[line].astype(np.int32),
                curveL=k[line].astype(np.int32), SV=SV, SF=SF, curves=[list(r) for r in rims], h=h)


def make_ctx(w):
    from pyhexopt import capi
    h = capi.HexOpt(w["V"], w["H"])
    nT, nL = w["idsT"].size, w["idsL"].size
    h.set_constraints(idsT=w["idsT"], normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=LAMDA0,
                      idsL=w["idsL"], axaL=np.zeros((nL, 3)), originL=np.zeros((nL, 3)), lamL=LAMDA0)
    surf = capi.Surface(w["SV"], w["SF"])
    curves = capi.Curves(w["SV"], w["curves"], np.ones(len(w["curves"]), dtype=np.uint8))
    return h, surf, curves


def run(h, surf, curves, w, mode, steps, lam):
    """`steps` loop iterations; returns seconds, the per-iteration logs and the final lamda."""
    logs = []
    h.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        h.rest_from_current()              # ts.V = ts.UV
        h.reference()                      # compute_referenceMesh at the current positions
        h.project_planes(surf, lam)        # tangent planes (dirty_graph_projection, REGULAR vertices)
        h.refresh_lines(curves, w["curveL"], lam)   # feature lines (LINE vertices)
        if mode == "fp64":
            E, Eq, Es, log, _ = h.slim_solve(1, weight_opt=WEIGHT_OPT)
            row = log[0]
        else:
            # the FP64 residual of the direction solve is measured on every 10th step
            E, log = h.opt32_solve(1, weight_opt=WEIGHT_OPT, residual_every=1 if k % 10 == 0 else 0)
            row = log[0]
            Eq, Es = row[1], row[2]
        logs.append(row)
        lam = min(LAMDA0 * max(Eq / Es * lam, 1.0), LAMDA_BOUND) if Es > 0 else LAMDA_BOUND
    h.synchronize()
    return time.perf_counter() - t0, np.array(logs), lam


def cpu_baseline(target_small):
    """The reference's own CPU loop iteration where its compiled build exists
    (oracle/_ref/slim_ref, built from the reference's sources; driven through the golden
    helper tests/golden/refio.py): projection_smooth's plane projection through igl's tree,
    then slim_m_opt (slim_precompute + one slim_solve2 iteration), single-threaded, on a
    bounded sample, a smaller mesh of the same workload; iterations/s scaled to 1M hexes.
    The whole driver process is timed (input, tree set-up and the final scaled-Jacobian pass
    included), and its loop leaves the feature lines out. Without that build: the CPU oracle's
    serial FP64 restatement of the same iteration (lines included)."""
    from pyhexopt import capi
    w = build_workload(target_small)
    V, H = w["V"], w["H"]
    ref_bin = os.path.join(REPO, "oracle", "_ref", "slim_ref")
    if os.path.exists(ref_bin):
        sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
        import refio
        ids = np.asarray(w["idsT"], np.int32)
        t0 = time.perf_counter()
        refio.run_slim(1, V, H, energy=capi.SYMMETRIC_DIRICHLET, n_outer=1, n_inner=1, lamT=LAMDA0,
                       lamT_bound=LAMDA_BOUND, idsT=ids, normalT=np.zeros((ids.size, 3)), disT=np.zeros(ids.size),
                       surfV=w["SV"], surfF=w["SF"], weight_opt=WEIGHT_OPT)
        dt = time.perf_counter() - t0
        return {"value": (1.0 / dt) * H.shape[0] / 1e6, "unit": "iterations/s (scaled to 1M hexes)", "cores": 1,
                "kind": "reference",
                "sample": "1 loop iteration of the compiled reference (projection_smooth + slim_m_opt: "
                          "compute_referenceMesh, slim_precompute, one slim_solve2 iteration; %d plane vertices, "
                          "no feature lines) on the %d-hex clipped torus (%.1f s for the whole driver run), "
                          "scaled by %d/1e6" % (ids.size, H.shape[0], dt, H.shape[0])}
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as O
    FN, VN, FEN = capi.surface_normals(w["SV"], w["SF"])     # host-only, part of the surface set-up
    circles = np.ones(len(w["curves"]), dtype=np.uint8)
    t0 = time.perf_counter()
    G = O.grad_ref(V, H)
    _, _, C, N, _ = O.signed_distance_pseudonormal(V[w["idsT"]], w["SV"], w["SF"], FN, VN, FEN)
    o, a = O.line_projection(w["SV"], w["curves"], circles, V[w["idsL"]], w["curveL"])
    O.slim_m_opt(V, H, G, V, 1, weight_opt=WEIGHT_OPT, lam_T=LAMDA0, lam_L=LAMDA0, idsT=w["idsT"], normalT=N,
                 disT=(N * C).sum(axis=1), idsL=w["idsL"], axaL=a, originL=o)
    dt = time.perf_counter() - t0
    return {"value": (1.0 / dt) * H.shape[0] / 1e6, "unit": "iterations/s (scaled to 1M hexes)", "cores": 1,
            "kind": "port",
            "sample": "1 loop iteration (reference operator, projections, FP64 SLIM iteration) of the CPU oracle "
                      "on the %d-hex clipped torus (%.1f s), scaled by %d/1e6" % (H.shape[0], dt, H.shape[0])}


def glue_measure(target, steps):
    """The padding untangle (hexopt_glue_setup / hexopt_glue_solve's SLIM iteration, FP64,
    CONFORMAL) on the exploded clipped-torus mesh: boundary vertices pinned, interior target
    jittered by 0.3 edge lengths, start copies from meshgen.exploded_start. Fixed iteration
    count; seconds per iteration and the FP64 element-operator kernel on the exploded mesh."""
    from pyhexopt import capi, meshgen as mg
    V, H, bnd = mg.torus_hex_mesh(target, R_MAJ, R_MIN, Z_CUT)
    fixed = np.zeros(V.shape[0], dtype=bool)
    fixed[bnd] = True
    edge = np.linalg.norm(V[H[0, 1]] - V[H[0, 0]])
    Vt = mg.jitter(V, ~fixed, 0.3 * edge, seed=1)
    start = mg.exploded_start(V, H, fixed, Vt, 0.2, seed=1)
    Ve, He, orig = mg.exploded_mesh(H, start)
    g = capi.HexOpt(Ve, He)
    g.glue_setup(Vt, H, orig, fixed)
    g.slim_solve(1, energy=capi.CONFORMAL)
    g.synchronize()
    t0 = time.perf_counter()
    E, Eq, Es, log, _ = g.slim_solve(steps, energy=capi.CONFORMAL)
    g.synchronize()
    secs = time.perf_counter() - t0
    ms_apply = g.time_kernel(0, 20)
    g.close()
    return {"hexes": int(H.shape[0]), "exploded_vertices": int(Ve.shape[0]), "iterations": steps,
            "ms_per_iteration": round(1e3 * secs / steps, 3), "energy": E,
            "fp64_operator_kernel_us": round(1e3 * ms_apply, 2)}


def octree_measure(reps=5):
    """The octree subdivision of octree_mesh (hexopt_octree_build: predicate, grading and
    pairing cascades, leaf-to-hex export) around the 40,000-triangle clipped torus at
    num_voxels 4096 and STOP extent 8 (3,717,184 leaves). Device time of the build, best of
    `reps`, and the leaf and node counts."""
    from pyhexopt import capi, meshgen as mg
    V, F, _ = mg.clipped_torus_surface(R_MAJ, R_MIN, Z_CUT, nu=200, nband=30, nring=20)
    lo, hi = V.min(0), V.max(0)
    vs = float((hi - lo).max()) / 4096
    grid = [1 << int(np.ceil(np.log2(np.ceil(e / vs)))) for e in (hi - lo)]
    mt = (lo + hi) / 2 - (lo + (lo + vs * np.array(grid, np.float64))) * 0.5
    s = capi.Surface(V, F)
    best = None
    for _ in range(reps):
        out = s.octree(grid, lo, vs, mt, 8)
        best = out["ms"] if best is None else min(best, out["ms"])
    return {"surface_triangles": int(F.shape[0]), "grid": [int(g) for g in grid], "leaves": int(len(out["leaves"])),
            "nodes": int(len(out["nodes"])), "levels": out["levels"], "predicate_evaluations": out["evaluated"],
            "ms_build": round(best, 3)}


def metro_measure(target):
    """Metro Hausdorff (hexopt_metro_hausdorff: vertex sampling both ways, closest points on
    the device) between the 40,000-triangle clipped torus and the boundary of its hex mesh
    (each boundary quad as two triangles), as the feature-alignment loop's Hausdorff check
    measures it. Wall time of the whole call: both surfaces' trees are built on the host."""
    from pyhexopt import capi, meshgen as mg
    V, H, _ = mg.torus_hex_mesh(target, R_MAJ, R_MIN, Z_CUT)
    T = mg.hex_boundary_triangles(H).astype(np.int32)
    SV, SF, _ = mg.clipped_torus_surface(R_MAJ, R_MIN, Z_CUT, nu=200, nband=30, nring=20)
    capi.metro_hausdorff(SV, SF, V, T)    # warm-up (device code loaded)
    t0 = time.perf_counter()
    diag, mx, mean = capi.metro_hausdorff(SV, SF, V, T)
    secs = time.perf_counter() - t0
    return {"surface_triangles": int(SF.shape[0]), "hex_boundary_triangles": int(T.shape[0]),
            "bbox_diagonal": diag, "max": mx, "mean": mean, "ms": round(1e3 * secs, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--hexes", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-glue", action="store_true")
    ap.add_argument("--no-octree", action="store_true")
    ap.add_argument("--no-metro", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see the design notes on "
              "how the mesh would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    w = build_workload(args.hexes)
    nV, nH = w["V"].shape[0], w["H"].shape[0]
    print("workload: %d hexes, %d vertices, %d plane + %d line vertices, %d surface triangles" %
          (nH, nV, w["idsT"].size, w["idsL"].size, w["SF"].shape[0]), file=sys.stderr, flush=True)

    res = {}
    for mode in ("fp64", "fp32"):
        h, surf, curves = make_ctx(w)
        _, _, lam = run(h, surf, curves, w, mode, args.warmup, LAMDA0)
        secs, logs, lam = run(h, surf, curves, w, mode, args.steps, lam)
        mn, avg = h.scaled_jacobian()
        X = h.get_positions()
        bnd = np.concatenate([w["idsT"], w["idsL"]])
        S, _, _, _ = surf.closest(X[bnd])
        which = 0 if mode == "fp64" else 1
        ms_apply = h.time_kernel(which, 20)
        res[mode] = dict(secs=secs, logs=logs, minsj=mn, avgsj=avg, ms_apply=ms_apply, maxdist=float(np.abs(S).max()),
                         lam=lam, X=X)
        print("%s: %.3f ms/iter, min SJ %.6f, max boundary distance %.3e, apply tile kernel %.1f us, "
              "final energy %.9g, lamda %.3g, steps %s" %
              (mode, 1e3 * secs / args.steps, mn, res[mode]["maxdist"], 1e3 * ms_apply, logs[-1][0], lam,
               np.array2string(logs[:, 4], precision=3)), file=sys.stderr, flush=True)
        h.close()
        surf.close()
        curves.close()

    r32, r64 = res["fp32"], res["fp64"]
    rV = nV / nH
    # algorithmic bytes per hex of the fused FP32 CG operator pass (cube reference operator):
    # K 8 tets x 6 x 4 B, connectivity as 8 16-bit tile-local ids, and per vertex (3 x 4 B
    # each) z and the previous direction read, the new direction and L p written
    bytes_per_hex32 = 192 + 16 + 4 * 12 * rV
    achieved = bytes_per_hex32 * nH / (r32["ms_apply"] * 1e-3) / 1e9
    traffic = measured_traffic(nH)
    value = args.steps / r32["secs"]
    diag = float(np.linalg.norm(w["SV"].max(axis=0) - w["SV"].min(axis=0)))
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
        "config": {"workload": "clipped torus R=1 r=0.35 z=+-0.2 (40k triangles, 4 rim feature curves), "
                               "uniform-level octree hex mesh, feature-alignment loop",
                   "hexes": nH, "vertices": nV, "plane_vertices": int(w["idsT"].size),
                   "line_vertices": int(w["idsL"].size), "surface_triangles": int(w["SF"].shape[0]),
                   "iteration": "compute_referenceMesh + plane/line projection + 1 SLIM iteration "
                                "(20-step CG, flip-avoiding line search), weight_opt 0.001"},
        "fp64_plain": {"iterations_per_s": round(args.steps / r64["secs"], 3),
                       "ms_per_step": round(1e3 * r64["secs"] / args.steps, 4), "min_scaled_jacobian": r64["minsj"],
                       "max_boundary_distance": r64["maxdist"], "apply_tile_kernel_us": round(1e3 * r64["ms_apply"], 2)},
        "fp32_adjoint": {"min_scaled_jacobian": r32["minsj"], "avg_scaled_jacobian": r32["avgsj"],
                         "max_fp64_relative_residual": float(r32["logs"][:, 7].max()),
                         "max_boundary_distance": r32["maxdist"],
                         "max_vertex_diff_vs_fp64": float(np.abs(r32["X"] - r64["X"]).max())},
        "speedup_fp32_vs_fp64": round((args.steps / r32["secs"]) / (args.steps / r64["secs"]), 3),
        "min_sj_abs_diff": abs(r32["minsj"] - r64["minsj"]),
        "max_boundary_distance_diff_over_bbox_diag": abs(r32["maxdist"] - r64["maxdist"]) / diag,
        "element_evals_per_s": round(value * nH, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "cg32_apply_pipe_kernel (persistent FP32 CG operator pass)",
                     "bytes_per_hex": round(bytes_per_hex32, 2), "algorithmic_bytes": round(bytes_per_hex32 * nH),
                     "traffic_source": "profiles/cg32_apply_pipe_torus1m_hbm_traffic.md" if traffic else None},
    }
    if not args.no_glue:
        out["padding_untangle"] = glue_measure(max(args.hexes // 8, 20000), 5)
    if not args.no_octree:
        out["octree_build"] = octree_measure()
    if not args.no_metro:
        out["metro_hausdorff"] = metro_measure(args.hexes)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(max(args.hexes // 16, 20000))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
