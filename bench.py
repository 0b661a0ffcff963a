"""Benchmark: 1M-hex optimiser iterations/s, FP32+adjoint vs FP64 plain, on one MI355X.

Workload (pyhexopt.workload): the torus R=1, r=0.35 cut by the planes z=+-0.2, 40,000
triangles in four patches meeting at four sharp rim circles; the hex mesh is the graded,
paired octree of that surface (leaves of 8 sizes, finest at the surface) restricted to the
leaves inside it, ~1.0M hexes. Boundary vertices near a rim are LINE vertices of that rim; the
others are REGULAR vertices of the patch of their nearest triangle.

One step = one iteration of the reference's feature-alignment loop (feature_alignment,
grid_hex_meshing.cpp:3783-3912, with dirty_graph_projection :3954-4082) on the device:
rest <- current, compute_referenceMesh (cube shapes, H_flag), tangent planes from each
REGULAR vertex's own patch tree, LINE vertices projected onto their rim, one SLIM iteration
(local step, 20-step CG global step, flip-avoiding line search), then
lamda = min(0.05 max(E_q / E_s lamda, 1), 1e16) with lamda starting at args.feature_weight
= 0.05 and opt.weight_opt = 0.001.

`value` is the FP32+adjoint optimiser's iterations/s: FP32 element arithmetic and CG, FP64
for the sign test, the step bound, the line-search energies and the CG dot products, and the
reference's 20 CG steps, which reproduce the FP64 path's truncated solve (the directions agree
to ~1e-7). The variant with the FP64-residual iterative refinement (18 CG steps, then 2 on the
FP64 residual) is timed and reported beside it. The FP64 plain (reference-faithful) path runs
the same loop from the same start. All three then continue, untimed, to `--iterations` loop
iterations in all, and the final min scaled Jacobian and metro Hausdorff distance of each are
reported with their differences, next to how far the FP64 loop itself moves from a start
perturbed by 1e-12 of the bounding-box diagonal. The per-result errors of the FP32 path
against FP64 (gradient, SLIM weights, step, direction, positions, energy) are measured on the
first iteration from the shared start.

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "feature-preserving-octree-hex-meshing_amd"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "cg32_apply_pipe_torus_octree1m_hbm_traffic.md")
WEIGHT_OPT = 0.001          # feature_alignment's opt.weight_opt = weight_opt * 0.001
LAMDA0, LAMDA_BOUND = 0.05, 1e16   # args.feature_weight, LAMDA_FEATURE_PROJECTION_BOUND


def measured_traffic(nH):
    """HBM bytes per launch of cg32_apply_pipe_kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a device copy of known size), when that summary
    was taken on a mesh of this many hexes; None otherwise."""
    try:
        kv = dict(l.split(":", 1) for l in open(TRAFFIC_PROFILE) if l.startswith(("traffic_bytes_per_launch:", "hexes:")))
        return int(kv["traffic_bytes_per_launch"]) if int(kv["hexes"]) == nH else None
    except (OSError, KeyError, ValueError):
        return None


def make_ctx(w):
    from pyhexopt import capi, workload
    h = capi.HexOpt(w["V"], w["H"])
    nT, nL = w["idsT"].size, w["idsL"].size
    h.set_constraints(idsT=w["idsT"], normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=LAMDA0,
                      idsL=w["idsL"], axaL=np.zeros((nL, 3)), originL=np.zeros((nL, 3)), lamL=LAMDA0)
    patches = workload.patch_surfaces(w)
    curves = capi.Curves(w["SV"], w["curves"], np.ones(len(w["curves"]), dtype=np.uint8))
    return h, patches, curves


def prepare(h, patches, curves, w, lam):
    """The loop body up to the SLIM iteration: reference operator at the current positions,
    patch planes and rim lines at weight lam."""
    h.rest_from_current()              # ts.V = ts.UV
    h.reference()                      # compute_referenceMesh(ts.UV, ..., md.H_flag, cube): every hex inside
    # the patch of each plane vertex is passed once; later calls reuse it
    h.project_patches(patches, None if getattr(h, "patches_set", False) else w["patchT"], lam)
    h.patches_set = True
    h.refresh_lines(curves, w["curveL"], lam)


def iterate(h, mode, solve, measure=False):
    """One SLIM iteration. mode "fp64": the plain FP64 path; "fp32": the FP32+adjoint path with
    solve = (CG steps, refinement CG steps). measure: the FP32 path also forms the FP64
    residual of its direction (log columns 7 and 9)."""
    if mode == "fp64":
        E, Eq, Es, log, _ = h.slim_solve(1, weight_opt=WEIGHT_OPT)
        return Eq, Es, log[0]
    E, log = h.opt32_solve(1, cg_iters=solve[0], weight_opt=WEIGHT_OPT, refine_iters=solve[1],
                           residual_every=1 if measure else 0)
    return log[0, 1], log[0, 2], log[0]


def run(h, patches, curves, w, mode, steps, lam, solve, measure_every=0):
    """`steps` loop iterations; returns seconds, the per-iteration logs and the final lamda.
    measure_every > 0: the FP64 residual of the FP32 direction on every that-many-th iteration
    (an extra FP64 gradient and operator pass; left out of timed runs)."""
    logs = []
    h.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        prepare(h, patches, curves, w, lam)
        Eq, Es, row = iterate(h, mode, solve, measure=measure_every > 0 and k % measure_every == 0)
        logs.append(row)
        lam = min(LAMDA0 * max(Eq / Es * lam, 1.0), LAMDA_BOUND) if Es > 0 else LAMDA_BOUND
    h.synchronize()
    return time.perf_counter() - t0, logs, lam


def first_iteration_errors(w, solve, refined):
    """FP32 path against FP64 from the shared start: dE/dx and the SLIM weight blocks at the
    same positions, then one iteration of each: the accepted step, the direction, the positions
    and the quality energy after it, and the FP64 residual of the FP32 direction. The same for
    the FP32 variant with the FP64-residual refinement (`refined` = (CG steps, refinement CG
    steps))."""
    from pyhexopt import capi
    res = {}
    ctx = {m: make_ctx(w) for m in ("fp64", "fp32", "fp32r")}
    try:
        h64 = ctx["fp64"][0]
        for m in ctx:
            prepare(*ctx[m], w, LAMDA0)
        X0 = h64.get_positions()
        # the FP32 adjoint's arithmetic error: both gradients at the same (FP32) positions
        X32 = X0.astype(np.float32)
        g64 = h64.energy_grad(X32.astype(np.float64), precision=64, weight_opt=WEIGHT_OPT)
        g32 = h64.energy_grad(X32, precision=32, weight_opt=WEIGHT_OPT).astype(np.float64)
        K64 = h64.debug_operator(64, weight_opt=WEIGHT_OPT)
        K32 = h64.debug_operator(32, weight_opt=WEIGHT_OPT)
        _, _, r64 = iterate(h64, "fp64", None)
        h32, h32r = ctx["fp32"][0], ctx["fp32r"][0]
        _, _, r32 = iterate(h32, "fp32", solve, measure=True)
        _, _, r32r = iterate(h32r, "fp32", refined, measure=True)
        X64, X32, X32r = h64.get_positions(), h32.get_positions(), h32r.get_positions()
        # directions from x' = x + step d
        d64 = (X64 - X0) / r64[4] if r64[4] > 0 else None

        def ddiff(X, r):
            if d64 is None or r[4] <= 0:
                return None
            return float(np.linalg.norm((X - X0) / r[4] - d64) / np.linalg.norm(d64))
        diag = float(np.linalg.norm(w["SV"].max(axis=0) - w["SV"].min(axis=0)))
        gscale = np.abs(g64).max()
        kscale = np.abs(K64).max()
        res = {
            # the quality energy after the FP32 path's iteration against the FP64 path's
            "energy_quality_rel": abs(r32[1] - r64[1]) / abs(r64[1]),
            "grad_rel_max": float(np.abs(g32 - g64).max() / gscale),
            "grad_rel_l2": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
            "weights_rel_max": float(np.abs(K32 - K64).max() / kscale),
            "weights_rel_l2": float(np.linalg.norm(K32 - K64) / np.linalg.norm(K64)),
            "step_fp32": float(r32[4]), "step_fp64": float(r64[4]),
            "step_bound_fp32": float(r32[3]), "step_bound_fp64": float(r64[3]),
            "direction_rel_diff": ddiff(X32, r32),
            "positions_max_over_diag": float(np.abs(X32 - X64).max() / diag),
            # |b64 - L d| / |b64| of the direction system L d = -dE/dx / 2 (FP64 sums), and the
            # FP32 CG's own estimate of it; both CGs stop at the same truncation, so this is
            # the truncation, not FP32 rounding (the directions agree to direction_rel_diff)
            "fp64_residual_of_fp32_direction": float(r32[7]),
            "fp32_cg_error": float(r32[6]),
            # the FP64 path's Eigen error |rhs - L x| / |rhs| of the position form
            "fp64_cg_error_position_form": float(r64[6]),
            "with_refinement": {
                "cg_iters": refined[0], "refinement_cg_iters": refined[1],
                "step": float(r32r[4]), "direction_rel_diff": ddiff(X32r, r32r),
                "positions_max_over_diag": float(np.abs(X32r - X64).max() / diag),
                "energy_quality_rel": abs(r32r[1] - r64[1]) / abs(r64[1]),
                "fp64_residual_before_refinement": float(r32r[7]),
                "fp64_residual_after_refinement": float(r32r[9])},
        }
    finally:
        for h, patches, curves in ctx.values():
            h.close()
            for p in patches:
                p.close()
            curves.close()
    return res


def boundary_triangles(quads):
    return np.concatenate([quads[:, [0, 1, 2]], quads[:, [2, 3, 0]]]).astype(np.int32)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(num_voxels):
    """The reference's own CPU loop iteration where its compiled build exists
    (oracle/_ref/slim_ref, built from the reference's sources; driven through the golden
    helper tests/golden/refio.py), on a bounded sample: the same workload at a coarser
    octree (num_voxels), single-threaded. The driver times the loop body alone
    (compute_referenceMesh, projection_smooth's plane projection through igl's tree,
    slim_m_opt: slim_precompute + one slim_solve2 iteration, the rim lines as constraints)
    with std::chrono; iterations/s are scaled linearly to the 1M-hex mesh by hex count.
    Without that build: the CPU oracle's serial FP64 restatement of the same iteration."""
    from pyhexopt import capi, workload
    w = workload.clipped_torus_octree(num_voxels)
    V, H = w["V"], w["H"]
    nL = w["idsL"].size
    circles = np.ones(len(w["curves"]), dtype=np.uint8)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as O
    o, a = O.line_projection(w["SV"], w["curves"], circles, V[w["idsL"]], w["curveL"])
    ref_bin = os.path.join(REPO, "oracle", "_ref", "slim_ref")
    if os.path.exists(ref_bin):
        sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
        import refio
        ids = np.asarray(w["idsT"], np.int32)
        t0 = time.perf_counter()
        r = refio.run_slim(1, V, H, energy=capi.SYMMETRIC_DIRICHLET, n_outer=1, n_inner=1, lamT=LAMDA0,
                           lamT_bound=LAMDA_BOUND, lamL=LAMDA0, idsT=ids, normalT=np.zeros((ids.size, 3)),
                           disT=np.zeros(ids.size), idsL=w["idsL"], axaL=a, originL=o, surfV=w["SV"],
                           surfF=w["SF"], weight_opt=WEIGHT_OPT)
        whole = time.perf_counter() - t0
        dt = float(r["t_iter"][0])
        return {"value": (1.0 / dt) * H.shape[0] / 1e6, "unit": "iterations/s (scaled to 1M hexes)", "cores": 1,
                "kind": "reference", "cpu": cpu_model(),
                "sample": "1 loop iteration of the compiled reference (compute_referenceMesh, projection_smooth, "
                          "slim_m_opt: slim_precompute + one slim_solve2 iteration; %d plane vertices against the "
                          "whole surface, %d rim vertices as line constraints) on the %d-hex octree mesh of the "
                          "same torus at num_voxels %d: %.2f s for the loop body (%.1f s for the whole driver "
                          "process), scaled by %d/1e6" % (ids.size, nL, H.shape[0], num_voxels, dt, whole,
                                                           H.shape[0])}
    FN, VN, FEN = capi.surface_normals(w["SV"], w["SF"])     # host-only, part of the surface set-up
    t0 = time.perf_counter()
    G = O.grad_ref(V, H)
    _, _, C, N, _ = O.signed_distance_pseudonormal(V[w["idsT"]], w["SV"], w["SF"], FN, VN, FEN)
    O.slim_m_opt(V, H, G, V, 1, weight_opt=WEIGHT_OPT, lam_T=LAMDA0, lam_L=LAMDA0, idsT=w["idsT"], normalT=N,
                 disT=(N * C).sum(axis=1), idsL=w["idsL"], axaL=a, originL=o)
    dt = time.perf_counter() - t0
    return {"value": (1.0 / dt) * H.shape[0] / 1e6, "unit": "iterations/s (scaled to 1M hexes)", "cores": 1,
            "kind": "port", "cpu": cpu_model(),
            "sample": "1 loop iteration (reference operator, projections, FP64 SLIM iteration) of the CPU oracle "
                      "on the %d-hex octree mesh at num_voxels %d (%.1f s), scaled by %d/1e6" %
                      (H.shape[0], num_voxels, dt, H.shape[0])}


def glue_measure(target, steps):
    """The padding untangle (hexopt_glue_setup / hexopt_glue_solve's SLIM iteration, FP64,
    CONFORMAL) on an exploded uniform clipped-torus mesh: boundary vertices pinned, interior
    target jittered by 0.3 edge lengths, start copies from meshgen.exploded_start. Fixed
    iteration count; seconds per iteration and the FP64 element-operator kernel."""
    from pyhexopt import capi, meshgen as mg
    V, H, bnd = mg.torus_hex_mesh(target, 1.0, 0.35, 0.2)
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
    from pyhexopt import capi, meshgen as mg, workload
    V, F, _, _ = workload.torus_surface()
    vs, lo, grid, mt = mg.octree_frame(V, 4096)
    s = capi.Surface(V, F)
    best = None
    for _ in range(reps):
        out = s.octree(grid, lo, vs, mt, 8)
        best = out["ms"] if best is None else min(best, out["ms"])
    s.close()
    return {"surface_triangles": int(F.shape[0]), "grid": [int(g) for g in grid], "leaves": int(len(out["leaves"])),
            "nodes": int(len(out["nodes"])), "levels": out["levels"], "predicate_evaluations": out["evaluated"],
            "ms_build": round(best, 3)}


def scale_measure(num_voxels, iterations, solve):
    """The scale workload (pyhexopt.workload.box_union_octree): the graded octree hex mesh of
    the union of 8 seeded rotated boxes (~203k triangles), its boundary vertices projected
    onto the union's outer triangles each iteration, FP64 plain and FP32+adjoint from the same
    start for `iterations` loop iterations each (the feature-alignment loop's constants, no
    feature curves). Reports both paths' ms per iteration, min scaled Jacobian, metro
    Hausdorff (vertex sampling) and the Hausdorff distance over 1M area-uniform samples of
    each surface, and the 256^3 voxel occupancy of the union."""
    from pyhexopt import capi, workload
    t0 = time.perf_counter()
    w = workload.box_union_octree(num_voxels, 1)
    nH = w["H"].shape[0]
    print("scale workload: %d hexes, %d vertices, %d boundary vertices, %d outer of %d triangles, built in %.1f s" %
          (nH, w["V"].shape[0], w["idsT"].size, w["SF"].shape[0], w["SF_soup"].shape[0], time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    T = boundary_triangles(w["quads"])
    out = {"hexes": int(nH), "vertices": int(w["V"].shape[0]), "boundary_vertices": int(w["idsT"].size),
           "surface_triangles": int(w["SF_soup"].shape[0]), "outer_triangles": int(w["SF"].shape[0]),
           "num_voxels": num_voxels, "iterations": iterations}
    for mode in ("fp64", "fp32"):
        h = capi.HexOpt(w["V"], w["H"])
        nT = w["idsT"].size
        h.set_constraints(idsT=w["idsT"], normalT=np.zeros((nT, 3)), disT=np.zeros(nT), lamT=LAMDA0)
        surf = capi.Surface(w["SV"], w["SF"])
        lam = LAMDA0
        h.synchronize()
        t0 = time.perf_counter()
        for k in range(iterations):
            h.rest_from_current()
            h.reference()
            h.project_patches([surf], w["patchT"] if k == 0 else None, lam)
            Eq, Es, _ = iterate(h, mode, solve)
            lam = min(LAMDA0 * max(Eq / Es * lam, 1.0), LAMDA_BOUND) if Es > 0 else LAMDA_BOUND
        h.synchronize()
        secs = time.perf_counter() - t0
        mn, avg = h.scaled_jacobian()
        X = h.get_positions()
        _, hmax, hmean = capi.metro_hausdorff(w["SV"], w["SF"], X, T)
        sampled = workload.sampled_hausdorff(w["SV"], w["SF"], X, T, n=1_000_000, seed=0)
        out[mode] = {"ms_per_iteration": round(1e3 * secs / iterations, 3), "min_scaled_jacobian": mn,
                     "avg_scaled_jacobian": avg, "metro_max": hmax, "metro_mean": hmean,
                     "sampled_hausdorff_1M": sampled}
        print("scale %s: %.3f ms/iter, min SJ %.6f, metro max %.4e, sampled max %.4e" %
              (mode, 1e3 * secs / iterations, mn, hmax, sampled["max"]), file=sys.stderr, flush=True)
        h.close()
        surf.close()
    out["speedup_fp32_vs_fp64"] = round(out["fp64"]["ms_per_iteration"] / out["fp32"]["ms_per_iteration"], 3)
    out["min_sj_abs_diff"] = abs(out["fp32"]["min_scaled_jacobian"] - out["fp64"]["min_scaled_jacobian"])
    # the 256^3 inside bitmap of the union (compute_sign(VoxelGrid), voxelization.h:221-272)
    SV = w["SV"]
    lo = SV.min(axis=0)
    sp = float((SV.max(axis=0) - lo).max()) / 256 * (1 + 1e-4)
    s = capi.Surface(SV, w["SF_soup"])
    s.voxel_parity(256, 256, 256, lo, sp)       # warm-up: column bins built on first use
    t0 = time.perf_counter()
    inside = s.voxel_parity(256, 256, 256, lo, sp)
    out["voxel_parity_256"] = {"ms": round(1e3 * (time.perf_counter() - t0), 2), "inside": int(inside.sum())}
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--iterations", type=int, default=100, help="loop iterations of the quality comparison")
    ap.add_argument("--num-voxels", type=int, default=448)
    ap.add_argument("--stop-extent", type=int, default=1)
    ap.add_argument("--cg-iters", type=int, default=20, help="FP32 CG steps of the direction solve (the reference's 20)")
    ap.add_argument("--refine", type=int, default=0, help="FP32 CG steps of the FP64-residual refinement")
    ap.add_argument("--refined-cg-iters", type=int, default=18,
                    help="the refinement variant reported beside: CG steps before its refinement")
    ap.add_argument("--refined-refine", type=int, default=2,
                    help="the refinement variant reported beside: its refinement CG steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-glue", action="store_true")
    ap.add_argument("--no-octree", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-scale", action="store_true")
    ap.add_argument("--scale-voxels", type=int, default=448)
    ap.add_argument("--scale-iterations", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see the design notes on "
              "how the mesh would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from pyhexopt import capi, workload
    solve = (args.cg_iters, args.refine)
    refined = (args.refined_cg_iters, args.refined_refine)
    t0 = time.perf_counter()
    w = workload.clipped_torus_octree(args.num_voxels, args.stop_extent)
    nV, nH = w["V"].shape[0], w["H"].shape[0]
    ext, cnt = w["extents"]
    print("workload: %d hexes (%s), %d vertices, %d plane + %d line vertices, %d surface triangles, "
          "built in %.1f s" % (nH, ", ".join("%d of extent %d" % (c, e) for e, c in zip(ext, cnt)), nV,
                                w["idsT"].size, w["idsL"].size, w["SF"].shape[0], time.perf_counter() - t0),
          file=sys.stderr, flush=True)

    errors = None if args.no_errors else first_iteration_errors(w, solve, refined)
    if errors:
        print("first-iteration errors: %s" % json.dumps(errors), file=sys.stderr, flush=True)

    res = {}
    T = boundary_triangles(w["quads"])
    total = max(args.iterations, args.warmup + args.steps)
    # the FP64 loop again from the start perturbed by 1e-12 of the bounding-box diagonal (three
    # seeds): how far the final min scaled Jacobian and Hausdorff distance move under last-bit
    # changes alone, the yardstick for the FP32 path's differences
    diag0 = float(np.linalg.norm(w["V"].max(axis=0) - w["V"].min(axis=0)))
    pert = []
    for seed in (9, 10, 11):
        h, patches, curves = make_ctx(w)
        h.set_positions(capi.CURRENT,
                        w["V"] + 1e-12 * diag0 * np.random.default_rng(seed).standard_normal(w["V"].shape))
        _, _, _ = run(h, patches, curves, w, "fp64", total, LAMDA0, None)
        mn_p, _ = h.scaled_jacobian()
        _, hmax_p, _ = capi.metro_hausdorff(w["SV"], w["SF"], h.get_positions(), T)
        h.close()
        for p in patches:
            p.close()
        curves.close()
        pert.append((mn_p, hmax_p))
        print("fp64 from a perturbed start (seed %d): min SJ %.6f, metro max %.6e" % (seed, mn_p, hmax_p),
              file=sys.stderr, flush=True)
    for mode in ("fp64", "fp32"):
        h, patches, curves = make_ctx(w)
        _, _, lam = run(h, patches, curves, w, mode, args.warmup, LAMDA0, solve)
        secs, logs, lam = run(h, patches, curves, w, mode, args.steps, lam, solve)
        _, rest, lam = run(h, patches, curves, w, mode, total - args.warmup - args.steps, lam, solve,
                           measure_every=10)
        logs = np.array(logs + rest)
        mn, avg, sj = h.scaled_jacobian(per_hex=True)
        q = np.quantile(sj, [0.001, 0.01])
        X = h.get_positions()
        diag, hmax, hmean = capi.metro_hausdorff(w["SV"], w["SF"], X, T)
        which = 0 if mode == "fp64" else 1
        ms_apply = h.time_kernel(which, 20)
        ms_grad = h.time_kernel(5 if mode == "fp64" else 4, 20)
        res[mode] = dict(secs=secs, logs=logs, minsj=mn, avgsj=avg, sjq=[float(v) for v in q], ms_apply=ms_apply,
                         ms_grad=ms_grad,
                         hmax=hmax, hmean=hmean, diag=diag, lam=lam, X=X)
        print("%s: %.3f ms/iter, after %d iterations min SJ %.6f, metro max %.6e mean %.6e, operator kernel "
              "%.1f us, energy+gradient kernel %.1f us, final energy %.9g, lamda %.3g" %
              (mode, 1e3 * secs / args.steps, total, mn, hmax, hmean, 1e3 * ms_apply, 1e3 * ms_grad, logs[-1][0],
               lam), file=sys.stderr, flush=True)
        h.close()
        for p in patches:
            p.close()
        curves.close()

    # the FP32 path once more with the FP64-residual refinement (the refined variant's CG steps
    # before it and its correction steps): what the refinement changes
    h, patches, curves = make_ctx(w)
    _, _, lam = run(h, patches, curves, w, "fp32", args.warmup, LAMDA0, refined)
    secs_r, _, lam = run(h, patches, curves, w, "fp32", args.steps, lam, refined)
    _, logs_r, lam = run(h, patches, curves, w, "fp32", total - args.warmup - args.steps, lam, refined,
                         measure_every=10)
    mn_r, _ = h.scaled_jacobian()
    _, hmax_r, _ = capi.metro_hausdorff(w["SV"], w["SF"], h.get_positions(), T)
    h.close()
    for p in patches:
        p.close()
    curves.close()
    print("fp32 with refinement (%d + %d CG steps): %.3f ms/iter, min SJ %.6f, metro max %.6e" %
          (refined[0], refined[1], 1e3 * secs_r / args.steps, mn_r, hmax_r), file=sys.stderr, flush=True)
    res32 = [l[7] for l in res["fp32"]["logs"] if l[7] >= 0]
    res32r = [(l[7], l[9]) for l in logs_r if l[9] >= 0]
    r32, r64 = res["fp32"], res["fp64"]
    rV = nV / nH
    # algorithmic bytes per hex of the fused FP32 CG operator pass (cube reference operator):
    # K 8 tets x 6 x 4 B, connectivity as 8 16-bit tile-local ids, and per vertex (3 x 4 B
    # each) z and the previous direction read, the new direction and L p written
    bytes_per_hex32 = 192 + 16 + 4 * 12 * rV
    achieved = bytes_per_hex32 * nH / (r32["ms_apply"] * 1e-3) / 1e9
    # the FP32 energy + gradient pass: 8 corner ids (32 B) and 8 reference factors (32 B) per
    # hex, positions read and gradient written (12 B each per vertex)
    grad_bytes_per_hex = 64 + 24 * rV
    grad_achieved = grad_bytes_per_hex * nH / (r32["ms_grad"] * 1e-3) / 1e9
    traffic = measured_traffic(nH)
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
        "dtype": "f32 (FP64 sign test, step bound, line-search energies, CG dot products)",
        "data": "synthetic",
        "config": {"workload": "clipped torus R=1 r=0.35 z=+-0.2 (40k triangles, 4 patches, 4 rim feature curves), "
                               "graded paired octree hex mesh (leaves inside the surface), feature-alignment loop",
                   "hexes": nH, "vertices": nV, "plane_vertices": int(w["idsT"].size),
                   "line_vertices": int(w["idsL"].size), "surface_triangles": int(w["SF"].shape[0]),
                   "num_voxels": args.num_voxels, "stop_extent": args.stop_extent,
                   "leaf_extents": {str(int(e)): int(c) for e, c in zip(ext, cnt)},
                   "iteration": "compute_referenceMesh + patch plane / rim line projection + 1 SLIM iteration "
                                "(20-step CG, flip-avoiding line search), weight_opt 0.001, lamda 0.05 -> 1e16",
                   "fp32_cg_iters": args.cg_iters, "fp32_refinement_cg_iters": args.refine,
                   "quality_iterations": total},
        "fp64_plain": {"iterations_per_s": round(args.steps / r64["secs"], 3),
                       "ms_per_step": round(1e3 * r64["secs"] / args.steps, 4), "min_scaled_jacobian": r64["minsj"],
                       "avg_scaled_jacobian": r64["avgsj"], "scaled_jacobian_quantiles_0.1%_1%": r64["sjq"],
                       "hausdorff_max": r64["hmax"],
                       "hausdorff_mean": r64["hmean"], "operator_kernel_us": round(1e3 * r64["ms_apply"], 2),
                       "energy_gradient_kernel_us": round(1e3 * r64["ms_grad"], 2)},
        "fp32_adjoint": {"min_scaled_jacobian": r32["minsj"], "avg_scaled_jacobian": r32["avgsj"],
                         "scaled_jacobian_quantiles_0.1%_1%": r32["sjq"],
                         "hausdorff_max": r32["hmax"], "hausdorff_mean": r32["hmean"],
                         "max_vertex_diff_vs_fp64": float(np.abs(r32["X"] - r64["X"]).max())},
        "speedup_fp32_vs_fp64": round((args.steps / r32["secs"]) / (args.steps / r64["secs"]), 3),
        "min_sj_abs_diff": abs(r32["minsj"] - r64["minsj"]),
        "hausdorff_diff_over_bbox_diag": abs(r32["hmax"] - r64["hmax"]) / r64["diag"],
        "fp32_fp64_residual_max": max(res32) if res32 else None,
        "fp32_with_refinement": {"cg_iters": refined[0], "refinement_cg_iters": refined[1],
                                 "ms_per_step": round(1e3 * secs_r / args.steps, 4), "min_scaled_jacobian": mn_r,
                                 "hausdorff_max": hmax_r,
                                 "min_sj_abs_diff_vs_fp64": abs(mn_r - r64["minsj"]),
                                 "hausdorff_diff_over_bbox_diag": abs(hmax_r - r64["hmax"]) / r64["diag"],
                                 "fp64_residual_before_after_max": [max(a for a, _ in res32r), max(b for _, b in res32r)]
                                 if res32r else None},
        "fp64_sensitivity": {"note": "FP64 loop from the start perturbed by 1e-12 of the bbox diagonal, 3 seeds: "
                                     "the largest difference from the unperturbed FP64 run",
                             "min_scaled_jacobian": [m for m, _ in pert],
                             "min_sj_abs_diff": max(abs(m - r64["minsj"]) for m, _ in pert),
                             "hausdorff_diff_over_bbox_diag": max(abs(x - r64["hmax"]) for _, x in pert) / r64["diag"]},
        "bbox_diagonal": r64["diag"],
        "first_iteration_errors_fp32_vs_fp64": errors,
        "element_evals_per_s": round(nH / (r32["ms_grad"] * 1e-3), 1),
        "element_evals_note": "one element-eval = energy + gradient of one hex (8 corner tets): the standalone "
                              "FP32 adjoint pass (hexopt_energy_grad's kernel), device time",
        "energy_gradient_roofline": {"bound": "hbm", "achieved": round(grad_achieved, 1), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": round(grad_achieved / HBM_PEAK_GBS, 4),
                                     "bytes_per_hex": round(grad_bytes_per_hex, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "cg32_apply_pipe_kernel (persistent FP32 CG operator pass)",
                     "bytes_per_hex": round(bytes_per_hex32, 2), "algorithmic_bytes": round(bytes_per_hex32 * nH),
                     "traffic_source": "profiles/cg32_apply_pipe_torus_octree1m_hbm_traffic.md" if traffic else None},
    }
    if not args.no_glue:
        out["padding_untangle"] = glue_measure(max(nH // 8, 20000), 5)
    if not args.no_octree:
        out["octree_build"] = octree_measure()
    if not args.no_scale:
        out["scale_box_union"] = scale_measure(args.scale_voxels, args.scale_iterations, solve)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(max(args.num_voxels // 4, 64))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
