/* hexopt.h -- C ABI of the MI355X hex-mesh optimisation path (libhexopt.so).
 *
 * Drop-in boundary for the reference's optimiser inner loop. Plain pointers and sizes only;
 * every array is host memory in the caller's numbering unless a function says otherwise.
 * Vertex arrays are nV x 3 row-major doubles (one row per vertex, like Tetralize_Set::V);
 * hexes are nH x 8 int32 in the reference's Hybrid::vs corner order; the 8 corner tets of
 * hex h are tets 8h..8h+7 in global_types.h:163-173 order (exactly how deformation(),
 * smoothing() and feature_alignment() build Tetralize_Set::T).
 *
 * Errors: every int-returning function returns HEXOPT_OK (0) or a negative code and leaves a
 * message readable with hexopt_last_error(). Nothing throws across the ABI. A context owns
 * one HIP stream on one device; calls on one context must not overlap (not thread-safe per
 * context, like the reference's single-threaded optimiser).
 */
#ifndef HEXOPT_H
#define HEXOPT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEXOPT_OK 0
#define HEXOPT_EINVAL (-1)
#define HEXOPT_EHIP (-2)
#define HEXOPT_EUNSUPPORTED (-3)

typedef struct hexopt_ctx hexopt_ctx;

/* SLIM_ENERGY values (global_types.h:56-64) */
enum { HEXOPT_ARAP = 0, HEXOPT_LOG_ARAP = 1, HEXOPT_SYMMETRIC_DIRICHLET = 2, HEXOPT_CONFORMAL = 3,
       HEXOPT_EXP_CONFORMAL = 4, HEXOPT_EXP_SYMMETRIC_DIRICHLET = 5 };

/* Which position buffer: REST = Tetralize_Set::V (defines the reference operator),
 * CURRENT = Tetralize_Set::UV (the iterate, SLIMData::V_o). */
enum { HEXOPT_REST = 0, HEXOPT_CURRENT = 1 };

typedef struct {
  int energy;          /* Tetralize_Set::energy_type */
  double weight_opt;   /* optimization::weight_opt: M = weight_opt, proximal = 1e-5 * weight_opt
                          (slim_m.cpp:1118-1122) */
  double exp_factor;   /* exponential energies; slim_precompute forces 1.0 (slim_m.cpp:1125) */
  int projection;      /* Tetralize_Set::projection: adds lamda_projection * |x_b - bc|^2 */
  double lamda_projection;
  double tet_mass;     /* per-tet M of the energy and the operator; 0: weight_opt. The igl::slim
                          variant that smooths the padded mesh (igl/slim.cpp:1241-1247, called by
                          optimization::slim_opt_igl) uses weight_opt / 2 with the same proximal
                          term 1e-5 * weight_opt */
} hexopt_params;

/* Soft constraints of Tetralize_Set / Feature_Constraints (global_types.h:583-650).
 * Vertex ids in the caller's numbering; positions/normals/axes are n x 3 row-major. */
typedef struct {
  int nb; const int32_t* b; const double* bc;                       /* ts.b, ts.bc */
  int nC; const int32_t* ids_C; const double* C; double lamda_C;     /* corners */
  int nT; const int32_t* ids_T; const double* normal_T; const double* dis_T; double lamda_T; /* planes */
  int nL; const int32_t* ids_L; const double* Axa_L; const double* origin_L; double lamda_L; /* lines */
} hexopt_constraints;

/* Per-iteration record of one SLIM iteration (slim_solve2 fixed branch, slim_m.cpp:1269-1288) */
typedef struct {
  double energy;        /* SLIMData::energy (normalised by mesh_area) */
  double e_quality;     /* SLIMData::engery_quality (unnormalised) */
  double e_soft;        /* SLIMData::engery_soft */
  double step_bound;    /* min(1, 0.8 t*) of flip_avoiding_line_search */
  double step;          /* accepted step of igl::line_search */
  int cg_iters;         /* iterations of the Eigen-style CG */
  double cg_error;      /* its relative residual estimate */
} hexopt_iter_log;

/* Creates a context on HIP device `device` for a hex mesh. V gives the positions used to
 * order vertices and hexes along a Morton curve (any representative positions will do).
 * Replaces the set-up part of slim_precompute (slim_m.cpp:1095-1132). */
hexopt_ctx* hexopt_create(int nV, const double* V, int nH, const int32_t* hexes, int device);
const char* hexopt_create_error(void);
void hexopt_destroy(hexopt_ctx* ctx);
const char* hexopt_last_error(hexopt_ctx* ctx);

int hexopt_set_positions(hexopt_ctx* ctx, int which, const double* V);
int hexopt_get_positions(hexopt_ctx* ctx, int which, double* V);
/* Device-side copy between the two position buffers (ts.V = ts.UV between slim_m_opt calls,
 * grid_hex_meshing.cpp:2956). */
int hexopt_copy_positions(hexopt_ctx* ctx, int from, int to);
/* Device-resident positions in the library's own layout (SoA x[d*nV+v], Morton numbering);
 * for callers that keep the iterate in HBM. */
double* hexopt_device_positions(hexopt_ctx* ctx, int which);

/* compute_referenceMesh (global_functions.cpp:3735-3806) on the REST positions followed by
 * igl::grad_ref(..., uniform=true) (grad.cpp:372-478). hflag: nullable nH bytes
 * (compute_referenceMesh's H_flag); cube: its `cube` argument. */
int hexopt_reference(hexopt_ctx* ctx, const uint8_t* hflag, int cube);

/* Soft constraints used by the next hexopt_slim_solve / energy / gradient calls. */
int hexopt_set_constraints(hexopt_ctx* ctx, const hexopt_constraints* cs);

/* optimization::slim_m_opt(ts, iters, -1) (optimization.cpp:883-900) minus the set-up:
 * slim_precompute's energy followed by `iters` SLIM iterations starting from CURRENT; the
 * result replaces CURRENT. log: nullable, `iters` entries. E/Eq/Es: nullable outputs
 * (optimization::energy, engery_quality, engery_soft). aL: nullable, nL entries (a_L). */
int hexopt_slim_solve(hexopt_ctx* ctx, const hexopt_params* prm, int iters, hexopt_iter_log* log, double* E,
                      double* Eq, double* Es, double* aL);

/* Padding untangle by gluing an exploded mesh (optimization::glueInternalNodes,
 * optimization.cpp:552-656). ctx holds the exploded mesh: hex h is hex h of the original mesh
 * (nO vertices, hexes H_orig, nH x 8) with its own copy of each corner; orig_of[v] (caller
 * numbering, nV entries) is the original vertex that vertex v copies, and corner k of context
 * hex h must copy corner k of original hex h. The copies of an original vertex flagged in
 * fixed_orig are pinned where ctx has them (known_value_post, slim_m.cpp:403-475); the others
 * are glued by lamda_glue |x_j - mean of the other copies|^2 (add_soft_constraints,
 * slim_m.cpp:611-645). Sets the reference operator: the original mesh's reference shapes
 * (V_orig, compute_referenceMesh with cube = false) scaled to the volumes of ctx's REST
 * positions (grad_tet_ref, grad.cpp:372-478). ctx must carry no other constraints. */
int hexopt_glue_setup(hexopt_ctx* ctx, int nO, const double* V_orig, const int32_t* H_orig, const int32_t* orig_of,
                      const uint8_t* fixed_orig);

typedef struct {
  double energy, e_quality, e_soft; /* after the iteration (the glue term is in e_soft) */
  double lamda_glue;                /* the lamda_glue the iteration ran with */
  double min_sj, avg_sj;            /* scaled Jacobian of the glued mesh after it */
  double step;                      /* accepted line-search step */
  int cg_iters;
} hexopt_glue_log;

/* slim_solve2's gluing branch (slim_m.cpp:1191-1268): SLIM iterations (energy from prm, the
 * reference uses CONFORMAL) with lamda_glue starting at 1e4 and set after each iteration to
 * min(1e4 max(E_quality / E_soft lamda_glue, 1), 1e16), until the glued mesh (each original
 * vertex at the mean of its copies) has a positive minimum scaled Jacobian or max_iters
 * (50 in the reference) iterations ran. log: nullable, max_iters entries. V_out: nullable,
 * nO x 3, the glued mesh. iters, min_sj: nullable. */
int hexopt_glue_solve(hexopt_ctx* ctx, const hexopt_params* prm, int max_iters, hexopt_glue_log* log, int* iters,
                      double* V_out, double* min_sj);

/* FP32 + adjoint optimiser (no single counterpart in the reference; it performs the same
 * SLIM iteration as hexopt_slim_solve in its adjoint form). Per iteration: one FP32 pass
 * over the elements gives the energy, -dE/dx / 2 (the hand-written adjoint), the SLIM
 * weights K and the Jacobi diagonal; the direction solves L d = -dE/dx / 2 (equal to the
 * SLIM global step taken from CURRENT) by FP32 CG with FP64 dot products. Iterative
 * refinement (refine_iters > 0): the right-hand side b64 = -dE/dx / 2 is recomputed by the
 * FP64 adjoint, the residual b64 - L d is formed with FP64 sums over the stored operator, and
 * an FP32 CG solve of L e = that residual gives d += e (it moves d toward the converged
 * solve; with the reference's truncated 20-step CG and no refinement, d reproduces the FP64
 * path's truncated iterate). The step bound (FP64 cubic), the corner-determinant sign test
 * and the line-search energies (per tet and summed) are FP64; the line search is the
 * reference's bisection. Symmetric Dirichlet only. */
typedef struct {
  int cg_iters;        /* FP32 CG iterations of the direction solve (the reference uses 20) */
  int refine_iters;    /* FP32 CG iterations of the FP64-residual correction (0: none) */
  int residual_every;  /* the FP64 residual is measured on every residual_every-th iteration and
                          the last (0: never); a correction always measures the one it starts
                          from */
} hexopt_opt32_options;

typedef struct {
  double energy, e_quality, e_soft, step_bound, step;
  int cg_iters;
  double cg_error;     /* relative residual estimate of the FP32 solve */
  double residual64;   /* |b64 - L d| / |b64| in FP64 before the correction (-1: not measured) */
  int inverted_tets;   /* corner tets with negative determinant at the iterate (FP64 test) */
  double residual64_refined; /* the same after the correction (-1: no correction, or not an
                                iteration residual_every measures) */
} hexopt_opt32_log;

int hexopt_opt32_solve(hexopt_ctx* ctx, const hexopt_params* prm, const hexopt_opt32_options* opt, int iters,
                       hexopt_opt32_log* log, double* E);

/* Diagnostic view of the SLIM weights (no reference counterpart; used to state the FP32
 * path's error): runs the local step at CURRENT in FP64 (update_weights_and_closest_rotations,
 * slim_m.cpp:229-377) or in FP32 (hexopt_opt32_solve's element pass) and returns, per hex in
 * caller order and per corner tet c, the 6 entries M W^T W (xx, yy, zz, xy, xz, yz) scaled by
 * the tet's squared reference factor: K[h*48 + c*6 + j]. Symmetric Dirichlet only. */
int hexopt_debug_operator(hexopt_ctx* ctx, const hexopt_params* prm, int precision, double* K);

/* Average duration (ms, HIP events on the context's stream) of one launch of a kernel:
 * 0 FP64 CG operator tile pass, 1 FP32 fused CG operator pass (cg32_apply, forming the previous
 * iteration's beta step as it does inside the CG loop), 2 FP64 local step
 * tile pass, 3 FP32 adjoint/local tile pass, 4 the FP32 energy + gradient pass of
 * hexopt_energy_grad (CURRENT rounded to FP32), 5 the same in FP64. Modes 0-3 need a prior
 * solve of the same precision (operands resident). */
int hexopt_time_kernel(hexopt_ctx* ctx, int which, int reps, double* ms_per_launch);

/* Unnormalised SLIM energy of CURRENT: quality part and soft part (compute_energy_total,
 * slim_m.cpp:716-733, with a_L = aL or 0 when aL is NULL). */
int hexopt_energy(hexopt_ctx* ctx, const hexopt_params* prm, const double* aL, double* Eq, double* Es);

/* Adjoint of the symmetric-Dirichlet SLIM energy (no counterpart in the reference, which
 * has no gradient): G = dE/dx at X, Hv = d^2E/dx^2 v. precision 64: X, v, G, Hv are host
 * double nV x 3; precision 32: host float nV x 3 with FP32 arithmetic. */
int hexopt_energy_grad(hexopt_ctx* ctx, const hexopt_params* prm, int precision, const void* X, const double* aL,
                       void* G);
int hexopt_hvp(hexopt_ctx* ctx, const hexopt_params* prm, int precision, const void* X, const void* v, void* Hv);

/* scaled_jacobian (global_functions.cpp:2309-2358) of the `which` positions: min and mean
 * of the per-hex minimum corner scaled Jacobian; per_hex nullable (caller order). */
int hexopt_scaled_jacobian(hexopt_ctx* ctx, int which, double* min_j, double* avg_j, double* per_hex);

/* Diagnostic view of the global system of the next SLIM iteration at CURRENT (no reference
 * counterpart; used by the parity tests): runs the local step, then returns rhs, the inverse
 * Jacobi diagonal and L p_in. Vectors are flattened like the reference's unknowns
 * [x(0..nV) y(0..nV) z(0..nV) a_L(0..nL)] in the caller's numbering; outputs nullable. On a
 * glue context L carries the glue rows at the current lamda_glue (1e4 after hexopt_glue_setup)
 * and inv_diag is 0 on pinned rows. */
int hexopt_debug_system(hexopt_ctx* ctx, const hexopt_params* prm, const double* p_in, double* rhs, double* inv_diag,
                        double* Lp);

/* ---- closest point on the input surface (igl::AABB + signed_distance_pseudonormal) ----
 * A surface is built once per input triangle mesh (build_aabb_tree,
 * grid_hex_meshing.cpp:4231-4248): the tree has igl::AABB::init's exact shape so the face
 * chosen on ties is the reference's. FN/VN/FEN are nullable (all or none): given, they are
 * used as is (igl's per_face_normals, per_vertex_normals(ANGLE), per_edge_normals(UNIFORM)
 * with FEN[f*9 + x*3 ..] = EN(EMAP(f + x*nf))); absent, the library computes them. */
typedef struct hexopt_surface hexopt_surface;
hexopt_surface* hexopt_surface_create(int nv, const double* V, int nf, const int32_t* F, const double* FN,
                                      const double* VN, const double* FEN, int device);
const char* hexopt_surface_error(void);
const char* hexopt_surface_last_error(hexopt_surface* s);
void hexopt_surface_destroy(hexopt_surface* s);
int hexopt_surface_normals(hexopt_surface* s, double* FN, double* VN, double* FEN);
/* Host-only: the three normal sets of build_aabb_tree for a triangle mesh (FN nf x 3,
 * VN nv x 3, FEN nf x 9 as above); needs no GPU. */
int hexopt_compute_surface_normals(int nv, const double* V, int nf, const int32_t* F, double* FN, double* VN,
                                   double* FEN);
/* igl::signed_distance_pseudonormal(P, V, F, tree, FN, VN, EN, EMAP, S, I, C, N)
 * (signed_distance.h:80-92, signed_distance.cpp:186-257): nq x 3 queries; outputs nullable. */
int hexopt_surface_closest(hexopt_surface* s, int nq, const double* P, double* S, int32_t* I, double* C, double* N);
/* projection_smooth (grid_hex_meshing.cpp:3760-3781) on the device: the plane vertices of the
 * context's constraints (ids_T) are projected at CURRENT; normal_T = pseudonormal,
 * dis_T = normal_T . closest point, and the plane weight becomes lamda_T. */
int hexopt_project_planes(hexopt_ctx* ctx, hexopt_surface* s, double lamda_T);

/* The patch loop of dirty_graph_projection (grid_hex_meshing.cpp:4034-4081) on the device:
 * plane vertex ids_T[i] (of hexopt_set_constraints, at CURRENT) is projected onto surface
 * patches[patch_of_T[i]] -- its patch's own tree (md.tri_Trees[i] for Qfg.Rs[i]), or the whole
 * input surface for a dirty vertex -- giving normal_T = pseudonormal and dis_T = normal_T .
 * closest point; the plane weight becomes lamda_T. patch_of_T has nT entries in ids_T order,
 * each in [0, npatch); it is checked and grouped again only when it changes between calls.
 * patch_of_T NULL reuses the assignment of the previous call (same npatch; the loop's usual
 * case, which then costs the host nothing per call); EINVAL when there is none. */
int hexopt_project_patches(hexopt_ctx* ctx, int npatch, hexopt_surface* const* patches, const int32_t* patch_of_T,
                           double lamda_T);

/* Metro Hausdorff distance between two triangle meshes with vertex sampling only, as the
 * reference's compute(const Mesh&, const Mesh&, bbox_diagonal, max, mean)
 * (metro_hausdorff.cpp:358-505; vcg Sampling, extern/vcg/sampling.h): the box of all vertices
 * of both meshes grown by 2% of its diagonal gives *bbox_diagonal; every face-referenced vertex
 * of each mesh is measured against the other mesh (exact point-triangle distance, samples at
 * or beyond the diagonal dropped); *max_dist is the larger of the two directions' maxima and
 * *mean_dist the larger of their means. On failure hexopt_surface_error() says why. */
int hexopt_metro_hausdorff(int nv0, const double* V0, int nf0, const int32_t* F0, int nv1, const double* V1, int nf1,
                           const int32_t* F1, int device, double* bbox_diagonal, double* max_dist, double* mean_dist);

/* ---- occupancy (octree subdivision predicate, +z ray parity, inside test of hexes) ----
 * All arrays are host arrays; cells are n x 4 int32 (x, y, z, extent) in grid units. */

/* should_subdivide of octree_mesh (grid_hex_meshing.cpp:502-517) for a batch of cells (one
 * BFS level of OctreeGrid::subdivide, octree.cpp:648-690): split[i] = 1 iff extent >
 * stop_extent and the closed box [mesh_transform + origin + voxel_size*(x,y,z),
 * + voxel_size*extent] touches the closed box of at least one facet
 * (GEO::MeshFacetsAABB::compute_bbox_facet_bbox_intersections). mesh_transform may be NULL
 * (compute_octree's predicate, voxelization.cpp:371-383, with stop_extent = 1). */
int hexopt_occ_classify_cells(hexopt_surface* s, int n, const int32_t* cells, const double* origin, double voxel_size,
                              const double* mesh_transform, int stop_extent, uint8_t* split);
/* compute_sign(M, aabb, VoxelGrid&) (voxelization.h:221-272): voxel centres
 * (i + 0.5) * spacing + origin; inside[(z*ny + y)*nx + x] = 1 where the signed count of +z
 * ray crossings below the centre is negative. ny <= 65535. */
int hexopt_occ_voxel_parity(hexopt_surface* s, int nx, int ny, int nz, const double* origin, double spacing,
                            uint8_t* inside);
/* compute_sign(M, aabb, OctreeGrid&, origin, spacing) (voxelization.cpp:101-163): the
 * octree cells' "inside" attribute (1.0f / 0.0f) for the given cells. */
int hexopt_occ_cell_parity(hexopt_surface* s, int n, const int32_t* cells, const double* origin, double spacing,
                           float* inside);
/* clean_hex_mesh's inside test (grid_hex_meshing.cpp:1937-1952 -> points_inside_mesh,
 * global_functions.cpp:4024-4048): pseudonormal signed distance of every hex's bounding-box
 * centre at REST or CURRENT (signed_dis, nH, nullable) and H_flag[h] = signed_dis < 0
 * (nH bytes, nullable), in the caller's hex order. */
int hexopt_hex_inside(hexopt_ctx* ctx, hexopt_surface* s, int which, double* signed_dis, uint8_t* H_flag);

/* ---- octree subdivision ----------------------------------------------------------------
 * Replaces, in octree_mesh (grid_hex_meshing.cpp:490-562), OctreeGrid(grid) (octree.cpp:17-117)
 * followed by OctreeGrid::subdivide(should_subdivide, graded, paired) (octree.cpp:648-690,
 * splitCell :501-593, makeCellGraded :598-627, pairing :569-589) and the leaf-to-hex export.
 * grid: cells per axis in voxel units, powers of two (next_pow2 of the surface extent);
 * should_subdivide is hexopt_occ_classify_cells' predicate with the same origin, voxel_size,
 * mesh_transform (nullable) and stop_extent. Runs level-synchronously on the surface's
 * device. The leaf set, the node positions and each leaf's corner positions equal the
 * reference's; leaves and nodes are numbered differently (leaves level by level, x fastest
 * within a level; nodes in lattice order). NULL on a bad grid, a level above 2^31 cells or a
 * device error: hexopt_octree_error() says why. */
typedef struct hexopt_octree hexopt_octree;
hexopt_octree* hexopt_octree_build(hexopt_surface* s, const int32_t* grid3, const double* origin, double voxel_size,
                                   const double* mesh_transform, int stop_extent, int graded, int paired);
const char* hexopt_octree_error(void);
void hexopt_octree_destroy(hexopt_octree* o);
/* out4 = { leaves, nodes, levels holding leaves, predicate evaluations }; *ms (nullable) the
 * device time of the last subdivision (build or refine) */
int hexopt_octree_sizes(const hexopt_octree* o, int64_t* out4, double* ms);
/* The targeted OctreeGrid::subdivide(should_subdivide, tb_subdivided_cells, graded, paired)
 * (octree.cpp:691-729) that the mesher's refinement loop runs on its existing octree
 * (grid_hex_meshing.cpp:518-521): each of the n listed cells (x, y, z, extent rows) that is a
 * leaf and passes the predicate at this stop_extent is split, with its cascades; its children
 * are not tested. Cells that are not leaves are skipped, as there; a row that is not a cell of
 * the grid is an error (HEXOPT_EINVAL). The octree's leaves, nodes and hexes are then those of
 * the refined tree (hexopt_octree_sizes / _copy). s must be on the octree's device. */
int hexopt_octree_refine(hexopt_octree* o, hexopt_surface* s, int stop_extent, int n, const int32_t* cells);
/* leaves[n][4] (x, y, z, extent), nodes[m][3] (integer positions), V[m][3]
 * (mesh_transform + origin + voxel_size * position, octree_mesh's mo.V) and H[n][8] (node
 * ids of each leaf's corners in the reference's corner order, octree_mesh's mo.Hs); any may
 * be NULL. */
int hexopt_octree_copy(const hexopt_octree* o, int32_t* leaves, int32_t* nodes, double* V, int32_t* H);

/* ---- feature-curve projection (LINE branch of dirty_graph_projection,
 * grid_hex_meshing.cpp:3961-4026, with point_line_projection, global_functions.cpp:3454-3465) ----
 * A curve set is the input surface's feature curves (mf.curve_vs / mf.circles): curve c is
 * the surface vertex ids curve_vs[curve_off[c] .. curve_off[c+1]) (at least one), closed
 * when circles[c]. V is the surface's nv x 3 vertices. */
typedef struct hexopt_curves hexopt_curves;
hexopt_curves* hexopt_curves_create(int nv, const double* V, int ncurves, const int32_t* curve_off,
                                    const int32_t* curve_vs, const uint8_t* circles, int device);
const char* hexopt_curves_error(void);
void hexopt_curves_destroy(hexopt_curves* cv);
/* For the nL LINE vertices ids_L of the context (positions at REST or CURRENT) on curves
 * curve_L (fc.V_ids): origin_L = closest point of the nearest curve segment (ties to the
 * lower segment index) and Axa_L = that segment's unit direction, both nL x 3 -- the
 * fc.origin_L / fc.Axa_L rows of the reference. */
int hexopt_project_lines(hexopt_ctx* ctx, hexopt_curves* cv, int which, int nL, const int32_t* ids_L,
                         const int32_t* curve_L, double* origin_L, double* Axa_L);
/* dirty_graph_projection's LINE branch feeding the next slim_m_opt, on the device: projects
 * the context's line vertices (ids_L of hexopt_set_constraints, at CURRENT) onto their
 * curves curve_L (nL entries, in ids_L order) and replaces origin_L / Axa_L and lamda_L in
 * the folded constraint rows, as hexopt_set_constraints would with the projected values.
 * Line vertices must carry no corner or plane constraint (HEXOPT_EUNSUPPORTED otherwise). */
int hexopt_refresh_lines(hexopt_ctx* ctx, hexopt_curves* cv, const int32_t* curve_L, double lamda_L);

/* No counterpart in the reference: the one entry point that allocates no device memory, for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). It evaluates hexopt_scaled_jacobian's scaled Jacobian with the same
 * kernels on device arrays the caller owns, in the caller's numbering: X the nV positions as
 * x[d*nV + v] (d = 0..2), conn the nH hexes as conn[k*nH + h] (corner k in the reference's
 * Hybrid::vs order, every id in [0, nV)); per_hex (nullable, nH doubles) each hex's minimum
 * corner scaled Jacobian; min_sum two doubles: the mesh minimum and the sum of the per-hex
 * minima (the mean is min_sum[1] / nH). Every pointer is device memory; the work is queued on
 * `stream` (a hipStream_t, NULL for the default stream) and nothing waits for it. The block
 * partials live in a fixed device array of 1 MiB per GPU, held until the process ends, which
 * caps nH at 16,777,216 (larger meshes: HEXOPT_EINVAL). Calls that overlap on one GPU share
 * that array and race. Errors: hexopt_noalloc_error(). */
int hexopt_scaled_jacobian_noalloc(int device, int nV, const double* X, int nH, const int32_t* conn, double* per_hex,
                                   double* min_sum, void* stream);
const char* hexopt_noalloc_error(void);

/* Blocks until the context's stream is idle. */
int hexopt_synchronize(hexopt_ctx* ctx);

/* Layout facts for tests/benchmarks: [nTiles, nShared, nSlots, max_colours, ref_mode,
 * corner_tiles]. corner_tiles counts the tiles whose scatter runs by corner index (no vertex
 * at the same corner index of two of the tile's hexes); the others run by hex colour. */
int hexopt_layout_info(hexopt_ctx* ctx, int32_t* out6);

#ifdef __cplusplus
}
#endif
#endif
