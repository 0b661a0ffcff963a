// kernels.h -- device data views and launch entry points of kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "layout.h"

namespace hexopt {

// Device view of a MeshLayout.
struct TileMesh {
  int nV, nH, nTiles, nShared, nSlots;
  const int* conn;
  const int* tile_hex;
  const uint16_t* tile_col;
  const int* tile_lv;
  const int* lv_gid;
  const int* lv_slot;
  const uint8_t* lv_first;
  const uint16_t* lidx;
  const int* sv_gid;
  const int* sv_off;
  const int* sv_slots;
  const uint8_t* tile_corner;
  const int* lv_dst;  // per local vertex: place among all (tile, vertex) pairs, vertex by vertex
  const int* v_dst;   // nV+1: each vertex's run of places
};

enum RefMode { REF_CUBE = 0, REF_GENERAL = 1 };

// Reference operator per corner tet: invs (cube) or B (general), see kernels.hip.
struct RefOp {
  int mode;
  const double* invs;  // [8][nH]
  const double* B;     // [8*9][nH]
};

// Soft constraints (Feature_Constraints + projection terms of Tetralize_Set).
struct SoftDev {
  // per-vertex folded form used by the linear operator and the gradient
  const int* cidx;     // nV: -1 or row in S/b/cline
  const double* S;     // rows*6: sum of constraint Hessian halves (xx yy zz xy xz yz)
  const double* b;     // rows*3: constraint rhs
  const int* cline;    // rows: -1 or feature-line index
  const int* cl;       // nV*2: (row in S/b or -1, feature-line index or -1) per vertex
  // entry lists used for the energy (the reference's own expression per entry)
  int nb, nC, nT, nL;
  const int* b_ids; const double* b_pos;
  const int* c_ids; const double* c_pos;
  const int* t_ids; const double* t_n; const double* t_d;
  const int* lv; const double* lt; const double* lo;
  double lamP, lamC, lamT, lamL;
  // padding glue (exploded mesh): groups of copies of one original vertex, CSR g_off/g_ids
  // (copies in the caller's order), g_of[v] the group of vertex v; g_sum holds 3 sums per
  // group of the vector being applied; rows of vertices with fixed[v] are left out of solves
  int nG;
  const int* g_of;
  const int* g_off;
  const int* g_ids;
  const double* g_sum;
  double lamG;
  const uint8_t* fixed;
};

void launch_ref_cube(const double* X, int nV, const int* conn, int nH, const uint8_t* hflag, int cube, double* invs,
                     int* n_general, hipStream_t st);
void launch_ref_general(const double* X, int nV, const int* conn, int nH, const uint8_t* hflag, int cube, double* B,
                        hipStream_t st, const double* Xvol = nullptr, int nVvol = 0, const int* connvol = nullptr);
// Glue (padding untangle): per-group sums of a SoA vector x (nV stride) into sd.g_sum layout
// gs[3g + d]; the glue energy lamG sum_j |y_j - mean of the other copies|^2 at y = cur + t (dst
// - cur) into out[0]; vertex means of the groups (SoA, nG stride); and the masked-solve helpers
// (fixed rows: invd == 0).
void launch_group_sum(const SoftDev& sd, int nV, const double* x, double* gs, hipStream_t st);
void launch_glue_energy(const SoftDev& sd, int nV, const double* cur, const double* dst, double step, double* part,
                        double* out, hipStream_t st);
void launch_group_mean(const SoftDev& sd, int nV, const double* x, double* mean, hipStream_t st);
// xfix = fixed rows of cur (others 0), x = free rows of cur (others 0)
void launch_glue_split(int nV, const double* invd, const double* cur, double* xfix, double* x, hipStream_t st);
// rhs = free rows of rhs - tmp (fixed rows 0)
void launch_glue_rhs(int nV, const double* invd, double* rhs, const double* tmp, hipStream_t st);
// x = cur on fixed rows
void launch_glue_restore(int nV, const double* invd, const double* cur, double* x, hipStream_t st);
void launch_local_step(const TileMesh& tm, const RefOp& ro, const double* X, double* K, int energy, double M,
                       double exp_f, double prox, const SoftDev& sd, double* rhs, double* invd, double* halo,
                       hipStream_t st);
void launch_apply(const TileMesh& tm, const RefOp& ro, const double* K, const double* p, double* out, double prox,
                  const SoftDev& sd, double* halo, double* part_tile, double* part_fix, double* part_line,
                  hipStream_t st);
void launch_cg_init(int n, const double* rhs, const double* tmp, double* r, double* pb, double* pr, int nb,
                    void* state, const double* invd, double* p, double* pabs, double tol, hipStream_t st);
void launch_cg_alpha(void* state, const double* p1, int n1, const double* p2, int n2, const double* p3, int n3,
                     hipStream_t st);
void launch_cg_update(int n, void* state, double* x, double* r, double* z, const double* p, const double* tmp,
                      const double* invd, double* prr, double* prz, int nb, double* p_out, hipStream_t st);
void launch_cg_zero_if(int n, void* state, double* x, int nb, hipStream_t st);
size_t cg_state_bytes();
void read_cg_state(const void* host_state, int* iters, double* tol_error, int* done);
void launch_max_step(const double* cur, const double* dst, int nV, const int* conn, int nH, double* part, double* out,
                     hipStream_t st);
void launch_energy(const RefOp& ro, const double* cur, const double* dst, double step, int nV, const int* conn, int nH,
                   int energy, double M, double exp_f, const SoftDev& sd, const double* aL, double* part_q,
                   double* part_s, int ns_blocks, double* out2, hipStream_t st);
void launch_step(int n, double* cur, const double* dst, double step, hipStream_t st);
void launch_grad(int prec, const TileMesh& tm, const RefOp& ro, const void* X, void* G, double M, const SoftDev& sd,
                 const double* aL, void* halo, hipStream_t st);
void launch_hvp(int prec, const TileMesh& tm, const RefOp& ro, const void* X, const void* Vd, void* Hv, double M,
                const SoftDev& sd, void* halo, hipStream_t st);
// FP32 optimiser (adjoint formulation of the SLIM step)
void launch_adj_local32(const TileMesh& tm, const RefOp& ro, const double* X, float* K, double M, double prox,
                        const SoftDev& sd, float* rhs, float* invd, float* halo, double* hex_part, int* n_inverted,
                        hipStream_t st);
void launch_apply32(const TileMesh& tm, const RefOp& ro, const float* K, const float* p, float* out, double prox,
                    const SoftDev& sd, float* halo, double* part_tile, double* part_fix, double* part_line,
                    hipStream_t st);
void launch_apply32_f64(const TileMesh& tm, const RefOp& ro, const float* K, const float* p, double* out, double prox,
                        const SoftDev& sd, double* halo, hipStream_t st);
void launch_cg_init32(int n, const float* rhs, const float* tmp, float* r, double* pb, double* pr, int nb, void* state,
                      const float* invd, float* p, double* pabs, double tol, hipStream_t st);
void launch_cg_update32(int n, void* state, float* x, float* r, float* z, const float* p, const float* tmp,
                        const float* invd, double* prr, double* prz, int nb, float* p_out, hipStream_t st);
void launch_cg_zero_if32(int n, void* state, float* x, int nb, hipStream_t st);
// Fused FP32 CG (Eigen's algorithm, FP64 dot products) on the CG operator of launch_apply32:
// launch_cg32_init sets x = 0, r = b, z = b / diag, p_prev = 0 and the state; each
// launch_cg32_iteration is one CG iteration reading z and p_prev, writing p_new and the
// updated x, r, z (zout may alias z); k is the CG iteration and last marks the final one (the
// operator pass of cube references forms the previous iteration's beta step, so the one-block
// beta pass runs only after the last update). Every tile writes E p of each of its vertices as a float4
// at lv_dst (halo holds one float4 per (tile, vertex) pair); the update sums vertex v's run
// v_dst[v] .. v_dst[v+1]. iso lists the vertices without hexes.
void launch_cg32_init(int nV, int nL, const SoftDev& sd, double prox, const float* b, const float* invd, float* x,
                      float* r, float* z, float* pprev, void* state, double* part_b, double* part_rz, double* part_q,
                      double tol, hipStream_t st);
// The operator pass of one iteration (the element part E p, the p.E p partials, p and the
// a_L rows of the line unknowns). halo holds a float4 per (tile, vertex) pair. Returns the number of
// partials, positive when the pass also set p of the vertices without hexes (iso) and
// negative when that is left to the caller.
struct UpdParts;
int launch_cg32_apply(const TileMesh& tm, const RefOp& ro, const float* K, const SoftDev& sd, const int* iso, int niso,
                      const float* z, const float* pprev, float* pnew, float* out, float* halo, void* state,
                      double* part, hipStream_t st, const UpdParts* fold = nullptr);
// The operator pass as launch_cg32_iteration runs it from the second CG iteration on: every
// block first forms CG iteration k's beta step from the update pass's partials (the five
// arrays, one entry per resident update block).
int launch_cg32_apply_folded(const TileMesh& tm, const RefOp& ro, const float* K, const SoftDev& sd, const int* iso,
                             int niso, const float* z, const float* pprev, float* pnew, float* out, float* halo,
                             void* state, double* part, const double* part_rr, const double* part_rz,
                             const double* part_qz, const double* part_bzq, const double* part_qq, int k,
                             hipStream_t st);
void launch_cg32_iteration(const TileMesh& tm, const RefOp& ro, const float* K, double prox, const SoftDev& sd,
                           const int* iso, int niso, const float* z, const float* pprev, float* pnew,
                           float* out, float* halo, const float* invd, float* x, float* r, float* zout, void* state,
                           double* part_tile, double* part_rr, double* part_rz, double* part_qz, double* part_bzq,
                           double* part_qq, int k, int last, hipStream_t st);
void launch_residual64(int n, const float* b, const double* Ld, float* r32, double* part, int nb, double* out,
                       hipStream_t st);
void launch_f2d(int n, const float* a, double* b, hipStream_t st);
void launch_d2f(int n, const double* a, float* b, hipStream_t st);
// b = -g / 2 on the vertex rows of g (the FP64 adjoint gradient, in place) and the a_L rows at
// a_L = 0: the FP64 right-hand side of the direction system (3 nV + nL entries)
void launch_rhs64_from_grad(int nV, double* g, const SoftDev& sd, const double* X, hipStream_t st);
// out2 = { |b - L d|^2, |b|^2 } in FP64; r32 (nullable) = FP32 copy of b - L d
void launch_residual64b(int n, const double* b, const double* Ld, float* r32, double* part_r, double* part_b, int nb,
                        double* out2, hipStream_t st);
void launch_add_inplace(int n, float* d, const float* e, int nb, hipStream_t st);
void launch_max_step_dir32(const double* cur, const float* d, int nV, const int* conn, int nH, double* part,
                           double* out, hipStream_t st);
void launch_energy_dir32(const RefOp& ro, const double* cur, const float* d, double step, int nV, const int* conn,
                         int nH, double M, const SoftDev& sd, const double* aL, double* part_q, double* part_s,
                         int ns_blocks, double* out2, hipStream_t st);
void launch_step_dir32(int n, double* cur, const float* d, double step, hipStream_t st);
void launch_sum_partials(const double* part, int n, double* out, hipStream_t st);

void launch_scaled_jacobian(const double* X, int nV, const int* conn, int nH, double* per_hex, double* pmin,
                            double* psum, double* out3, hipStream_t st);

}  // namespace hexopt
