// hexopt.cpp -- host side of libhexopt: the context, the SLIM iteration driver and the C ABI
// declared in include/hexopt.h.
//
// The iteration follows slim_solve2's fixed-iteration branch (slim_m.cpp:1269-1288):
//   local step (update_weights_and_closest_rotations)      -> launch_local_step
//   global step (solve_weighted_arap: Eigen CG, Jacobi,    -> launch_apply + cg kernels,
//                tol 1e-8, <= 20 iterations, warm start)     all state on the device
//   igl::flip_avoiding_line_search + igl::line_search      -> max_step / energy kernels,
//                                                             the accept test on the host
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hexopt.h"
#include "kernels.h"
#include "layout.h"
#include "bvh.h"
#include "occupancy.h"
#include "lines.h"
#include "octree.h"

namespace hexopt {
namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t ensure(size_t nb) {
    if (nb <= bytes && p) return hipSuccess;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    if (nb == 0) nb = 8;
    hipError_t e = hipMalloc(&p, nb);
    if (e == hipSuccess) bytes = nb;
    return e;
  }
  template <class T> T* as() const { return (T*)p; }
};

}  // namespace
}  // namespace hexopt

using namespace hexopt;

struct hexopt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  MeshLayout L;
  std::string err;
  // layout on the device
  DevBuf d_conn, d_tile_hex, d_tile_col, d_tile_lv, d_lv_gid, d_lv_slot, d_lv_first, d_lidx, d_tile_corner, d_sv_gid,
      d_sv_off, d_sv_slots, d_lv_dst, d_v_dst;
  DevBuf d_iso;  // vertices without hexes
  DevBuf d_torder;       // plane queries sorted by vertex (Morton) order
  DevBuf d_cgsave;       // CG state of the last direction solve, read back with the line search
  int n_iso = 0;
  TileMesh tm{};
  // positions (SoA, Morton numbering)
  DevBuf d_rest, d_cur;
  // reference operator
  DevBuf d_invs, d_B, d_hflag, d_count;
  int ref_mode = -1;
  // SLIM state
  DevBuf d_K, d_rhs, d_invd, d_x, d_r, d_z, d_p, d_tmp, d_halo, d_cg, d_part[8], d_scal;
  // constraints
  DevBuf d_cidx, d_cl, d_S, d_b, d_cline, d_bids, d_bpos, d_cids, d_cpos, d_tids, d_tn, d_td, d_lv, d_lt, d_lo, d_rhs_a,
      d_invd_a, d_aL;
  SoftDev sd{};
  int nL = 0;
  // padding glue (hexopt_glue_setup): groups of copies, pinned copies, the original mesh
  bool glue = false;
  int nO = 0;
  double lamG = 0;
  DevBuf d_g_of, d_g_off, d_g_ids, d_gsum, d_gfixed, d_Vo, d_Ho, d_xfix, d_gmean;
  std::vector<double> h_lt;  // nL*3 (caller copy for rhs/diag of the a_L rows)
  // adjoint scratch
  DevBuf d_X, d_G, d_v;
  // tangent-plane refresh: folded rows without the plane terms
  DevBuf d_S0, d_b0;
  DevBuf d_trow;  // per constraint row: its plane (T) index, or -1
  int nrows = 0;
  bool planes_refreshable = false;
  std::vector<int> h_tids;          // plane rows' vertices (Morton numbering), ids_T order
  // per-patch projection: the last patch assignment, its rows grouped by patch (each group in
  // vertex order) and the groups' offsets
  std::vector<int32_t> h_patch;
  std::vector<int> patch_off;
  std::vector<PatchSeg> h_segs;  // kept until the next call: the source of an async upload
  DevBuf d_porder, d_pC, d_pstat, d_segs;
  bool lines_refreshable = false;
  // FP32 optimiser state
  DevBuf f_K, f_rhs, f_invd, f_d, f_r, f_z, f_p, f_p2, f_tmp, f_e, f_halo, d_Ld64, d_halo64b, d_ninv, d_b64;
};

namespace {

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
      return HEXOPT_EHIP;                                                                   \
    }                                                                                       \
  } while (0)

int fail(hexopt_ctx* c, int code, const char* msg) {
  c->err = msg;
  return code;
}

template <class T>
int upload(hexopt_ctx* c, DevBuf& b, const std::vector<T>& v) {
  HIPCHK(c, b.ensure(v.size() * sizeof(T)));
  if (!v.empty()) HIPCHK(c, hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return HEXOPT_OK;
}

// caller rows (nV x 3) -> SoA Morton order
void to_soa(const MeshLayout& L, const double* V, std::vector<double>& out) {
  out.resize((size_t)3 * L.nV);
  for (int i = 0; i < L.nV; ++i)
    for (int d = 0; d < 3; ++d) out[(size_t)d * L.nV + i] = V[(size_t)L.vperm[i] * 3 + d];
}
void from_soa(const MeshLayout& L, const double* S, double* V) {
  for (int i = 0; i < L.nV; ++i)
    for (int d = 0; d < 3; ++d) V[(size_t)L.vperm[i] * 3 + d] = S[(size_t)d * L.nV + i];
}

int nblocks(long n) { long b = (n + 255) / 256; if (b < 1) b = 1; if (b > 2048) b = 2048; return (int)b; }

RefOp refop(hexopt_ctx* c) { return RefOp{c->ref_mode, c->d_invs.as<double>(), c->d_B.as<double>()}; }

// Scratch sized for the current mesh and constraints.
int ensure_state(hexopt_ctx* c) {
  const int nV = c->L.nV, nH = c->L.nH;
  const size_t n = (size_t)3 * nV + c->nL;
  HIPCHK(c, c->d_K.ensure((size_t)48 * nH * sizeof(double)));
  for (DevBuf* b : {&c->d_rhs, &c->d_invd, &c->d_x, &c->d_r, &c->d_z, &c->d_p, &c->d_tmp})
    HIPCHK(c, b->ensure(n * sizeof(double)));
  HIPCHK(c, c->d_halo.ensure((size_t)6 * (c->L.nSlots + 1) * sizeof(double)));
  HIPCHK(c, c->d_cg.ensure(cg_state_bytes()));
  const size_t np = (size_t)std::max(std::max(std::max(c->L.nTiles, nblocks(n)), (nH + 255) / 256 + 1),
                                     std::max(2048, (nV + c->nL + 255) / 256 + 1)) + 64;
  const size_t npf = (size_t)(c->L.nShared + 255) / 256 + 1;
  for (int i = 0; i < 8; ++i) HIPCHK(c, c->d_part[i].ensure(std::max(np, npf) * sizeof(double)));
  HIPCHK(c, c->d_scal.ensure(16 * sizeof(double)));
  return HEXOPT_OK;
}

SoftDev soft_with(hexopt_ctx* c, const hexopt_params* prm);

// per-tet M: weight_opt unless the caller gives tet_mass (the igl::slim variant's weight_opt / 2)
double tet_mass(const hexopt_params* prm) { return prm->tet_mass > 0 ? prm->tet_mass : prm->weight_opt; }

int energy_now(hexopt_ctx* c, const hexopt_params* prm, const double* dst, double step, const double* aL, double* eq,
               double* es) {
  const int nsb = nblocks(c->sd.nb + c->sd.nC + c->sd.nT + c->sd.nL);
  double* out = c->d_scal.as<double>();
  launch_energy(refop(c), c->d_cur.as<double>(), dst, step, c->L.nV, c->tm.conn, c->L.nH, prm->energy, tet_mass(prm),
                prm->exp_factor, c->sd, aL, c->d_part[0].as<double>(), c->d_part[1].as<double>(), nsb, out, c->stream);
  double h[3] = {0, 0, 0};
  if (c->glue)
    launch_glue_energy(soft_with(c, prm), c->L.nV, c->d_cur.as<double>(), dst, step, c->d_part[2].as<double>(), out + 2,
                       c->stream);
  HIPCHK(c, hipMemcpyAsync(h, out, (c->glue ? 3 : 2) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *eq = h[0];
  *es = h[1] + h[2];
  return HEXOPT_OK;
}

SoftDev soft_with(hexopt_ctx* c, const hexopt_params* prm) {
  SoftDev s = c->sd;
  s.lamP = prm->lamda_projection;
  if (!prm->projection) s.nb = 0;
  if (c->glue) {
    s.nG = c->nO;
    s.g_of = c->d_g_of.as<int>();
    s.g_off = c->d_g_off.as<int>();
    s.g_ids = c->d_g_ids.as<int>();
    s.g_sum = c->d_gsum.as<double>();
    s.lamG = c->lamG;
    s.fixed = c->d_gfixed.as<uint8_t>();
  }
  return s;
}

}  // namespace

extern "C" {

const char* hexopt_create_error(void) { return g_create_error.c_str(); }

hexopt_ctx* hexopt_create(int nV, const double* V, int nH, const int32_t* hexes, int device) {
  g_create_error.clear();
  if (nV <= 0 || nH < 0 || !V || (nH > 0 && !hexes)) { g_create_error = "invalid mesh arguments"; return nullptr; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { g_create_error = "no HIP device"; return nullptr; }
  if (device < 0 || device >= ndev) { g_create_error = "device index out of range"; return nullptr; }
  hexopt_ctx* c = new hexopt_ctx();
  c->device = device;
  const char* e = nullptr;
  if (!build_layout(nV, V, nH, hexes, c->L, &e)) { g_create_error = e; delete c; return nullptr; }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    g_create_error = "hipStreamCreate failed";
    delete c;
    return nullptr;
  }
  const MeshLayout& L = c->L;
  int rc = HEXOPT_OK;
  rc |= upload(c, c->d_conn, L.conn);
  rc |= upload(c, c->d_tile_hex, L.tile_hex);
  rc |= upload(c, c->d_tile_col, L.tile_col);
  rc |= upload(c, c->d_tile_lv, L.tile_lv);
  rc |= upload(c, c->d_lv_gid, L.lv_gid);
  rc |= upload(c, c->d_lv_slot, L.lv_slot);
  rc |= upload(c, c->d_lv_first, L.lv_first);
  rc |= upload(c, c->d_lidx, L.lidx);
  rc |= upload(c, c->d_tile_corner, L.tile_corner);
  rc |= upload(c, c->d_lv_dst, L.lv_dst);
  rc |= upload(c, c->d_v_dst, L.v_dst);
  rc |= upload(c, c->d_sv_gid, L.sv_gid);
  rc |= upload(c, c->d_sv_off, L.sv_off);
  rc |= upload(c, c->d_sv_slots, L.sv_slots);
  {
    std::vector<int> iso;
    for (int i = 0; i < L.nShared; ++i)
      if (L.sv_off[i + 1] == L.sv_off[i]) iso.push_back(L.sv_gid[i]);
    c->n_iso = (int)iso.size();
    rc |= upload(c, c->d_iso, iso);
  }
  std::vector<double> soa;
  to_soa(L, V, soa);
  rc |= upload(c, c->d_rest, soa);
  rc |= upload(c, c->d_cur, soa);
  if (rc != HEXOPT_OK || hipStreamSynchronize(c->stream) != hipSuccess) {
    g_create_error = "device upload failed: " + c->err;
    hexopt_destroy(c);
    return nullptr;
  }
  c->tm = TileMesh{L.nV, L.nH, L.nTiles, L.nShared, L.nSlots, c->d_conn.as<int>(), c->d_tile_hex.as<int>(),
                   c->d_tile_col.as<uint16_t>(), c->d_tile_lv.as<int>(), c->d_lv_gid.as<int>(),
                   c->d_lv_slot.as<int>(), c->d_lv_first.as<uint8_t>(), c->d_lidx.as<uint16_t>(),
                   c->d_sv_gid.as<int>(), c->d_sv_off.as<int>(), c->d_sv_slots.as<int>(),
                   c->d_tile_corner.as<uint8_t>(), c->d_lv_dst.as<int>(), c->d_v_dst.as<int>()};
  hexopt_constraints none{};
  if (hexopt_set_constraints(c, &none) != HEXOPT_OK) {
    g_create_error = c->err;
    hexopt_destroy(c);
    return nullptr;
  }
  return c;
}

void hexopt_destroy(hexopt_ctx* c) {
  if (!c) return;
  if (c->stream) { (void)hipStreamSynchronize(c->stream); }
  hipStream_t s = c->stream;
  delete c;  // frees device buffers
  if (s) (void)hipStreamDestroy(s);
}

const char* hexopt_last_error(hexopt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hexopt_set_positions(hexopt_ctx* c, int which, const double* V) {
  if (!c || !V || (which != HEXOPT_REST && which != HEXOPT_CURRENT)) return HEXOPT_EINVAL;
  std::vector<double> soa;
  to_soa(c->L, V, soa);
  DevBuf& b = which == HEXOPT_REST ? c->d_rest : c->d_cur;
  HIPCHK(c, hipMemcpyAsync(b.p, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HEXOPT_OK;
}

int hexopt_get_positions(hexopt_ctx* c, int which, double* V) {
  if (!c || !V || (which != HEXOPT_REST && which != HEXOPT_CURRENT)) return HEXOPT_EINVAL;
  std::vector<double> soa((size_t)3 * c->L.nV);
  DevBuf& b = which == HEXOPT_REST ? c->d_rest : c->d_cur;
  HIPCHK(c, hipMemcpyAsync(soa.data(), b.p, soa.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  from_soa(c->L, soa.data(), V);
  return HEXOPT_OK;
}

int hexopt_copy_positions(hexopt_ctx* c, int from, int to) {
  if (!c || (from != HEXOPT_REST && from != HEXOPT_CURRENT) || (to != HEXOPT_REST && to != HEXOPT_CURRENT))
    return HEXOPT_EINVAL;
  if (from == to) return HEXOPT_OK;
  DevBuf& a = from == HEXOPT_REST ? c->d_rest : c->d_cur;
  DevBuf& b = to == HEXOPT_REST ? c->d_rest : c->d_cur;
  HIPCHK(c, hipMemcpyAsync(b.p, a.p, (size_t)3 * c->L.nV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return HEXOPT_OK;
}

double* hexopt_device_positions(hexopt_ctx* c, int which) {
  if (!c) return nullptr;
  return which == HEXOPT_REST ? c->d_rest.as<double>() : c->d_cur.as<double>();
}

int hexopt_reference(hexopt_ctx* c, const uint8_t* hflag, int cube) {
  if (!c) return HEXOPT_EINVAL;
  const int nH = c->L.nH, nV = c->L.nV;
  const uint8_t* dfl = nullptr;
  if (hflag) {
    std::vector<uint8_t> hf(nH);
    for (int i = 0; i < nH; ++i) hf[i] = hflag[c->L.hperm[i]];
    if (upload(c, c->d_hflag, hf) != HEXOPT_OK) return HEXOPT_EHIP;
    dfl = c->d_hflag.as<uint8_t>();
  }
  HIPCHK(c, c->d_invs.ensure((size_t)8 * nH * sizeof(double) + 8));
  HIPCHK(c, c->d_count.ensure(sizeof(int)));
  HIPCHK(c, hipMemsetAsync(c->d_count.p, 0, sizeof(int), c->stream));
  launch_ref_cube(c->d_rest.as<double>(), nV, c->tm.conn, nH, dfl, cube, c->d_invs.as<double>(), c->d_count.as<int>(),
                  c->stream);
  int ngen = 0;
  HIPCHK(c, hipMemcpyAsync(&ngen, c->d_count.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ngen == 0) {
    c->ref_mode = REF_CUBE;
  } else {
    HIPCHK(c, c->d_B.ensure((size_t)72 * nH * sizeof(double) + 8));
    launch_ref_general(c->d_rest.as<double>(), nV, c->tm.conn, nH, dfl, cube, c->d_B.as<double>(), c->stream);
    c->ref_mode = REF_GENERAL;
  }
  HIPCHK(c, hipGetLastError());
  return HEXOPT_OK;
}

int hexopt_set_constraints(hexopt_ctx* c, const hexopt_constraints* cs) {
  if (!c || !cs) return HEXOPT_EINVAL;
  const MeshLayout& L = c->L;
  const int nV = L.nV;
  auto chk = [&](int n, const int32_t* ids) {
    for (int i = 0; i < n; ++i)
      if (ids[i] < 0 || ids[i] >= nV) return false;
    return true;
  };
  if (cs->nb < 0 || cs->nC < 0 || cs->nT < 0 || cs->nL < 0) return fail(c, HEXOPT_EINVAL, "negative constraint count");
  if ((cs->nb && (!cs->b || !cs->bc || !chk(cs->nb, cs->b))) || (cs->nC && (!cs->ids_C || !cs->C || !chk(cs->nC, cs->ids_C))) ||
      (cs->nT && (!cs->ids_T || !cs->normal_T || !cs->dis_T || !chk(cs->nT, cs->ids_T))) ||
      (cs->nL && (!cs->ids_L || !cs->Axa_L || !cs->origin_L || !chk(cs->nL, cs->ids_L))))
    return fail(c, HEXOPT_EINVAL, "invalid constraint arrays");
  // per-vertex folded rows (projection rows are folded in at solve time only when enabled,
  // so the projection part is kept in its own rows: index space [0,nV) for the rest,
  // the projection contribution is added through the b list (see hexopt_slim_solve)).
  std::vector<int> cidx(nV, -1);
  std::vector<double> S, b;
  std::vector<int> cline;
  auto row = [&](int v) {
    if (cidx[v] < 0) {
      cidx[v] = (int)cline.size();
      cline.push_back(-1);
      for (int k = 0; k < 6; ++k) S.push_back(0);
      for (int k = 0; k < 3; ++k) b.push_back(0);
    }
    return cidx[v];
  };
  // rows numbered in the device's vertex order (whatever order the caller lists the
  // constraints in), so the vertex passes, which walk the vertices in that order, read the
  // 48-byte rows of neighbouring vertices from neighbouring places; the values of a row do
  // not depend on its number
  {
    std::vector<char> has(nV, 0);
    for (int i = 0; i < cs->nC; ++i) has[L.vinv[cs->ids_C[i]]] = 1;
    for (int i = 0; i < cs->nT; ++i) has[L.vinv[cs->ids_T[i]]] = 1;
    for (int i = 0; i < cs->nL; ++i) has[L.vinv[cs->ids_L[i]]] = 1;
    for (int v = 0; v < nV; ++v)
      if (has[v]) row(v);
  }
  for (int i = 0; i < cs->nC; ++i) {
    int r = row(L.vinv[cs->ids_C[i]]);
    for (int d = 0; d < 3; ++d) { S[r * 6 + d] += cs->lamda_C; b[r * 3 + d] += cs->lamda_C * cs->C[i * 3 + d]; }
  }
  // rows of the plane vertices exist before the plane terms are added, so S0/b0 (the rows
  // without the planes) can be kept for the on-device plane refresh
  for (int i = 0; i < cs->nT; ++i) row(L.vinv[cs->ids_T[i]]);
  for (int i = 0; i < cs->nL; ++i) row(L.vinv[cs->ids_L[i]]);
  std::vector<double> S0_pre = S, b0_pre = b;
  std::vector<int> tcount(nV, 0);
  bool dupT = false;
  for (int i = 0; i < cs->nT; ++i) dupT |= (tcount[L.vinv[cs->ids_T[i]]]++ > 0);
  for (int i = 0; i < cs->nT; ++i) {
    int r = row(L.vinv[cs->ids_T[i]]);
    const double* n = &cs->normal_T[i * 3];
    const double l = cs->lamda_T;
    S[r * 6 + 0] += l * n[0] * n[0]; S[r * 6 + 1] += l * n[1] * n[1]; S[r * 6 + 2] += l * n[2] * n[2];
    S[r * 6 + 3] += l * n[0] * n[1]; S[r * 6 + 4] += l * n[0] * n[2]; S[r * 6 + 5] += l * n[1] * n[2];
    for (int d = 0; d < 3; ++d) b[r * 3 + d] += l * n[d] * cs->dis_T[i];
  }
  std::vector<int> lv(cs->nL);
  std::vector<double> rhs_a(cs->nL), invd_a(cs->nL);
  for (int i = 0; i < cs->nL; ++i) {
    int v = L.vinv[cs->ids_L[i]];
    int r = row(v);
    if (cline[r] >= 0) return fail(c, HEXOPT_EUNSUPPORTED, "vertex with two feature-line constraints");
    cline[r] = i;
    lv[i] = v;
    double s = 0, tt = 0;
    for (int d = 0; d < 3; ++d) {
      S[r * 6 + d] += cs->lamda_L;
      b[r * 3 + d] += cs->lamda_L * cs->origin_L[i * 3 + d];
      S0_pre[r * 6 + d] += cs->lamda_L;
      b0_pre[r * 3 + d] += cs->lamda_L * cs->origin_L[i * 3 + d];
      s += -cs->Axa_L[i * 3 + d] * cs->origin_L[i * 3 + d];
      tt += cs->Axa_L[i * 3 + d] * cs->Axa_L[i * 3 + d];
    }
    rhs_a[i] = cs->lamda_L * s;
    double dg = cs->lamda_L * tt;
    invd_a[i] = dg != 0 ? 1.0 / dg : 1.0;
  }
  // projection rows: kept separate (enabled per solve) as extra rows appended after the
  // feature rows; b_ids/b_pos feed the energy, S/b of these rows are added when enabled.
  std::vector<int> bids(cs->nb), cids(cs->nC), tids(cs->nT);
  for (int i = 0; i < cs->nb; ++i) bids[i] = L.vinv[cs->b[i]];
  for (int i = 0; i < cs->nC; ++i) cids[i] = L.vinv[cs->ids_C[i]];
  for (int i = 0; i < cs->nT; ++i) tids[i] = L.vinv[cs->ids_T[i]];
  auto vec = [](const double* p, size_t n) { return std::vector<double>(p ? p : (const double*)nullptr, p ? p + n : nullptr); };
  c->nL = cs->nL;
  int rc = HEXOPT_OK;
  rc |= upload(c, c->d_cidx, cidx);
  rc |= upload(c, c->d_S0, S0_pre);
  rc |= upload(c, c->d_b0, b0_pre);
  c->planes_refreshable = !dupT;
  c->nrows = (int)cline.size();
  if (!dupT) {
    std::vector<int> trow(std::max<size_t>(1, cline.size()), -1);
    for (int i = 0; i < cs->nT; ++i) trow[cidx[L.vinv[cs->ids_T[i]]]] = i;
    rc |= upload(c, c->d_trow, trow);
  }
  {
    std::vector<char> other(nV, 0);
    for (int i = 0; i < cs->nC; ++i) other[L.vinv[cs->ids_C[i]]] = 1;
    for (int i = 0; i < cs->nT; ++i) other[L.vinv[cs->ids_T[i]]] = 1;
    c->lines_refreshable = true;
    for (int i = 0; i < cs->nL; ++i) c->lines_refreshable &= !other[L.vinv[cs->ids_L[i]]];
  }
  rc |= upload(c, c->d_S, S);
  rc |= upload(c, c->d_b, b);
  rc |= upload(c, c->d_cline, cline);
  {
    std::vector<int> cl((size_t)2 * nV, -1);
    for (int v = 0; v < nV; ++v) {
      cl[(size_t)2 * v] = cidx[v];
      cl[(size_t)2 * v + 1] = cidx[v] >= 0 ? cline[cidx[v]] : -1;
    }
    rc |= upload(c, c->d_cl, cl);
  }
  rc |= upload(c, c->d_bids, bids);
  rc |= upload(c, c->d_bpos, vec(cs->bc, (size_t)cs->nb * 3));
  rc |= upload(c, c->d_cids, cids);
  rc |= upload(c, c->d_cpos, vec(cs->C, (size_t)cs->nC * 3));
  rc |= upload(c, c->d_tids, tids);
  c->h_tids = tids;
  c->h_patch.clear();
  {
    // plane queries in the vertices' Morton order (for coherent closest-point traversals)
    std::vector<int> order(cs->nT);
    for (int i = 0; i < cs->nT; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tids[a] < tids[b]; });
    rc |= upload(c, c->d_torder, order);
  }
  rc |= upload(c, c->d_tn, vec(cs->normal_T, (size_t)cs->nT * 3));
  rc |= upload(c, c->d_td, vec(cs->dis_T, (size_t)cs->nT));
  rc |= upload(c, c->d_lv, lv);
  rc |= upload(c, c->d_lt, vec(cs->Axa_L, (size_t)cs->nL * 3));
  rc |= upload(c, c->d_lo, vec(cs->origin_L, (size_t)cs->nL * 3));
  rc |= upload(c, c->d_rhs_a, rhs_a);
  rc |= upload(c, c->d_invd_a, invd_a);
  HIPCHK(c, c->d_aL.ensure(((size_t)cs->nL + 1) * sizeof(double)));
  if (rc != HEXOPT_OK) return HEXOPT_EHIP;
  SoftDev& s = c->sd;
  s.cidx = c->d_cidx.as<int>(); s.S = c->d_S.as<double>(); s.b = c->d_b.as<double>(); s.cline = c->d_cline.as<int>();
  s.cl = c->d_cl.as<int>();
  s.nb = cs->nb; s.nC = cs->nC; s.nT = cs->nT; s.nL = cs->nL;
  s.b_ids = c->d_bids.as<int>(); s.b_pos = c->d_bpos.as<double>();
  s.c_ids = c->d_cids.as<int>(); s.c_pos = c->d_cpos.as<double>();
  s.t_ids = c->d_tids.as<int>(); s.t_n = c->d_tn.as<double>(); s.t_d = c->d_td.as<double>();
  s.lv = c->d_lv.as<int>(); s.lt = c->d_lt.as<double>(); s.lo = c->d_lo.as<double>();
  s.lamP = 0; s.lamC = cs->lamda_C; s.lamT = cs->lamda_T; s.lamL = cs->lamda_L;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HEXOPT_OK;
}

// The projection rows (ts.projection: lamda_projection |x_b - bc|^2, slim_m.cpp:600-609)
// belong to the padding optimisers, which this build does not run.
static int check_projection(hexopt_ctx* c, const hexopt_params* prm) {
  if (!prm->projection || c->sd.nb == 0) return HEXOPT_OK;
  return fail(c, HEXOPT_EUNSUPPORTED, "projection constraints (ts.projection) are not supported by this build");
}

int hexopt_slim_solve(hexopt_ctx* c, const hexopt_params* prm, int iters, hexopt_iter_log* log, double* E, double* Eq,
                      double* Es, double* aL_out) {
  if (!c || !prm || iters < 0) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  if (prm->energy < 0 || prm->energy > 5) return fail(c, HEXOPT_EINVAL, "unknown energy");
  int rc = check_projection(c, prm);
  if (rc != HEXOPT_OK) return rc;
  if (c->glue && (c->sd.nC || c->sd.nT || c->nL || prm->projection))
    return fail(c, HEXOPT_EUNSUPPORTED, "the glue takes no feature or projection constraints (slim_m.cpp:647)");
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = ensure_state(c)) != HEXOPT_OK) return rc;
  const int nV = c->L.nV, nH = c->L.nH, nL = c->nL;
  const int n = 3 * nV + nL;
  const double M = tet_mass(prm), prox = 0.00001 * prm->weight_opt, mesh_area = M * 8.0 * nH;
  const double exp_f = prm->exp_factor;
  hipStream_t st = c->stream;
  SoftDev sd = soft_with(c, prm);
  RefOp ro = refop(c);
  double* aL = c->d_aL.as<double>();
  if (nL > 0) HIPCHK(c, hipMemsetAsync(aL, 0, nL * sizeof(double), st));
  // slim_precompute: energy at the start point with a_L = 0
  double eq = 0, es = 0;
  if ((rc = energy_now(c, prm, nullptr, 0.0, aL, &eq, &es)) != HEXOPT_OK) return rc;
  double energy = (eq + es) / mesh_area;
  const int nb = nblocks(n);
  double* x = c->d_x.as<double>();
  double* cur = c->d_cur.as<double>();
  for (int it = 0; it < iters; ++it) {
    // local step: K, rhs, Jacobi diagonal
    launch_local_step(c->tm, ro, cur, c->d_K.as<double>(), prm->energy, M, exp_f, prox, sd, c->d_rhs.as<double>(),
                      c->d_invd.as<double>(), c->d_halo.as<double>(), st);
    if (nL > 0) {
      HIPCHK(c, hipMemcpyAsync(c->d_rhs.as<double>() + 3 * (size_t)nV, c->d_rhs_a.p, nL * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
      HIPCHK(c, hipMemcpyAsync(c->d_invd.as<double>() + 3 * (size_t)nV, c->d_invd_a.p, nL * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
      HIPCHK(c, hipMemsetAsync(x + 3 * (size_t)nV, 0, nL * sizeof(double), st));
    }
    double* pt = c->d_part[2].as<double>();
    double* pf = c->d_part[3].as<double>();
    double* pl2 = c->d_part[4].as<double>();
    const int nfix = (c->L.nShared + 255) / 256, nline = (nL + 255) / 256;
    double* gsum = c->glue ? c->d_gsum.as<double>() : nullptr;
    if (c->glue) {
      // known_value_post (slim_m.cpp:403-475): the pinned rows leave the system. Its right-hand
      // side is rhs - L x_pinned on the free rows, and the CG runs on the free rows from the
      // free part of the current positions (pinned rows of r, z and p stay 0)
      double* xfix = c->d_xfix.as<double>();
      launch_glue_split(nV, c->d_invd.as<double>(), cur, xfix, x, st);
      launch_group_sum(sd, nV, xfix, gsum, st);
      launch_apply(c->tm, ro, c->d_K.as<double>(), xfix, c->d_tmp.as<double>(), prox, sd, c->d_halo.as<double>(), pt,
                   pf, pl2, st);
      launch_glue_rhs(nV, c->d_invd.as<double>(), c->d_rhs.as<double>(), c->d_tmp.as<double>(), st);
      launch_group_sum(sd, nV, x, gsum, st);
    } else {
      // CG warm start at the current positions (a_L guess 0)
      HIPCHK(c, hipMemcpyAsync(x, cur, 3 * (size_t)nV * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    launch_apply(c->tm, ro, c->d_K.as<double>(), x, c->d_tmp.as<double>(), prox, sd, c->d_halo.as<double>(), pt, pf,
                 pl2, st);
    launch_cg_init(n, c->d_rhs.as<double>(), c->d_tmp.as<double>(), c->d_r.as<double>(), c->d_part[0].as<double>(),
                   c->d_part[1].as<double>(), nb, c->d_cg.p, c->d_invd.as<double>(), c->d_p.as<double>(),
                   c->d_part[5].as<double>(), 1e-8, st);
    for (int k = 0; k < 20; ++k) {
      if (c->glue) launch_group_sum(sd, nV, c->d_p.as<double>(), gsum, st);
      launch_apply(c->tm, ro, c->d_K.as<double>(), c->d_p.as<double>(), c->d_tmp.as<double>(), prox, sd,
                   c->d_halo.as<double>(), pt, pf, pl2, st);
      launch_cg_alpha(c->d_cg.p, pt, c->L.nTiles, pf, nfix, pl2, nline, st);
      launch_cg_update(n, c->d_cg.p, x, c->d_r.as<double>(), c->d_z.as<double>(), c->d_p.as<double>(),
                       c->d_tmp.as<double>(), c->d_invd.as<double>(), c->d_part[6].as<double>(),
                       c->d_part[7].as<double>(), nb, c->d_p.as<double>(), st);
    }
    launch_cg_zero_if(n, c->d_cg.p, x, nb, st);
    if (c->glue) launch_glue_restore(nV, c->d_invd.as<double>(), cur, x, st);
    if (nL > 0 && !prm->projection)
      HIPCHK(c, hipMemcpyAsync(aL, x + 3 * (size_t)nV, nL * sizeof(double), hipMemcpyDeviceToDevice, st));
    // flip-avoiding step bound and, outside the glue, the energy at the start of the line
    // search, read back together
    double* scal = c->d_scal.as<double>();
    double q0 = 0, s0 = 0;
    launch_max_step(cur, x, nV, c->tm.conn, nH, c->d_part[0].as<double>(), scal + 4, st);
    if (!c->glue)
      launch_energy(ro, cur, nullptr, 0.0, nV, c->tm.conn, nH, prm->energy, M, prm->exp_factor, c->sd, aL,
                    c->d_part[0].as<double>(), c->d_part[1].as<double>(),
                    nblocks(c->sd.nb + c->sd.nC + c->sd.nT + c->sd.nL), scal, st);
    std::vector<unsigned char> cgs(cg_state_bytes());
    double hs[5] = {0, 0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(hs, scal, 5 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(cgs.data(), c->d_cg.p, cgs.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(st));
    const double tmin = hs[4];
    int cg_iters = 0, cg_done = 0;
    double cg_err = 0;
    read_cg_state(cgs.data(), &cg_iters, &cg_err, &cg_done);
    double step = std::min(1., tmin * 0.8);
    const double bound = step;
    // igl::line_search (line_search.cpp:10-56)
    if (c->glue) {
      if ((rc = energy_now(c, prm, nullptr, 0.0, aL, &q0, &s0)) != HEXOPT_OK) return rc;
    } else {
      q0 = hs[0];
      s0 = hs[1];
    }
    const double old_energy = q0 + s0;
    double new_energy = old_energy;
    int cur_iter = 0;
    while (new_energy >= old_energy) {
      if (cur_iter > 20) step = 0;
      double q1, s1;
      if ((rc = energy_now(c, prm, x, step, aL, &q1, &s1)) != HEXOPT_OK) return rc;
      const double cur_e = q1 + s1;
      eq = q1;
      es = s1;
      if (std::isnan(cur_e) || std::isinf(cur_e) || cur_e > old_energy) {
        step /= 2;
      } else {
        launch_step(3 * nV, cur, x, step, st);
        new_energy = cur_e;
        break;
      }
      if (step == 0) break;
      cur_iter++;
    }
    energy = new_energy / mesh_area;
    if (log) {
      hexopt_iter_log& lg = log[it];
      lg.energy = energy; lg.e_quality = eq; lg.e_soft = es; lg.step_bound = bound; lg.step = step;
      lg.cg_iters = cg_iters; lg.cg_error = cg_err;
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  if (E) *E = energy;
  if (Eq) *Eq = eq;
  if (Es) *Es = es;
  if (aL_out && nL > 0) HIPCHK(c, hipMemcpy(aL_out, aL, nL * sizeof(double), hipMemcpyDeviceToHost));
  return HEXOPT_OK;
}

int hexopt_glue_setup(hexopt_ctx* c, int nO, const double* V_orig, const int32_t* H_orig, const int32_t* orig_of,
                      const uint8_t* fixed_orig) {
  if (!c || nO <= 0 || !V_orig || !H_orig || !orig_of || !fixed_orig) return HEXOPT_EINVAL;
  const MeshLayout& L = c->L;
  const int nV = L.nV, nH = L.nH;
  if (nH == 0) return fail(c, HEXOPT_EINVAL, "mesh without hexes");
  for (int v = 0; v < nV; ++v)
    if (orig_of[v] < 0 || orig_of[v] >= nO) return fail(c, HEXOPT_EINVAL, "orig_of out of range");
  for (long i = 0; i < 8L * nH; ++i)
    if (H_orig[i] < 0 || H_orig[i] >= nO) return fail(c, HEXOPT_EINVAL, "H_orig out of range");
  // hex h of the context must be hex h of the original mesh, corner by corner
  for (int h = 0; h < nH; ++h)
    for (int k = 0; k < 8; ++k)
      if (orig_of[L.vperm[L.conn[(size_t)k * nH + h]]] != H_orig[(size_t)L.hperm[h] * 8 + k])
        return fail(c, HEXOPT_EINVAL, "exploded hex corners do not copy the original hex's corners");
  // groups: the copies of each original vertex, in the caller's order
  std::vector<int> cnt(nO + 1, 0), g_of(nV), g_ids(nV);
  std::vector<uint8_t> fx(nV);
  for (int v = 0; v < nV; ++v) cnt[orig_of[v] + 1]++;
  for (int o = 0; o < nO; ++o) cnt[o + 1] += cnt[o];
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int v = 0; v < nV; ++v) {
    const int o = orig_of[v];
    g_ids[pos[o]++] = L.vinv[v];
    g_of[L.vinv[v]] = o;
    fx[L.vinv[v]] = fixed_orig[o] ? 1 : 0;
  }
  std::vector<double> vo((size_t)3 * nO);
  for (int o = 0; o < nO; ++o)
    for (int d = 0; d < 3; ++d) vo[(size_t)d * nO + o] = V_orig[(size_t)o * 3 + d];
  std::vector<int> ho((size_t)8 * nH);
  for (int h = 0; h < nH; ++h)
    for (int k = 0; k < 8; ++k) ho[(size_t)k * nH + h] = H_orig[(size_t)L.hperm[h] * 8 + k];
  int rc = HEXOPT_OK;
  rc |= upload(c, c->d_g_of, g_of);
  rc |= upload(c, c->d_g_off, cnt);
  rc |= upload(c, c->d_g_ids, g_ids);
  rc |= upload(c, c->d_gfixed, fx);
  rc |= upload(c, c->d_Vo, vo);
  rc |= upload(c, c->d_Ho, ho);
  if (rc != HEXOPT_OK) return HEXOPT_EHIP;
  HIPCHK(c, c->d_gsum.ensure((size_t)3 * nO * sizeof(double) + 8));
  HIPCHK(c, c->d_gmean.ensure((size_t)3 * nO * sizeof(double) + 8));
  HIPCHK(c, c->d_xfix.ensure((size_t)3 * nV * sizeof(double) + 8));
  // glueInternalNodes' reference (optimization.cpp:578-582): compute_referenceMesh of the
  // original mesh with cube = false and every H_flag set; grad_tet_ref scales it by the
  // corner-tet volumes of the exploded start positions (REST)
  HIPCHK(c, c->d_B.ensure((size_t)72 * nH * sizeof(double) + 8));
  launch_ref_general(c->d_Vo.as<double>(), nO, c->d_Ho.as<int>(), nH, nullptr, 0, c->d_B.as<double>(), c->stream,
                     c->d_rest.as<double>(), nV, c->tm.conn);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ref_mode = REF_GENERAL;
  c->glue = true;
  c->nO = nO;
  c->lamG = 1e4;
  return HEXOPT_OK;
}

int hexopt_glue_solve(hexopt_ctx* c, const hexopt_params* prm, int max_iters, hexopt_glue_log* log, int* iters_out,
                      double* V_out, double* min_sj) {
  if (!c || !prm || max_iters < 0) return HEXOPT_EINVAL;
  if (!c->glue) return fail(c, HEXOPT_EINVAL, "hexopt_glue_setup has not been called");
  const int nH = c->L.nH, nO = c->nO;
  // slim_solve2's gluing branch (slim_m.cpp:1191-1268): lamda_glue from its lower bound
  const double LMIN = 1e+4, LMAX = 1e+16;
  c->lamG = LMIN;
  SoftDev sd = soft_with(c, prm);
  double mj = -1, aj = 0;
  int it = 0;
  int rc = HEXOPT_OK;
  while (it < max_iters) {
    hexopt_iter_log lg{};
    double E = 0, Eq = 0, Es = 0;
    const double lam = c->lamG;
    if ((rc = hexopt_slim_solve(c, prm, 1, &lg, &E, &Eq, &Es, nullptr)) != HEXOPT_OK) return rc;
    ++it;
    // the glued mesh: each original vertex at the mean of its copies
    launch_group_mean(sd, c->L.nV, c->d_cur.as<double>(), c->d_gmean.as<double>(), c->stream);
    c->lamG = std::min(LMIN * std::max(Eq / Es * c->lamG, 1.0), LMAX);
    double* out = c->d_scal.as<double>();
    launch_scaled_jacobian(c->d_gmean.as<double>(), nO, c->d_Ho.as<int>(), nH, nullptr, c->d_part[0].as<double>(),
                           c->d_part[1].as<double>(), out, c->stream);
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    mj = h[0];
    aj = h[1] / nH;
    if (log) {
      hexopt_glue_log& g = log[it - 1];
      g.energy = E; g.e_quality = Eq; g.e_soft = Es; g.lamda_glue = lam; g.min_sj = mj; g.avg_sj = aj;
      g.step = lg.step; g.cg_iters = lg.cg_iters;
    }
    if (mj > 0) break;
  }
  if (iters_out) *iters_out = it;
  if (min_sj) *min_sj = mj;
  if (V_out) {
    if (it == 0) launch_group_mean(sd, c->L.nV, c->d_cur.as<double>(), c->d_gmean.as<double>(), c->stream);
    std::vector<double> m((size_t)3 * nO);
    HIPCHK(c, hipMemcpyAsync(m.data(), c->d_gmean.p, m.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int o = 0; o < nO; ++o)
      for (int d = 0; d < 3; ++d) V_out[(size_t)o * 3 + d] = m[(size_t)d * nO + o];
  }
  return HEXOPT_OK;
}

int hexopt_debug_system(hexopt_ctx* c, const hexopt_params* prm, const double* p_in, double* rhs_out,
                        double* invd_out, double* Lp_out) {
  if (!c || !prm || !p_in) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  int rc = ensure_state(c);
  if (rc) return rc;
  const int nV = c->L.nV, nL = c->nL;
  const size_t n = (size_t)3 * nV + nL;
  const double M = tet_mass(prm), prox = 0.00001 * prm->weight_opt;
  hipStream_t st = c->stream;
  SoftDev sd = soft_with(c, prm);
  launch_local_step(c->tm, refop(c), c->d_cur.as<double>(), c->d_K.as<double>(), prm->energy, M, prm->exp_factor, prox,
                    sd, c->d_rhs.as<double>(), c->d_invd.as<double>(), c->d_halo.as<double>(), st);
  if (nL > 0) {
    HIPCHK(c, hipMemcpyAsync(c->d_rhs.as<double>() + 3 * (size_t)nV, c->d_rhs_a.p, nL * sizeof(double),
                             hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->d_invd.as<double>() + 3 * (size_t)nV, c->d_invd_a.p, nL * sizeof(double),
                             hipMemcpyDeviceToDevice, st));
  }
  // p_in: caller numbering, flattened [x(0..nV) y z a_L]
  std::vector<double> h(n), o(n);
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < 3; ++d) h[(size_t)d * nV + i] = p_in[(size_t)d * nV + c->L.vperm[i]];
  for (int i = 0; i < nL; ++i) h[3 * (size_t)nV + i] = p_in[3 * (size_t)nV + i];
  HIPCHK(c, hipMemcpyAsync(c->d_p.p, h.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  if (c->glue) launch_group_sum(sd, nV, c->d_p.as<double>(), c->d_gsum.as<double>(), st);
  launch_apply(c->tm, refop(c), c->d_K.as<double>(), c->d_p.as<double>(), c->d_tmp.as<double>(), prox, sd,
               c->d_halo.as<double>(), c->d_part[2].as<double>(), c->d_part[3].as<double>(), c->d_part[4].as<double>(),
               st);
  auto back = [&](DevBuf& b, double* out) -> int {
    if (!out) return HEXOPT_OK;
    HIPCHK(c, hipMemcpyAsync(o.data(), b.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < 3; ++d) out[(size_t)d * nV + c->L.vperm[i]] = o[(size_t)d * nV + i];
    for (int i = 0; i < nL; ++i) out[3 * (size_t)nV + i] = o[3 * (size_t)nV + i];
    return HEXOPT_OK;
  };
  HIPCHK(c, hipGetLastError());
  if ((rc = back(c->d_rhs, rhs_out))) return rc;
  if ((rc = back(c->d_invd, invd_out))) return rc;
  return back(c->d_tmp, Lp_out);
}

static int ensure_state32(hexopt_ctx* c) {
  const int nH = c->L.nH;
  const size_t n = (size_t)3 * c->L.nV + c->nL;
  HIPCHK(c, c->f_K.ensure((size_t)48 * nH * sizeof(float)));
  for (DevBuf* b : {&c->f_rhs, &c->f_invd, &c->f_d, &c->f_r, &c->f_z, &c->f_p, &c->f_p2, &c->f_tmp, &c->f_e})
    HIPCHK(c, b->ensure(n * sizeof(float)));
  // the FP32 local pass keeps 6 floats per halo slot, the FP32 CG a float4 per (tile, vertex) pair
  const size_t nh32 = std::max((size_t)6 * (c->L.nSlots + 1), (size_t)4 * (c->L.lv_gid.size() + 1));
  HIPCHK(c, c->f_halo.ensure(nh32 * sizeof(float)));
  HIPCHK(c, c->d_Ld64.ensure(n * sizeof(double)));
  HIPCHK(c, c->d_b64.ensure(n * sizeof(double)));
  HIPCHK(c, c->d_ninv.ensure(sizeof(int)));
  HIPCHK(c, c->d_cgsave.ensure(cg_state_bytes()));
  return HEXOPT_OK;
}

// FP32 CG on L x = b from x = 0 (Eigen's algorithm, FP64 dot products), `iters` iterations,
// four launches per iteration (five with feature lines) and no host round trip. Uses f_r,
// f_z, f_p, f_p2 (the two direction buffers), f_tmp and f_halo.
static int cg32(hexopt_ctx* c, const SoftDev& sd, double prox, const float* b, float* x, int iters, int* it_out,
                double* err_out) {
  const int nV = c->L.nV, nL = c->nL;
  if (c->L.nTiles == 0) return fail(c, HEXOPT_EINVAL, "mesh without hexes");
  hipStream_t st = c->stream;
  RefOp ro = refop(c);
  float* P[2] = {c->f_p.as<float>(), c->f_p2.as<float>()};
  float* z = c->f_z.as<float>();
  float* r = c->f_r.as<float>();
  launch_cg32_init(nV, nL, sd, prox, b, c->f_invd.as<float>(), x, r, z, P[0], c->d_cg.p, c->d_part[0].as<double>(),
                   c->d_part[1].as<double>(), c->d_part[3].as<double>(), 1e-8, st);
  for (int k = 0; k < iters; ++k)
    launch_cg32_iteration(c->tm, ro, c->f_K.as<float>(), prox, sd, c->d_iso.as<int>(), c->n_iso, z,
                          P[k & 1], P[(k + 1) & 1], c->f_tmp.as<float>(), c->f_halo.as<float>(), c->f_invd.as<float>(),
                          x, r, z, c->d_cg.p, c->d_part[2].as<double>(), c->d_part[4].as<double>(),
                          c->d_part[5].as<double>(), c->d_part[3].as<double>(), c->d_part[6].as<double>(),
                          c->d_part[7].as<double>(), k, k == iters - 1, st);
  HIPCHK(c, hipGetLastError());
  if (it_out || err_out) {
    std::vector<unsigned char> cgs(cg_state_bytes());
    HIPCHK(c, hipMemcpyAsync(cgs.data(), c->d_cg.p, cgs.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    int done = 0, its = 0;
    double err = 0;
    read_cg_state(cgs.data(), &its, &err, &done);
    if (it_out) *it_out = its;
    if (err_out) *err_out = err;
  }
  return HEXOPT_OK;
}

int hexopt_opt32_solve(hexopt_ctx* c, const hexopt_params* prm, const hexopt_opt32_options* opt, int iters,
                       hexopt_opt32_log* log, double* E_out) {
  if (!c || !prm || !opt || iters < 0) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  if (prm->energy != HEXOPT_SYMMETRIC_DIRICHLET)
    return fail(c, HEXOPT_EUNSUPPORTED, "the FP32 adjoint optimiser supports SYMMETRIC_DIRICHLET");
  if (c->glue) return fail(c, HEXOPT_EUNSUPPORTED, "the FP32 adjoint optimiser does not run the glue");
  if (prm->projection && c->sd.nb > 0) return fail(c, HEXOPT_EUNSUPPORTED, "projection constraints");
  if (opt->cg_iters < 1 || opt->refine_iters < 0 || opt->residual_every < 0)
    return fail(c, HEXOPT_EINVAL, "bad iteration counts");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_state(c);
  if (rc) return rc;
  if ((rc = ensure_state32(c))) return rc;
  const int nV = c->L.nV, nH = c->L.nH, nL = c->nL;
  const int n = 3 * nV + nL;
  const double M = tet_mass(prm), prox = 0.00001 * prm->weight_opt, mesh_area = M * 8.0 * nH;
  hipStream_t st = c->stream;
  SoftDev sd = soft_with(c, prm);
  RefOp ro = refop(c);
  double* cur = c->d_cur.as<double>();
  double* scal = c->d_scal.as<double>();
  const int nsb = nblocks(c->sd.nb + c->sd.nC + c->sd.nT + c->sd.nL);
  double energy = 0;
  for (int it = 0; it < iters; ++it) {
    // one pass over the elements: energy, adjoint rhs -dE/dx/2, K, Jacobi diagonal
    HIPCHK(c, hipMemsetAsync(c->d_ninv.p, 0, sizeof(int), st));
    launch_adj_local32(c->tm, ro, cur, c->f_K.as<float>(), M, prox, sd, c->f_rhs.as<float>(), c->f_invd.as<float>(),
                       c->f_halo.as<float>(), c->d_part[0].as<double>(), c->d_ninv.as<int>(), st);
    // direction: L d = -dE/dx / 2 by FP32 CG, then one FP64-residual refinement solve
    float* d = c->f_d.as<float>();
    if ((rc = cg32(c, sd, prox, c->f_rhs.as<float>(), d, opt->cg_iters, nullptr, nullptr))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_cgsave.p, c->d_cg.p, cg_state_bytes(), hipMemcpyDeviceToDevice, st));
    const int nb = nblocks(n);
    const bool every = opt->residual_every > 0 && (it % opt->residual_every == 0 || it == iters - 1);
    const bool measure = opt->refine_iters > 0 || every;
    double* b64 = c->d_b64.as<double>();
    if (measure) {
      // FP64 right-hand side -dE/dx / 2 from the FP64 adjoint gradient at CURRENT (a_L = 0),
      // then r = b64 - L d with FP64 sums over the stored operator
      if (nL > 0) HIPCHK(c, hipMemsetAsync(c->d_aL.p, 0, (size_t)nL * sizeof(double), st));
      launch_grad(64, c->tm, ro, cur, b64, M, sd, c->d_aL.as<double>(), c->d_halo.p, st);
      launch_rhs64_from_grad(nV, b64, sd, cur, st);
      launch_apply32_f64(c->tm, ro, c->f_K.as<float>(), d, c->d_Ld64.as<double>(), prox, sd, c->d_halo.as<double>(),
                         st);
      launch_residual64b(n, b64, c->d_Ld64.as<double>(), c->f_r.as<float>(), c->d_part[1].as<double>(),
                         c->d_part[2].as<double>(), nb, scal + 6, st);
    }
    if (opt->refine_iters > 0) {
      // FP32 copy of the FP64 residual -> correction e by FP32 CG, d += e (the copy keeps the
      // right-hand side apart from the CG's own r, z and p)
      HIPCHK(c, hipMemcpyAsync(c->d_Ld64.p, c->f_r.p, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st));
      float* e = c->f_e.as<float>();
      if ((rc = cg32(c, sd, prox, (const float*)c->d_Ld64.p, e, opt->refine_iters, nullptr, nullptr))) return rc;
      launch_add_inplace(n, d, e, nb, st);
      if (every) {
        launch_apply32_f64(c->tm, ro, c->f_K.as<float>(), d, c->d_Ld64.as<double>(), prox, sd,
                           c->d_halo.as<double>(), st);
        launch_residual64b(n, b64, c->d_Ld64.as<double>(), nullptr, c->d_part[1].as<double>(),
                           c->d_part[2].as<double>(), nb, scal + 8, st);
      }
    }
    // a_L is the line unknowns' solution (as slim_m.cpp:501-503), used by the soft energy
    double* aL = c->d_aL.as<double>();
    if (nL > 0) launch_f2d(nL, d + 3 * (size_t)nV, aL, st);
    // FP64 flip-free bound, then the reference's bisection on FP64 energies along the FP32
    // direction
    launch_max_step_dir32(cur, d, nV, c->tm.conn, nH, c->d_part[0].as<double>(), scal + 4, st);
    launch_energy_dir32(ro, cur, d, 0.0, nV, c->tm.conn, nH, M, sd, aL, c->d_part[0].as<double>(),
                        c->d_part[1].as<double>(), nsb, scal, st);
    double hs[10];
    int ninv = 0;
    std::vector<unsigned char> cgs(cg_state_bytes());
    HIPCHK(c, hipMemcpyAsync(hs, scal, 10 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&ninv, c->d_ninv.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(cgs.data(), c->d_cgsave.p, cgs.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(st));
    int cg_it = 0, cg_done = 0;
    double cg_err = 0;
    read_cg_state(cgs.data(), &cg_it, &cg_err, &cg_done);
    // |b64| of the direction system, from the FP64 right-hand side
    const double rhs_norm = std::sqrt(hs[7]);
    const double res64 = hs[6];
    const double tmin = hs[4];
    double h2[2] = {hs[0], hs[1]};
    const double old_energy = h2[0] + h2[1];
    double step = std::min(1., tmin * 0.8);
    const double bound = step;
    double new_energy = old_energy, eq = h2[0], es = h2[1];
    int cur_iter = 0;
    while (new_energy >= old_energy) {
      if (cur_iter > 20) step = 0;
      launch_energy_dir32(ro, cur, d, step, nV, c->tm.conn, nH, M, sd, aL, c->d_part[0].as<double>(),
                          c->d_part[1].as<double>(), nsb, scal, st);
      HIPCHK(c, hipMemcpyAsync(h2, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCHK(c, hipStreamSynchronize(st));
      const double ce = h2[0] + h2[1];
      if (std::isnan(ce) || std::isinf(ce) || ce > old_energy) {
        step /= 2;
      } else {
        launch_step_dir32(3 * nV, cur, d, step, st);
        new_energy = ce;
        eq = h2[0];
        es = h2[1];
        break;
      }
      if (step == 0) break;
      cur_iter++;
    }
    energy = new_energy / mesh_area;
    if (log) {
      hexopt_opt32_log& lg = log[it];
      lg.energy = energy; lg.e_quality = eq; lg.e_soft = es; lg.step_bound = bound; lg.step = step;
      lg.cg_iters = cg_it; lg.cg_error = cg_err;
      lg.residual64 = !measure ? -1.0 : (rhs_norm > 0 ? std::sqrt(res64) / rhs_norm : 0.0);
      lg.residual64_refined = !(every && opt->refine_iters > 0) ? -1.0
                              : (rhs_norm > 0 ? std::sqrt(hs[8]) / rhs_norm : 0.0);
      lg.inverted_tets = ninv;
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  if (E_out) *E_out = energy;
  return HEXOPT_OK;
}

int hexopt_debug_operator(hexopt_ctx* c, const hexopt_params* prm, int precision, double* K_out) {
  if (!c || !prm || !K_out || (precision != 32 && precision != 64)) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  if (prm->energy != HEXOPT_SYMMETRIC_DIRICHLET) return fail(c, HEXOPT_EUNSUPPORTED, "SYMMETRIC_DIRICHLET only");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_state(c);
  if (rc) return rc;
  if ((rc = ensure_state32(c))) return rc;
  const int nH = c->L.nH;
  const double M = tet_mass(prm), prox = 0.00001 * prm->weight_opt;
  hipStream_t st = c->stream;
  SoftDev sd = soft_with(c, prm);
  std::vector<double> K64;
  std::vector<float> K32;
  if (precision == 64) {
    launch_local_step(c->tm, refop(c), c->d_cur.as<double>(), c->d_K.as<double>(), prm->energy, M, prm->exp_factor,
                      prox, sd, c->d_rhs.as<double>(), c->d_invd.as<double>(), c->d_halo.as<double>(), st);
    K64.resize((size_t)48 * nH);
    HIPCHK(c, hipMemcpyAsync(K64.data(), c->d_K.p, K64.size() * 8, hipMemcpyDeviceToHost, st));
  } else {
    HIPCHK(c, hipMemsetAsync(c->d_ninv.p, 0, sizeof(int), st));
    launch_adj_local32(c->tm, refop(c), c->d_cur.as<double>(), c->f_K.as<float>(), M, prox, sd, c->f_rhs.as<float>(),
                       c->f_invd.as<float>(), c->f_halo.as<float>(), c->d_part[0].as<double>(), c->d_ninv.as<int>(),
                       st);
    K32.resize((size_t)48 * nH);
    HIPCHK(c, hipMemcpyAsync(K32.data(), c->f_K.p, K32.size() * 4, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  // value m of hex h sits at (m / W * nH + h) * W + m % W with W = 16 / sizeof(value)
  const int W = precision == 64 ? 2 : 4;
  for (int h = 0; h < nH; ++h)
    for (int m = 0; m < 48; ++m) {
      const size_t at = ((size_t)(m / W) * nH + h) * W + m % W;
      K_out[(size_t)c->L.hperm[h] * 48 + m] = precision == 64 ? K64[at] : (double)K32[at];
    }
  return HEXOPT_OK;
}

int hexopt_time_kernel(hexopt_ctx* c, int which, int reps, double* ms_per_launch) {
  if (!c || !ms_per_launch || reps < 1) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  int rc = ensure_state(c);
  if (rc) return rc;
  if ((rc = ensure_state32(c))) return rc;
  hipStream_t st = c->stream;
  RefOp ro = refop(c);
  SoftDev sd = c->sd;
  TileMesh tm = c->tm;
  tm.nShared = 0;  // the tile kernel alone (the fixup pass is timed by the profiler)
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  auto one = [&]() {
    switch (which) {
      case 0:
        launch_apply(tm, ro, c->d_K.as<double>(), c->d_p.as<double>(), c->d_tmp.as<double>(), 1e-5, sd,
                     c->d_halo.as<double>(), c->d_part[2].as<double>(), nullptr, nullptr, st);
        break;
      case 1:
        // as inside the CG loop after its first iteration: the previous beta step formed from
        // the update partials of the last solve (a zero threshold keeps every pass running)
        (void)launch_cg32_apply_folded(c->tm, ro, c->f_K.as<float>(), sd, c->d_iso.as<int>(), c->n_iso,
                                       c->f_z.as<float>(), c->f_p.as<float>(), c->f_p2.as<float>(),
                                       c->f_tmp.as<float>(), c->f_halo.as<float>(), c->d_cg.p,
                                       c->d_part[2].as<double>(), c->d_part[4].as<double>(),
                                       c->d_part[5].as<double>(), c->d_part[3].as<double>(),
                                       c->d_part[6].as<double>(), c->d_part[7].as<double>(), 0, st);
        break;
      case 2:
        launch_local_step(tm, ro, c->d_cur.as<double>(), c->d_K.as<double>(), HEXOPT_SYMMETRIC_DIRICHLET, 1.0, 1.0,
                          1e-5, sd, c->d_rhs.as<double>(), c->d_invd.as<double>(), c->d_halo.as<double>(), st);
        break;
      case 4:
        // the adjoint energy + gradient pass alone, FP32 positions in, FP32 dE/dx out
        launch_grad(32, tm, ro, c->d_X.p, c->d_G.p, 1.0, sd, c->d_aL.as<double>(), c->f_halo.p, st);
        break;
      case 5:
        // the same in FP64, on the FP64 positions
        launch_grad(64, tm, ro, c->d_cur.p, c->d_b64.p, 1.0, sd, c->d_aL.as<double>(), c->d_halo.p, st);
        break;
      default:
        launch_adj_local32(tm, ro, c->d_cur.as<double>(), c->f_K.as<float>(), 1.0, 1e-5, sd, c->f_rhs.as<float>(),
                           c->f_invd.as<float>(), c->f_halo.as<float>(), c->d_part[0].as<double>(),
                           c->d_ninv.as<int>(), st);
    }
  };
  sd.nL = 0;
  if (which == 4) {
    HIPCHK(c, c->d_X.ensure((size_t)3 * c->L.nV * sizeof(float)));
    HIPCHK(c, c->d_G.ensure((size_t)3 * c->L.nV * sizeof(float)));
    launch_d2f(3 * c->L.nV, c->d_cur.as<double>(), c->d_X.as<float>(), st);
  }
  if (which == 1) {
    // a fresh CG state (not converged) on the current FP32 system, so the pass runs in full
    HIPCHK(c, hipMemsetAsync(c->d_cg.p, 0, cg_state_bytes(), st));
    launch_cg32_init(c->L.nV, c->nL, sd, 1e-5, c->f_rhs.as<float>(), c->f_invd.as<float>(), c->f_d.as<float>(),
                     c->f_r.as<float>(), c->f_z.as<float>(), c->f_p.as<float>(), c->d_cg.p, c->d_part[0].as<double>(),
                     c->d_part[1].as<double>(), c->d_part[3].as<double>(), 0.0, st);
  }
  one();  // warm
  HIPCHK(c, hipEventRecord(e0, st));
  for (int i = 0; i < reps; ++i) one();
  HIPCHK(c, hipEventRecord(e1, st));
  HIPCHK(c, hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_launch = ms / reps;
  HIPCHK(c, hipGetLastError());
  return HEXOPT_OK;
}

}  // extern "C"

struct hexopt_surface {
  int device = 0;
  hipStream_t stream = nullptr;
  SurfaceTree T;
  int nv = 0, nf = 0;
  std::vector<double> FN, VN, FEN;
  DevBuf dV, dF, dbox, dl, dr, dp, dFN, dVN, dFEN, drec, dstat, dP, dS, dI, dC, dN;
  SurfaceDev view{};
  // occupancy: column bins (built on first use), cell and flag staging
  bool bins_ready = false;
  ColumnBins bins;
  DevBuf dboff, dbfid, dbfbox, dcells, dflags;
  DevBuf dep;  // metro: the faces' edge-plane records
  ColumnBinsDev bview{};
  std::string err;
};

namespace {
thread_local std::string g_surface_error;
template <class T>
int sup(hexopt_surface* s, DevBuf& b, const T* p, size_t n) {
  if (b.ensure(n * sizeof(T) + 8) != hipSuccess) return HEXOPT_EHIP;
  if (n && hipMemcpyAsync(b.p, p, n * sizeof(T), hipMemcpyHostToDevice, s->stream) != hipSuccess) return HEXOPT_EHIP;
  return HEXOPT_OK;
}
}  // namespace

extern "C" {

const char* hexopt_surface_error(void) { return g_surface_error.c_str(); }

hexopt_surface* hexopt_surface_create(int nv, const double* V, int nf, const int32_t* F, const double* FN,
                                      const double* VN, const double* FEN, int device) {
  g_surface_error.clear();
  if (nv <= 0 || nf <= 0 || !V || !F) { g_surface_error = "invalid surface"; return nullptr; }
  if ((FN || VN || FEN) && !(FN && VN && FEN)) { g_surface_error = "give all of FN, VN, FEN or none"; return nullptr; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) { g_surface_error = "no such HIP device"; return nullptr; }
  hexopt_surface* s = new hexopt_surface();
  s->device = device;
  s->nv = nv;
  s->nf = nf;
  const char* e = nullptr;
  if (!build_surface_tree(nv, V, nf, F, s->T, &e)) { g_surface_error = e; delete s; return nullptr; }
  if (s->T.depth > kBvhMaxDepth) { g_surface_error = "surface tree deeper than the traversal stack"; delete s; return nullptr; }
  if (FN) {
    s->FN.assign(FN, FN + (size_t)nf * 3);
    s->VN.assign(VN, VN + (size_t)nv * 3);
    s->FEN.assign(FEN, FEN + (size_t)nf * 9);
  } else {
    surface_normals(nv, V, nf, F, s->FN, s->VN, s->FEN);
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess) {
    g_surface_error = "hipStreamCreate failed";
    delete s;
    return nullptr;
  }
  int rc = 0;
  rc |= sup(s, s->dV, V, (size_t)nv * 3);
  rc |= sup(s, s->dF, F, (size_t)nf * 3);
  rc |= sup(s, s->dbox, s->T.box.data(), s->T.box.size());
  rc |= sup(s, s->dl, s->T.left.data(), s->T.left.size());
  rc |= sup(s, s->dr, s->T.right.data(), s->T.right.size());
  rc |= sup(s, s->dp, s->T.prim.data(), s->T.prim.size());
  rc |= sup(s, s->dFN, s->FN.data(), s->FN.size());
  rc |= sup(s, s->dVN, s->VN.data(), s->VN.size());
  rc |= sup(s, s->dFEN, s->FEN.data(), s->FEN.size());
  std::vector<double> rec;
  pack_node_records(s->T, V, F, rec);
  rc |= sup(s, s->drec, rec.data(), rec.size());
  if (rc || s->dstat.ensure(sizeof(int)) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) {
    g_surface_error = "device upload failed";
    hexopt_surface_destroy(s);
    return nullptr;
  }
  s->view = SurfaceDev{nv, nf, (int)s->T.prim.size(), s->T.depth, s->dV.as<double>(), s->dF.as<int>(),
                       s->dbox.as<double>(), s->dl.as<int>(), s->dr.as<int>(), s->dp.as<int>(), s->dFN.as<double>(),
                       s->dVN.as<double>(), s->dFEN.as<double>(), s->drec.as<double>()};
  return s;
}

void hexopt_surface_destroy(hexopt_surface* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  hipStream_t st = s->stream;
  delete s;
  if (st) (void)hipStreamDestroy(st);
}

const char* hexopt_surface_last_error(hexopt_surface* s) { return s ? s->err.c_str() : "null surface"; }

int hexopt_compute_surface_normals(int nv, const double* V, int nf, const int32_t* F, double* FN, double* VN,
                                   double* FEN) {
  if (nv <= 0 || nf <= 0 || !V || !F || !FN || !VN || !FEN) return HEXOPT_EINVAL;
  for (long i = 0; i < 3L * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) return HEXOPT_EINVAL;
  std::vector<double> a, b, c;
  surface_normals(nv, V, nf, F, a, b, c);
  std::memcpy(FN, a.data(), a.size() * sizeof(double));
  std::memcpy(VN, b.data(), b.size() * sizeof(double));
  std::memcpy(FEN, c.data(), c.size() * sizeof(double));
  return HEXOPT_OK;
}

int hexopt_surface_normals(hexopt_surface* s, double* FN, double* VN, double* FEN) {
  if (!s) return HEXOPT_EINVAL;
  if (FN) std::memcpy(FN, s->FN.data(), s->FN.size() * sizeof(double));
  if (VN) std::memcpy(VN, s->VN.data(), s->VN.size() * sizeof(double));
  if (FEN) std::memcpy(FEN, s->FEN.data(), s->FEN.size() * sizeof(double));
  return HEXOPT_OK;
}

int hexopt_surface_closest(hexopt_surface* s, int nq, const double* P, double* S, int32_t* I, double* C, double* N) {
  if (!s || nq < 0 || (nq > 0 && !P)) return HEXOPT_EINVAL;
  if (nq == 0) return HEXOPT_OK;
  hipStream_t st = s->stream;
  HIPCHK(s, s->dP.ensure((size_t)nq * 24));
  HIPCHK(s, s->dS.ensure((size_t)nq * 8));
  HIPCHK(s, s->dI.ensure((size_t)nq * 4));
  HIPCHK(s, s->dC.ensure((size_t)nq * 24));
  HIPCHK(s, s->dN.ensure((size_t)nq * 24));
  HIPCHK(s, hipMemcpyAsync(s->dP.p, P, (size_t)nq * 24, hipMemcpyHostToDevice, st));
  HIPCHK(s, hipMemsetAsync(s->dstat.p, 0, sizeof(int), st));
  launch_closest(s->view, nq, s->dP.as<double>(), 0, 0, nullptr, s->dS.as<double>(), s->dI.as<int>(), s->dC.as<double>(),
                 s->dN.as<double>(), nullptr, s->dstat.as<int>(), st);
  HIPCHK(s, hipGetLastError());
  int stat = 0;
  HIPCHK(s, hipMemcpyAsync(&stat, s->dstat.p, sizeof(int), hipMemcpyDeviceToHost, st));
  if (S) HIPCHK(s, hipMemcpyAsync(S, s->dS.p, (size_t)nq * 8, hipMemcpyDeviceToHost, st));
  if (I) HIPCHK(s, hipMemcpyAsync(I, s->dI.p, (size_t)nq * 4, hipMemcpyDeviceToHost, st));
  if (C) HIPCHK(s, hipMemcpyAsync(C, s->dC.p, (size_t)nq * 24, hipMemcpyDeviceToHost, st));
  if (N) HIPCHK(s, hipMemcpyAsync(N, s->dN.p, (size_t)nq * 24, hipMemcpyDeviceToHost, st));
  HIPCHK(s, hipStreamSynchronize(st));
  if (stat) return (s->err = "traversal stack overflow", HEXOPT_EUNSUPPORTED);
  return HEXOPT_OK;
}

int hexopt_project_planes(hexopt_ctx* c, hexopt_surface* s, double lamda_T) {
  if (!c || !s) return HEXOPT_EINVAL;
  if (s->device != c->device) return fail(c, HEXOPT_EINVAL, "surface and context on different devices");
  if (!c->planes_refreshable) return fail(c, HEXOPT_EUNSUPPORTED, "duplicate plane vertices");
  const int nT = c->sd.nT;
  if (nT == 0) return HEXOPT_OK;
  hipStream_t st = c->stream;
  HIPCHK(c, s->dC.ensure((size_t)nT * 24));
  HIPCHK(c, hipMemsetAsync(s->dstat.p, 0, sizeof(int), st));
  // closest point C and pseudonormal N of every plane vertex at CURRENT (projection_smooth,
  // grid_hex_meshing.cpp:3760-3781): normal_T = N, dis_T = N . C
  launch_closest(s->view, nT, c->d_cur.as<double>(), 1, c->L.nV, c->d_tids.as<int>(), nullptr, nullptr,
                 s->dC.as<double>(), const_cast<double*>(c->sd.t_n), nullptr, s->dstat.as<int>(), st,
                 c->d_torder.as<int>());
  launch_refold_planes(c->sd, c->d_trow.as<int>(), c->nrows, s->dC.as<double>(), const_cast<double*>(c->sd.t_d), c->d_S0.as<double>(),
                       c->d_b0.as<double>(), const_cast<double*>(c->sd.S), const_cast<double*>(c->sd.b), lamda_T, st);
  c->sd.lamT = lamda_T;
  // no host round trip: the walk's stack holds the tree's depth plus two frames, and trees
  // deeper than the stack limit are refused at creation, so its overflow flag stays clear
  HIPCHK(c, hipGetLastError());
  return HEXOPT_OK;
}

int hexopt_project_patches(hexopt_ctx* c, int npatch, hexopt_surface* const* patches, const int32_t* patch_of_T,
                           double lamda_T) {
  if (!c || npatch <= 0 || !patches) return HEXOPT_EINVAL;
  for (int p = 0; p < npatch; ++p) {
    if (!patches[p]) return fail(c, HEXOPT_EINVAL, "null patch surface");
    if (patches[p]->device != c->device) return fail(c, HEXOPT_EINVAL, "surface and context on different devices");
  }
  if (!c->planes_refreshable) return fail(c, HEXOPT_EUNSUPPORTED, "duplicate plane vertices");
  const int nT = c->sd.nT;
  if (nT == 0) return HEXOPT_OK;
  hipStream_t st = c->stream;
  if (!patch_of_T && ((int)c->h_patch.size() != nT || (int)c->patch_off.size() != npatch + 1))
    return fail(c, HEXOPT_EINVAL, "null patch_of_T and no previous assignment for these planes and patches");
  if (patch_of_T && ((int)c->h_patch.size() != nT || (int)c->patch_off.size() != npatch + 1 ||
                     std::memcmp(c->h_patch.data(), patch_of_T, (size_t)nT * sizeof(int32_t)) != 0)) {
    // rows grouped by patch, each group in the vertices' Morton order (coherent walks)
    for (int i = 0; i < nT; ++i)
      if (patch_of_T[i] < 0 || patch_of_T[i] >= npatch) return fail(c, HEXOPT_EINVAL, "patch index out of range");
    std::vector<int> order(nT);
    for (int i = 0; i < nT; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return patch_of_T[a] != patch_of_T[b] ? patch_of_T[a] < patch_of_T[b] : c->h_tids[a] < c->h_tids[b];
    });
    std::vector<int> off(npatch + 1, 0);
    for (int i = 0; i < nT; ++i) off[patch_of_T[i] + 1]++;
    for (int p = 0; p < npatch; ++p) off[p + 1] += off[p];
    int rc = upload(c, c->d_porder, order);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(st));
    c->h_patch.assign(patch_of_T, patch_of_T + nT);
    c->patch_off = off;
  }
  HIPCHK(c, c->d_pC.ensure((size_t)nT * 24));
  HIPCHK(c, c->d_pstat.ensure(sizeof(int)));
  HIPCHK(c, hipMemsetAsync(c->d_pstat.p, 0, sizeof(int), st));
  // each patch's rows against that patch's tree (dirty_graph_projection's loop over Qfg.Rs),
  // all patches in one launch: normal_T = pseudonormal N, dis_T = N . closest point
  std::vector<PatchSeg> segs;
  int nblocks = 0, max_depth = 1;
  for (int p = 0; p < npatch; ++p) {
    const int n = c->patch_off[p + 1] - c->patch_off[p];
    if (n == 0) continue;
    PatchSeg g{};
    g.s = patches[p]->view;
    g.block0 = nblocks;
    g.q0 = c->patch_off[p];
    g.n = n;
    segs.push_back(g);
    nblocks += (n + kBvhBlock - 1) / kBvhBlock;
    max_depth = std::max(max_depth, patches[p]->view.depth);
  }
  if (segs.size() != c->h_segs.size() ||
      std::memcmp(segs.data(), c->h_segs.data(), segs.size() * sizeof(PatchSeg)) != 0) {
    HIPCHK(c, hipStreamSynchronize(st));  // the previous upload from h_segs has finished
    c->h_segs = segs;
    HIPCHK(c, c->d_segs.ensure(std::max<size_t>(1, segs.size()) * sizeof(PatchSeg)));
    HIPCHK(c, hipMemcpyAsync(c->d_segs.p, c->h_segs.data(), segs.size() * sizeof(PatchSeg), hipMemcpyHostToDevice,
                             st));
  }
  launch_closest_segs(c->d_segs.as<PatchSeg>(), (int)segs.size(), nblocks, max_depth, c->d_cur.as<double>(), 1,
                      c->L.nV, c->d_tids.as<int>(), c->d_porder.as<int>(), c->d_pC.as<double>(),
                      const_cast<double*>(c->sd.t_n), c->d_pstat.as<int>(), st);
  launch_refold_planes(c->sd, c->d_trow.as<int>(), c->nrows, c->d_pC.as<double>(), const_cast<double*>(c->sd.t_d), c->d_S0.as<double>(),
                       c->d_b0.as<double>(), const_cast<double*>(c->sd.S), const_cast<double*>(c->sd.b), lamda_T, st);
  c->sd.lamT = lamda_T;
  HIPCHK(c, hipGetLastError());
  return HEXOPT_OK;
}

static int surface_bins(hexopt_surface* s) {
  if (s->bins_ready) return HEXOPT_OK;
  std::vector<double> V((size_t)s->nv * 3);
  std::vector<int> F((size_t)s->nf * 3);
  HIPCHK(s, hipMemcpyAsync(V.data(), s->dV.p, V.size() * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipMemcpyAsync(F.data(), s->dF.p, F.size() * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  build_column_bins(s->nv, V.data(), s->nf, F.data(), s->bins);
  const ColumnBins& B = s->bins;
  if (sup(s, s->dboff, B.off.data(), B.off.size()) || sup(s, s->dbfid, B.fid.data(), B.fid.size()) ||
      sup(s, s->dbfbox, B.fbox.data(), B.fbox.size()))
    return (s->err = "device upload failed", HEXOPT_EHIP);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  s->bview = ColumnBinsDev{{B.lo[0], B.lo[1]}, {B.hi[0], B.hi[1]}, {B.inv[0], B.inv[1]}, {B.nb[0], B.nb[1]},
                           s->dboff.as<int>(), s->dbfid.as<int>(), s->dbfbox.as<double>()};
  s->bins_ready = true;
  return HEXOPT_OK;
}

static int upload_cells(hexopt_surface* s, int n, const int32_t* cells) {
  for (long i = 0; i < n; ++i)
    if (cells[4 * i + 3] <= 0) return (s->err = "cell extent must be positive", HEXOPT_EINVAL);
  HIPCHK(s, s->dcells.ensure((size_t)n * 16));
  HIPCHK(s, hipMemcpyAsync(s->dcells.p, cells, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
  return HEXOPT_OK;
}

int hexopt_occ_classify_cells(hexopt_surface* s, int n, const int32_t* cells, const double* origin, double voxel_size,
                              const double* mesh_transform, int stop_extent, uint8_t* split) {
  if (!s || n < 0 || (n > 0 && (!cells || !split)) || !origin) return HEXOPT_EINVAL;
  if (n == 0) return HEXOPT_OK;
  int rc = upload_cells(s, n, cells);
  if (rc) return rc;
  HIPCHK(s, s->dflags.ensure((size_t)n));
  double base[3];
  for (int d = 0; d < 3; ++d) base[d] = (mesh_transform ? mesh_transform[d] : 0.0) + origin[d];
  launch_occ_classify(s->view, n, s->dcells.as<int>(), base, voxel_size, stop_extent, s->dflags.as<uint8_t>(), s->stream);
  HIPCHK(s, hipGetLastError());
  HIPCHK(s, hipMemcpyAsync(split, s->dflags.p, (size_t)n, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return HEXOPT_OK;
}

int hexopt_occ_voxel_parity(hexopt_surface* s, int nx, int ny, int nz, const double* origin, double spacing,
                            uint8_t* inside) {
  if (!s || nx < 0 || ny < 0 || nz < 0 || !origin) return HEXOPT_EINVAL;
  const size_t nvox = (size_t)nx * ny * nz;
  if (nvox == 0) return HEXOPT_OK;
  if (!inside || ny > 65535) return HEXOPT_EINVAL;
  int rc = surface_bins(s);
  if (rc) return rc;
  HIPCHK(s, s->dflags.ensure(nvox));
  launch_occ_voxel_parity(s->view, s->bview, nx, ny, nz, origin, spacing, s->dflags.as<uint8_t>(), s->stream);
  HIPCHK(s, hipGetLastError());
  HIPCHK(s, hipMemcpyAsync(inside, s->dflags.p, nvox, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return HEXOPT_OK;
}

int hexopt_occ_cell_parity(hexopt_surface* s, int n, const int32_t* cells, const double* origin, double spacing,
                           float* inside) {
  if (!s || n < 0 || (n > 0 && (!cells || !inside)) || !origin) return HEXOPT_EINVAL;
  if (n == 0) return HEXOPT_OK;
  int rc = surface_bins(s);
  if (rc) return rc;
  rc = upload_cells(s, n, cells);
  if (rc) return rc;
  HIPCHK(s, s->dflags.ensure((size_t)n * 4));
  launch_occ_cell_parity(s->view, s->bview, n, s->dcells.as<int>(), origin, spacing, s->dflags.as<float>(), s->stream);
  HIPCHK(s, hipGetLastError());
  HIPCHK(s, hipMemcpyAsync(inside, s->dflags.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return HEXOPT_OK;
}

int hexopt_hex_inside(hexopt_ctx* c, hexopt_surface* s, int which, double* signed_dis, uint8_t* H_flag) {
  if (!c || !s || (which != HEXOPT_REST && which != HEXOPT_CURRENT)) return HEXOPT_EINVAL;
  if (s->device != c->device) return fail(c, HEXOPT_EINVAL, "surface and context on different devices");
  const int nH = c->L.nH;
  if (nH == 0) return HEXOPT_OK;
  hipStream_t st = c->stream;
  HIPCHK(c, s->dP.ensure((size_t)nH * 24));
  HIPCHK(c, s->dS.ensure((size_t)nH * 8));
  HIPCHK(c, hipMemsetAsync(s->dstat.p, 0, sizeof(int), st));
  launch_hex_centres((which == HEXOPT_REST ? c->d_rest : c->d_cur).as<double>(), c->L.nV, c->tm.conn, nH,
                     s->dP.as<double>(), st);
  launch_closest(s->view, nH, s->dP.as<double>(), 0, 0, nullptr, s->dS.as<double>(), nullptr, nullptr, nullptr, nullptr,
                 s->dstat.as<int>(), st);
  HIPCHK(c, hipGetLastError());
  std::vector<double> S(nH);
  int stat = 0;
  HIPCHK(c, hipMemcpyAsync(S.data(), s->dS.p, (size_t)nH * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(&stat, s->dstat.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  if (stat) return fail(c, HEXOPT_EUNSUPPORTED, "traversal stack overflow");
  // internal hex i is the caller's hex hperm[i]
  for (int i = 0; i < nH; ++i) {
    const int h = c->L.hperm[i];
    if (signed_dis) signed_dis[h] = S[i];
    if (H_flag) H_flag[h] = S[i] < 0 ? 1 : 0;
  }
  return HEXOPT_OK;
}

struct hexopt_octree {
  int device = 0;
  OctreeDev* dev = nullptr;
  OctreeOut r;
  ~hexopt_octree() {
    if (dev) {
      (void)hipSetDevice(device);
      octree_destroy(dev);
    }
  }
};

namespace {
thread_local std::string g_octree_error;
}

const char* hexopt_octree_error(void) { return g_octree_error.c_str(); }

hexopt_octree* hexopt_octree_build(hexopt_surface* s, const int32_t* grid3, const double* origin, double voxel_size,
                                   const double* mesh_transform, int stop_extent, int graded, int paired) {
  g_octree_error.clear();
  if (!s || !grid3 || !origin) {
    g_octree_error = "null argument";
    return nullptr;
  }
  if (hipSetDevice(s->device) != hipSuccess) {
    g_octree_error = "hipSetDevice failed";
    return nullptr;
  }
  double base[3];
  for (int d = 0; d < 3; ++d) base[d] = (mesh_transform ? mesh_transform[d] : 0.0) + origin[d];
  const int grid[3] = {grid3[0], grid3[1], grid3[2]};
  std::unique_ptr<hexopt_octree> o(new (std::nothrow) hexopt_octree);
  if (!o) {
    g_octree_error = "out of host memory";
    return nullptr;
  }
  o->device = s->device;
  std::string err;
  try {
    o->dev = octree_create(grid, base, voxel_size, graded != 0, paired != 0, err);
    if (o->dev && octree_timed(*o->dev, &s->view, stop_extent, 0, nullptr, s->stream, o->r, err)) return o.release();
  } catch (const std::exception& e) {
    err = std::string("octree: ") + e.what();
  }
  g_octree_error = err;
  return nullptr;
}

int hexopt_octree_refine(hexopt_octree* o, hexopt_surface* s, int stop_extent, int n, const int32_t* cells) {
  g_octree_error.clear();
  if (!o || !s || n < 0 || (n > 0 && !cells)) return (g_octree_error = "bad argument", HEXOPT_EINVAL);
  if (s->device != o->device) return (g_octree_error = "surface and octree are on different devices", HEXOPT_EINVAL);
  if (hipSetDevice(o->device) != hipSuccess) return (g_octree_error = "hipSetDevice failed", HEXOPT_EHIP);
  std::string err;
  try {
    if (octree_timed(*o->dev, &s->view, stop_extent, n, cells, s->stream, o->r, err)) return HEXOPT_OK;
  } catch (const std::exception& e) {
    err = std::string("octree: ") + e.what();
  }
  g_octree_error = err;
  return err.find("is not a cell") != std::string::npos ? HEXOPT_EINVAL : HEXOPT_EHIP;
}

void hexopt_octree_destroy(hexopt_octree* o) { delete o; }

int hexopt_octree_sizes(const hexopt_octree* o, int64_t* out4, double* ms) {
  if (!o || !out4) return HEXOPT_EINVAL;
  out4[0] = (int64_t)(o->r.leaves.size() / 4);
  out4[1] = (int64_t)(o->r.nodes.size() / 3);
  out4[2] = o->r.levels;
  out4[3] = o->r.evaluated;
  if (ms) *ms = o->r.ms;
  return HEXOPT_OK;
}

int hexopt_octree_copy(const hexopt_octree* o, int32_t* leaves, int32_t* nodes, double* V, int32_t* H) {
  if (!o) return HEXOPT_EINVAL;
  if (leaves) std::copy(o->r.leaves.begin(), o->r.leaves.end(), leaves);
  if (nodes) std::copy(o->r.nodes.begin(), o->r.nodes.end(), nodes);
  if (V) std::copy(o->r.V.begin(), o->r.V.end(), V);
  if (H) std::copy(o->r.hexes.begin(), o->r.hexes.end(), H);
  return HEXOPT_OK;
}

struct hexopt_curves {
  int device = 0;
  hipStream_t stream = nullptr;
  int nv = 0, ncurves = 0;
  std::vector<int> off;
  DevBuf dV, doff, dvs, dcirc, drows, dcid, dorig, daxis;
  std::vector<int32_t> cid_host;  // what dcid holds (empty: unknown)
  std::string err;
};

namespace {
thread_local std::string g_curves_error;
}

const char* hexopt_curves_error(void) { return g_curves_error.c_str(); }

hexopt_curves* hexopt_curves_create(int nv, const double* V, int ncurves, const int32_t* curve_off,
                                    const int32_t* curve_vs, const uint8_t* circles, int device) {
  g_curves_error.clear();
  if (nv <= 0 || ncurves <= 0 || !V || !curve_off || !curve_vs || !circles) { g_curves_error = "invalid curves"; return nullptr; }
  if (curve_off[0] != 0) { g_curves_error = "curve_off[0] must be 0"; return nullptr; }
  for (int c = 0; c < ncurves; ++c)
    if (curve_off[c + 1] <= curve_off[c]) { g_curves_error = "every curve needs at least one vertex"; return nullptr; }
  for (int k = 0; k < curve_off[ncurves]; ++k)
    if (curve_vs[k] < 0 || curve_vs[k] >= nv) { g_curves_error = "curve vertex out of range"; return nullptr; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) { g_curves_error = "no such HIP device"; return nullptr; }
  hexopt_curves* cv = new hexopt_curves();
  cv->device = device;
  cv->nv = nv;
  cv->ncurves = ncurves;
  cv->off.assign(curve_off, curve_off + ncurves + 1);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&cv->stream) != hipSuccess) {
    g_curves_error = "hipStreamCreate failed";
    delete cv;
    return nullptr;
  }
  const size_t nk = (size_t)curve_off[ncurves];
  bool ok = cv->dV.ensure((size_t)nv * 24) == hipSuccess && cv->doff.ensure((size_t)(ncurves + 1) * 4) == hipSuccess &&
            cv->dvs.ensure(nk * 4) == hipSuccess && cv->dcirc.ensure((size_t)ncurves) == hipSuccess;
  ok = ok && hipMemcpyAsync(cv->dV.p, V, (size_t)nv * 24, hipMemcpyHostToDevice, cv->stream) == hipSuccess;
  ok = ok && hipMemcpyAsync(cv->doff.p, curve_off, (size_t)(ncurves + 1) * 4, hipMemcpyHostToDevice, cv->stream) == hipSuccess;
  ok = ok && hipMemcpyAsync(cv->dvs.p, curve_vs, nk * 4, hipMemcpyHostToDevice, cv->stream) == hipSuccess;
  ok = ok && hipMemcpyAsync(cv->dcirc.p, circles, (size_t)ncurves, hipMemcpyHostToDevice, cv->stream) == hipSuccess;
  ok = ok && hipStreamSynchronize(cv->stream) == hipSuccess;
  if (!ok) {
    g_curves_error = "device upload failed";
    hexopt_curves_destroy(cv);
    return nullptr;
  }
  return cv;
}

void hexopt_curves_destroy(hexopt_curves* cv) {
  if (!cv) return;
  if (cv->stream) (void)hipStreamSynchronize(cv->stream);
  hipStream_t st = cv->stream;
  delete cv;
  if (st) (void)hipStreamDestroy(st);
}

int hexopt_project_lines(hexopt_ctx* c, hexopt_curves* cv, int which, int nL, const int32_t* ids_L,
                         const int32_t* curve_L, double* origin_L, double* Axa_L) {
  if (!c || !cv || (which != HEXOPT_REST && which != HEXOPT_CURRENT) || nL < 0) return HEXOPT_EINVAL;
  if (nL == 0) return HEXOPT_OK;
  if (!ids_L || !curve_L || !origin_L || !Axa_L) return HEXOPT_EINVAL;
  if (cv->device != c->device) return fail(c, HEXOPT_EINVAL, "curves and context on different devices");
  std::vector<int> rows(nL);
  for (int i = 0; i < nL; ++i) {
    if (ids_L[i] < 0 || ids_L[i] >= c->L.nV) return fail(c, HEXOPT_EINVAL, "ids_L out of range");
    if (curve_L[i] < 0 || curve_L[i] >= cv->ncurves) return fail(c, HEXOPT_EINVAL, "curve_L out of range");
    rows[i] = c->L.vinv[ids_L[i]];
  }
  hipStream_t st = c->stream;
  HIPCHK(c, cv->drows.ensure((size_t)nL * 4));
  HIPCHK(c, cv->dcid.ensure((size_t)nL * 4));
  HIPCHK(c, cv->dorig.ensure((size_t)nL * 24));
  HIPCHK(c, cv->daxis.ensure((size_t)nL * 24));
  HIPCHK(c, hipMemcpyAsync(cv->drows.p, rows.data(), (size_t)nL * 4, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(cv->dcid.p, curve_L, (size_t)nL * 4, hipMemcpyHostToDevice, st));
  cv->cid_host.clear();
  CurvesDev view{cv->dV.as<double>(), cv->doff.as<int>(), cv->dvs.as<int>(), cv->dcirc.as<uint8_t>()};
  launch_project_lines(view, (which == HEXOPT_REST ? c->d_rest : c->d_cur).as<double>(), c->L.nV, nL,
                       cv->drows.as<int>(), cv->dcid.as<int>(), cv->dorig.as<double>(), cv->daxis.as<double>(), st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(origin_L, cv->dorig.p, (size_t)nL * 24, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(Axa_L, cv->daxis.p, (size_t)nL * 24, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  return HEXOPT_OK;
}

int hexopt_refresh_lines(hexopt_ctx* c, hexopt_curves* cv, const int32_t* curve_L, double lamda_L) {
  if (!c || !cv) return HEXOPT_EINVAL;
  const int nL = c->sd.nL;
  if (nL == 0) return HEXOPT_OK;
  if (!curve_L) return HEXOPT_EINVAL;
  if (cv->device != c->device) return fail(c, HEXOPT_EINVAL, "curves and context on different devices");
  if (!c->lines_refreshable) return fail(c, HEXOPT_EUNSUPPORTED, "line vertices with corner or plane terms");
  for (int i = 0; i < nL; ++i)
    if (curve_L[i] < 0 || curve_L[i] >= cv->ncurves) return fail(c, HEXOPT_EINVAL, "curve_L out of range");
  hipStream_t st = c->stream;
  void* const cid_before = cv->dcid.p;
  HIPCHK(c, cv->dcid.ensure((size_t)nL * 4));
  if (cv->dcid.p != cid_before) cv->cid_host.clear();  // a new buffer holds nothing yet
  HIPCHK(c, cv->dorig.ensure((size_t)nL * 24));
  HIPCHK(c, cv->daxis.ensure((size_t)nL * 24));
  // the curve of each line vertex, uploaded only when it changed (the feature-alignment loop
  // passes the same list every iteration), so the refresh needs no host round trip
  if (cv->cid_host.size() != (size_t)nL || !std::equal(curve_L, curve_L + nL, cv->cid_host.begin())) {
    HIPCHK(c, hipStreamSynchronize(st));
    HIPCHK(c, hipMemcpy(cv->dcid.p, curve_L, (size_t)nL * 4, hipMemcpyHostToDevice));
    cv->cid_host.assign(curve_L, curve_L + nL);
  }
  CurvesDev view{cv->dV.as<double>(), cv->doff.as<int>(), cv->dvs.as<int>(), cv->dcirc.as<uint8_t>()};
  launch_project_lines(view, c->d_cur.as<double>(), c->L.nV, nL, c->sd.lv, cv->dcid.as<int>(), cv->dorig.as<double>(),
                       cv->daxis.as<double>(), st);
  c->sd.lamL = lamda_L;
  launch_refold_lines(c->sd, cv->dorig.as<double>(), cv->daxis.as<double>(), c->d_S0.as<double>(),
                      c->d_b0.as<double>(), c->d_rhs_a.as<double>(), c->d_invd_a.as<double>(), st);
  HIPCHK(c, hipGetLastError());
  return HEXOPT_OK;
}

// One metro direction (vcg Sampling::Hausdorff with vertex sampling): every face-referenced
// vertex of (V, F), in index order, against surface `to` (whose mesh is Vt, Ft), each at
// vcg's edge-plane point-to-face distance; the distances are added one by one in that order.
static int metro_direction(hexopt_surface* to, int nvt, const double* Vt, int nft, const int32_t* Ft, int nv,
                           const double* V, int nf, const int32_t* F, double upper, double out3[3]) {
  std::vector<char> ref(nv, 0);
  for (long i = 0; i < 3L * nf; ++i) ref[F[i]] = 1;
  std::vector<double> P;
  P.reserve((size_t)nv * 3);
  for (int v = 0; v < nv; ++v)
    if (ref[v]) P.insert(P.end(), V + (size_t)v * 3, V + (size_t)v * 3 + 3);
  const int nq = (int)(P.size() / 3);
  std::vector<double> ep;
  build_ep_faces(nvt, Vt, nft, Ft, ep);
  hipStream_t st = to->stream;
  HIPCHK(to, to->dep.ensure(ep.size() * 8 + 8));
  HIPCHK(to, to->dP.ensure(P.size() * 8 + 8));
  HIPCHK(to, to->dS.ensure((size_t)nq * 8 + 8));
  HIPCHK(to, hipMemcpyAsync(to->dep.p, ep.data(), ep.size() * 8, hipMemcpyHostToDevice, st));
  HIPCHK(to, hipMemcpyAsync(to->dP.p, P.data(), P.size() * 8, hipMemcpyHostToDevice, st));
  HIPCHK(to, hipMemsetAsync(to->dstat.p, 0, sizeof(int), st));
  launch_metro_ep(to->view, to->dep.as<double>(), nq, to->dP.as<double>(), upper, to->dS.as<double>(),
                  to->dstat.as<int>(), st);
  HIPCHK(to, hipGetLastError());
  int stat = 0;
  std::vector<double> dist(nq);
  if (nq) HIPCHK(to, hipMemcpyAsync(dist.data(), to->dS.p, (size_t)nq * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(to, hipMemcpyAsync(&stat, to->dstat.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(to, hipStreamSynchronize(st));
  if (stat) return (to->err = "traversal stack overflow", HEXOPT_EUNSUPPORTED);
  // vcg's AddSample adds the samples one by one in vertex order (a serial chain of additions
  // that no parallel sum reproduces bit for bit): done here on the host, where the chain runs
  // at a few ns per sample; samples at the upper bound found no face and are skipped
  double mx = -HUGE_VAL, sm = 0, cn = 0;
  for (int i = 0; i < nq; ++i) {
    const double d = dist[i];
    if (d == upper) continue;
    if (d > mx) mx = d;
    sm += d;
    cn += 1;
  }
  out3[0] = mx;
  out3[1] = sm;
  out3[2] = cn;
  return HEXOPT_OK;
}

int hexopt_metro_hausdorff(int nv0, const double* V0, int nf0, const int32_t* F0, int nv1, const double* V1, int nf1,
                           const int32_t* F1, int device, double* bbox_diagonal, double* max_dist, double* mean_dist) {
  g_surface_error.clear();
  if (nv0 <= 0 || nf0 <= 0 || nv1 <= 0 || nf1 <= 0 || !V0 || !F0 || !V1 || !F1) return HEXOPT_EINVAL;
  // box of all vertices of both meshes, grown by 2% of its diagonal on every side
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int m = 0; m < 2; ++m) {
    const double* V = m ? V1 : V0;
    const int nv = m ? nv1 : nv0;
    for (int i = 0; i < nv; ++i)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], V[(size_t)i * 3 + d]);
        hi[d] = std::max(hi[d], V[(size_t)i * 3 + d]);
      }
  }
  auto diag = [&]() {
    const double a = hi[0] - lo[0], b = hi[1] - lo[1], c = hi[2] - lo[2];
    return std::sqrt(a * a + b * b + c * c);
  };
  const double off = diag() * 0.02;
  for (int d = 0; d < 3; ++d) { lo[d] -= off; hi[d] += off; }
  const double D = diag();
  hexopt_surface* s0 = hexopt_surface_create(nv0, V0, nf0, F0, nullptr, nullptr, nullptr, device);
  if (!s0) return HEXOPT_EINVAL;
  hexopt_surface* s1 = hexopt_surface_create(nv1, V1, nf1, F1, nullptr, nullptr, nullptr, device);
  if (!s1) { hexopt_surface_destroy(s0); return HEXOPT_EINVAL; }
  double fwd[3] = {0, 0, 0}, bwd[3] = {0, 0, 0};
  int rc = metro_direction(s1, nv1, V1, nf1, F1, nv0, V0, nf0, F0, D, fwd);
  if (rc == HEXOPT_OK) rc = metro_direction(s0, nv0, V0, nf0, F0, nv1, V1, nf1, F1, D, bwd);
  if (rc != HEXOPT_OK) g_surface_error = s0->err.empty() ? s1->err : s0->err;
  hexopt_surface_destroy(s0);
  hexopt_surface_destroy(s1);
  if (rc != HEXOPT_OK) return rc;
  if (bbox_diagonal) *bbox_diagonal = D;
  if (max_dist) *max_dist = std::max(fwd[0], bwd[0]);
  if (mean_dist) *mean_dist = std::max(bwd[1] / bwd[2], fwd[1] / fwd[2]);
  return HEXOPT_OK;
}

int hexopt_energy(hexopt_ctx* c, const hexopt_params* prm, const double* aL, double* Eq, double* Es) {
  if (!c || !prm) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  int rc = ensure_state(c);
  if (rc) return rc;
  double* d_aL = c->d_aL.as<double>();
  if (c->nL > 0) {
    if (aL) HIPCHK(c, hipMemcpyAsync(d_aL, aL, c->nL * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else HIPCHK(c, hipMemsetAsync(d_aL, 0, c->nL * sizeof(double), c->stream));
  }
  return energy_now(c, prm, nullptr, 0.0, d_aL, Eq, Es);
}

static int adj_common(hexopt_ctx* c, int precision, const void* X, const void* v) {
  const size_t es = precision == 64 ? 8 : 4;
  const int nV = c->L.nV;
  HIPCHK(c, c->d_X.ensure((size_t)3 * nV * es));
  HIPCHK(c, c->d_G.ensure((size_t)3 * nV * es));
  HIPCHK(c, c->d_v.ensure((size_t)3 * nV * es));
  HIPCHK(c, c->d_halo.ensure((size_t)6 * (c->L.nSlots + 1) * sizeof(double)));
  auto soa = [&](const void* src, DevBuf& dst) -> int {
    std::vector<unsigned char> tmp((size_t)3 * nV * es);
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < 3; ++d)
        std::memcpy(&tmp[((size_t)d * nV + i) * es], (const unsigned char*)src + ((size_t)c->L.vperm[i] * 3 + d) * es, es);
    HIPCHK(c, hipMemcpyAsync(dst.p, tmp.data(), tmp.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HEXOPT_OK;
  };
  int rc = soa(X, c->d_X);
  if (rc) return rc;
  if (v) rc = soa(v, c->d_v);
  return rc;
}

static int adj_out(hexopt_ctx* c, int precision, void* out) {
  const size_t es = precision == 64 ? 8 : 4;
  const int nV = c->L.nV;
  std::vector<unsigned char> tmp((size_t)3 * nV * es);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), c->d_G.p, tmp.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < 3; ++d)
      std::memcpy((unsigned char*)out + ((size_t)c->L.vperm[i] * 3 + d) * es, &tmp[((size_t)d * nV + i) * es], es);
  return HEXOPT_OK;
}

int hexopt_energy_grad(hexopt_ctx* c, const hexopt_params* prm, int precision, const void* X, const double* aL, void* G) {
  if (!c || !prm || !X || !G || (precision != 32 && precision != 64)) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  if (prm->energy != HEXOPT_SYMMETRIC_DIRICHLET) return fail(c, HEXOPT_EUNSUPPORTED, "adjoint supports SYMMETRIC_DIRICHLET");
  int rc = adj_common(c, precision, X, nullptr);
  if (rc) return rc;
  double* d_aL = c->d_aL.as<double>();
  if (c->nL > 0) {
    if (aL) HIPCHK(c, hipMemcpyAsync(d_aL, aL, c->nL * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else HIPCHK(c, hipMemsetAsync(d_aL, 0, c->nL * sizeof(double), c->stream));
  }
  SoftDev sd = soft_with(c, prm);
  launch_grad(precision, c->tm, refop(c), c->d_X.p, c->d_G.p, tet_mass(prm), sd, d_aL, c->d_halo.p, c->stream);
  HIPCHK(c, hipGetLastError());
  return adj_out(c, precision, G);
}

int hexopt_hvp(hexopt_ctx* c, const hexopt_params* prm, int precision, const void* X, const void* v, void* Hv) {
  if (!c || !prm || !X || !v || !Hv || (precision != 32 && precision != 64)) return HEXOPT_EINVAL;
  if (c->ref_mode < 0) return fail(c, HEXOPT_EINVAL, "hexopt_reference has not been called");
  if (prm->energy != HEXOPT_SYMMETRIC_DIRICHLET) return fail(c, HEXOPT_EUNSUPPORTED, "adjoint supports SYMMETRIC_DIRICHLET");
  int rc = adj_common(c, precision, X, v);
  if (rc) return rc;
  SoftDev sd = soft_with(c, prm);
  launch_hvp(precision, c->tm, refop(c), c->d_X.p, c->d_v.p, c->d_G.p, tet_mass(prm), sd, c->d_halo.p, c->stream);
  HIPCHK(c, hipGetLastError());
  return adj_out(c, precision, Hv);
}

int hexopt_scaled_jacobian(hexopt_ctx* c, int which, double* min_j, double* avg_j, double* per_hex) {
  if (!c || (which != HEXOPT_REST && which != HEXOPT_CURRENT)) return HEXOPT_EINVAL;
  int rc = ensure_state(c);
  if (rc) return rc;
  const int nH = c->L.nH;
  DevBuf ph;
  if (per_hex) HIPCHK(c, ph.ensure((size_t)nH * sizeof(double) + 8));
  double* out = c->d_scal.as<double>();
  launch_scaled_jacobian((which == HEXOPT_REST ? c->d_rest : c->d_cur).as<double>(), c->L.nV, c->tm.conn, nH,
                         per_hex ? ph.as<double>() : nullptr, c->d_part[0].as<double>(), c->d_part[1].as<double>(), out,
                         c->stream);
  double h[2] = {1, 0};
  if (nH > 0) HIPCHK(c, hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> tmp(per_hex ? nH : 0);
  if (per_hex && nH) HIPCHK(c, hipMemcpyAsync(tmp.data(), ph.p, nH * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (min_j) *min_j = nH ? h[0] : 1.0;
  if (avg_j) *avg_j = nH ? h[1] / nH : 0.0;
  if (per_hex)
    for (int i = 0; i < nH; ++i) per_hex[c->L.hperm[i]] = tmp[i];
  return HEXOPT_OK;
}

int hexopt_synchronize(hexopt_ctx* c) {
  if (!c) return HEXOPT_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HEXOPT_OK;
}

int hexopt_layout_info(hexopt_ctx* c, int32_t* o) {
  if (!c || !o) return HEXOPT_EINVAL;
  o[0] = c->L.nTiles; o[1] = c->L.nShared; o[2] = c->L.nSlots; o[3] = c->L.max_colours; o[4] = c->ref_mode;
  o[5] = 0;
  for (uint8_t f : c->L.tile_corner) o[5] += f;
  return HEXOPT_OK;
}

}  // extern "C"
