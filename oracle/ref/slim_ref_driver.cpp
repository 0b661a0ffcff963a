// slim_ref_driver.cpp -- runs the reference's own hex-optimisation code (compiled from the
// sources under the reference tree, see oracle/ref/Makefile) on a small binary input and
// writes its outputs, so the tests can pin the CPU restatement and the HIP path to it.
// Test infrastructure only; built in the development container, never shipped.
//
// Modes
//   0  one optimization::slim_m_opt-equivalent call (slim_precompute + slim_solve2, fixed
//      iteration branch, slim_m.cpp:1095-1132,1269-1288) with given soft constraints; the
//      iterations are run one slim_solve2(data,1) call at a time so every iteration is logged.
//   1  the deformation loop body (grid_hex_meshing.cpp:2945-2959) with 1 SLIM iteration per
//      outer iteration: compute_referenceMesh(cube) -> tangent-plane projection of the given
//      surface vertices via igl::signed_distance_pseudonormal -> slim_m_opt(ts,1,-1) ->
//      lambda_T update.
//   2  closest point + pseudonormal only (igl::AABB + normals as build_aabb_tree,
//      grid_hex_meshing.cpp:4231-4248).
//   5  the padding smoothing: grid_hex_meshing_bijective::smoothing (grid_hex_meshing.cpp:
//      2979-3056) without feature types: per outer iteration compute_referenceMesh(cube) and
//      optimization::slim_opt_igl (optimization.cpp:967-1003), i.e. the igl::slim variant
//      (igl/slim.cpp:1190-1287), run one logged igl::slim_solve(data, 1) at a time.
//   4  the padding untangle: optimization::glueInternalNodes' set-up of the exploded mesh
//      (optimization.cpp:552-628) and slim_solve2's gluing branch (slim_m.cpp:1191-1268),
//      its loop written out here around the reference's own functions so that every
//      iteration is logged, then checked against one slim_solve2(data, -1, 1) call.
//   6  point_line_projection (global_functions.cpp:3454-3465), the segment projection of the
//      feature-line branch, on given (segment, point) triples.
//   7  the feature-alignment loop body (grid_hex_meshing.cpp:3851-3870) with every boundary
//      vertex clean: per outer iteration dirty_graph_projection (:3954-4082) -- CORNER rows
//      keep their input targets, LINE rows are projected onto their curve (the LINE branch,
//      :3967-4019, written out around the reference's point_line_projection), REGULAR rows
//      onto their own patch's tree (:4050-4080) -- then compute_referenceMesh(cube, H_flag),
//      slim_m_opt(ts, 1, -1) and lamda_T = lamda_L = lamda_C updated together (:3866).
//      Extra input after the surface: the patch of each surface face, the patch of each plane
//      row, the curves (vertex lists into the surface, circle flags) and the curve of each
//      line row.
#include "slim_m.h"
#include "global_functions.h"
#include "igl/flip_avoiding_line_search.h"
#include "igl/slim.h"
#include "igl/AABB.h"
#include "igl/per_face_normals.h"
#include "igl/per_vertex_normals.h"
#include "igl/per_edge_normals.h"
#include "igl/signed_distance.h"
#include <chrono>
#include <cstdio>
#include <cstdint>
#include <vector>
#include <string>

void update_weights_and_closest_rotations(SLIMData& s, const Eigen::MatrixXd& V, const Eigen::MatrixXi& F,
                                          Eigen::MatrixXd& uv);
void solve_weighted_arap(SLIMData& s, const Eigen::MatrixXd& V, const Eigen::MatrixXi& F, Eigen::MatrixXd& uv,
                         Eigen::VectorXi& soft_b_p, Eigen::MatrixXd& soft_bc_p);
double compute_energy_total(SLIMData& s, Eigen::MatrixXd& V_new);
void build_linear_system(SLIMData& s, Eigen::SparseMatrix<double>& L);

namespace {
struct In {
  FILE* f;
  template <class T> T get() { T v; if (fread(&v, sizeof(T), 1, f) != 1) { fprintf(stderr, "short read\n"); exit(3); } return v; }
  template <class T> std::vector<T> arr(size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(3); }
    return v;
  }
};
struct Out {
  FILE* f;
  void put(const std::string& name, const void* p, size_t bytes, char kind) {
    uint32_t ln = (uint32_t)name.size();
    fwrite(&ln, 4, 1, f);
    fwrite(name.data(), 1, ln, f);
    fwrite(&kind, 1, 1, f);
    uint64_t nb = bytes;
    fwrite(&nb, 8, 1, f);
    if (bytes) fwrite(p, 1, bytes, f);
  }
  void putd(const std::string& n, const std::vector<double>& v) { put(n, v.data(), v.size() * 8, 'd'); }
  void puti(const std::string& n, const std::vector<int>& v) { put(n, v.data(), v.size() * 4, 'i'); }
};

Eigen::MatrixXd rows3(const std::vector<double>& v, int n) {
  Eigen::MatrixXd M(n, 3);
  for (int i = 0; i < n; ++i) for (int d = 0; d < 3; ++d) M(i, d) = v[(size_t)i * 3 + d];
  return M;
}
std::vector<double> flat3(const Eigen::MatrixXd& M) {
  std::vector<double> v((size_t)M.rows() * 3);
  for (int i = 0; i < M.rows(); ++i) for (int d = 0; d < 3; ++d) v[(size_t)i * 3 + d] = M(i, d);
  return v;
}

struct Surface {
  Eigen::MatrixXd V, FN, VN, EN;
  Eigen::MatrixXi F, E;
  Eigen::VectorXi EMAP;
  igl::AABB<Eigen::MatrixXd, 3> tree;
  void build() {
    tree.init(V, F);
    igl::per_face_normals(V, F, FN);
    igl::per_vertex_normals(V, F, igl::PER_VERTEX_NORMALS_WEIGHTING_TYPE_ANGLE, FN, VN);
    igl::per_edge_normals(V, F, igl::PER_EDGE_NORMALS_WEIGHTING_TYPE_UNIFORM, FN, EN, E, EMAP);
  }
};
}  // namespace

// Input: nO, nH, energy, weight_opt, V (nO x 3, the glued target BV), H (nH x 8), fixed
// (nO, the V_Boundary_flag), cubes (nH x 8 x 3, the exploded start positions).
int glue_mode(In& in, Out& out) {
  const int nO = in.get<int32_t>(), nH = in.get<int32_t>(), energy = in.get<int32_t>();
  const double weight_opt = in.get<double>();
  auto V = in.arr<double>((size_t)nO * 3);
  auto H = in.arr<int32_t>((size_t)nH * 8);
  auto fixed = in.arr<int32_t>(nO);
  auto cubes = in.arr<double>((size_t)nH * 24);
  Eigen::MatrixXd BV = rows3(V, nO);
  Mesh mesh;
  mesh.type = Mesh_type::Hex;
  mesh.V = BV.transpose();
  for (int h = 0; h < nH; ++h) {
    Hybrid hy; hy.id = h;
    for (int i = 0; i < 8; ++i) hy.vs.push_back(H[(size_t)h * 8 + i]);
    mesh.Hs.push_back(hy);
  }
  std::vector<bool> V_Boundary_flag(nO);
  for (int v = 0; v < nO; ++v) V_Boundary_flag[v] = fixed[v] != 0;
  // glueInternalNodes (optimization.cpp:552-628)
  Tetralize_Set ts;
  std::vector<int> nhex(nO, 0);
  for (int h = 0; h < nH; ++h) for (int i = 0; i < 8; ++i) nhex[H[(size_t)h * 8 + i]]++;
  std::vector<uint32_t> V_ranges;
  int32_t nv = 0;
  for (int v = 0; v < nO; ++v) { nv += nhex[v]; V_ranges.push_back(nv - 1); }
  ts.V_ranges = V_ranges;
  ts.V.resize(nv, 3);
  ts.T.resize(nH * 8, 4);
  std::vector<int> copy_of((size_t)nH * 8);  // exploded vertex of (hex, corner)
  Eigen::Vector4i t;
  for (int h = 0; h < nH; ++h) {
    const std::vector<uint32_t>& vs = mesh.Hs[h].vs;
    for (int i = 0; i < 8; ++i) {
      for (int j = 0; j < 4; ++j) t[j] = V_ranges[vs[hex_tetra_table[i][j]]];
      ts.T.row(h * 8 + i) = t;
    }
    for (int i = 0; i < 8; ++i) {
      for (int d = 0; d < 3; ++d) ts.V(V_ranges[vs[i]], d) = cubes[((size_t)h * 8 + i) * 3 + d];
      copy_of[(size_t)h * 8 + i] = V_ranges[vs[i]];
      V_ranges[vs[i]]--;
    }
  }
  std::vector<bool> H_inout_tag(nH, true);
  std::vector<uint32_t> Hids;
  for (int h = 0; h < nH; ++h) Hids.push_back(h);
  compute_referenceMesh(BV, mesh.Hs, H_inout_tag, Hids, ts.RT, false);
  ts.mappedV.resize(ts.V.rows(), -1);
  ts.mappedV_reverse.resize(ts.V.rows(), -1);
  nv = 0;
  for (uint32_t i = 0; i < ts.V_ranges.size(); i++) {
    if (V_Boundary_flag[i]) continue;
    int start = 0, end = ts.V_ranges[i];
    if (i != 0) start = ts.V_ranges[i - 1] + 1;
    for (int j = start; j < end + 1; j++) { ts.mappedV[j] = nv; ts.mappedV_reverse[nv++] = j; }
  }
  ts.known_value_post = true;
  ts.post_index = nv;
  ts.post_Variables.resize(ts.V.rows() - nv, 3);
  for (uint32_t i = 0; i < ts.V_ranges.size(); i++) {
    if (!V_Boundary_flag[i]) continue;
    int start = 0, end = ts.V_ranges[i];
    if (i != 0) start = ts.V_ranges[i - 1] + 1;
    for (int j = start; j < end + 1; j++) {
      ts.mappedV[j] = nv;
      ts.mappedV_reverse[nv] = j;
      ts.post_Variables.row(nv - ts.post_index) = BV.row(i);
      nv++;
    }
  }
  V_ranges = ts.V_ranges;
  Eigen::MatrixXd tempV = ts.V;
  nv = 0;
  int id = 0;
  for (uint32_t i = 0; i < ts.V_ranges.size(); i++) if (!V_Boundary_flag[i]) {
    int start = 0, end = V_ranges[i];
    if (i != 0) start = V_ranges[i - 1] + 1;
    nv += end - start + 1;
    ts.V_ranges[id++] = nv - 1;
  }
  for (uint32_t i = 0; i < ts.V_ranges.size(); i++) if (V_Boundary_flag[i]) {
    int start = 0, end = V_ranges[i];
    if (i != 0) start = V_ranges[i - 1] + 1;
    nv += end - start + 1;
    ts.V_ranges[id++] = nv - 1;
  }
  for (int i = 0; i < ts.V.rows(); i++) ts.V.row(i) = tempV.row(ts.mappedV_reverse[i]);
  for (int i = 0; i < ts.T.rows(); i++) for (int j = 0; j < 4; j++) ts.T(i, j) = ts.mappedV[ts.T(i, j)];
  ts.UV = ts.V;
  ts.O_Vranges = V_ranges;
  ts.global = false;
  ts.projection = false;
  ts.lamda_projection = 1e+3;
  ts.glue = true;
  ts.record_Sequence = true;
  ts.known_value_post = true;
  ts.lamda_glue = 1e4;
  ts.m = mesh;
  ts.energy_type = (SLIM_ENERGY)energy;
  // optimization::slim_m_opt (optimization.cpp:883-900)
  SLIMData data;
  data.stop_threshold = 1.e-7;
  data.soft_const_p = 1e1;
  data.exp_factor = 5.0;
  data.weight_opt = weight_opt;
  slim_precompute(ts.V, ts.T, ts.UV, data, ts.energy_type, ts);
  std::vector<double> e0 = {data.energy, data.engery_quality, data.engery_soft};
  SLIMData whole = data;
  // the first iteration's assembled system (weights at the start point, lamda_glue 1e4):
  // L and rhs over all unknowns, the sliced right-hand side of the free rows and the CG guess
  {
    SLIMData d2 = data;
    d2.ts.lamda_glue = 1e4;
    Eigen::MatrixXd uv = d2.V_o;
    update_weights_and_closest_rotations(d2, d2.V, d2.F, uv);
    Eigen::SparseMatrix<double> Lm;
    build_linear_system(d2, Lm);
    std::vector<double> trip;
    for (int k = 0; k < Lm.outerSize(); ++k)
      for (Eigen::SparseMatrix<double>::InnerIterator it(Lm, k); it; ++it) {
        trip.push_back(it.row()); trip.push_back(it.col()); trip.push_back(it.value());
      }
    out.putd("L_triplets", trip);
    std::vector<double> rh(d2.rhs.data(), d2.rhs.data() + d2.rhs.size());
    out.putd("rhs", rh);
  }
  {
    std::vector<int> ri((size_t)nH * 8);
    for (int k = 0; k < nH * 8; ++k) ri[k] = ts.mappedV[copy_of[k]];
    out.puti("ref_index", ri);
    std::vector<int> pi = {ts.post_index};
    out.puti("post_index", pi);
  }
  // slim_solve2's gluing branch, one logged iteration at a time
  const double LMIN = 1e+4, LMAX = 1e+16;
  data.ts.lamda_glue = LMIN;
  Mesh_Quality mq;
  std::vector<double> log;
  int iter = 0;
  do {
    Eigen::MatrixXd mV = data.ts.m.V.transpose();
    compute_referenceMesh(mV, data.ts.m.Hs, H_inout_tag, Hids, data.ts.RT, false);
    Eigen::MatrixXd dest_res;
    data.ts.V = dest_res = data.V_o;
    update_weights_and_closest_rotations(data, data.V, data.F, dest_res);
    solve_weighted_arap(data, data.V, data.F, dest_res, data.b, data.bc);
    std::function<double(Eigen::MatrixXd&)> ce = [&](Eigen::MatrixXd& a) { return compute_energy_total(data, a); };
    const double lam = data.ts.lamda_glue;
    data.energy = igl::flip_avoiding_line_search(data.F, data.V_o, dest_res, ce, data.energy * data.mesh_area) /
                  data.mesh_area;
    iter++;
    data.ts.m.V.resize(3, data.ts.O_Vranges.size());
    Eigen::MatrixXd tv = data.V_o;
    for (int i = 0; i < data.ts.V.rows(); i++) tv.row(i) = data.V_o.row(data.ts.mappedV[i]);
    for (uint32_t i = 0; i < data.ts.O_Vranges.size(); i++) {
      int start = 0, end = data.ts.O_Vranges[i];
      if (i != 0) start = data.ts.O_Vranges[i - 1] + 1;
      Eigen::Vector3d v; v.setZero();
      for (int j = start; j <= end; j++) v += tv.row(j);
      v /= (end - start + 1);
      data.ts.m.V.col(i) = v;
    }
    data.ts.lamda_glue = std::min(LMIN * std::max(data.engery_quality / data.engery_soft * data.ts.lamda_glue, 1.0), LMAX);
    scaled_jacobian(data.ts.m, mq);
    log.push_back(data.energy); log.push_back(data.engery_quality); log.push_back(data.engery_soft);
    log.push_back(lam); log.push_back(mq.min_Jacobian);
    if (mq.min_Jacobian > 0) break;
  } while (iter < 50);
  slim_solve2(whole, -1, 1, false);
  double dmax = (whole.V_o - data.V_o).cwiseAbs().maxCoeff();
  if (dmax != 0) { fprintf(stderr, "glue loop differs from slim_solve2 by %g\n", dmax); return 4; }
  // exploded positions by (hex, corner), and the glued mesh
  std::vector<double> UVx((size_t)nH * 24);
  for (int k = 0; k < nH * 8; ++k)
    for (int d = 0; d < 3; ++d) UVx[(size_t)k * 3 + d] = data.V_o(ts.mappedV[copy_of[k]], d);
  std::vector<double> Vg((size_t)nO * 3);
  for (int i = 0; i < nO; ++i) for (int d = 0; d < 3; ++d) Vg[(size_t)i * 3 + d] = data.ts.m.V(d, i);
  out.putd("E0", e0);
  out.putd("log", log);
  out.putd("UVx", UVx);
  out.putd("Vglued", Vg);
  std::vector<double> q = {mq.min_Jacobian, mq.ave_Jacobian, (double)iter, data.ts.lamda_glue};
  out.putd("SJ", q);
  fclose(out.f);
  return 0;
}

// Input: nV, nH, outer, inner, weight_opt, V (nV x 3), H (nH x 8)
int smooth_mode(In& in, Out& out) {
  const int nV = in.get<int32_t>(), nH = in.get<int32_t>(), outer = in.get<int32_t>(), inner = in.get<int32_t>();
  const double weight_opt = in.get<double>();
  auto V = in.arr<double>((size_t)nV * 3);
  auto H = in.arr<int32_t>((size_t)nH * 8);
  Mesh m;
  m.type = Mesh_type::Hex;
  m.V.resize(3, nV);
  for (int i = 0; i < nV; ++i) for (int d = 0; d < 3; ++d) m.V(d, i) = V[(size_t)i * 3 + d];
  for (int h = 0; h < nH; ++h) {
    Hybrid hy; hy.id = h;
    for (int i = 0; i < 8; ++i) hy.vs.push_back(H[(size_t)h * 8 + i]);
    m.Hs.push_back(hy);
  }
  std::vector<bool> H_inout_tag(nH, true);
  std::vector<uint32_t> Hids;
  for (int h = 0; h < nH; ++h) Hids.push_back(h);
  Tetralize_Set ts;
  ts.V = m.V.transpose();
  ts.T.resize(nH * 8, 4);
  for (int h = 0; h < nH; ++h)
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 4; ++j) ts.T(h * 8 + i, j) = H[(size_t)h * 8 + hex_tetra_table[i][j]];
  ts.energy_type = SYMMETRIC_DIRICHLET;
  ts.UV = ts.V;
  ts.projection = false;
  ts.global = true;
  ts.glue = false;
  ts.lamda_glue = 0;
  ts.lamda_region = 0;
  ts.record_Sequence = false;
  ts.b.resize(0); ts.bc.resize(0, 3);
  ts.lamda_b = 0;
  std::vector<double> log, e0;
  for (int o = 0; o < outer; ++o) {
    compute_referenceMesh(ts.V, m.Hs, H_inout_tag, Hids, ts.RT, true);
    // optimization::slim_opt_igl with empty feature, region and group sets
    igl::SLIMData sData;
    sData.weight_opt = weight_opt;
    sData.soft_const_p = ts.lamda_b;
    sData.exp_factor = 5.0;
    sData.lamda_C = sData.lamda_T = sData.lamda_L = 0;
    sData.lamda_region = ts.lamda_region;
    sData.lamda_glue = ts.lamda_glue;
    Eigen::VectorXi ids_C, ids_L, ids_T, regionb;
    Eigen::MatrixXd C(0, 3), Axa(0, 3), Org(0, 3), nT(0, 3), regionbc(0, 3);
    Eigen::VectorXd dT;
    igl::slim_precompute(ts.V, ts.T, ts.V, sData, igl::SLIMData::SYMMETRIC_DIRICHLET, ts.b, ts.bc, ids_C, C, ids_L,
                         Axa, Org, ids_T, nT, dT, regionb, regionbc, false, ts.global, ts.RT);
    e0.push_back(sData.energy);
    for (int k = 0; k < inner; ++k) {
      igl::slim_solve(sData, 1);
      log.push_back(sData.energy);
    }
    ts.V = sData.V_o;
    m.V = ts.V.transpose();
  }
  out.putd("E0", e0);
  out.putd("log", log);
  out.putd("UV", flat3(ts.V));
  Mesh_Quality mq;
  scaled_jacobian(m, mq);
  std::vector<double> q = {mq.min_Jacobian, mq.ave_Jacobian};
  out.putd("SJ", q);
  fclose(out.f);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
  In in{fopen(argv[1], "rb")};
  if (!in.f) return 2;
  Out out{fopen(argv[2], "wb")};
  int mode = in.get<int32_t>();

  if (mode == 2) {
    Surface s;
    int nv = in.get<int32_t>(), nf = in.get<int32_t>(), nq = in.get<int32_t>();
    s.V = rows3(in.arr<double>((size_t)nv * 3), nv);
    auto F = in.arr<int32_t>((size_t)nf * 3);
    s.F.resize(nf, 3);
    for (int i = 0; i < nf; ++i) for (int d = 0; d < 3; ++d) s.F(i, d) = F[(size_t)i * 3 + d];
    Eigen::MatrixXd P = rows3(in.arr<double>((size_t)nq * 3), nq);
    s.build();
    Eigen::VectorXd S; Eigen::VectorXi I; Eigen::MatrixXd C, N;
    igl::signed_distance_pseudonormal(P, s.V, s.F, s.tree, s.FN, s.VN, s.EN, s.EMAP, S, I, C, N);
    std::vector<double> Sv(S.data(), S.data() + nq);
    std::vector<int> Iv(I.data(), I.data() + nq);
    out.putd("S", Sv); out.puti("I", Iv); out.putd("C", flat3(C)); out.putd("N", flat3(N));
    out.putd("FN", flat3(s.FN)); out.putd("VN", flat3(s.VN));
    std::vector<double> fen((size_t)nf * 9);
    for (int f = 0; f < nf; ++f)
      for (int x = 0; x < 3; ++x)
        for (int d = 0; d < 3; ++d) fen[(size_t)f * 9 + x * 3 + d] = s.EN(s.EMAP(nf * x + f), d);
    out.putd("FEN", fen);
    fclose(out.f);
    return 0;
  }

  if (mode == 4) return glue_mode(in, out);
  if (mode == 5) return smooth_mode(in, out);
  if (mode == 6) {
    // point_line_projection (global_functions.cpp:3454-3465) on n (segment, point) triples:
    // the foot pv and the parameter t
    const int n = in.get<int32_t>();
    auto A = in.arr<double>((size_t)n * 3), B = in.arr<double>((size_t)n * 3), P = in.arr<double>((size_t)n * 3);
    std::vector<double> pv((size_t)n * 3), tv(n);
    for (int i = 0; i < n; ++i) {
      Eigen::Vector3d v1(A[i * 3], A[i * 3 + 1], A[i * 3 + 2]), v2(B[i * 3], B[i * 3 + 1], B[i * 3 + 2]);
      Eigen::Vector3d v(P[i * 3], P[i * 3 + 1], P[i * 3 + 2]), f;
      double t = 0;
      point_line_projection(v1, v2, v, f, t);
      for (int d = 0; d < 3; ++d) pv[(size_t)i * 3 + d] = f[d];
      tv[i] = t;
    }
    out.putd("pv", pv);
    out.putd("t", tv);
    fclose(out.f);
    return 0;
  }

  int nV = in.get<int32_t>(), nH = in.get<int32_t>();
  int energy = in.get<int32_t>(), n_outer = in.get<int32_t>(), n_inner = in.get<int32_t>();
  int cube = in.get<int32_t>(), projection = in.get<int32_t>();
  double weight_opt = in.get<double>(), lamT = in.get<double>(), lamT_bound = in.get<double>();
  double lamC = in.get<double>(), lamL = in.get<double>(), lamProj = in.get<double>();
  auto V = in.arr<double>((size_t)nV * 3);
  auto H = in.arr<int32_t>((size_t)nH * 8);
  auto hflag = in.arr<int32_t>(nH);
  int nb = in.get<int32_t>();
  auto b = in.arr<int32_t>(nb);
  auto bc = in.arr<double>((size_t)nb * 3);
  int nTc = in.get<int32_t>();
  auto idsT = in.arr<int32_t>(nTc);
  auto normT = in.arr<double>((size_t)nTc * 3);
  auto disT = in.arr<double>(nTc);
  int nC = in.get<int32_t>();
  auto idsC = in.arr<int32_t>(nC);
  auto Cc = in.arr<double>((size_t)nC * 3);
  int nL = in.get<int32_t>();
  auto idsL = in.arr<int32_t>(nL);
  auto axa = in.arr<double>((size_t)nL * 3);
  auto org = in.arr<double>((size_t)nL * 3);
  Surface surf;
  int nsv = in.get<int32_t>(), nsf = in.get<int32_t>();
  if (nsv > 0) {
    surf.V = rows3(in.arr<double>((size_t)nsv * 3), nsv);
    auto F = in.arr<int32_t>((size_t)nsf * 3);
    surf.F.resize(nsf, 3);
    for (int i = 0; i < nsf; ++i) for (int d = 0; d < 3; ++d) surf.F(i, d) = F[(size_t)i * 3 + d];
    surf.build();
  }

  // Mesh and Tetralize_Set exactly as deformation() builds them (grid_hex_meshing.cpp:2890-2935)
  Mesh m;
  m.type = Mesh_type::Hex;
  m.V.resize(3, nV);
  for (int i = 0; i < nV; ++i) for (int d = 0; d < 3; ++d) m.V(d, i) = V[(size_t)i * 3 + d];
  for (int h = 0; h < nH; ++h) {
    Hybrid hy; hy.id = h;
    for (int i = 0; i < 8; ++i) hy.vs.push_back(H[(size_t)h * 8 + i]);
    m.Hs.push_back(hy);
  }
  Tetralize_Set ts;
  ts.V = m.V.transpose();
  ts.T.resize(nH * 8, 4);
  for (int h = 0; h < nH; ++h)
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 4; ++j) ts.T(h * 8 + i, j) = H[(size_t)h * 8 + hex_tetra_table[i][j]];
  Feature_Constraints fc;
  fc.lamda_C = lamC; fc.lamda_L = lamL; fc.lamda_T = lamT;
  fc.ids_T.resize(nTc); fc.normal_T.resize(nTc, 3); fc.dis_T.resize(nTc);
  for (int i = 0; i < nTc; ++i) {
    fc.ids_T(i) = idsT[i];
    for (int d = 0; d < 3; ++d) fc.normal_T(i, d) = normT[(size_t)i * 3 + d];
    fc.dis_T(i) = disT[i];
  }
  fc.ids_C.resize(nC); fc.C.resize(nC, 3);
  for (int i = 0; i < nC; ++i) { fc.ids_C(i) = idsC[i]; for (int d = 0; d < 3; ++d) fc.C(i, d) = Cc[(size_t)i * 3 + d]; }
  fc.num_a = nL; fc.ids_L.resize(nL); fc.Axa_L.resize(nL, 3); fc.origin_L.resize(nL, 3);
  for (int i = 0; i < nL; ++i) {
    fc.ids_L(i) = idsL[i];
    for (int d = 0; d < 3; ++d) { fc.Axa_L(i, d) = axa[(size_t)i * 3 + d]; fc.origin_L(i, d) = org[(size_t)i * 3 + d]; }
  }
  ts.energy_type = (SLIM_ENERGY)energy;
  ts.UV = ts.V;
  ts.fc = fc;
  ts.projection = projection != 0;
  ts.lamda_projection = lamProj;
  ts.global = true;
  ts.glue = false;
  ts.lamda_glue = 0;
  ts.lamda_region = 0;
  ts.b.resize(nb); ts.bc.resize(nb, 3);
  for (int i = 0; i < nb; ++i) { ts.b(i) = b[i]; for (int d = 0; d < 3; ++d) ts.bc(i, d) = bc[(size_t)i * 3 + d]; }

  std::vector<bool> H_flag(nH, true);
  for (int h = 0; h < nH; ++h) H_flag[h] = hflag[h] != 0;
  std::vector<uint32_t> Hids;
  for (int h = 0; h < nH; ++h) Hids.push_back(h);

  std::vector<double> log;
  std::vector<double> Vtrace;
  if (mode == 0 || mode == 3) {
    compute_referenceMesh(ts.V, m.Hs, H_flag, Hids, ts.RT, cube != 0);
    SLIMData data;
    data.stop_threshold = 1.e-7;
    data.soft_const_p = 1e1;
    data.exp_factor = 5.0;
    data.weight_opt = weight_opt;
    slim_precompute(ts.V, ts.T, ts.UV, data, ts.energy_type, ts);
    // G per tet from the assembled Dx/Dy/Dz
    int nT = nH * 8;
    std::vector<double> G((size_t)nT * 12, 0.0);
    const Eigen::SparseMatrix<double>* D[3] = {&data.Dx, &data.Dy, &data.Dz};
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < D[c]->outerSize(); ++k)
        for (Eigen::SparseMatrix<double>::InnerIterator it(*D[c], k); it; ++it) {
          int t = it.row(), col = it.col();
          for (int j = 0; j < 4; ++j)
            if (ts.T(t, j) == col) G[(size_t)t * 12 + c * 4 + j] += it.value();
        }
    out.putd("G", G);
    std::vector<double> e0 = {data.energy, data.engery_quality, data.engery_soft};
    out.putd("E0", e0);
    // local step at the start point (W and R per tet, row-major 3x3)
    {
      SLIMData d2 = data;
      Eigen::MatrixXd uv = d2.V_o;
      update_weights_and_closest_rotations(d2, d2.V, d2.F, uv);
      std::vector<double> W((size_t)nT * 9), R((size_t)nT * 9);
      const Eigen::VectorXd* Wc[9] = {&d2.W_11, &d2.W_12, &d2.W_13, &d2.W_21, &d2.W_22, &d2.W_23, &d2.W_31, &d2.W_32, &d2.W_33};
      for (int t = 0; t < nT; ++t) {
        for (int k = 0; k < 9; ++k) W[(size_t)t * 9 + k] = (*Wc[k])(t);
        // Ri(t, b + 3c) = R(b, c)
        for (int bb = 0; bb < 3; ++bb)
          for (int c = 0; c < 3; ++c) R[(size_t)t * 9 + bb * 3 + c] = d2.Ri(t, bb + 3 * c);
      }
      out.putd("W0", W);
      out.putd("R0", R);
      if (mode == 3) {
        // the assembled global system of the first iteration (small meshes only)
        Eigen::SparseMatrix<double> Lm;
        build_linear_system(d2, Lm);
        std::vector<double> trip;
        for (int k = 0; k < Lm.outerSize(); ++k)
          for (Eigen::SparseMatrix<double>::InnerIterator it(Lm, k); it; ++it) {
            trip.push_back(it.row()); trip.push_back(it.col()); trip.push_back(it.value());
          }
        out.putd("L_triplets", trip);
        std::vector<double> rh(d2.rhs.data(), d2.rhs.data() + d2.rhs.size());
        out.putd("rhs", rh);
      }
    }
    for (int it = 0; it < n_inner; ++it) {
      slim_solve2(data, 1, -1, false);
      log.push_back(data.energy); log.push_back(data.engery_quality); log.push_back(data.engery_soft);
      auto f = flat3(data.V_o);
      Vtrace.insert(Vtrace.end(), f.begin(), f.end());
    }
    std::vector<double> aL(data.a_L.data(), data.a_L.data() + data.a_L.size());
    out.putd("aL", aL);
    ts.UV = data.V_o;
  } else if (mode == 1) {
    // ts.fc.ids_T lists the vertices to project (projection_smooth, grid_hex_meshing.cpp:3760-3781)
    double LAMDA0 = lamT;
    std::vector<double> t_iter;
    for (int it = 0; it < n_outer; ++it) {
      // wall time of the loop body alone: reference mesh, projection, slim_m_opt
      const auto t0 = std::chrono::steady_clock::now();
      compute_referenceMesh(ts.V, m.Hs, H_flag, Hids, ts.RT, cube != 0);
      if (nsv > 0 && nTc > 0) {
        Eigen::MatrixXd Ps(nTc, 3);
        for (int i = 0; i < nTc; ++i) Ps.row(i) = ts.V.row(fc.ids_T(i));
        Eigen::VectorXd S; Eigen::VectorXi I; Eigen::MatrixXd C, N;
        igl::signed_distance_pseudonormal(Ps, surf.V, surf.F, surf.tree, surf.FN, surf.VN, surf.EN, surf.EMAP, S, I, C, N);
        for (int j = 0; j < nTc; ++j) {
          fc.normal_T.row(j) = N.row(j);
          fc.dis_T[j] = N.row(j).dot(C.row(j));
        }
      }
      ts.fc = fc;
      SLIMData data;
      data.stop_threshold = 1.e-7;
      data.soft_const_p = 1e1;
      data.exp_factor = 5.0;
      data.weight_opt = weight_opt;
      slim_precompute(ts.V, ts.T, ts.UV, data, ts.energy_type, ts);
      for (int k = 0; k < n_inner; ++k) slim_solve2(data, 1, -1, false);
      ts.UV = data.V_o;
      ts.V = ts.UV;
      t_iter.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      log.push_back(data.energy); log.push_back(data.engery_quality); log.push_back(data.engery_soft);
      log.push_back(fc.lamda_T);
      fc.lamda_T = std::min(LAMDA0 * std::max(data.engery_quality / data.engery_soft * fc.lamda_T, 1.0), lamT_bound);
      auto f = flat3(ts.UV);
      Vtrace.insert(Vtrace.end(), f.begin(), f.end());
    }
    out.putd("t_iter", t_iter);
  } else if (mode == 7) {
    auto fpatch = in.arr<int32_t>(nsf);
    auto patchT = in.arr<int32_t>(nTc);
    const int ncurve = in.get<int32_t>();
    auto clen = in.arr<int32_t>(ncurve);
    std::vector<std::vector<int>> curves(ncurve);
    for (int c = 0; c < ncurve; ++c) {
      auto cv = in.arr<int32_t>(clen[c]);
      curves[c].assign(cv.begin(), cv.end());
    }
    auto circles = in.arr<int32_t>(ncurve);
    auto curveL = in.arr<int32_t>(nL);
    // one tree per patch over that patch's faces (the tri_Trees of Qfg.Rs)
    int npatch = 0;
    for (int f = 0; f < nsf; ++f) npatch = std::max(npatch, fpatch[f] + 1);
    std::vector<Surface> ps(npatch);
    for (int p = 0; p < npatch; ++p) {
      std::vector<int> fl;
      for (int f = 0; f < nsf; ++f) if (fpatch[f] == p) fl.push_back(f);
      std::vector<int> remap(nsv, -1);
      int nu = 0;
      for (int f : fl) for (int k = 0; k < 3; ++k) if (remap[surf.F(f, k)] < 0) remap[surf.F(f, k)] = nu++;
      ps[p].V.resize(nu, 3);
      for (int v = 0; v < nsv; ++v) if (remap[v] >= 0) ps[p].V.row(remap[v]) = surf.V.row(v);
      ps[p].F.resize((int)fl.size(), 3);
      for (size_t i = 0; i < fl.size(); ++i) for (int k = 0; k < 3; ++k) ps[p].F((int)i, k) = remap[surf.F(fl[i], k)];
      ps[p].build();
    }
    double LAMDA0 = lamT;
    fc.lamda_C = fc.lamda_L = fc.lamda_T = LAMDA0;
    std::vector<double> t_iter;
    for (int it = 0; it < n_outer; ++it) {
      const auto t0 = std::chrono::steady_clock::now();
      // LINE rows: nearest segment of the vertex's curve (ties to the lower segment, as the
      // sort of (distance, index) pairs leaves them)
      for (int i = 0; i < nL; ++i) {
        const Eigen::Vector3d v = ts.UV.row(fc.ids_L(i)).transpose();
        const std::vector<int>& curve = curves[curveL[i]];
        int curve_len = (int)curve.size();
        if (!circles[curveL[i]]) curve_len--;
        std::vector<Eigen::Vector3d> pvs, tangents;
        std::vector<std::pair<double, uint32_t>> dis_ids;
        for (int j = 0; j < curve_len; ++j) {
          const int p0 = curve[j], p1 = curve[(j + 1) % curve.size()];
          Eigen::Vector3d pv;
          double t;
          point_line_projection(surf.V.row(p0).transpose(), surf.V.row(p1).transpose(), v, pv, t);
          const Eigen::Vector3d tangent = (surf.V.row(p1) - surf.V.row(p0)).transpose().normalized();
          dis_ids.push_back(std::make_pair((v - pv).norm(), (uint32_t)pvs.size()));
          pvs.push_back(pv);
          tangents.push_back(tangent);
        }
        std::sort(dis_ids.begin(), dis_ids.end());
        const uint32_t k = dis_ids[0].second;
        fc.origin_L.row(i) = pvs[k].transpose();
        fc.Axa_L.row(i) = tangents[k].transpose();
      }
      // REGULAR rows onto their own patch
      for (int p = 0; p < npatch; ++p) {
        std::vector<int> rows;
        for (int j = 0; j < nTc; ++j) if (patchT[j] == p) rows.push_back(j);
        if (rows.empty()) continue;
        Eigen::MatrixXd Ps((int)rows.size(), 3);
        for (size_t j = 0; j < rows.size(); ++j) Ps.row((int)j) = ts.UV.row(fc.ids_T(rows[j]));
        Eigen::VectorXd S; Eigen::VectorXi I; Eigen::MatrixXd C, N;
        igl::signed_distance_pseudonormal(Ps, ps[p].V, ps[p].F, ps[p].tree, ps[p].FN, ps[p].VN, ps[p].EN, ps[p].EMAP,
                                          S, I, C, N);
        for (size_t j = 0; j < rows.size(); ++j) {
          fc.normal_T.row(rows[j]) = N.row((int)j);
          fc.dis_T[rows[j]] = N.row((int)j).dot(C.row((int)j));
        }
      }
      ts.fc = fc;
      compute_referenceMesh(ts.UV, m.Hs, H_flag, Hids, ts.RT, true);
      SLIMData data;
      data.stop_threshold = 1.e-7;
      data.soft_const_p = 1e1;
      data.exp_factor = 5.0;
      data.weight_opt = weight_opt;
      slim_precompute(ts.V, ts.T, ts.UV, data, ts.energy_type, ts);
      slim_solve2(data, 1, -1, false);
      ts.UV = data.V_o;
      ts.V = ts.UV;
      t_iter.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      log.push_back(data.energy); log.push_back(data.engery_quality); log.push_back(data.engery_soft);
      log.push_back(fc.lamda_C);
      fc.lamda_T = fc.lamda_L = fc.lamda_C =
          std::min(LAMDA0 * std::max(data.engery_quality / data.engery_soft * fc.lamda_C, 1.0), lamT_bound);
      auto f = flat3(ts.UV);
      Vtrace.insert(Vtrace.end(), f.begin(), f.end());
    }
    out.putd("t_iter", t_iter);
  }
  out.putd("log", log);
  out.putd("Vtrace", Vtrace);
  out.putd("UV", flat3(ts.UV));
  Mesh_Quality mq;
  m.V = ts.UV.transpose();
  scaled_jacobian(m, mq);
  std::vector<double> q = {mq.min_Jacobian, mq.ave_Jacobian};
  out.putd("SJ", q);
  std::vector<double> hj(mq.H_Js.data(), mq.H_Js.data() + mq.H_Js.size());
  out.putd("SJ_hex", hj);
  fclose(out.f);
  return 0;
}
