// slim_m_opt_hip.cpp -- drop-in body of optimization::slim_m_opt (optimization.h:32,
// optimization.cpp:883-900) that runs the SLIM iterations on the GPU through libhexopt's C ABI
// (include/hexopt.h). A maintainer compiles this file in place of that member's definition in
// optimization.cpp; everything else of the reference stays as it is.
//
// The reference builds SLIMData from ts (slim_precompute, slim_m.cpp:1095-1132), runs
// slim_solve2(data, iter, type) (fixed-iteration branch, slim_m.cpp:1269-1288) and writes
// ts.UV and the three energies back. Here the same inputs are handed to one hexopt context:
//  - ts.T's 8 corner tets per hex give the hexes (ts.T(8h + c, 0) is corner c of hex h,
//    global_types.h:163-173);
//  - ts.RT, which the caller filled by compute_referenceMesh(ts.V, Hs, H_flag, Hids, RT, cube),
//    is read back as that call's H_flag: a hex whose 8 reference tets are corner tets of a
//    cube was flagged, the others carry their own shape (global_functions.cpp:3746-3806), and
//    hexopt_reference recomputes both kinds on the device from ts.V;
//  - ts.fc's corner, plane and line constraints become hexopt_constraints.
// The context is kept between calls while the connectivity is unchanged (one optimisation
// loop), as the reference keeps ts. Errors are reported the way the reference reports a
// failed optimisation to its callers: the positions are left unchanged, energy = -1, and the
// library's message goes to stderr.
#include <cmath>
#include <cstdio>
#include <vector>

#include "optimization.h"
#include "hexopt.h"

namespace {

typedef Eigen::Matrix<double, Eigen::Dynamic, 3, Eigen::RowMajor> RowMat;

hexopt_ctx* g_ctx = nullptr;
Eigen::MatrixXi g_T;

// Are the 4 x 3 reference tet's three edges from its first vertex orthogonal and of one length?
bool cube_corner(const Eigen::MatrixXd& R) {
  if (R.rows() != 4 || R.cols() != 3) return false;
  Eigen::Matrix3d E;
  for (int k = 0; k < 3; ++k) E.row(k) = R.row(k + 1) - R.row(0);
  const Eigen::Matrix3d G = E * E.transpose();
  const double e2 = G(0, 0);
  if (!(e2 > 0)) return false;
  const double tol = 1e-9 * e2;
  return std::abs(G(1, 1) - e2) <= tol && std::abs(G(2, 2) - e2) <= tol && std::abs(G(0, 1)) <= tol &&
         std::abs(G(0, 2)) <= tol && std::abs(G(1, 2)) <= tol;
}

bool ok(int rc, const char* what) {
  if (rc == HEXOPT_OK) return true;
  std::fprintf(stderr, "slim_m_opt (GPU): %s failed (%d): %s\n", what, rc, g_ctx ? hexopt_last_error(g_ctx) : "");
  return false;
}

}  // namespace

void optimization::slim_m_opt(Tetralize_Set& ts, const uint32_t iter, const int type, bool verbose) {
  (void)type;
  (void)verbose;
  energy = engery_quality = engery_soft = -1;
  const int nV = (int)ts.V.rows(), nH = (int)ts.T.rows() / 8;
  if (!g_ctx || g_T.rows() != ts.T.rows() || g_T != ts.T) {
    if (g_ctx) hexopt_destroy(g_ctx);
    g_ctx = nullptr;
    std::vector<int32_t> H((size_t)8 * nH);
    for (int h = 0; h < nH; ++h)
      for (int c = 0; c < 8; ++c) H[(size_t)8 * h + c] = ts.T(8 * h + c, 0);
    RowMat V0 = ts.V;
    g_ctx = hexopt_create(nV, V0.data(), nH, H.data(), /*device*/ 0);
    if (!g_ctx) {
      std::fprintf(stderr, "slim_m_opt (GPU): hexopt_create failed: %s\n", hexopt_create_error());
      return;
    }
    g_T = ts.T;
  }
  RowMat Vr = ts.V, Ur = ts.UV;
  if (!ok(hexopt_set_positions(g_ctx, HEXOPT_REST, Vr.data()), "set_positions") ||
      !ok(hexopt_set_positions(g_ctx, HEXOPT_CURRENT, Ur.data()), "set_positions"))
    return;
  // compute_referenceMesh's H_flag, read back from the reference tets it produced
  std::vector<uint8_t> hflag(nH, 1);
  bool all = true;
  if ((int)ts.RT.size() == 8 * nH)
    for (int h = 0; h < nH; ++h) {
      for (int c = 0; c < 8 && hflag[h]; ++c) hflag[h] = cube_corner(ts.RT[(size_t)8 * h + c]);
      all = all && hflag[h];
    }
  if (!ok(hexopt_reference(g_ctx, all ? nullptr : hflag.data(), /*cube*/ 1), "reference")) return;

  const Feature_Constraints& fc = ts.fc;
  RowMat C = fc.C, NT = fc.normal_T, AL = fc.Axa_L, OL = fc.origin_L;
  std::vector<int32_t> idsC(fc.ids_C.data(), fc.ids_C.data() + fc.ids_C.size());
  std::vector<int32_t> idsT(fc.ids_T.data(), fc.ids_T.data() + fc.ids_T.size());
  std::vector<int32_t> idsL(fc.ids_L.data(), fc.ids_L.data() + fc.ids_L.size());
  std::vector<double> disT(fc.dis_T.data(), fc.dis_T.data() + fc.dis_T.size());
  hexopt_constraints cs = {};
  cs.nC = (int)idsC.size(); cs.ids_C = idsC.data(); cs.C = C.data(); cs.lamda_C = fc.lamda_C;
  cs.nT = (int)idsT.size(); cs.ids_T = idsT.data(); cs.normal_T = NT.data(); cs.dis_T = disT.data();
  cs.lamda_T = fc.lamda_T;
  cs.nL = (int)idsL.size(); cs.ids_L = idsL.data(); cs.Axa_L = AL.data(); cs.origin_L = OL.data();
  cs.lamda_L = fc.lamda_L;
  if (!ok(hexopt_set_constraints(g_ctx, &cs), "set_constraints")) return;

  // slim_precompute forces exp_factor to 1.0 (slim_m.cpp:1125); ts.projection is not run
  hexopt_params prm = {(int)ts.energy_type, weight_opt, 1.0, 0, ts.lamda_projection, 0.0};
  std::vector<double> aL(cs.nL + 1);
  double E = -1, Eq = -1, Es = -1;
  if (!ok(hexopt_slim_solve(g_ctx, &prm, (int)iter, nullptr, &E, &Eq, &Es, aL.data()), "slim_solve")) return;
  if (!ok(hexopt_get_positions(g_ctx, HEXOPT_CURRENT, Ur.data()), "get_positions")) return;
  ts.UV = Ur;
  energy = E;
  engery_quality = Eq;
  engery_soft = Es;
}
