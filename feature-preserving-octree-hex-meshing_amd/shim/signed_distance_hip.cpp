// signed_distance_hip.cpp -- drop-in for the closest-point queries of the reference
// (igl::signed_distance_pseudonormal with a prebuilt tree, signed_distance.h:80-92,
// signed_distance.cpp:186-257), running them on the GPU through libhexopt's C ABI
// (include/hexopt.h). A maintainer compiles this file into the reference and calls
// hexopt_shim::signed_distance_pseudonormal where the reference calls
// igl::signed_distance_pseudonormal: projection_smooth (grid_hex_meshing.cpp:3772),
// dirty_graph_projection (:4042,4071) and points_inside_mesh (global_functions.cpp:4047).
// The arguments and outputs are igl's; the outputs are igl's bit for bit (the device walk
// replays igl's tree walk, ties included), given the same surface, normals and queries.
//
// The device surface is built from V, F and the caller's normals (FN, VN, and the per-face
// edge normals EN(EMAP(f + x nf)), which build_aabb_tree computes once per tree) on the
// first call for a surface and kept while V and F keep their contents: the same tree serves
// every iteration of the reference's loops. igl's own tree argument is not read (the library
// builds the same tree shape from V and F). Errors leave S, I, C, N empty and the library's
// message goes to stderr, as the reference reports nothing either.
#include <cstdint>
#include <cstdio>
#include <vector>

#include <Eigen/Dense>
#include "igl/AABB.h"
#include "hexopt.h"

namespace hexopt_shim {
namespace {

typedef Eigen::Matrix<double, Eigen::Dynamic, 3, Eigen::RowMajor> RowMat;

struct Cached {
  hexopt_surface* s = nullptr;
  int64_t nv = -1, nf = -1;
  uint64_t hash = 0;
};
Cached g_surf;

// FNV-1a over the bytes of V and F: a surface is rebuilt only when its contents change
uint64_t content_hash(const Eigen::MatrixXd& V, const Eigen::MatrixXi& F) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  };
  mix(V.data(), (size_t)V.size() * sizeof(double));
  mix(F.data(), (size_t)F.size() * sizeof(int));
  return h;
}

hexopt_surface* surface_for(const Eigen::MatrixXd& V, const Eigen::MatrixXi& F, const Eigen::MatrixXd& FN,
                            const Eigen::MatrixXd& VN, const Eigen::MatrixXd& EN, const Eigen::VectorXi& EMAP) {
  const uint64_t h = content_hash(V, F);
  if (g_surf.s && g_surf.nv == V.rows() && g_surf.nf == F.rows() && g_surf.hash == h) return g_surf.s;
  if (g_surf.s) hexopt_surface_destroy(g_surf.s);
  g_surf = Cached{};
  const int nv = (int)V.rows(), nf = (int)F.rows();
  RowMat Vr = V, FNr = FN, VNr = VN;
  Eigen::Matrix<double, Eigen::Dynamic, 9, Eigen::RowMajor> FEN(nf, 9);
  Eigen::Matrix<int32_t, Eigen::Dynamic, 3, Eigen::RowMajor> Fr = F.cast<int32_t>();
  for (int f = 0; f < nf; ++f)
    for (int x = 0; x < 3; ++x)
      for (int d = 0; d < 3; ++d) FEN(f, 3 * x + d) = EN(EMAP(f + x * nf), d);
  hexopt_surface* s = hexopt_surface_create(nv, Vr.data(), nf, Fr.data(), FNr.data(), VNr.data(), FEN.data(), 0);
  if (!s) {
    fprintf(stderr, "hexopt_surface_create: %s\n", hexopt_surface_error());
    return nullptr;
  }
  g_surf.s = s;
  g_surf.nv = nv;
  g_surf.nf = nf;
  g_surf.hash = h;
  return s;
}

}  // namespace

void signed_distance_pseudonormal(const Eigen::MatrixXd& P, const Eigen::MatrixXd& V, const Eigen::MatrixXi& F,
                                  const igl::AABB<Eigen::MatrixXd, 3>& /*tree*/, const Eigen::MatrixXd& FN,
                                  const Eigen::MatrixXd& VN, const Eigen::MatrixXd& EN, const Eigen::VectorXi& EMAP,
                                  Eigen::VectorXd& S, Eigen::VectorXi& I, Eigen::MatrixXd& C, Eigen::MatrixXd& N) {
  const int nq = (int)P.rows();
  hexopt_surface* s = surface_for(V, F, FN, VN, EN, EMAP);
  if (!s) {
    S.resize(0); I.resize(0); C.resize(0, 3); N.resize(0, 3);
    return;
  }
  RowMat Pr = P, Cr(nq, 3), Nr(nq, 3);
  Eigen::VectorXd Sv(nq);
  std::vector<int32_t> Iv(nq);
  const int rc = hexopt_surface_closest(s, nq, Pr.data(), Sv.data(), Iv.data(), Cr.data(), Nr.data());
  if (rc != HEXOPT_OK) {
    fprintf(stderr, "hexopt_surface_closest (%d): %s\n", rc, hexopt_surface_last_error(s));
    S.resize(0); I.resize(0); C.resize(0, 3); N.resize(0, 3);
    return;
  }
  S = Sv;
  I.resize(nq);
  for (int i = 0; i < nq; ++i) I(i) = Iv[i];
  C = Cr;
  N = Nr;
}

}  // namespace hexopt_shim
