// sd_shim_driver.cpp -- the closest-point drop-in side by side with the reference: builds
// igl's tree and normals the way build_aabb_tree does (grid_hex_meshing.cpp:4231-4248:
// igl::AABB::init, per_face_normals, per_vertex_normals by angle, per_edge_normals uniform),
// then calls igl::signed_distance_pseudonormal (signed_distance.cpp:186-257, compiled from the
// reference's libigl) and hexopt_shim::signed_distance_pseudonormal (the package's
// shim/signed_distance_hip.cpp, linked with libhexopt.so) on the same queries, twice (the
// second call reuses the shim's cached surface), and writes both results.
// Input: int32 nv, nf, nq; double V[nv][3]; int32 F[nf][3]; double P[nq][3].
// Output (refio format): ref_S, ref_I, ref_C, ref_N, shim_S, shim_I, shim_C, shim_N (the
// shim's second call), shim1_S ... (its first call).
// Test infrastructure only; built in the development container (oracle/ref/Makefile).
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "igl/AABB.h"
#include "igl/per_edge_normals.h"
#include "igl/per_face_normals.h"
#include "igl/per_vertex_normals.h"
#include "igl/signed_distance.h"

namespace hexopt_shim {
void signed_distance_pseudonormal(const Eigen::MatrixXd& P, const Eigen::MatrixXd& V, const Eigen::MatrixXi& F,
                                  const igl::AABB<Eigen::MatrixXd, 3>& tree, const Eigen::MatrixXd& FN,
                                  const Eigen::MatrixXd& VN, const Eigen::MatrixXd& EN, const Eigen::VectorXi& EMAP,
                                  Eigen::VectorXd& S, Eigen::VectorXi& I, Eigen::MatrixXd& C, Eigen::MatrixXd& N);
}

namespace {
void put(FILE* f, const std::string& name, const void* p, size_t bytes, char kind) {
  uint32_t ln = (uint32_t)name.size();
  fwrite(&ln, 4, 1, f);
  fwrite(name.data(), 1, ln, f);
  fwrite(&kind, 1, 1, f);
  uint64_t nb = bytes;
  fwrite(&nb, 8, 1, f);
  if (bytes) fwrite(p, 1, bytes, f);
}
void put_all(FILE* f, const std::string& pre, const Eigen::VectorXd& S, const Eigen::VectorXi& I,
             const Eigen::MatrixXd& C, const Eigen::MatrixXd& N) {
  std::vector<double> c((size_t)C.rows() * 3), n((size_t)N.rows() * 3);
  for (int i = 0; i < C.rows(); ++i)
    for (int d = 0; d < 3; ++d) { c[(size_t)i * 3 + d] = C(i, d); n[(size_t)i * 3 + d] = N(i, d); }
  std::vector<int> iv(I.data(), I.data() + I.size());
  put(f, pre + "_S", S.data(), (size_t)S.size() * 8, 'd');
  put(f, pre + "_I", iv.data(), iv.size() * 4, 'i');
  put(f, pre + "_C", c.data(), c.size() * 8, 'd');
  put(f, pre + "_N", n.data(), n.size() * 8, 'd');
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t nv = 0, nf = 0, nq = 0;
  if (fread(&nv, 4, 1, in) != 1 || fread(&nf, 4, 1, in) != 1 || fread(&nq, 4, 1, in) != 1) return 3;
  std::vector<double> Vb((size_t)nv * 3), Pb((size_t)nq * 3);
  std::vector<int32_t> Fb((size_t)nf * 3);
  if (fread(Vb.data(), 8, Vb.size(), in) != Vb.size() || fread(Fb.data(), 4, Fb.size(), in) != Fb.size() ||
      fread(Pb.data(), 8, Pb.size(), in) != Pb.size())
    return 3;
  fclose(in);
  Eigen::MatrixXd V(nv, 3), P(nq, 3);
  Eigen::MatrixXi F(nf, 3);
  for (int i = 0; i < nv; ++i) for (int d = 0; d < 3; ++d) V(i, d) = Vb[(size_t)i * 3 + d];
  for (int i = 0; i < nf; ++i) for (int d = 0; d < 3; ++d) F(i, d) = Fb[(size_t)i * 3 + d];
  for (int i = 0; i < nq; ++i) for (int d = 0; d < 3; ++d) P(i, d) = Pb[(size_t)i * 3 + d];
  igl::AABB<Eigen::MatrixXd, 3> tree;
  Eigen::MatrixXd FN, VN, EN;
  Eigen::MatrixXi E;
  Eigen::VectorXi EMAP;
  tree.init(V, F);
  igl::per_face_normals(V, F, FN);
  igl::per_vertex_normals(V, F, igl::PER_VERTEX_NORMALS_WEIGHTING_TYPE_ANGLE, FN, VN);
  igl::per_edge_normals(V, F, igl::PER_EDGE_NORMALS_WEIGHTING_TYPE_UNIFORM, FN, EN, E, EMAP);
  FILE* out = fopen(argv[2], "wb");
  if (!out) return 2;
  Eigen::VectorXd S; Eigen::VectorXi I; Eigen::MatrixXd C, N;
  igl::signed_distance_pseudonormal(P, V, F, tree, FN, VN, EN, EMAP, S, I, C, N);
  put_all(out, "ref", S, I, C, N);
  hexopt_shim::signed_distance_pseudonormal(P, V, F, tree, FN, VN, EN, EMAP, S, I, C, N);
  put_all(out, "shim1", S, I, C, N);
  hexopt_shim::signed_distance_pseudonormal(P, V, F, tree, FN, VN, EN, EMAP, S, I, C, N);
  put_all(out, "shim", S, I, C, N);
  fclose(out);
  return 0;
}
