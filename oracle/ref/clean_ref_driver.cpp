// clean_ref_driver.cpp -- runs the flag passes of the reference's clean_hex_mesh
// (grid_meshing/grid_hex_meshing.cpp:1932-1964) on a hex mesh with given inside flags:
// tagging_uneven_element (:1983-2005), re_indexing_connectivity, clean_non_manifold_ve
// (:2006-2080) and drop_small_pieces (:2081-2124), all the reference's own code compiled from
// its sources (oracle/ref/Makefile). Writes the flags after each pass.
// Input: nV, nH (int32), H (nH x 8 int32), flags (nH int32). Output (refio format): "tag",
// "manifold", "final" (nH doubles, 0 / 1), then the medial-surface tags of the final flags
// (:1966-1981): "medial_v" (nV doubles, 0 / 1) and "medial_f" (4 vertex ids per tagged face,
// in face order).
// Test infrastructure only; built in the development container, never shipped.
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include "grid_meshing/grid_hex_meshing.h"

namespace {
void put(FILE* f, const std::string& name, const std::vector<bool>& flags) {
  std::vector<double> v(flags.begin(), flags.end());
  uint32_t ln = (uint32_t)name.size();
  fwrite(&ln, 4, 1, f);
  fwrite(name.data(), 1, ln, f);
  char kind = 'd';
  fwrite(&kind, 1, 1, f);
  uint64_t nb = v.size() * 8;
  fwrite(&nb, 8, 1, f);
  if (nb) fwrite(v.data(), 1, nb, f);
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t nV = 0, nH = 0;
  if (fread(&nV, 4, 1, in) != 1 || fread(&nH, 4, 1, in) != 1) return 3;
  std::vector<int32_t> H((size_t)nH * 8), fl(nH);
  if (fread(H.data(), 4, H.size(), in) != H.size() || fread(fl.data(), 4, fl.size(), in) != fl.size()) return 3;
  fclose(in);

  Mesh_Domain md;
  Mesh& mi = md.mesh_entire;
  mi.type = Mesh_type::Hex;
  mi.V = Eigen::MatrixXd::Zero(3, nV);
  for (int v = 0; v < nV; ++v) {
    Hybrid_V hv;
    hv.id = v;
    hv.v = {0.0, 0.0, 0.0};
    mi.Vs.push_back(hv);
  }
  for (int h = 0; h < nH; ++h) {
    Hybrid hy;
    hy.id = h;
    for (int i = 0; i < 8; ++i) hy.vs.push_back(H[(size_t)h * 8 + i]);
    mi.Hs.push_back(hy);
  }
  build_connectivity(mi);
  md.H_flag.assign(nH, false);
  for (int h = 0; h < nH; ++h) md.H_flag[h] = fl[h] != 0;

  grid_hex_meshing_bijective g;
  FILE* out = fopen(argv[2], "wb");
  if (!out) return 2;
  g.tagging_uneven_element(mi, md.H_flag);
  put(out, "tag", md.H_flag);
  re_indexing_connectivity(mi, md.H_flag, md.mesh_subA, md.V_map, md.V_map_reverse, md.H_map, md.H_map_reverse);
  if (md.mesh_subA.Hs.size()) {
    Eigen::VectorXd signed_dis = Eigen::VectorXd::Zero(nH);
    g.clean_non_manifold_ve(mi, md.mesh_subA, md.V_map, md.V_map_reverse, md.H_map, md.H_map_reverse, signed_dis,
                            md.H_flag);
    put(out, "manifold", md.H_flag);
    g.drop_small_pieces(md);
  } else {
    put(out, "manifold", md.H_flag);
  }
  put(out, "final", md.H_flag);
  // the medial-surface tags that end clean_hex_mesh (grid_hex_meshing.cpp:1966-1981), on the
  // reference's own faces of the whole mesh; written as vertex flags and the tagged faces'
  // vertex ids in face order
  std::vector<bool> vmed(nV, false);
  std::vector<double> fmed;
  for (auto& f : mi.Fs) {
    if ((f.boundary && md.H_flag[f.neighbor_hs[0]]) ||
        (!f.boundary && (md.H_flag[f.neighbor_hs[0]] != md.H_flag[f.neighbor_hs[1]]))) {
      f.on_medial_surface = true;
      for (auto& vid : f.vs) {
        mi.Vs[vid].on_medial_surface = true;
        vmed[vid] = true;
        fmed.push_back(vid);
      }
    }
  }
  put(out, "medial_v", vmed);
  {
    uint32_t ln = 8;
    fwrite(&ln, 4, 1, out);
    fwrite("medial_f", 1, 8, out);
    char kind = 'd';
    fwrite(&kind, 1, 1, out);
    uint64_t nb = fmed.size() * 8;
    fwrite(&nb, 8, 1, out);
    if (nb) fwrite(fmed.data(), 1, nb, out);
  }
  fclose(out);
  return 0;
}
