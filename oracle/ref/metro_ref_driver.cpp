// metro_ref_driver.cpp -- golden-vector driver for the reference's metro Hausdorff distance
// (compute(const Mesh&, const Mesh&, bbox_diagonal, max, mean), metro_hausdorff.cpp:358-505),
// linked against the reference's own metro_hausdorff.cpp compiled where it lies with VCG's
// bundled Eigen. Test infrastructure only: run in the development container.
//
// in.bin:  int nv0, nf0; double V0[nv0*3]; int F0[nf0*3]; int nv1, nf1; V1; F1
// out.bin: records (uint32 name length, name, 'd', uint64 byte count, data) for
//          "bbox_diagonal", "max", "mean"
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include "metro_hausdorff.h"

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

static bool read_mesh(FILE* f, Mesh& m) {
  int nv = 0, nf = 0;
  if (!rd(f, &nv, 4) || !rd(f, &nf, 4)) return false;
  std::vector<double> V((size_t)nv * 3);
  std::vector<int> F((size_t)nf * 3);
  if (!rd(f, V.data(), V.size() * 8) || !rd(f, F.data(), F.size() * 4)) return false;
  m.V.resize(3, nv);
  for (int i = 0; i < nv; ++i)
    for (int d = 0; d < 3; ++d) m.V(d, i) = V[(size_t)i * 3 + d];
  m.Fs.resize(nf);
  for (int i = 0; i < nf; ++i) {
    m.Fs[i].id = i;
    m.Fs[i].vs = {(uint32_t)F[(size_t)i * 3], (uint32_t)F[(size_t)i * 3 + 1], (uint32_t)F[(size_t)i * 3 + 2]};
  }
  return true;
}

static void rec(FILE* f, const char* name, double v) {
  uint32_t ln = (uint32_t)strlen(name);
  uint64_t nb = 8;
  fwrite(&ln, 4, 1, f);
  fwrite(name, 1, ln, f);
  fputc('d', f);
  fwrite(&nb, 8, 1, f);
  fwrite(&v, 8, 1, f);
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* fi = fopen(argv[1], "rb");
  if (!fi) return 3;
  Mesh m0, m1;
  m0.type = m1.type = 0;
  if (!read_mesh(fi, m0) || !read_mesh(fi, m1)) return 4;
  fclose(fi);
  double diag = 0, mx = 0, mean = 0;
  // the const overload is the one stop_criterior_satisfied calls (grid_hex_meshing.cpp:4154);
  // with non-const meshes overload resolution would pick the ratio/threshold variant
  const Mesh& c0 = m0;
  const Mesh& c1 = m1;
  compute(c0, c1, diag, mx, mean);
  FILE* fo = fopen(argv[2], "wb");
  if (!fo) return 5;
  rec(fo, "bbox_diagonal", diag);
  rec(fo, "max", mx);
  rec(fo, "mean", mean);
  fclose(fo);
  return 0;
}
