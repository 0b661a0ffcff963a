// io_ref_driver.cpp -- golden-vector driver for the reference's file readers and writers
// (io.cpp, class h_io). Test infrastructure: links the reference's io.cpp compiled where it
// lies; never shipped, never run on a GPU machine.
//
//   io_ref write <vtk|mesh|obj|mobj|mobjc|fgraph> <in.bin> <out-file>
//   io_ref read  <vtk|mesh|obj|fgraph> <file> <out.bin>
//
// in.bin (little-endian): int32 type, nV, dim, nC, ragged, k, nIdx, has_boundary; then
// double V[nV][dim]; int32 off[nC+1] when ragged; int32 idx[nIdx]; uint8 boundary[nV] when
// has_boundary. For fgraph: double angle; int32 orphan, orphan_single, nCorners, nEdges, nOther;
// int32 corners[nCorners]; int32 pairs[nEdges][2]; int32 other[nOther][2] (edges that are
// in the mesh but not flagged, interleaved with the flagged ones to exercise the filter).
//
// out.bin of a mesh read: int32 type, nV, nC, nIdx; double V[nV][3]; int32 off[nC+1];
// int32 idx[nIdx]. Of an fgraph read: int32 ok; double angle; int32 orphan, orphan_single,
// nCorners, nEdges; int32 corners[]; int32 pairs[][2].
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "io.h"

namespace {

struct Reader {
  FILE* f;
  explicit Reader(const char* p) : f(fopen(p, "rb")) {
    if (!f) throw std::runtime_error(std::string("cannot open ") + p);
  }
  ~Reader() { fclose(f); }
  template <class T>
  T get() {
    T x;
    if (fread(&x, sizeof x, 1, f) != 1) throw std::runtime_error("short input");
    return x;
  }
  template <class T>
  std::vector<T> arr(size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short input");
    return v;
  }
};

struct Writer {
  FILE* f;
  explicit Writer(const char* p) : f(fopen(p, "wb")) {
    if (!f) throw std::runtime_error(std::string("cannot open ") + p);
  }
  ~Writer() { fclose(f); }
  template <class T>
  void put(T x) { fwrite(&x, sizeof x, 1, f); }
  template <class T>
  void arr(const std::vector<T>& v) {
    if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), f);
  }
};

int write_mode(const std::string& fmt, const char* in, const char* out) {
  Reader r(in);
  h_io io;
  if (fmt == "fgraph") {
    Mesh_Feature mf;
    mf.angle_threshold = r.get<double>();
    mf.orphan_curve = r.get<int32_t>() != 0;
    mf.orphan_curve_single = r.get<int32_t>() != 0;
    const int nc = r.get<int32_t>(), ne = r.get<int32_t>(), no = r.get<int32_t>();
    auto corners = r.arr<int32_t>(nc);
    auto pairs = r.arr<int32_t>(2 * (size_t)ne);
    auto other = r.arr<int32_t>(2 * (size_t)no);
    for (int c : corners) mf.corners.push_back((uint32_t)c);
    // flagged and unflagged edges interleaved: flagged edge i goes before unflagged edge i
    int ie = 0, io_ = 0;
    while (ie < ne || io_ < no) {
      for (int pass = 0; pass < 2; ++pass) {
        const bool flagged = pass == 0;
        if (flagged ? ie >= ne : io_ >= no) continue;
        const int32_t* p = flagged ? &pairs[2 * ie++] : &other[2 * io_++];
        Hybrid_E e;
        e.id = (uint32_t)mf.tri.Es.size();
        e.vs = {(uint32_t)p[0], (uint32_t)p[1]};
        mf.tri.Es.push_back(e);
        mf.E_feature_flag.push_back(flagged);
      }
    }
    io.write_feature_Graph_FGRAPH(mf, out);
    return 0;
  }
  const int type = r.get<int32_t>(), nV = r.get<int32_t>(), dim = r.get<int32_t>(), nC = r.get<int32_t>();
  const int ragged = r.get<int32_t>(), k = r.get<int32_t>(), nIdx = r.get<int32_t>(), hb = r.get<int32_t>();
  auto V = r.arr<double>((size_t)nV * dim);
  std::vector<int32_t> off;
  if (ragged) off = r.arr<int32_t>(nC + 1);
  else
    for (int c = 0; c <= nC; ++c) off.push_back(c * k);
  auto idx = r.arr<int32_t>(nIdx);
  std::vector<uint8_t> bnd = hb ? r.arr<uint8_t>(nV) : std::vector<uint8_t>(nV, 0);
  if (fmt == "mobj" || fmt == "mobjc") {
    const bool col = fmt == "mobjc";
    Eigen::MatrixXd M(col ? dim : nV, col ? nV : dim);
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < dim; ++d) (col ? M(d, i) : M(i, d)) = V[(size_t)i * dim + d];
    Eigen::MatrixXi F(col ? k : nC, col ? nC : k);
    for (int c = 0; c < nC; ++c)
      for (int j = 0; j < k; ++j) (col ? F(j, c) : F(c, j)) = idx[(size_t)c * k + j];
    io.write_mesh_OBJ(M, F, col, out);
    return 0;
  }
  Mesh m;
  m.type = (short)type;
  m.V.resize(3, nV);
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < 3; ++d) m.V(d, i) = V[(size_t)i * dim + d];
  m.Vs.resize(nV);
  for (int i = 0; i < nV; ++i) {
    m.Vs[i].id = i;
    m.Vs[i].boundary = bnd[i] != 0;
    m.Vs[i].v = {m.V(0, i), m.V(1, i), m.V(2, i)};
  }
  const bool faces = fmt == "obj" || type == Mesh_type::Tri || type == Mesh_type::Qua || type == Mesh_type::Hyb ||
                     (fmt == "mesh" && type == Mesh_type::HSur);
  for (int c = 0; c < nC; ++c) {
    std::vector<uint32_t> vs;
    for (int j = off[c]; j < off[c + 1]; ++j) vs.push_back((uint32_t)idx[j]);
    if (faces) {
      Hybrid_F f;
      f.id = c;
      f.vs = vs;
      m.Fs.push_back(f);
    } else {
      Hybrid h;
      h.id = c;
      h.vs = vs;
      m.Hs.push_back(h);
    }
  }
  if (fmt == "vtk") io.write_hybrid_mesh_VTK(m, out);
  else if (fmt == "mesh") io.write_hybrid_mesh_MESH(m, out);
  else if (fmt == "obj") io.write_hybrid_mesh_OBJ(m, out);
  else throw std::runtime_error("unknown format " + fmt);
  return 0;
}

int read_mode(const std::string& fmt, const char* path, const char* out) {
  h_io io;
  Writer w(out);
  if (fmt == "fgraph") {
    Mesh_Feature mf;
    const bool ok = io.read_feature_Graph_FGRAPH(mf, path);
    w.put<int32_t>(ok);
    w.put<double>(mf.angle_threshold);
    w.put<int32_t>(mf.orphan_curve);
    w.put<int32_t>(mf.orphan_curve_single);
    w.put<int32_t>((int32_t)mf.IN_corners.size());
    w.put<int32_t>((int32_t)mf.IN_v_pairs.size());
    for (uint32_t c : mf.IN_corners) w.put<int32_t>((int32_t)c);
    for (auto& p : mf.IN_v_pairs) {
      w.put<int32_t>((int32_t)p[0]);
      w.put<int32_t>((int32_t)p[1]);
    }
    return 0;
  }
  Mesh m;
  m.type = -2;  // the VTK and OBJ readers leave it to the caller
  bool ok = true;
  if (fmt == "vtk") ok = io.read_hybrid_mesh_VTK(m, path);
  else if (fmt == "mesh") ok = io.read_hybrid_mesh_MESH(m, path);
  else if (fmt == "obj") io.read_hybrid_mesh_OBJ(m, path);
  else throw std::runtime_error("unknown format " + fmt);
  if (!ok) throw std::runtime_error("reader failed");
  const bool faces = fmt == "obj";
  const int nV = (int)m.V.cols();
  const int nC = faces ? (int)m.Fs.size() : (int)m.Hs.size();
  std::vector<int32_t> off(1, 0), idx;
  for (int c = 0; c < nC; ++c) {
    const auto& vs = faces ? m.Fs[c].vs : m.Hs[c].vs;
    for (uint32_t v : vs) idx.push_back((int32_t)v);
    off.push_back((int32_t)idx.size());
  }
  w.put<int32_t>(m.type);
  w.put<int32_t>(nV);
  w.put<int32_t>(nC);
  w.put<int32_t>((int32_t)idx.size());
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < 3; ++d) w.put<double>(m.V(d, i));
  w.arr(off);
  w.arr(idx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: io_ref write|read <format> <in> <out>\n");
    return 2;
  }
  try {
    const std::string mode = argv[1], fmt = argv[2];
    if (mode == "write") return write_mode(fmt, argv[3], argv[4]);
    if (mode == "read") return read_mode(fmt, argv[3], argv[4]);
    fprintf(stderr, "unknown mode %s\n", argv[1]);
    return 2;
  } catch (const std::exception& e) {
    fprintf(stderr, "io_ref: %s\n", e.what());
    return 1;
  }
}
