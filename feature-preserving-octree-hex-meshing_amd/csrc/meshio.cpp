// meshio.cpp -- VTK, Medit .mesh, OBJ and .fgraph files (declared in hexio.h). Host only.
//
// Writers format numbers with std::to_chars, which is specified to print what C's printf
// prints for the same conversion: %g with precision 6 is what an std::ostream with default
// flags writes, so the bytes match the reference's writers. Readers take the whole file into
// memory, cut it into NUL-terminated lines and parse the long sections (points, cells) on
// several threads with strtod/strtol, the conversions sscanf's %lf/%d/%i use. Large meshes
// are formatted on several threads into per-range buffers written out in order.
#include "../../include/hexio.h"

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

struct hexio_mesh {
  int type = HEXIO_HEX;
  std::vector<double> V;     // [nV][3]
  std::vector<int32_t> off;  // nCells + 1
  std::vector<int32_t> idx;
};

struct hexio_fgraph {
  double angle = 0;
  int oc = 1, ocs = 1;
  std::vector<int32_t> corners;
  std::vector<int32_t> pairs;
};

namespace {

thread_local std::string g_err;

bool fail(const std::string& m) {
  g_err = m;
  return false;
}

// ---- threading -------------------------------------------------------------------------
constexpr size_t kLinesPerThread = 1 << 15;

int thread_count(size_t n) {
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  size_t t = std::min<size_t>(std::min<unsigned>(hc, 16), (n + kLinesPerThread - 1) / kLinesPerThread);
  return (int)std::max<size_t>(t, 1);
}

// fn(tid, begin, end) over [0, n) cut into contiguous ranges, one per thread
template <class F>
void parallel_ranges(size_t n, int nt, F&& fn) {
  if (nt <= 1) {
    fn(0, (size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    size_t b = n * t / nt, e = n * (t + 1) / nt;
    th.emplace_back([&fn, t, b, e] { fn(t, b, e); });
  }
  for (auto& x : th) x.join();
}

// ---- output ------------------------------------------------------------------------------
struct Out {
  FILE* f = nullptr;
  bool ok = true;
  explicit Out(const char* path) : f(std::fopen(path, "wb")) {}
  ~Out() {
    if (f) std::fclose(f);
  }
  void put(const char* s, size_t n) {
    if (ok && n && std::fwrite(s, 1, n, f) != n) ok = false;
  }
  void put(const std::string& s) { put(s.data(), s.size()); }
  bool close() {
    if (!f) return false;
    bool good = ok && std::fclose(f) == 0;
    f = nullptr;
    return good;
  }
};

inline void put_g(std::string& s, double x) {  // %g, 6 significant digits
  char b[32];
  auto r = std::to_chars(b, b + sizeof b, x, std::chars_format::general, 6);
  s.append(b, r.ptr);
}

template <class I>
inline void put_i(std::string& s, I x) {
  char b[24];
  auto r = std::to_chars(b, b + sizeof b, x);
  s.append(b, r.ptr);
}

// lines [0, n) made by line(i, buf), formatted on several threads, written in order
template <class F>
void emit_lines(Out& o, size_t n, F&& line) {
  const int nt = thread_count(n);
  std::vector<std::string> part(nt);
  parallel_ranges(n, nt, [&](int t, size_t b, size_t e) {
    std::string& s = part[t];
    s.reserve((e - b) * 48);
    for (size_t i = b; i < e; ++i) line(i, s);
  });
  for (auto& s : part) o.put(s);
}

struct CellView {
  int nC, k;
  const int32_t* off;
  const int32_t* idx;
  size_t begin(size_t c) const { return off ? (size_t)off[c] : c * (size_t)k; }
  size_t size(size_t c) const { return off ? (size_t)(off[c + 1] - off[c]) : (size_t)k; }
};

bool check_cells(int nV, const CellView& cv) {
  if (nV < 0 || cv.nC < 0) return fail("negative size");
  if (!cv.off && cv.k < 0) return fail("negative cell size");
  if (cv.nC > 0 && !cv.idx) return fail("cell ids missing");
  if (cv.off) {
    if (cv.off[0] != 0) return fail("cell offsets must start at 0");
    for (int c = 0; c < cv.nC; ++c)
      if (cv.off[c + 1] < cv.off[c]) return fail("cell offsets must not decrease");
  }
  const size_t len = cv.nC > 0 ? cv.begin(cv.nC - 1) + cv.size(cv.nC - 1) : 0;
  for (size_t j = 0; j < len; ++j)
    if (cv.idx[j] < 0 || cv.idx[j] >= nV)
      return fail("cell vertex id " + std::to_string(cv.idx[j]) + " outside the " + std::to_string(nV) +
                  " vertices");
  return true;
}

bool check_open(Out& o, const char* path) {
  if (!o.f) return fail(std::string("cannot open ") + path + " for writing: " + std::strerror(errno));
  return true;
}

bool finish(Out& o, const char* path) {
  if (!o.close()) return fail(std::string("write to ") + path + " failed");
  return true;
}

// ---- input -------------------------------------------------------------------------------
// The file cut into NUL-terminated lines (the '\n' replaced, a '\r' before it kept, as
// std::getline keeps it).
struct Lines {
  std::string data;
  std::vector<size_t> start;
  size_t size() const { return start.size(); }
  char* operator[](size_t i) { return &data[start[i]]; }
};

bool load_lines(const char* path, Lines& L) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::string& d = L.data;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) d.append(buf, n);
  bool bad = std::ferror(f);
  std::fclose(f);
  if (bad) return fail(std::string("read of ") + path + " failed");
  L.start.clear();
  size_t p = 0;
  while (p < d.size()) {
    L.start.push_back(p);
    const void* nl = std::memchr(&d[p], '\n', d.size() - p);
    if (!nl) break;
    size_t q = (const char*)nl - d.data();
    d[q] = '\0';
    p = q + 1;
  }
  d.push_back('\0');  // the last line's terminator when the file does not end with '\n'
  return true;
}

inline const char* skip_ws(const char* s) {
  while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\v' || *s == '\f' || *s == '\n') ++s;
  return s;
}

// sscanf's %s: the next word (after white space); empty when the line has none
inline const char* word(const char* s, std::string& w) {
  s = skip_ws(s);
  const char* e = s;
  while (*e && !(*e == ' ' || *e == '\t' || *e == '\r' || *e == '\v' || *e == '\f' || *e == '\n')) ++e;
  w.assign(s, e);
  return e;
}

inline bool read_double(const char*& s, double& x) {
  char* e;
  x = std::strtod(s, &e);
  if (e == s) return false;
  s = e;
  return true;
}

inline bool read_long(const char*& s, long& x, int base) {
  char* e;
  x = std::strtol(s, &e, base);
  if (e == s) return false;
  s = e;
  return true;
}

std::string at_line(const char* path, size_t i) { return std::string(path) + ":" + std::to_string(i + 1); }

// n lines from `first` of 3 numbers each into V; false (message set) at the first bad line
bool parse_points(Lines& L, size_t first, size_t n, const char* path, std::vector<double>& V) {
  if (first + n > L.size()) return fail(std::string(path) + ": file ends inside the point list");
  V.resize(3 * n);
  const int nt = thread_count(n);
  std::vector<size_t> bad(nt, SIZE_MAX);
  parallel_ranges(n, nt, [&](int t, size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const char* s = L[first + i];
      if (!read_double(s, V[3 * i]) || !read_double(s, V[3 * i + 1]) || !read_double(s, V[3 * i + 2])) {
        bad[t] = i;
        return;
      }
    }
  });
  for (size_t b : bad)
    if (b != SIZE_MAX) return fail(at_line(path, first + b) + ": expected three numbers");
  return true;
}

bool check_ids(const hexio_mesh& m, const char* path) {
  const long nV = (long)(m.V.size() / 3);
  for (size_t i = 0; i < m.idx.size(); ++i)
    if (m.idx[i] < 0 || m.idx[i] >= nV)
      return fail(std::string(path) + ": cell vertex id " + std::to_string(m.idx[i]) + " outside the " +
                  std::to_string(nV) + " vertices");
  return true;
}

int uniform_type(const hexio_mesh& m, int if8, int if4, int other) {
  const size_t nC = m.off.size() - 1;
  bool all8 = true, all4 = true;
  for (size_t c = 0; c < nC; ++c) {
    const int k = m.off[c + 1] - m.off[c];
    all8 &= k == 8;
    all4 &= k == 4;
  }
  if (nC > 0 && all8) return if8;
  if (nC > 0 && all4) return if4;
  return other;
}

// ---- VTK -----------------------------------------------------------------------------------
hexio_mesh* read_vtk(const char* path) {
  Lines L;
  if (!load_lines(path, L)) return nullptr;
  std::string w;
  size_t i = 0;
  long np = -1;
  // "POINTS n <type>" within the first 10 lines (io.cpp:11-19)
  for (; i < L.size() && i < 10; ++i) {
    const char* s = word(L[i], w);
    long n;
    if (w != "POINTS" || !read_long(s, n, 10)) continue;
    std::string w2;
    word(s, w2);
    if (w2.empty()) continue;
    if (n < 0) {
      fail(at_line(path, i) + ": negative point count");
      return nullptr;
    }
    np = n;
    ++i;
    break;
  }
  if (np < 0) {
    fail(std::string(path) + ": no \"POINTS n type\" line in the first 10 lines");
    return nullptr;
  }
  auto m = std::make_unique<hexio_mesh>();
  if (!parse_points(L, i, (size_t)np, path, m->V)) return nullptr;
  i += np;
  // first "CELLS n m" line after the points (io.cpp:37-44)
  long nc = -1;
  for (; i < L.size(); ++i) {
    const char* s = word(L[i], w);
    long n, t;
    if (w == "CELLS" && read_long(s, n, 10) && read_long(s, t, 10)) {
      if (n < 0) {
        fail(at_line(path, i) + ": negative cell count");
        return nullptr;
      }
      nc = n;
      ++i;
      break;
    }
  }
  if (nc < 0) {
    fail(std::string(path) + ": no \"CELLS n size\" line after the points");
    return nullptr;
  }
  if ((size_t)i + nc > L.size()) {
    fail(std::string(path) + ": file ends inside the cell list");
    return nullptr;
  }
  // one cell per line, "k v0 .. v(k-1)" with k in 4, 5, 6, 8 (io.cpp:49-92)
  std::vector<int32_t> tmp((size_t)nc * 8), kk(nc);
  const int nt = thread_count(nc);
  std::vector<size_t> bad(nt, SIZE_MAX);
  parallel_ranges(nc, nt, [&](int t, size_t b, size_t e) {
    for (size_t c = b; c < e; ++c) {
      const char* s = L[i + c];
      long k, v;
      if (!read_long(s, k, 10) || !(k == 4 || k == 5 || k == 6 || k == 8)) {
        bad[t] = c;
        return;
      }
      for (int j = 0; j < k; ++j) {
        if (!read_long(s, v, 10)) {
          bad[t] = c;
          return;
        }
        tmp[8 * c + j] = (int32_t)v;
      }
      kk[c] = (int32_t)k;
    }
  });
  for (size_t b : bad)
    if (b != SIZE_MAX) {
      fail(at_line(path, i + b) + ": expected a cell line \"k ids\" with k in 4, 5, 6, 8");
      return nullptr;
    }
  m->off.resize(nc + 1);
  m->off[0] = 0;
  for (long c = 0; c < nc; ++c) m->off[c + 1] = m->off[c] + kk[c];
  m->idx.resize(m->off[nc]);
  for (long c = 0; c < nc; ++c) std::copy(&tmp[8 * c], &tmp[8 * c] + kk[c], &m->idx[m->off[c]]);
  if (!check_ids(*m, path)) return nullptr;
  m->type = uniform_type(*m, HEXIO_HEX, HEXIO_TET, HEXIO_HYB);
  return m.release();
}

int write_vtk(const char* path, int type, int nV, const double* V, const CellView& cv, const uint8_t* boundary) {
  if (!check_cells(nV, cv)) return 1;
  if (nV > 0 && !V) return fail("vertex positions missing"), 1;
  Out o(path);
  if (!check_open(o, path)) return 1;
  std::string h;
  h += "# vtk DataFile Version 2.0\nmesh vtk data - converted from .off\nASCII\nDATASET UNSTRUCTURED_GRID\n";
  h += "POINTS ";
  put_i(h, nV);
  h += " double\n";
  o.put(h);
  emit_lines(o, nV, [&](size_t i, std::string& s) {
    put_g(s, V[3 * i]);
    s += ' ';
    put_g(s, V[3 * i + 1]);
    s += ' ';
    put_g(s, V[3 * i + 2]);
    s += '\n';
  });
  // cells: TRI/QUA faces at the first face's size, HYB faces at their own sizes, anything
  // else as volume cells at the first cell's size (io.cpp:118-169)
  const bool hyb = type == HEXIO_HYB;
  const size_t k0 = cv.nC > 0 ? cv.size(0) : 0;
  uint64_t total = 0;
  if (hyb)
    for (int c = 0; c < cv.nC; ++c) total += cv.size(c) + 1;
  else
    total = (uint64_t)cv.nC * (k0 + 1);
  h.clear();
  h += "CELLS ";
  put_i(h, cv.nC);
  h += ' ';
  put_i(h, total);
  h += '\n';
  o.put(h);
  emit_lines(o, cv.nC, [&](size_t c, std::string& s) {
    const size_t k = hyb ? cv.size(c) : k0;
    const int32_t* ids = cv.idx + cv.begin(c);
    s += ' ';
    put_i(s, k);
    s += ' ';
    for (size_t j = 0; j < k; ++j) {
      put_i(s, ids[j]);
      s += ' ';
    }
    s += '\n';
  });
  const char* ct = type == HEXIO_TRI ? "5\n" : type == HEXIO_QUA ? "9\n" : hyb ? "7\n" : type == HEXIO_TET ? "10\n" : "12\n";
  h.clear();
  h += "CELL_TYPES ";
  put_i(h, cv.nC);
  h += '\n';
  o.put(h);
  {
    std::string s;
    const size_t lc = std::strlen(ct);
    s.reserve(lc * cv.nC);
    for (int c = 0; c < cv.nC; ++c) s.append(ct, lc);
    o.put(s);
  }
  h.clear();
  h += "POINT_DATA ";
  put_i(h, nV);
  h += "\nSCALARS fixed int\nLOOKUP_TABLE default\n";
  o.put(h);
  {
    std::string s((size_t)2 * nV, '\n');
    for (int i = 0; i < nV; ++i) s[2 * i] = boundary && boundary[i] ? '1' : '0';
    o.put(s);
  }
  return finish(o, path) ? 0 : 1;
}

// ---- Medit .mesh ---------------------------------------------------------------------------
hexio_mesh* read_mesh(const char* path) {
  Lines L;
  if (!load_lines(path, L)) return nullptr;
  std::string w;
  size_t i = 0;
  // a section header: the keyword, then its count on the same line or on the next
  auto header = [&](size_t& li, long& n) -> bool {
    const char* s = word(L[li], w);
    if (read_long(s, n, 10)) return true;
    if (li + 1 >= L.size()) return false;
    const char* t = L[++li];
    return read_long(t, n, 10);
  };
  long nv = -1;
  for (; i < L.size(); ++i) {
    word(L[i], w);
    if (w != "Vertices") continue;
    if (!header(i, nv) || nv < 0) {
      fail(at_line(path, i) + ": bad vertex count");
      return nullptr;
    }
    ++i;
    break;
  }
  if (nv < 0) {
    fail(std::string(path) + ": no Vertices section");
    return nullptr;
  }
  auto m = std::make_unique<hexio_mesh>();
  if (!parse_points(L, i, (size_t)nv, path, m->V)) return nullptr;
  i += nv;
  long nc = -1;
  int k = 0;
  for (; i < L.size(); ++i) {
    word(L[i], w);
    if (w != "Hexahedra" && w != "Tetrahedra") continue;
    k = w == "Hexahedra" ? 8 : 4;
    if (!header(i, nc) || nc < 0) {
      fail(at_line(path, i) + ": bad cell count");
      return nullptr;
    }
    ++i;
    break;
  }
  if (nc < 0) {
    fail(std::string(path) + ": no Hexahedra or Tetrahedra section after the vertices");
    return nullptr;
  }
  if ((size_t)i + nc > L.size()) {
    fail(std::string(path) + ": file ends inside the cell list");
    return nullptr;
  }
  m->type = k == 8 ? HEXIO_HEX : HEXIO_TET;
  m->off.resize(nc + 1);
  for (long c = 0; c <= nc; ++c) m->off[c] = (int32_t)(c * k);
  m->idx.resize((size_t)nc * k);
  const int nt = thread_count(nc);
  std::vector<size_t> bad(nt, SIZE_MAX);
  parallel_ranges(nc, nt, [&](int t, size_t b, size_t e) {
    for (size_t c = b; c < e; ++c) {
      const char* s = L[i + c];
      for (int j = 0; j < k; ++j) {
        long v;
        if (!read_long(s, v, 10)) {
          bad[t] = c;
          return;
        }
        m->idx[c * k + j] = (int32_t)(v - 1);
      }
    }
  });
  for (size_t b : bad)
    if (b != SIZE_MAX) {
      fail(at_line(path, i + b) + ": expected " + std::to_string(k) + " vertex ids");
      return nullptr;
    }
  if (!check_ids(*m, path)) return nullptr;
  return m.release();
}

int write_mesh(const char* path, int type, int nV, const double* V, const CellView& cv) {
  if (!check_cells(nV, cv)) return 1;
  if (nV > 0 && !V) return fail("vertex positions missing"), 1;
  const bool tri = type == HEXIO_TRI || type == HEXIO_HSUR, hex = type == HEXIO_HEX;
  if (tri)
    for (int c = 0; c < cv.nC; ++c)
      if (cv.size(c) < 3) return fail("a triangle needs 3 vertex ids"), 1;
  Out o(path);
  if (!check_open(o, path)) return 1;
  std::string h = "MeshVersionFormatted 1\nDimension 3\nVertices ";
  put_i(h, nV);
  h += '\n';
  o.put(h);
  emit_lines(o, nV, [&](size_t i, std::string& s) {
    put_g(s, V[3 * i]);
    s += ' ';
    put_g(s, V[3 * i + 1]);
    s += ' ';
    put_g(s, V[3 * i + 2]);
    s += " 0\n";
  });
  if (tri || hex) {
    h = tri ? "Triangles\n" : "Hexahedra\n";
    put_i(h, cv.nC);
    h += '\n';
    o.put(h);
    emit_lines(o, cv.nC, [&](size_t c, std::string& s) {
      const int32_t* ids = cv.idx + cv.begin(c);
      const size_t n = tri ? 3 : cv.size(c);
      for (size_t j = 0; j < n; ++j) {
        put_i(s, (int64_t)ids[j] + 1);
        s += ' ';
      }
      s += "0\n";
    });
  }
  o.put("End", 3);
  return finish(o, path) ? 0 : 1;
}

// ---- OBJ -------------------------------------------------------------------------------------
// One "f" word: i, i/t, i//n or i/t/n, as readOBJ tries them with sscanf (readOBJ.cpp:145-170).
// Returns 0 on a bad word, else 1 and which of t and n were present.
int face_word(const char* s, long& i, bool& has_t, bool& has_n) {
  has_t = has_n = false;
  if (!read_long(s, i, 10)) return 0;
  if (*s != '/') return 1;
  const char* p = s + 1;
  long t;
  if (read_long(p, t, 10)) {
    has_t = true;
    if (*p == '/') {
      const char* q = p + 1;
      long n;
      if (read_long(q, n, 10)) has_n = true;
    }
    return 1;
  }
  if (*p == '/') {
    const char* q = p + 1;
    long n;
    if (read_long(q, n, 10)) has_n = true;
  }
  return 1;
}

hexio_mesh* read_obj(const char* path) {
  Lines L;
  if (!load_lines(path, L)) return nullptr;
  auto m = std::make_unique<hexio_mesh>();
  std::vector<int32_t> fidx;
  std::vector<int32_t> foff(1, 0);
  std::string type, w;
  long nv = 0;
  int vdim = -1;
  for (size_t li = 0; li < L.size(); ++li) {
    const char* line = L[li];
    word(line, type);
    if (type.empty()) continue;
    // the rest of the line starts strlen(type) characters into it, as in readOBJ.cpp:58
    const char* l = line + std::min(type.size(), std::strlen(line));
    if (type == "v") {
      double x[4];
      int cnt = 0;
      const char* s = l;
      while (cnt < 4 && read_double(s, x[cnt])) ++cnt;
      if (cnt != 3 && cnt != 4) {
        fail(at_line(path, li) + ": a vertex needs 3 or 4 coordinates");
        return nullptr;
      }
      if (vdim >= 0 && vdim != cnt) {
        fail(at_line(path, li) + ": vertices with 3 and with 4 coordinates are mixed");
        return nullptr;
      }
      vdim = cnt;
      m->V.insert(m->V.end(), x, x + 3);
      ++nv;
    } else if (type == "vn" || type == "vt") {
      double x[3];
      int cnt = 0;
      const char* s = l;
      while (cnt < 3 && read_double(s, x[cnt])) ++cnt;
      if (type == "vn" ? cnt != 3 : (cnt != 2 && cnt != 3)) {
        fail(at_line(path, li) + (type == "vn" ? ": a normal needs 3 coordinates" :
                                                  ": texture coordinates need 2 or 3 numbers"));
        return nullptr;
      }
    } else if (type == "f") {
      const char* s = l;
      size_t nf = 0, nt = 0, nn = 0;
      for (;;) {
        s = word(s, w);
        if (w.empty()) break;
        long i;
        bool ht, hn;
        if (!face_word(w.c_str(), i, ht, hn)) {
          fail(at_line(path, li) + ": face has an invalid element \"" + w + "\"");
          return nullptr;
        }
        const long id = i < 0 ? i + nv : i - 1;
        fidx.push_back((int32_t)id);
        ++nf;
        nt += ht;
        nn += hn;
      }
      const bool ok = nf > 0 && (nn == 0 || nn == nf) && (nt == 0 || nt == nf);
      if (!ok) {
        fail(at_line(path, li) + ": face has an invalid format");
        return nullptr;
      }
      foff.push_back((int32_t)fidx.size());
    }
    // '#', 'g', 's' lines, usemtl and mtllib are ignored; other lines are skipped too (readOBJ
    // warns about them)
  }
  const size_t nF = foff.size() - 1;
  int k = nF > 0 ? foff[1] - foff[0] : 0;
  for (size_t f = 0; f < nF; ++f)
    if (foff[f + 1] - foff[f] != k) {
      fail(std::string(path) + ": faces have different numbers of vertices");
      return nullptr;
    }
  m->off.swap(foff);
  m->idx.swap(fidx);
  if (!check_ids(*m, path)) return nullptr;
  m->type = k == 3 ? HEXIO_TRI : k == 4 ? HEXIO_QUA : HEXIO_HYB;
  return m.release();
}

int write_obj(const char* path, int nV, int dim, const double* V, const CellView& cv, bool trailing_space) {
  if (!check_cells(nV, cv)) return 1;
  if (dim < 0) return fail("negative dimension"), 1;
  if (nV > 0 && dim > 0 && !V) return fail("vertex positions missing"), 1;
  Out o(path);
  if (!check_open(o, path)) return 1;
  emit_lines(o, nV, [&](size_t i, std::string& s) {
    s += 'v';
    for (int d = 0; d < dim; ++d) {
      s += ' ';
      put_g(s, V[(size_t)dim * i + d]);
    }
    if (trailing_space) s += ' ';
    s += '\n';
  });
  emit_lines(o, cv.nC, [&](size_t c, std::string& s) {
    const int32_t* ids = cv.idx + cv.begin(c);
    const size_t n = cv.size(c);
    s += 'f';
    for (size_t j = 0; j < n; ++j) {
      s += ' ';
      put_i(s, (int64_t)ids[j] + 1);
    }
    s += '\n';
  });
  return finish(o, path) ? 0 : 1;
}

// ---- .fgraph -------------------------------------------------------------------------------
hexio_fgraph* read_fgraph(const char* path) {
  Lines L;
  if (!load_lines(path, L)) return nullptr;
  auto g = std::make_unique<hexio_fgraph>();
  size_t li = 0;
  auto next = [&](const char*& s) -> bool {
    if (li >= L.size()) return false;
    s = L[li++];
    return true;
  };
  const char* s;
  long a, b;
  if (!next(s) || !read_double(s, g->angle) || !read_long(s, a, 0) || !read_long(s, b, 0)) {
    fail(at_line(path, 0) + ": expected \"angle_threshold orphan_curve orphan_curve_single\"");
    return nullptr;
  }
  g->oc = a != 0;
  g->ocs = b != 0;
  long nc, ne;
  if (!next(s) || !read_long(s, nc, 0) || !read_long(s, ne, 0) || nc < 0 || ne < 0) {
    fail(at_line(path, 1) + ": expected \"n_corners n_edges\"");
    return nullptr;
  }
  if (li + nc + ne > L.size()) {
    fail(std::string(path) + ": file ends before its " + std::to_string(nc) + " corners and " +
         std::to_string(ne) + " edges");
    return nullptr;
  }
  g->corners.resize(nc);
  for (long i = 0; i < nc; ++i) {
    next(s);
    if (!read_long(s, a, 0)) {
      fail(at_line(path, li - 1) + ": expected a corner id");
      return nullptr;
    }
    g->corners[i] = (int32_t)(uint32_t)a;
  }
  g->pairs.resize(2 * ne);
  for (long i = 0; i < ne; ++i) {
    next(s);
    if (!read_long(s, a, 0) || !read_long(s, b, 0)) {
      fail(at_line(path, li - 1) + ": expected an edge \"v0 v1\"");
      return nullptr;
    }
    g->pairs[2 * i] = (int32_t)(uint32_t)a;
    g->pairs[2 * i + 1] = (int32_t)(uint32_t)b;
  }
  return g.release();
}

int write_fgraph(const char* path, double angle, int oc, int ocs, int nc, const int32_t* corners, int ne,
                 const int32_t* pairs) {
  if (nc < 0 || ne < 0) return fail("negative size"), 1;
  if ((nc > 0 && !corners) || (ne > 0 && !pairs)) return fail("corner or edge ids missing"), 1;
  Out o(path);
  if (!check_open(o, path)) return 1;
  std::string s;
  put_g(s, angle);
  s += oc ? " 1 " : " 0 ";
  s += ocs ? "1\n" : "0\n";
  put_i(s, nc);
  s += ' ';
  put_i(s, ne);
  s += '\n';
  for (int i = 0; i < nc; ++i) {
    put_i(s, (uint32_t)corners[i]);
    s += '\n';
  }
  for (int i = 0; i < ne; ++i) {
    put_i(s, (uint32_t)pairs[2 * i]);
    s += ' ';
    put_i(s, (uint32_t)pairs[2 * i + 1]);
    s += '\n';
  }
  o.put(s);
  return finish(o, path) ? 0 : 1;
}

template <class F>
auto guarded(F&& f, decltype(f()) on_error) -> decltype(f()) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception& e) {
    g_err = std::string("hexio: ") + e.what();
  } catch (...) {
    g_err = "hexio: unknown error";
  }
  return on_error;
}

}  // namespace

extern "C" {

const char* hexio_error(void) { return g_err.c_str(); }

hexio_mesh* hexio_read_vtk(const char* path) {
  return guarded([&] { return read_vtk(path); }, (hexio_mesh*)nullptr);
}
hexio_mesh* hexio_read_mesh(const char* path) {
  return guarded([&] { return read_mesh(path); }, (hexio_mesh*)nullptr);
}
hexio_mesh* hexio_read_obj(const char* path) {
  return guarded([&] { return read_obj(path); }, (hexio_mesh*)nullptr);
}

int hexio_mesh_sizes(const hexio_mesh* m, int64_t* out4) {
  if (!m || !out4) return fail("null argument"), 1;
  out4[0] = (int64_t)(m->V.size() / 3);
  out4[1] = (int64_t)m->off.size() - 1;
  out4[2] = (int64_t)m->idx.size();
  out4[3] = m->type;
  return 0;
}

int hexio_mesh_copy(const hexio_mesh* m, double* V, int32_t* off, int32_t* idx) {
  if (!m) return fail("null mesh"), 1;
  if (V) std::copy(m->V.begin(), m->V.end(), V);
  if (off) std::copy(m->off.begin(), m->off.end(), off);
  if (idx) std::copy(m->idx.begin(), m->idx.end(), idx);
  return 0;
}

void hexio_mesh_free(hexio_mesh* m) { delete m; }

int hexio_write_vtk(const char* path, int type, int nV, const double* V, int nC, int k, const int32_t* off,
                    const int32_t* idx, const uint8_t* boundary) {
  return guarded([&] { return write_vtk(path, type, nV, V, CellView{nC, k, off, idx}, boundary); }, 1);
}

int hexio_write_mesh(const char* path, int type, int nV, const double* V, int nC, int k, const int32_t* off,
                     const int32_t* idx) {
  return guarded([&] { return write_mesh(path, type, nV, V, CellView{nC, k, off, idx}); }, 1);
}

int hexio_write_obj(const char* path, int nV, const double* V, int nF, int k, const int32_t* off,
                    const int32_t* idx) {
  return guarded([&] { return write_obj(path, nV, 3, V, CellView{nF, k, off, idx}, false); }, 1);
}

int hexio_write_matrix_obj(const char* path, int nV, int dim, const double* V, int nF, int k, const int32_t* F) {
  return guarded([&] { return write_obj(path, nV, dim, V, CellView{nF, k, nullptr, F}, true); }, 1);
}

hexio_fgraph* hexio_read_fgraph(const char* path) {
  return guarded([&] { return read_fgraph(path); }, (hexio_fgraph*)nullptr);
}

int hexio_fgraph_info(const hexio_fgraph* g, hexio_fgraph_header* out) {
  if (!g || !out) return fail("null argument"), 1;
  out->angle_threshold = g->angle;
  out->orphan_curve = g->oc;
  out->orphan_curve_single = g->ocs;
  out->n_corners = (int32_t)g->corners.size();
  out->n_edges = (int32_t)(g->pairs.size() / 2);
  return 0;
}

int hexio_fgraph_copy(const hexio_fgraph* g, int32_t* corners, int32_t* pairs) {
  if (!g) return fail("null graph"), 1;
  if (corners) std::copy(g->corners.begin(), g->corners.end(), corners);
  if (pairs) std::copy(g->pairs.begin(), g->pairs.end(), pairs);
  return 0;
}

void hexio_fgraph_free(hexio_fgraph* g) { delete g; }

int hexio_write_fgraph(const char* path, double angle_threshold, int orphan_curve, int orphan_curve_single,
                       int n_corners, const int32_t* corners, int n_edges, const int32_t* pairs) {
  return guarded([&] {
    return write_fgraph(path, angle_threshold, orphan_curve, orphan_curve_single, n_corners, corners, n_edges,
                        pairs);
  }, 1);
}

}  // extern "C"
