// hexclean.cpp -- the flag passes of clean_hex_mesh (declared in hexclean.h). Host only.
//
// Each pass is a serial sweep with in-place updates, so its result depends on the order in
// which hexes, vertices, faces and edges are visited. The connectivity below numbers faces
// and edges the way build_connectivity does for a hex mesh (global_functions.cpp:121-183):
// faces by their sorted vertex-id quadruple (the first occurrence's corner order is kept),
// edges by their sorted vertex-id pair, and every neighbour list in ascending id order
// (:225-262). The sub-mesh of flagged hexes keeps hex and vertex order (re_indexing_
// connectivity, global_functions.cpp:664-698), so it is built straight from the flagged
// hexes with the original vertex ids: the compaction is monotone and leaves every sort
// order unchanged.
#include "../../include/hexclean.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

// hex_face_table (global_types.h:154-162): the 6 faces of a hex in local corner ids.
const int kHexFace[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {0, 3, 7, 4}, {3, 2, 6, 7}, {1, 5, 6, 2}};
// hex_tetra_table (global_types.h:164-174): the corner tetrahedron at each corner.
const int kHexTet[8][4] = {{0, 3, 4, 1}, {1, 0, 5, 2}, {2, 1, 6, 3}, {3, 2, 7, 0},
                           {4, 7, 5, 0}, {5, 4, 6, 1}, {6, 5, 7, 2}, {7, 6, 4, 3}};

struct Fail {};

void fail(const std::string& m) {
  g_err = m;
  throw Fail();
}

// Connectivity of a hex mesh: hex -> faces, face -> hexes, face -> edges, edge -> faces,
// vertex -> faces, and the boundary flags of faces, edges and vertices.
struct Conn {
  int64_t nH = 0, nF = 0, nE = 0;
  std::vector<int32_t> hv;                // nH x 8 vertex ids
  std::vector<int32_t> hfs;               // nH x 6 face ids, local face order
  std::vector<int32_t> fh;                // nF x 2 neighbouring hexes, ascending, -1 when absent
  std::vector<uint8_t> fbnd;              // face has one hex
  std::vector<int32_t> fvs;               // nF x 4 corners of the first occurrence
  std::vector<int32_t> fes;               // nF x 4 edge ids, edge j from corner j to j+1
  std::vector<uint8_t> ebnd, vbnd;        // edge / vertex on a boundary face
  std::vector<int64_t> ef_off, vf_off;    // CSR offsets
  std::vector<int32_t> ef, vf;            // edge -> faces, vertex -> faces, ascending face id
};

// Sort keys packed into integers: a face's sorted vertex quadruple in 128 bits, an edge's
// sorted pair in 64, each followed by its slot, so ties keep the slot order the reference's
// tuple sort gives them.
struct FaceKey {
  unsigned __int128 k;
  int64_t id;  // 6 * hex + local face
  bool operator<(const FaceKey& o) const { return k != o.k ? k < o.k : id < o.id; }
};

struct EdgeKey {
  uint64_t k;
  int64_t slot;  // 4 * face + local edge
  bool operator<(const EdgeKey& o) const { return k != o.k ? k < o.k : slot < o.slot; }
};

// std::sort on up to 16 threads: sorted chunks, then rounds of pairwise merges.
template <class T>
void psort(std::vector<T>& a) {
  const size_t n = a.size();
  const size_t parts = std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (n < (1u << 16) || parts == 1) {
    std::sort(a.begin(), a.end());
    return;
  }
  std::vector<size_t> b(parts + 1);
  for (size_t i = 0; i <= parts; ++i) b[i] = n * i / parts;
  std::vector<std::thread> th;
  for (size_t i = 0; i < parts; ++i) th.emplace_back([&, i] { std::sort(a.begin() + b[i], a.begin() + b[i + 1]); });
  for (auto& t : th) t.join();
  std::vector<T> tmp(n);
  for (size_t w = 1; w < parts; w *= 2) {
    th.clear();
    for (size_t i = 0; i < parts; i += 2 * w) {
      const size_t lo = b[i], mid = b[std::min(i + w, parts)], hi = b[std::min(i + 2 * w, parts)];
      th.emplace_back([&, lo, mid, hi] {
        std::merge(a.begin() + lo, a.begin() + mid, a.begin() + mid, a.begin() + hi, tmp.begin() + lo);
      });
    }
    for (auto& t : th) t.join();
    a.swap(tmp);
  }
}

void check_hexes(int64_t nV, int64_t nH, const int32_t* H) {
  if (nV < 0 || nH < 0) fail("negative size");
  if (nH > 0 && !H) fail("H is null");
  if (nV > INT32_MAX || nH > INT32_MAX / 6) fail("mesh too large for 32-bit face ids");
  for (int64_t h = 0; h < nH; ++h) {
    const int32_t* c = H + 8 * h;
    for (int k = 0; k < 8; ++k) {
      if (c[k] < 0 || c[k] >= nV) fail("hex " + std::to_string(h) + ": vertex id out of range");
      for (int l = 0; l < k; ++l)
        if (c[l] == c[k]) fail("hex " + std::to_string(h) + ": repeated vertex");
    }
  }
}

// build_connectivity for the hexes listed in hexes (ids into H). full = false stops after
// the hex <-> face relation (all tagging_uneven_element and drop_small_pieces use).
void build(Conn& c, int64_t nV, const int32_t* H, const std::vector<int32_t>& hexes, bool full) {
  c.nH = (int64_t)hexes.size();
  c.hv.resize(8 * c.nH);
  for (int64_t i = 0; i < c.nH; ++i) std::memcpy(&c.hv[8 * i], H + 8 * (int64_t)hexes[i], 8 * sizeof(int32_t));
  std::vector<FaceKey> t(6 * c.nH);
  for (int64_t h = 0; h < c.nH; ++h)
    for (int j = 0; j < 6; ++j) {
      uint32_t v[4];
      for (int q = 0; q < 4; ++q) v[q] = (uint32_t)c.hv[8 * h + kHexFace[j][q]];
      std::sort(v, v + 4);
      unsigned __int128 k = 0;
      for (int q = 0; q < 4; ++q) k = (k << 32) | v[q];
      t[6 * h + j] = {k, 6 * h + j};
    }
  psort(t);
  c.hfs.assign(6 * c.nH, -1);
  c.fbnd.clear();
  c.fvs.clear();
  int64_t F = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    if (i == 0 || t[i].k != t[i - 1].k) {
      ++F;
      c.fbnd.push_back(1);
      const int64_t h = t[i].id / 6;
      const int j = (int)(t[i].id % 6);
      for (int q = 0; q < 4; ++q) c.fvs.push_back(c.hv[8 * h + kHexFace[j][q]]);
    } else {
      c.fbnd[F - 1] = 0;
    }
    c.hfs[t[i].id] = (int32_t)(F - 1);
  }
  c.nF = F;
  std::vector<FaceKey>().swap(t);
  // face -> hexes, hexes ascending (global_functions.cpp:226-229)
  c.fh.assign(2 * F, -1);
  for (int64_t h = 0; h < c.nH; ++h)
    for (int j = 0; j < 6; ++j) {
      int32_t* s = &c.fh[2 * c.hfs[6 * h + j]];
      if (s[0] < 0) s[0] = (int32_t)h;
      else if (s[1] < 0) s[1] = (int32_t)h;
      else fail("a face is shared by more than two hexes");
    }
  if (!full) return;
  // edges (global_functions.cpp:155-176)
  std::vector<EdgeKey> e(4 * F);
  for (int64_t f = 0; f < F; ++f)
    for (int j = 0; j < 4; ++j) {
      int32_t a = c.fvs[4 * f + j], b = c.fvs[4 * f + (j + 1) % 4];
      if (a > b) std::swap(a, b);
      e[4 * f + j] = {((uint64_t)(uint32_t)a << 32) | (uint32_t)b, 4 * f + j};
    }
  psort(e);
  c.fes.assign(4 * F, -1);
  int64_t E = 0;
  for (size_t i = 0; i < e.size(); ++i) {
    if (i == 0 || e[i].k != e[i - 1].k) ++E;
    c.fes[e[i].slot] = (int32_t)(E - 1);
  }
  c.nE = E;
  std::vector<EdgeKey>().swap(e);
  // boundary edges and vertices: those of boundary faces (:177-182)
  c.ebnd.assign(E, 0);
  c.vbnd.assign(nV, 0);
  for (int64_t f = 0; f < F; ++f)
    if (c.fbnd[f])
      for (int j = 0; j < 4; ++j) {
        c.ebnd[c.fes[4 * f + j]] = 1;
        c.vbnd[c.fvs[4 * f + j]] = 1;
      }
  // edge -> faces and vertex -> faces, faces ascending (:231-235)
  c.ef_off.assign(E + 1, 0);
  c.vf_off.assign(nV + 1, 0);
  for (int64_t s = 0; s < 4 * F; ++s) {
    ++c.ef_off[c.fes[s] + 1];
    ++c.vf_off[c.fvs[s] + 1];
  }
  for (int64_t i = 0; i < E; ++i) c.ef_off[i + 1] += c.ef_off[i];
  for (int64_t i = 0; i < nV; ++i) c.vf_off[i + 1] += c.vf_off[i];
  c.ef.resize(4 * F);
  c.vf.resize(4 * F);
  std::vector<int64_t> pe(c.ef_off.begin(), c.ef_off.end() - 1), pv(c.vf_off.begin(), c.vf_off.end() - 1);
  for (int64_t f = 0; f < F; ++f)
    for (int j = 0; j < 4; ++j) {
      c.ef[pe[c.fes[4 * f + j]]++] = (int32_t)f;
      c.vf[pv[c.fvs[4 * f + j]]++] = (int32_t)f;
    }
}

std::vector<int32_t> flagged(int64_t nH, const uint8_t* H_flag) {
  std::vector<int32_t> ids;
  for (int64_t h = 0; h < nH; ++h)
    if (H_flag[h]) ids.push_back((int32_t)h);
  return ids;
}

int64_t count(int64_t nH, const uint8_t* H_flag) {
  int64_t n = 0;
  for (int64_t h = 0; h < nH; ++h) n += H_flag[h] != 0;
  return n;
}

// Eigen's 3x3 determinant (cofactors along the first column) of the columns b-a, c-a, d-a.
double det_cols(const double* a, const double* b, const double* c, const double* d) {
  double m[3][3];
  for (int r = 0; r < 3; ++r) {
    m[r][0] = b[r] - a[r];
    m[r][1] = c[r] - a[r];
    m[r][2] = d[r] - a[r];
  }
  auto h = [&](int x, int y, int z) { return m[0][x] * (m[1][y] * m[2][z] - m[1][z] * m[2][y]); };
  return h(0, 1, 2) - h(1, 0, 2) + h(2, 0, 1);
}

void tagging(int64_t nV, int64_t nH, const int32_t* H, uint8_t* flag) {
  std::vector<int32_t> all(nH);
  for (int64_t h = 0; h < nH; ++h) all[h] = (int32_t)h;
  Conn c;
  build(c, nV, H, all, false);
  for (int64_t i = 0; i < nH; ++i) {
    int bn = 0, in = 0;
    const bool me = flag[i] != 0;
    for (int j = 0; j < 6; ++j) {
      const int32_t* s = &c.fh[2 * c.hfs[6 * i + j]];
      if (s[1] < 0) {
        ++bn;
      } else {
        const bool a = flag[s[0]] != 0, b = flag[s[1]] != 0;
        if (me ? (!a || !b) : (a || b)) ++bn;
        else ++in;
      }
    }
    if (bn == 5 && in == 1) flag[i] = !me;
  }
}

int index_in(const std::vector<int32_t>& v, int32_t x) {
  return (int)(std::find(v.begin(), v.end(), x) - v.begin());
}

void non_manifold(int64_t nV, int64_t nH, const int32_t* H, uint8_t* flag) {
  Conn c;
  std::vector<int32_t> hmap = flagged(nH, flag);  // H_map_reverse: local hex -> hex
  build(c, nV, H, hmap, true);
  std::vector<uint8_t> v_tag, e_tag, f_tag;
  std::vector<int32_t> bfs, pool, next;
  while (true) {
    bool changed = false;
    v_tag.assign(nV, 0);
    for (int64_t v = 0; v < nV; ++v) {  // vertices of the sub-mesh in order
      if (!c.vbnd[v] || v_tag[v]) continue;
      v_tag[v] = 1;
      bfs.clear();
      for (int64_t k = c.vf_off[v]; k < c.vf_off[v + 1]; ++k)
        if (c.fbnd[c.vf[k]]) bfs.push_back(c.vf[k]);
      f_tag.assign(bfs.size(), 0);
      // fan of bfs[0]: boundary faces at v reached over shared edges
      pool.assign(1, bfs[0]);
      bool non_manifold = false;
      while (true) {
        next.clear();
        for (int32_t f : pool) {
          const int i = index_in(bfs, f);
          if (f_tag[i]) continue;
          f_tag[i] = 1;
          for (int j = 0; j < 4; ++j) {
            const int32_t eid = c.fes[4 * (int64_t)f + j];
            for (int64_t k = c.ef_off[eid]; k < c.ef_off[eid + 1]; ++k) {
              const int l = index_in(bfs, c.ef[k]);
              if (l < (int)bfs.size() && !f_tag[l]) next.push_back(c.ef[k]);
            }
          }
        }
        if (!next.empty()) {
          pool.swap(next);
        } else {
          non_manifold = std::find(f_tag.begin(), f_tag.end(), 0) != f_tag.end();
          break;
        }
      }
      if (non_manifold) {
        for (size_t i = 0; i < bfs.size(); ++i)
          if (!f_tag[i]) {
            const int32_t nh = c.fh[2 * (int64_t)bfs[i]];
            flag[hmap[nh]] = 0;
            for (int q = 0; q < 8; ++q) v_tag[c.hv[8 * (int64_t)nh + q]] = 1;
          }
        changed = true;
      }
    }
    if (!changed) {
      e_tag.assign(c.nE, 0);
      for (int64_t e = 0; e < c.nE; ++e) {
        if (!c.ebnd[e] || e_tag[e]) continue;
        e_tag[e] = 1;
        bfs.clear();
        for (int64_t k = c.ef_off[e]; k < c.ef_off[e + 1]; ++k)
          if (c.fbnd[c.ef[k]]) bfs.push_back(c.ef[k]);
        if (bfs.size() != 2) {
          changed = true;
          const int32_t nh = c.fh[2 * (int64_t)bfs[0]];
          flag[hmap[nh]] = 0;
          for (int j = 0; j < 6; ++j) {
            const int64_t f = c.hfs[6 * (int64_t)nh + j];
            for (int q = 0; q < 4; ++q) e_tag[c.fes[4 * f + q]] = 1;
          }
        }
      }
      if (!changed) break;
    }
    hmap = flagged(nH, flag);
    c = Conn();
    build(c, nV, H, hmap, true);
  }
}

void drop_small(int64_t nV, int64_t nH, const int32_t* H, uint8_t* flag) {
  std::vector<int32_t> hmap = flagged(nH, flag);
  if (hmap.empty()) return;
  Conn c;
  build(c, nV, H, hmap, false);
  const int64_t n = c.nH;
  std::vector<int32_t> piece(n, -1), list;
  std::vector<int64_t> sizes;
  int64_t first = 0;
  while (true) {
    const int32_t g = (int32_t)sizes.size();
    list.assign(1, (int32_t)first);
    for (size_t i = 0; i < list.size(); ++i) {  // the list grows while it is walked
      const int32_t h = list[i];
      if (piece[h] >= 0) continue;
      piece[h] = g;
      for (int j = 0; j < 6; ++j) {
        const int32_t* s = &c.fh[2 * (int64_t)c.hfs[6 * (int64_t)h + j]];
        for (int q = 0; q < 2; ++q)
          if (s[q] >= 0 && s[q] != h && piece[s[q]] < 0) list.push_back(s[q]);
      }
    }
    sizes.push_back((int64_t)list.size());
    while (first < n && piece[first] >= 0) ++first;
    if (first == n) break;
  }
  if (sizes.size() < 2) return;
  int32_t which = 0;
  for (size_t g = 1; g < sizes.size(); ++g)
    if (sizes[g] > sizes[which]) which = (int32_t)g;
  std::memset(flag, 0, nH);
  for (int64_t i = 0; i < n; ++i)
    if (piece[i] == which) flag[hmap[i]] = 1;
}

// Sets the error message without letting an allocation failure escape (the message is then
// left empty).
void set_err(const char* msg) noexcept {
  try {
    g_err = msg;
  } catch (...) {
    g_err.clear();
  }
}

// Runs body and turns every exception into a -1 return with the message in g_err, so that
// nothing is thrown across the C ABI.
template <class F>
int64_t catching(F&& body) noexcept {
  try {
    return body();
  } catch (const Fail&) {
    return -1;
  } catch (const std::bad_alloc&) {
    set_err("out of host memory");
    return -1;
  } catch (const std::exception& e) {  // e.g. a sort thread that could not start
    set_err(e.what());
    return -1;
  } catch (...) {
    set_err("unknown exception");
    return -1;
  }
}

template <class F>
int64_t guarded(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag, F&& body) {
  return catching([&]() -> int64_t {
    check_hexes(nV, nH, H);
    if (nH > 0 && !H_flag) fail("H_flag is null");
    body();
    return count(nH, H_flag);
  });
}

}  // namespace

extern "C" {

const char* hexclean_error(void) { return g_err.c_str(); }

int64_t hexclean_reorder_hex_mesh(int64_t nV, const double* V, int64_t nH, int32_t* H, int64_t* flipped) {
  return catching([&]() -> int64_t {
    check_hexes(nV, nH, H);
    if (nH > 0 && !V) fail("V is null");
    int64_t n = 0;
    for (int64_t h = 0; h < nH; ++h) {
      int32_t* c = H + 8 * h;
      double vol = 0;
      for (int j = 0; j < 8; ++j)
        vol += det_cols(V + 3 * (int64_t)c[kHexTet[j][0]], V + 3 * (int64_t)c[kHexTet[j][1]],
                        V + 3 * (int64_t)c[kHexTet[j][2]], V + 3 * (int64_t)c[kHexTet[j][3]]);
      if (vol < 0) {
        std::reverse(c, c + 4);
        std::reverse(c + 4, c + 8);
        ++n;
      }
    }
    if (flipped) *flipped = n;
    return nH;
  });
}

int64_t hexclean_tagging_uneven_element(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag) {
  return guarded(nV, nH, H, H_flag, [&] { tagging(nV, nH, H, H_flag); });
}

int64_t hexclean_clean_non_manifold_ve(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag) {
  return guarded(nV, nH, H, H_flag, [&] { non_manifold(nV, nH, H, H_flag); });
}

int64_t hexclean_drop_small_pieces(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag) {
  return guarded(nV, nH, H, H_flag, [&] { drop_small(nV, nH, H, H_flag); });
}

int64_t hexclean_medial_surface(int64_t nV, int64_t nH, const int32_t* H, const uint8_t* H_flag, uint8_t* v_flag,
                                int32_t* quads) {
  return catching([&]() -> int64_t {
    check_hexes(nV, nH, H);
    if (nH > 0 && !H_flag) fail("H_flag is null");
    // the faces of the whole mesh in build_connectivity's numbering, then the loop that ends
    // clean_hex_mesh (grid_hex_meshing.cpp:1966-1981) over them in that order
    Conn c;
    std::vector<int32_t> all(nH);
    for (int64_t h = 0; h < nH; ++h) all[h] = (int32_t)h;
    build(c, nV, H, all, false);
    if (v_flag) std::memset(v_flag, 0, (size_t)nV);
    int64_t n = 0;
    for (int64_t f = 0; f < c.nF; ++f) {
      const int32_t a = c.fh[2 * f], b = c.fh[2 * f + 1];
      const bool on = c.fbnd[f] ? H_flag[a] != 0 : (H_flag[a] != 0) != (H_flag[b] != 0);
      if (!on) continue;
      for (int q = 0; q < 4; ++q) {
        if (v_flag) v_flag[c.fvs[4 * f + q]] = 1;
        if (quads) quads[4 * n + q] = c.fvs[4 * f + q];
      }
      ++n;
    }
    return n;
  });
}

int64_t hexclean_clean_flags(int64_t nV, int64_t nH, const int32_t* H, uint8_t* H_flag) {
  return guarded(nV, nH, H, H_flag, [&] {
    tagging(nV, nH, H, H_flag);
    if (count(nH, H_flag) == 0) return;  // "no elements inside the object"
    non_manifold(nV, nH, H, H_flag);
    drop_small(nV, nH, H, H_flag);
  });
}

}  // extern "C"
