// layout.cpp -- host construction of the tiled hex layout (see layout.h).
#include "layout.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

namespace hexopt {
namespace {

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit of 63
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

struct Box {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
};

uint64_t morton(const double p[3], const Box& b) {
  uint64_t k = 0;
  for (int d = 0; d < 3; ++d) {
    double ext = b.hi[d] - b.lo[d];
    double t = ext > 0 ? (p[d] - b.lo[d]) / ext : 0.0;
    t = std::min(std::max(t, 0.0), 1.0);
    uint64_t q = (uint64_t)(t * 2097151.0);
    k |= spread3(q) << d;
  }
  return k;
}

}  // namespace

bool build_layout(int nV, const double* V, int nH, const int* H, MeshLayout& L, const char** err) {
  L = MeshLayout();
  L.nV = nV;
  L.nH = nH;
  if (nV <= 0 || nH < 0) { *err = "empty mesh"; return false; }
  for (long i = 0; i < 8L * nH; ++i)
    if (H[i] < 0 || H[i] >= nV) { *err = "hex corner index out of range"; return false; }
  Box b;
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = std::min(b.lo[d], V[(size_t)i * 3 + d]);
      b.hi[d] = std::max(b.hi[d], V[(size_t)i * 3 + d]);
    }
  // vertices along the Morton curve
  std::vector<uint64_t> vk(nV);
  for (int i = 0; i < nV; ++i) vk[i] = morton(&V[(size_t)i * 3], b);
  L.vperm.resize(nV);
  std::iota(L.vperm.begin(), L.vperm.end(), 0);
  std::stable_sort(L.vperm.begin(), L.vperm.end(), [&](int a, int c) { return vk[a] < vk[c]; });
  L.vinv.resize(nV);
  for (int i = 0; i < nV; ++i) L.vinv[L.vperm[i]] = i;
  // hexes along the Morton curve of their centroids
  std::vector<uint64_t> hk(nH);
  for (int h = 0; h < nH; ++h) {
    double c[3] = {0, 0, 0};
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) c[d] += V[(size_t)H[(size_t)h * 8 + k] * 3 + d];
    for (int d = 0; d < 3; ++d) c[d] *= 0.125;
    hk[h] = morton(c, b);
  }
  std::vector<int> horder(nH);
  std::iota(horder.begin(), horder.end(), 0);
  std::stable_sort(horder.begin(), horder.end(), [&](int a, int c) { return hk[a] < hk[c]; });

  // greedy tiles with in-tile greedy colouring
  std::vector<int> hexcol(nH, 0);
  std::vector<int> tile_of_pos;  // per position in horder: tile
  L.tile_hex.push_back(0);
  {
    std::unordered_map<int, uint32_t> vmask;  // vertex -> colours used in current tile
    vmask.reserve(2 * kMaxLV);
    int count = 0;
    for (int pos = 0; pos < nH; ++pos) {
      int h = horder[pos];
      const int* hv = &H[(size_t)h * 8];
      int newv = 0;
      uint32_t used = 0;
      for (int k = 0; k < 8; ++k) {
        auto it = vmask.find(hv[k]);
        if (it == vmask.end()) {
          bool dup = false;
          for (int j = 0; j < k; ++j) dup |= (hv[j] == hv[k]);
          if (!dup) ++newv;
        } else {
          used |= it->second;
        }
      }
      int col = used == 0xffffffffu ? 32 : __builtin_ctz(~used);
      bool close = count == kTileHex || (int)vmask.size() + newv > kMaxLV || col >= kMaxCol;
      if (close && count > 0) {
        L.tile_hex.push_back(pos);
        vmask.clear();
        count = 0;
        used = 0;
        col = 0;
      }
      for (int k = 0; k < 8; ++k) vmask[hv[k]] |= (1u << col);
      hexcol[h] = col;
      ++count;
    }
    if (count > 0) L.tile_hex.push_back(nH);
  }
  L.nTiles = (int)L.tile_hex.size() - 1;
  // sort hexes by colour inside each tile; record colour offsets
  L.hperm.resize(nH);
  L.tile_col.assign((size_t)L.nTiles * (kMaxCol + 1), 0);
  for (int t = 0; t < L.nTiles; ++t) {
    int s = L.tile_hex[t], e = L.tile_hex[t + 1];
    std::vector<int> hs(horder.begin() + s, horder.begin() + e);
    std::stable_sort(hs.begin(), hs.end(), [&](int a, int c) { return hexcol[a] < hexcol[c]; });
    for (int i = s; i < e; ++i) L.hperm[i] = hs[i - s];
    int ncol = 0;
    for (int h : hs) ncol = std::max(ncol, hexcol[h] + 1);
    L.max_colours = std::max(L.max_colours, ncol);
    uint16_t* off = &L.tile_col[(size_t)t * (kMaxCol + 1)];
    for (int c = 0; c <= kMaxCol; ++c) {
      int n = 0;
      for (int h : hs) n += hexcol[h] < c;
      off[c] = (uint16_t)n;
    }
  }
  // vertices renumbered tile by tile: each takes the next id in the first tile that holds it
  // (in Morton order within the tile), and vertices in no hex come last (Morton order). A
  // tile's vertices are then one contiguous run of ids plus the halo vertices of earlier
  // tiles, so its vertex loads and stores coalesce.
  {
    std::vector<int> nid(nV, -1);
    int next = 0;
    std::vector<int> fresh;
    for (int t = 0; t < L.nTiles; ++t) {
      fresh.clear();
      for (int i = L.tile_hex[t]; i < L.tile_hex[t + 1]; ++i)
        for (int k = 0; k < 8; ++k) {
          const int v = H[(size_t)L.hperm[i] * 8 + k];
          if (nid[v] == -1) { nid[v] = -2; fresh.push_back(v); }
        }
      std::sort(fresh.begin(), fresh.end(), [&](int a, int c) { return L.vinv[a] < L.vinv[c]; });
      for (int v : fresh) nid[v] = next++;
    }
    for (int i = 0; i < nV; ++i)
      if (nid[L.vperm[i]] == -1) nid[L.vperm[i]] = next++;
    for (int v = 0; v < nV; ++v) {
      L.vinv[v] = nid[v];
      L.vperm[nid[v]] = v;
    }
  }
  // connectivity in new numbering, corner-major
  L.conn.resize((size_t)8 * nH);
  for (int i = 0; i < nH; ++i)
    for (int k = 0; k < 8; ++k) L.conn[(size_t)k * nH + i] = L.vinv[H[(size_t)L.hperm[i] * 8 + k]];
  // local vertices per tile and ownership
  std::vector<int> ntiles_of_v(nV, 0), last_tile(nV, -1);
  for (int t = 0; t < L.nTiles; ++t)
    for (int i = L.tile_hex[t]; i < L.tile_hex[t + 1]; ++i)
      for (int k = 0; k < 8; ++k) {
        int v = L.conn[(size_t)k * nH + i];
        if (last_tile[v] != t) { last_tile[v] = t; ++ntiles_of_v[v]; }
      }
  L.lidx.resize((size_t)8 * nH);
  L.tile_lv.push_back(0);
  std::vector<int> slot_count(nV, 0);
  std::vector<int> vslot_start(nV + 1, 0);
  for (int v = 0; v < nV; ++v) vslot_start[v + 1] = vslot_start[v] + (ntiles_of_v[v] > 1 ? ntiles_of_v[v] : 0);
  L.v_dst.assign(nV + 1, 0);
  for (int v = 0; v < nV; ++v) L.v_dst[v + 1] = L.v_dst[v] + ntiles_of_v[v];
  std::vector<int> dst_count(nV, 0);
  L.nSlots = vslot_start[nV];
  L.sv_slots.assign(L.nSlots, -1);
  std::vector<int> local(nV, -1);
  for (int t = 0; t < L.nTiles; ++t) {
    // the tile's vertices in ascending id order
    std::vector<int> lv;
    for (int i = L.tile_hex[t]; i < L.tile_hex[t + 1]; ++i)
      for (int k = 0; k < 8; ++k) {
        int v = L.conn[(size_t)k * nH + i];
        if (local[v] < 0) { local[v] = 0; lv.push_back(v); }
      }
    if ((int)lv.size() > kMaxLV) { *err = "tile local vertex overflow"; return false; }
    std::sort(lv.begin(), lv.end());
    for (size_t j = 0; j < lv.size(); ++j) local[lv[j]] = (int)j;
    for (int i = L.tile_hex[t]; i < L.tile_hex[t + 1]; ++i)
      for (int k = 0; k < 8; ++k) L.lidx[(size_t)k * nH + i] = (uint16_t)local[L.conn[(size_t)k * nH + i]];
    {
      std::vector<uint8_t> seen(lv.size(), 0);
      uint8_t ok = 1;
      for (int i = L.tile_hex[t]; i < L.tile_hex[t + 1]; ++i)
        for (int k = 0; k < 8; ++k) {
          uint8_t& m = seen[L.lidx[(size_t)k * nH + i]];
          if (m & (1u << k)) ok = 0;
          m |= (uint8_t)(1u << k);
        }
      L.tile_corner.push_back(ok);
    }
    for (int v : lv) {
      L.lv_gid.push_back(v);
      L.lv_dst.push_back(L.v_dst[v] + dst_count[v]++);
      if (ntiles_of_v[v] > 1) {
        // slots are numbered vertex by vertex (each vertex's in tile order), so the fold of a
        // vertex reads one contiguous run
        const int s = vslot_start[v] + slot_count[v];
        L.lv_slot.push_back(s);
        L.lv_first.push_back(slot_count[v] == 0 ? 1 : 0);
        L.sv_slots[vslot_start[v] + slot_count[v]++] = s;
      } else {
        L.lv_slot.push_back(-1);
        L.lv_first.push_back(1);
      }
      local[v] = -1;
    }
    L.tile_lv.push_back((int)L.lv_gid.size());
  }
  // fixup list: shared vertices (their slots) and isolated vertices (no slot)
  L.sv_off.push_back(0);
  for (int v = 0; v < nV; ++v) {
    if (ntiles_of_v[v] > 1 || ntiles_of_v[v] == 0) {
      L.sv_gid.push_back(v);
      L.sv_off.push_back(vslot_start[v + 1]);
    }
  }
  // sv_off must index sv_slots per listed vertex; rebuild compactly
  {
    std::vector<int> off(1, 0), slots;
    for (int v : L.sv_gid) {
      for (int j = vslot_start[v]; j < vslot_start[v + 1]; ++j) slots.push_back(L.sv_slots[j]);
      off.push_back((int)slots.size());
    }
    L.sv_off.swap(off);
    L.sv_slots.swap(slots);
  }
  L.nShared = (int)L.sv_gid.size();
  return true;
}

}  // namespace hexopt
