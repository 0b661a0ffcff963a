// layout.h -- device data layout of a hex mesh for the tiled, atomic-free scatter kernels.
//
// Vertices and hexes are renumbered along a Morton (Z-order) curve so that the 8-corner
// gathers of neighbouring hexes hit the same cache lines. Hexes are then cut into tiles of
// at most kTileHex consecutive hexes; one workgroup owns one tile. Inside a tile the hexes
// are sorted by a greedy colour (no two hexes of one colour share a vertex), so a tile can
// accumulate its per-corner contributions into LDS colour by colour without atomics.
// A vertex whose incident hexes all lie in one tile is written directly by that tile; a
// vertex touched by several tiles gets one halo slot per touching tile and a fixup pass sums
// the slots in a fixed order. Every result is therefore bitwise reproducible.
#pragma once
#include <cstdint>
#include <vector>

namespace hexopt {

constexpr int kTileHex = 256;   // hexes per tile = threads per workgroup
constexpr int kMaxLV = 1024;    // local (LDS) vertices per tile
constexpr int kMaxCol = 16;     // colours per tile

struct MeshLayout {
  int nV = 0, nH = 0, nTiles = 0, nShared = 0, nSlots = 0;
  std::vector<int> vperm;       // new vertex id -> caller's vertex id
  std::vector<int> vinv;        // caller's vertex id -> new id
  std::vector<int> hperm;       // new hex id -> caller's hex id
  std::vector<int> conn;        // [8][nH] new vertex ids (corner-major, hex fastest)
  std::vector<int> tile_hex;    // nTiles+1 hex ranges
  std::vector<uint16_t> tile_col;  // nTiles*(kMaxCol+1) colour offsets relative to tile start
  std::vector<int> tile_lv;     // nTiles+1 local-vertex ranges
  std::vector<int> lv_gid;      // local vertex -> new vertex id
  std::vector<int> lv_slot;     // -1: written directly; else halo slot
  std::vector<uint8_t> lv_first;  // 1: the vertex is this tile's alone, or this tile holds its
                                  //    first halo slot (the one tile that writes per-vertex data)
  std::vector<uint16_t> lidx;   // [8][nH] local vertex of each corner within its tile
  std::vector<int> lv_dst;      // per local vertex: its place among all (tile, vertex) pairs,
                                // numbered vertex by vertex (each vertex's in tile order)
  std::vector<int> v_dst;       // nV+1: vertex v's run of places in that numbering
  std::vector<uint8_t> tile_corner;  // 1: no vertex of the tile sits at the same corner index of
                                     // two of its hexes, so corner k of every hex can be added in
                                     // phase k of the tile's scatter
  std::vector<int> sv_gid;      // vertices finished by the fixup pass (shared or isolated)
  std::vector<int> sv_off;      // nShared+1 ranges into sv_slots
  std::vector<int> sv_slots;    // halo slots, ascending tile order; slots are numbered vertex by
                                // vertex, so sv_slots[k] == k
  int max_colours = 0;
};

// Builds the layout from caller-numbered hexes (nH*8, reference corner order) and vertex
// positions (nV*3 row-major, used only for the Morton keys). Returns false on invalid input.
bool build_layout(int nV, const double* V, int nH, const int* H, MeshLayout& L, const char** err);

}  // namespace hexopt
