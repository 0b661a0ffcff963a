// octree_ref_driver.cpp -- golden-vector driver for the reference's octree subdivision
// (grid_meshing/octree.cpp: OctreeGrid, splitCell with its grading and pairing cascades).
// Test infrastructure: links octree.cpp compiled where it lies; never shipped, never run on a
// GPU machine.
//
//   octree_ref <in.bin> <out.bin>
//
// in.bin (little-endian): int32 gx, gy, gz, stop_extent, graded, paired, nV, nF; double
// origin[3], voxel_size, mesh_transform[3]; double V[nV][3]; int32 F[nF][3]; optionally int32
// stop2, nList and int32 list[nList][4]: after the subdivision, the targeted subdivide of
// octree.cpp:691-729 over the leaves at those (x, y, z, extent), with the predicate at stop2.
// out.bin: int32 nLeaves, nNodes, is_graded, is_paired; int32 leaves[nLeaves][4] (x, y, z,
// extent of each leaf, in cell order); int32 nodes[nNodes][3] (node positions, in node order);
// int32 corners[nLeaves][8] (node ids of each leaf's corners, the hexes of octree_mesh).
//
// OctreeGrid::subdivide (octree.cpp:648-690) also writes through geogram's logger, which needs
// the geogram library; its queue loop is therefore restated below around the reference's own
// splitCell (octree.cpp:501-593). The predicate restates octree_mesh's should_subdivide
// (grid_meshing/grid_hex_meshing.cpp:502-517) with the facet-box test of
// GEO::MeshFacetsAABB::compute_bbox_facet_bbox_intersections (closed intervals, geogram
// basic/geometry.h bboxes_overlap) over the facets binned in a uniform grid: a facet box and a
// cell box that touch share the bin of the touching coordinate, since both map it through the
// same expression, so the answer is the every-facet answer.
// The subdivision's wall time (queue loop and splits, not the binning) goes to stdout.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <queue>
#include <stdexcept>
#include <vector>

#include "grid_meshing/octree.h"

namespace {

// splitCell is private in OctreeGrid; an explicit instantiation may name it (access checks do
// not apply there), which hands its member pointer to the driver without touching the header.
using SplitFn = int (OctreeGrid::*)(int, bool, bool);
SplitFn split_cell_ptr();
template <SplitFn P>
struct SplitCellAccess {
  friend SplitFn split_cell_ptr() { return P; }
};
template struct SplitCellAccess<&OctreeGrid::splitCell>;

template <class T>
std::vector<T> read_arr(FILE* f, size_t n) {
  std::vector<T> v(n);
  if (n && fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short input");
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: octree_ref in.bin out.bin\n");
    return 2;
  }
  try {
    FILE* f = fopen(argv[1], "rb");
    if (!f) throw std::runtime_error("cannot open input");
    auto hdr = read_arr<int32_t>(f, 8);
    auto prm = read_arr<double>(f, 7);
    const int nV = hdr[6], nF = hdr[7];
    auto V = read_arr<double>(f, (size_t)nV * 3);
    auto F = read_arr<int32_t>(f, (size_t)nF * 3);
    int32_t tail[2] = {0, 0};
    const bool refine = fread(tail, sizeof(int32_t), 2, f) == 2;
    auto list = refine ? read_arr<int32_t>(f, (size_t)tail[1] * 4) : std::vector<int32_t>();
    fclose(f);
    const int stop_extent = hdr[3];
    const bool graded = hdr[4] != 0, paired = hdr[5] != 0;
    const double origin[3] = {prm[0], prm[1], prm[2]}, voxel_size = prm[3];
    const double mesh_transform[3] = {prm[4], prm[5], prm[6]};

    std::vector<double> fb((size_t)nF * 6);
    for (int t = 0; t < nF; ++t)
      for (int d = 0; d < 3; ++d) {
        double mn = V[(size_t)F[3 * t] * 3 + d], mx = mn;
        for (int c = 1; c < 3; ++c) {
          const double v = V[(size_t)F[3 * t + c] * 3 + d];
          mn = std::min(mn, v);
          mx = std::max(mx, v);
        }
        fb[6 * t + d] = mn;
        fb[6 * t + 3 + d] = mx;
      }
    // uniform bins over the facet boxes' bounding box
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int t = 0; t < nF; ++t)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], fb[6 * t + d]);
        hi[d] = std::max(hi[d], fb[6 * t + 3 + d]);
      }
    const int nb = std::max(1, std::min(128, (int)std::cbrt((double)nF)));
    double w[3];
    for (int d = 0; d < 3; ++d) w[d] = hi[d] > lo[d] ? (hi[d] - lo[d]) / nb : 1.0;
    auto bin = [&](double x, int d) {
      const double t = std::floor((x - lo[d]) / w[d]);
      return t < 0 ? 0 : t >= nb ? nb - 1 : (int)t;
    };
    std::vector<std::vector<int>> bins((size_t)nb * nb * nb);
    for (int t = 0; t < nF; ++t) {
      int b0[3], b1[3];
      for (int d = 0; d < 3; ++d) {
        b0[d] = bin(fb[6 * t + d], d);
        b1[d] = bin(fb[6 * t + 3 + d], d);
      }
      for (int k = b0[2]; k <= b1[2]; ++k)
        for (int j = b0[1]; j <= b1[1]; ++j)
          for (int i = b0[0]; i <= b1[0]; ++i) bins[((size_t)k * nb + j) * nb + i].push_back(t);
    }
    std::vector<int> stamp(nF, -1);
    int query = 0;
    int stop = stop_extent;
    auto should_subdivide = [&](int x, int y, int z, int extent) {
      if (extent <= stop) return false;
      const int p[3] = {x, y, z};
      double mn[3], mx[3];
      for (int d = 0; d < 3; ++d) {
        mn[d] = mesh_transform[d] + origin[d] + voxel_size * p[d];
        mx[d] = mn[d] + voxel_size * extent;
      }
      for (int d = 0; d < 3; ++d)
        if (mx[d] < lo[d] || hi[d] < mn[d]) return false;
      int b0[3], b1[3];
      for (int d = 0; d < 3; ++d) {
        b0[d] = bin(mn[d], d);
        b1[d] = bin(mx[d], d);
      }
      ++query;
      for (int k = b0[2]; k <= b1[2]; ++k)
        for (int j = b0[1]; j <= b1[1]; ++j)
          for (int i = b0[0]; i <= b1[0]; ++i)
            for (int t : bins[((size_t)k * nb + j) * nb + i]) {
              if (stamp[t] == query) continue;
              stamp[t] = query;
              bool ov = true;
              for (int d = 0; d < 3 && ov; ++d) ov = !(mx[d] < fb[6 * t + d] || fb[6 * t + 3 + d] < mn[d]);
              if (ov) return true;
            }
      return false;
    };

    OctreeGrid octree(Eigen::Vector3i(hdr[0], hdr[1], hdr[2]));
    // the queue of OctreeGrid::subdivide (octree.cpp:648-684), maxCells unlimited
    const auto t0 = std::chrono::steady_clock::now();
    std::queue<int> pending;
    for (int i = 0; i < octree.numCells(); ++i)
      if (octree.cellIsLeaf(i)) pending.push(i);
    while (!pending.empty()) {
      const int id = pending.front();
      pending.pop();
      const int extent = octree.cellExtent(id);
      const Eigen::Vector3i pos = octree.cellCornerPos(id, 0);
      if (should_subdivide(pos[0], pos[1], pos[2], extent)) {
        if (extent == 1) continue;
        if (octree.cellIsLeaf(id)) (octree.*split_cell_ptr())(id, graded, paired);
        for (int k = 0; k < 8; ++k) pending.push(octree.m_Cells[id].firstChild + k);
      }
    }
    if (refine) {
      // the targeted variant (octree.cpp:691-722): the listed leaves, no children queued
      stop = tail[0];
      std::vector<int> ids;
      for (int i = 0; i < tail[1]; ++i)
        for (int c = 0; c < octree.numCells(); ++c) {
          if (!octree.cellIsLeaf(c) || octree.cellExtent(c) != list[4 * i + 3]) continue;
          const Eigen::Vector3i p = octree.cellCornerPos(c, 0);
          if (p[0] == list[4 * i] && p[1] == list[4 * i + 1] && p[2] == list[4 * i + 2]) ids.push_back(c);
        }
      std::queue<int> listed;
      for (int c : ids)
        if (octree.cellIsLeaf(c)) listed.push(c);
      while (!listed.empty()) {
        const int id = listed.front();
        listed.pop();
        const int extent = octree.cellExtent(id);
        const Eigen::Vector3i pos = octree.cellCornerPos(id, 0);
        if (should_subdivide(pos[0], pos[1], pos[2], extent) && extent != 1 && octree.cellIsLeaf(id))
          (octree.*split_cell_ptr())(id, graded, paired);
      }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("subdivide %.6f s, %d cells, %d nodes\n", secs, octree.numCells(), octree.numNodes());

    std::vector<int32_t> leaves, corners;
    for (int c = 0; c < octree.numCells(); ++c) {
      if (!octree.cellIsLeaf(c)) continue;
      const Eigen::Vector3i p = octree.cellCornerPos(c, 0);
      leaves.insert(leaves.end(), {p[0], p[1], p[2], octree.cellExtent(c)});
      for (int k = 0; k < 8; ++k) corners.push_back(octree.cellCornerId(c, k));
    }
    FILE* o = fopen(argv[2], "wb");
    if (!o) throw std::runtime_error("cannot open output");
    const int32_t head[4] = {(int32_t)(leaves.size() / 4), octree.numNodes(), octree.is2to1Graded() ? 1 : 0,
                             octree.isPaired() ? 1 : 0};
    fwrite(head, sizeof head, 1, o);
    fwrite(leaves.data(), sizeof(int32_t), leaves.size(), o);
    for (int n = 0; n < octree.numNodes(); ++n) {
      const Eigen::Vector3i p = octree.nodePos(n);
      const int32_t q[3] = {p[0], p[1], p[2]};
      fwrite(q, sizeof q, 1, o);
    }
    fwrite(corners.data(), sizeof(int32_t), corners.size(), o);
    fclose(o);
    return 0;
  } catch (const std::exception& e) {
    fprintf(stderr, "octree_ref: %s\n", e.what());
    return 1;
  }
}
