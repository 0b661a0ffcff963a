// octree.h -- level-synchronous octree subdivision on the GPU: the leaf set and the octree
// hex mesh of OctreeGrid::subdivide(should_subdivide, graded, paired) started from the root
// cells of OctreeGrid(grid) (grid_meshing/octree.cpp:17-117, 501-690;
// grid_meshing/grid_hex_meshing.cpp:490-562).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "bvh.h"

namespace hexopt {

struct OctreeOut {
  std::vector<int32_t> leaves;  // n x 4: x, y, z, extent (voxel units), level by level
  std::vector<int32_t> nodes;   // m x 3: integer node positions
  std::vector<double> V;        // m x 3: base + voxel_size * node position
  std::vector<int32_t> hexes;   // n x 8: node ids of each leaf's corners, reference corner order
  int levels = 0;               // number of levels holding leaves (deepest leaf level + 1)
  long evaluated = 0;           // predicate evaluations
  double ms = 0;                // device time of the subdivision (hipEvents on its stream)
};

// Device state of one octree: a dense byte grid of split flags per level. grid: cells per axis
// in voxels, each a power of two; base[d] = mesh_transform[d] + origin[d]. A cell is split by
// the predicate when its extent exceeds stop_extent (and 1) and its closed box touches a facet
// box; grading makes the 6 face and 12 edge neighbours of every split cell exist at its level;
// pairing splits all siblings of a split cell (all root cells for a root).
struct OctreeDev;
OctreeDev* octree_create(const int grid[3], const double base[3], double voxel_size, bool graded, bool paired,
                         std::string& err);
void octree_destroy(OctreeDev* o);
// OctreeGrid::subdivide(predicate, graded, paired) of a fresh octree (octree.cpp:648-690: its
// queue starts from the root cells) when cells is null; the targeted variant over n listed cells (x, y, z, extent rows, host
// memory; octree.cpp:691-729: a listed cell that is a leaf and passes the predicate is split,
// its children are not tested) otherwise. Then the leaves, nodes and hexes of the result into
// out (ms: device time of the subdivision). False with a message on a bad grid or cell, a
// level above 2^31 cells or a device error.
bool octree_timed(OctreeDev& o, const SurfaceDev* s, int stop_extent, int n, const int* cells, hipStream_t st,
                  OctreeOut& out, std::string& err);

}  // namespace hexopt
