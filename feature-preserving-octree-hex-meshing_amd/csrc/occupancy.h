// occupancy.h -- octree subdivision predicate and +z ray-parity inside tests on a triangle
// surface (grid_hex_meshing.cpp:502-517, voxelization.cpp:101-163, voxelization.h:221-272).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "bvh.h"

namespace hexopt {

// Facets binned by their closed xy box on a uniform nb[0] x nb[1] grid spanning the xy box of
// the facet vertices. column_bin is monotone in x, so a point inside a facet's closed xy box
// always falls in one of the facet's bins; a point outside [lo, hi] lies in no facet box.
__host__ __device__ inline int column_bin(double x, double lo, double inv, int nb) {
  const double t = (x - lo) * inv;
  if (!(t >= 0.0)) return 0;
  if (t >= (double)nb) return nb - 1;
  return (int)t;
}

struct ColumnBins {
  double lo[2] = {0, 0}, hi[2] = {0, 0}, inv[2] = {0, 0};
  int nb[2] = {1, 1};
  std::vector<int> off;      // nb0*nb1 + 1
  std::vector<int> fid;      // facet ids, ascending within a bin
  std::vector<double> fbox;  // nf*4: xmin, xmax, ymin, ymax (geogram get_facet_bbox)
};
void build_column_bins(int nv, const double* V, int nf, const int* F, ColumnBins& B);

struct ColumnBinsDev {
  double lo[2], hi[2], inv[2];
  int nb[2];
  const int* off;
  const int* fid;
  const double* fbox;
};

// Cells are n x 4 ints (x, y, z, extent). base[d] = mesh_transform[d] + origin[d] (the
// reference's left-to-right sum); split[i] = extent > stop_extent and the cell box touches
// at least one facet box.
void launch_occ_classify(const SurfaceDev& s, int n, const int* cells, const double base[3], double voxel_size,
                         int stop_extent, uint8_t* split, hipStream_t st);
// inside[(z*ny + y)*nx + x] = 1 when the voxel centre is inside (VoxelGrid variant).
void launch_occ_voxel_parity(const SurfaceDev& s, const ColumnBinsDev& b, int nx, int ny, int nz,
                             const double origin[3], double spacing, uint8_t* inside, hipStream_t st);
// inside[i] = 1.0f when the centre of cell i is inside (OctreeGrid variant).
void launch_occ_cell_parity(const SurfaceDev& s, const ColumnBinsDev& b, int n, const int* cells,
                            const double origin[3], double spacing, float* inside, hipStream_t st);
// Hex bounding-box centres (clean_hex_mesh) as nH x 3 rows, from SoA positions.
void launch_hex_centres(const double* X, int nV, const int* conn, int nH, double* P, hipStream_t st);

}  // namespace hexopt
