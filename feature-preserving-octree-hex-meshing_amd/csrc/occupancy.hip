// occupancy.hip -- occupancy of octree cells and voxels against a triangle surface, FP64.
// Built without fused multiply-add so every expression rounds as in the reference's CPU
// build; the expressions below keep the reference's operand order.
//
// Subdivision predicate (should_subdivide, grid_hex_meshing.cpp:502-517): does the closed
// cell box touch the closed box of any facet (GEO::MeshFacetsAABB::
// compute_bbox_facet_bbox_intersections, mesh_AABB.h:116-123,214-237; bboxes_overlap,
// basic/geometry.h:612-622)? One lane per cell walks the surface tree depth first and stops
// at the first facet box it touches. Node boxes are exact unions of their facets' boxes, so
// the answer does not depend on the tree's shape.
//
// Ray parity (compute_sign for VoxelGrid, voxelization.h:221-272, and for OctreeGrid,
// voxelization.cpp:101-163): a vertical line through the query's xy crosses facet f when
// point_in_triangle_2d (voxelization.cpp:57-93, SOS orientation) accepts it, at height
// z_f = u p1z + v p2z + w p3z with sign s_f = orient_2d_inexact of the facet's xy triangle.
// With S = sum of s_f over crossings with z_f < z_query:
//   VoxelGrid: inside = S < 0, literally.
//   OctreeGrid: the reference sorts crossings by (z, s), keeps the ones where the running sum
//   steps from 0 to -1 or from -1 to 0, and calls the cell inside when an odd number of kept
//   heights lie below the centre. Every step is +-1, so the kept crossings below the centre
//   are exactly the crossings of the level -1/2 by the running sum over that prefix, and
//   their count is odd iff the prefix sum is <= -1: inside = S < 0 as well. Neither needs a
//   sort.
// Candidate facets come from the xy bin holding the query point (a facet that accepts the
// point has it inside its closed xy box); each candidate then passes the reference's
// closed-box prefilter before the 2D test.
#include <hip/hip_runtime.h>
#include "occupancy.h"

namespace hexopt {
namespace {

constexpr int kOccBlock = 128;
constexpr int kColLanes = 64;   // one wave of columns along x
constexpr int kColCap = 32;     // crossings cached per column in LDS
constexpr int kColZChunk = 64;  // z slices per block

__device__ __forceinline__ bool overlap_closed(const double* nb, const double mn[3], const double mx[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (mx[d] < nb[d]) return false;
    if (mn[d] > nb[3 + d]) return false;
  }
  return true;
}

__global__ __launch_bounds__(kOccBlock) void occ_classify_kernel(SurfaceDev s, int n, const int* __restrict__ cells,
                                                                  double b0, double b1, double b2, double vs,
                                                                  int stop_extent, uint8_t* __restrict__ split) {
  // depth-first, at most one pending sibling per level plus the node being expanded:
  // stack_frames(depth) x kOccBlock entries of dynamic LDS (see launch_occ_classify)
  extern __shared__ int stack_mem[];
  int (*stack)[kOccBlock] = reinterpret_cast<int (*)[kOccBlock]>(stack_mem);
  const int i = blockIdx.x * kOccBlock + threadIdx.x;
  if (i >= n) return;
  const int x = cells[4 * i], y = cells[4 * i + 1], z = cells[4 * i + 2], ext = cells[4 * i + 3];
  if (ext <= stop_extent) { split[i] = 0; return; }
  double mn[3], mx[3];
  mn[0] = b0 + vs * x;
  mn[1] = b1 + vs * y;
  mn[2] = b2 + vs * z;
#pragma unroll
  for (int d = 0; d < 3; ++d) mx[d] = mn[d] + vs * ext;
  const int t = threadIdx.x;
  int sp = 0;
  stack[sp++][t] = 0;
  uint8_t hit = 0;
  while (sp > 0) {
    const int node = stack[--sp][t];
    if (!overlap_closed(s.box + (size_t)node * 6, mn, mx)) continue;
    if (s.prim[node] >= 0) { hit = 1; break; }
    // depth-first; at most one pending sibling per level plus the node being expanded
    const int r = s.right[node], l = s.left[node];
    if (r >= 0) stack[sp++][t] = r;
    if (l >= 0) stack[sp++][t] = l;
  }
  split[i] = hit;
}

// orientation() of voxelization.cpp:62-72
__device__ __forceinline__ int sos_orientation(double x1, double y1, double x2, double y2, double& area2) {
  area2 = y1 * x2 - x1 * y2;
  if (area2 > 0) return 1;
  if (area2 < 0) return -1;
  if (y2 > y1) return 1;
  if (y2 < y1) return -1;
  if (x1 > x2) return 1;
  if (x1 < x2) return -1;
  return 0;
}

// intersect_ray_z (voxelization.h:195-217) for facet f at (qx, qy): crossing sign, height in z
__device__ int crossing(const SurfaceDev& s, int f, double qx, double qy, double& z) {
  const int i1 = s.F[3 * f], i2 = s.F[3 * f + 1], i3 = s.F[3 * f + 2];
  const double* p1 = s.V + (size_t)i1 * 3;
  const double* p2 = s.V + (size_t)i2 * 3;
  const double* p3 = s.V + (size_t)i3 * 3;
  // point_in_triangle_2d
  const double x1 = p1[0] - qx, x2 = p2[0] - qx, x3 = p3[0] - qx;
  const double y1 = p1[1] - qy, y2 = p2[1] - qy, y3 = p3[1] - qy;
  double a, b, c;
  const int sa = sos_orientation(x2, y2, x3, y3, a);
  if (sa == 0) return 0;
  const int sb = sos_orientation(x3, y3, x1, y1, b);
  if (sb != sa) return 0;
  const int sc = sos_orientation(x1, y1, x2, y2, c);
  if (sc != sa) return 0;
  const double sum = a + b + c;
  a /= sum;
  b /= sum;
  c /= sum;
  z = a * p1[2] + b * p2[2] + c * p3[2];
  // orient_2d_inexact(p1, p2, p3) (voxelization.h:169-182)
  const double a11 = p2[0] - p1[0], a12 = p2[1] - p1[1];
  const double a21 = p3[0] - p1[0], a22 = p3[1] - p1[1];
  const double delta = a11 * a22 - a12 * a21;
  return delta > 0 ? 1 : (delta < 0 ? -1 : 0);
}

// Facets of the bin holding (qx, qy); empty when the point lies outside every facet box.
__device__ __forceinline__ bool bin_range(const ColumnBinsDev& b, double qx, double qy, int& beg, int& end) {
  if (qx < b.lo[0] || qx > b.hi[0] || qy < b.lo[1] || qy > b.hi[1]) return false;
  const int bx = column_bin(qx, b.lo[0], b.inv[0], b.nb[0]);
  const int by = column_bin(qy, b.lo[1], b.inv[1], b.nb[1]);
  const int k = by * b.nb[0] + bx;
  beg = b.off[k];
  end = b.off[k + 1];
  return true;
}

// Sum of crossing signs below cz over the candidate facets of (qx, qy); the query box of the
// prefilter is [bx0, bx1] x [by0, by1] in xy (its z range spans the whole surface).
__device__ int signed_crossings_below(const SurfaceDev& s, const ColumnBinsDev& b, double qx, double qy, double cz,
                                     double bx0, double bx1, double by0, double by1) {
  int beg, end;
  if (!bin_range(b, qx, qy, beg, end)) return 0;
  int S = 0;
  for (int k = beg; k < end; ++k) {
    const int f = b.fid[k];
    const double* fb = b.fbox + (size_t)f * 4;
    if (bx1 < fb[0] || bx0 > fb[1] || by1 < fb[2] || by0 > fb[3]) continue;
    double z;
    const int sg = crossing(s, f, qx, qy, z);
    if (sg != 0 && z < cz) S += sg;
  }
  return S;
}

__global__ __launch_bounds__(kColLanes) void occ_voxel_kernel(SurfaceDev s, ColumnBinsDev b, int nx, int ny, int nz,
                                                              double o0, double o1, double o2, double sp,
                                                              uint8_t* __restrict__ inside) {
  __shared__ double zs[kColCap][kColLanes];
  __shared__ signed char ss[kColCap][kColLanes];
  const int t = threadIdx.x;
  const int x = blockIdx.x * kColLanes + t;
  const int y = blockIdx.y;
  const int z0 = blockIdx.z * kColZChunk;
  const int z1 = min(nz, z0 + kColZChunk);
  if (x >= nx) return;
  // voxel_center (voxelization.h:84-91): (i + 0.5) * spacing + origin
  const double cx = (x + 0.5) * sp + o0;
  const double cy = (y + 0.5) * sp + o1;
  int cnt = 0;
  int beg = 0, end = 0;
  if (bin_range(b, cx, cy, beg, end)) {
    for (int k = beg; k < end; ++k) {
      const int f = b.fid[k];
      const double* fb = b.fbox + (size_t)f * 4;
      if (cx < fb[0] || cx > fb[1] || cy < fb[2] || cy > fb[3]) continue;
      double z;
      const int sg = crossing(s, f, cx, cy, z);
      if (sg == 0) continue;
      if (cnt < kColCap) {
        zs[cnt][t] = z;
        ss[cnt][t] = (signed char)sg;
      }
      ++cnt;
    }
  }
  const size_t plane = (size_t)nx * ny;
  uint8_t* out = inside + (size_t)y * nx + x;
  if (cnt <= kColCap) {
    // insertion sort by height, then one sweep up the column
    for (int i = 1; i < cnt; ++i) {
      const double zv = zs[i][t];
      const signed char sv = ss[i][t];
      int j = i - 1;
      while (j >= 0 && zs[j][t] > zv) {
        zs[j + 1][t] = zs[j][t];
        ss[j + 1][t] = ss[j][t];
        --j;
      }
      zs[j + 1][t] = zv;
      ss[j + 1][t] = sv;
    }
    int i = 0, S = 0;
    for (int z = z0; z < z1; ++z) {
      const double cz = (z + 0.5) * sp + o2;
      for (; i < cnt && zs[i][t] < cz; ++i) S += ss[i][t];
      out[(size_t)z * plane] = (uint8_t)(S < 0 ? 1 : 0);
    }
  } else {
    // more crossings than the cache holds: recount per voxel
    for (int z = z0; z < z1; ++z) {
      const double cz = (z + 0.5) * sp + o2;
      const int S = signed_crossings_below(s, b, cx, cy, cz, cx, cx, cy, cy);
      out[(size_t)z * plane] = (uint8_t)(S < 0 ? 1 : 0);
    }
  }
}

__global__ __launch_bounds__(kOccBlock) void occ_cell_kernel(SurfaceDev s, ColumnBinsDev b, int n,
                                                             const int* __restrict__ cells, double o0, double o1,
                                                             double o2, double sp, float* __restrict__ inside) {
  const int i = blockIdx.x * kOccBlock + threadIdx.x;
  if (i >= n) return;
  const int x = cells[4 * i], y = cells[4 * i + 1], z = cells[4 * i + 2], ext = cells[4 * i + 3];
  // voxelization.cpp:115-128
  const double bx0 = o0 + sp * x, by0 = o1 + sp * y;
  const double bx1 = bx0 + sp * ext, by1 = by0 + sp * ext;
  const double half = 0.5 * sp * ext;
  const double cx = bx0 + half, cy = by0 + half;
  const double cz = o2 + sp * z + half;
  const int S = signed_crossings_below(s, b, cx, cy, cz, bx0, bx1, by0, by1);
  inside[i] = S < 0 ? 1.0f : 0.0f;
}

__global__ void hex_centre_kernel(const double* __restrict__ X, int nV, const int* __restrict__ conn, int nH,
                                  double* __restrict__ P) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nH) return;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int v = conn[(size_t)k * nH + h];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double p = X[(size_t)d * nV + v];
      mn[d] = fmin(mn[d], p);
      mx[d] = fmax(mx[d], p);
    }
  }
  // (VV.rowwise().maxCoeff() + VV.rowwise().minCoeff()) / 2 (grid_hex_meshing.cpp:1943)
#pragma unroll
  for (int d = 0; d < 3; ++d) P[(size_t)h * 3 + d] = (mx[d] + mn[d]) / 2;
}

}  // namespace

void launch_occ_classify(const SurfaceDev& s, int n, const int* cells, const double base[3], double voxel_size,
                         int stop_extent, uint8_t* split, hipStream_t st) {
  if (n <= 0) return;
  const size_t lds = (size_t)stack_frames(s.depth) * kOccBlock * sizeof(int);
  hipLaunchKernelGGL(occ_classify_kernel, dim3((n + kOccBlock - 1) / kOccBlock), dim3(kOccBlock), lds, st, s, n, cells,
                     base[0], base[1], base[2], voxel_size, stop_extent, split);
}

void launch_occ_voxel_parity(const SurfaceDev& s, const ColumnBinsDev& b, int nx, int ny, int nz,
                             const double origin[3], double spacing, uint8_t* inside, hipStream_t st) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  const dim3 grid((nx + kColLanes - 1) / kColLanes, ny, (nz + kColZChunk - 1) / kColZChunk);
  hipLaunchKernelGGL(occ_voxel_kernel, grid, dim3(kColLanes), 0, st, s, b, nx, ny, nz, origin[0], origin[1], origin[2],
                     spacing, inside);
}

void launch_occ_cell_parity(const SurfaceDev& s, const ColumnBinsDev& b, int n, const int* cells,
                            const double origin[3], double spacing, float* inside, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(occ_cell_kernel, dim3((n + kOccBlock - 1) / kOccBlock), dim3(kOccBlock), 0, st, s, b, n, cells,
                     origin[0], origin[1], origin[2], spacing, inside);
}

void launch_hex_centres(const double* X, int nV, const int* conn, int nH, double* P, hipStream_t st) {
  if (nH <= 0) return;
  hipLaunchKernelGGL(hex_centre_kernel, dim3((nH + 255) / 256), dim3(256), 0, st, X, nV, conn, nH, P);
}

}  // namespace hexopt
