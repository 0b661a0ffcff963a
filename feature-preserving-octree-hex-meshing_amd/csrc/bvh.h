// bvh.h -- closest point + pseudonormal on a triangle surface (igl::AABB +
// igl::signed_distance_pseudonormal semantics) for the tangent-plane projection.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace hexopt {

// Tree in igl::AABB's shape, flattened: node i has box[6*i..] = (min xyz, max xyz),
// children left/right (-1 if absent) and prim (face id at a leaf, else -1).
struct SurfaceTree {
  std::vector<double> box;
  std::vector<int> left, right, prim;
  int depth = 0;
};

bool build_surface_tree(int nv, const double* V, int nf, const int* F, SurfaceTree& T, const char** err);

// Node records for the closest-point walk, kNodeRec doubles (128 B) per node, so that one
// node visit is one dependent fetch. Inner node: its left child's box, then its right
// child's box (12 doubles). Leaf: its face's corners a, b, c (9 doubles). In both, the int32
// words 24..26 (doubles 12..13) hold left, right, prim; the rest is zero.
constexpr int kNodeRec = 16;
void pack_node_records(const SurfaceTree& T, const double* V, const int* F, std::vector<double>& rec);
void surface_normals(int nv, const double* V, int nf, const int* F, std::vector<double>& FN, std::vector<double>& VN,
                     std::vector<double>& FEN);

struct SurfaceDev {
  int nv, nf, nnodes, depth;
  const double* V;    // nv*3 row-major
  const int* F;       // nf*3
  const double* box;  // nnodes*6
  const int* left;
  const int* right;
  const int* prim;
  const double* FN;   // nf*3
  const double* VN;   // nv*3
  const double* FEN;  // nf*9
  const double* rec;  // nnodes*kNodeRec (pack_node_records)
};

constexpr int kBvhBlock = 128;
constexpr int kBvhMaxDepth = 48;
// traversal stack entries per lane for a tree of `depth` levels (root = level 1): a
// depth-first walk holds at most one frame per level, plus one spare; never above
// kBvhMaxDepth + 2 (surfaces with deeper trees are refused at creation)
__host__ __device__ inline int stack_frames(int depth) {
  const int d = depth < 1 ? 1 : depth;
  return d + 2 < kBvhMaxDepth + 2 ? d + 2 : kBvhMaxDepth + 2;
}

// Queries: P (SoA x[d*stride + i] when soa, else nq*3 rows), optional index list `ids`
// (query i uses point ids[i]). Outputs (nullable): S, I, C (nq*3), N (nq*3), sqrD. status:
// set to 1 if any traversal exceeded the stack depth (never for trees of depth <= 48).
// `order` (nullable, a permutation of the queries) sets which query each lane takes: queries
// of nearby points in neighbouring lanes walk the same tree paths.
void launch_closest(const SurfaceDev& s, int nq, const double* P, int soa, int stride, const int* ids, double* S, int* I,
                    double* C, double* N, double* sqrD, int* status, hipStream_t st, const int* order = nullptr);

// Several surfaces in one launch: segment j's queries are order[q0 .. q0 + n) against surface
// s, taken by blocks block0 .. block0 + ceil(n / kBvhBlock) (segments in increasing block0,
// the first at block 0). Writes C and N (indexed by the query, as launch_closest does).
struct PatchSeg {
  SurfaceDev s;
  int block0, q0, n, pad;
};
void launch_closest_segs(const PatchSeg* segs, int nseg, int nblocks, int max_depth, const double* P, int soa,
                         int stride, const int* ids, const int* order, double* C, double* N, int* status,
                         hipStream_t st);

// Per-face edge-plane data of vcg's tri::UpdateEdges<>::Set (vcg/complex/algorithms/update/
// component_ep.h:48-73), kEPRec doubles per face: the three edges scaled by 1 / the normal's
// dominant component (9), the twice-normalised normal (3), the plane offset (1), the dominant
// axis 0/1/2 (1), padding (2). Metro's point-to-face distance (PointDistanceEP) reads them.
constexpr int kEPRec = 16;
void build_ep_faces(int nv, const double* V, int nf, const int* F, std::vector<double>& ep);
// Metro's sampling distance of each query (vcg Sampling::AddSample, extern/vcg/sampling.h:
// 224-258): min(upper, min over faces of PointDistanceEP), through the surface tree. dist[i]
// == upper when no face is closer.
void launch_metro_ep(const SurfaceDev& s, const double* ep, int nq, const double* P, double upper, double* dist,
                     int* status, hipStream_t st);

struct SoftDev;
void launch_refold_planes(const SoftDev& sd, const int* trow, int nrows, const double* Cp, double* t_d,
                          const double* S0, const double* b0, double* S, double* b, double lam, hipStream_t st);

}  // namespace hexopt
