// octree.hip -- level-synchronous octree subdivision (see octree.h).
//
// The reference grows its octree from a FIFO queue (OctreeGrid::subdivide, octree.cpp:648-690):
// a popped cell that satisfies the predicate is split and its children are queued, and every
// split cascades (splitCell, octree.cpp:501-593) into grading splits of coarser face and edge
// neighbours (makeCellGraded, :598-627) and pairing splits of the split cell's siblings, or of
// every root cell (:569-589). Two facts make the result independent of that order:
//   - a cell is tested by the predicate exactly when every ancestor passed it, so the set P of
//     predicate splits is fixed by the predicate alone;
//   - every cascade rule reads "if C is split then D is split" with D at C's level or coarser,
//     so the split set is the least set that holds P and is closed under those rules.
// The build therefore runs in three level-synchronous phases on the device:
//   A. top-down, one predicate launch per level over the children of the previous level's
//      predicate splits (compacted lists), marking P in one dense byte grid per level;
//   B. bottom-up, per level: pairing (a parent with one split child gets all eight split),
//      then every split cell marks its parent and, graded, the parents of its 6 face and 12
//      edge neighbours one level up;
//   C. leaves (cells whose parent is split and that are not split themselves), the octree
//      nodes (leaf corners marked in a dense lattice at the finest leaf level, numbered by an
//      exclusive scan) and each leaf's 8 corner ids (grid_hex_meshing.cpp:527-562).
// All flag writes store the value 1 to bytes, so concurrent writers need no atomics.
#include "octree.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <memory>

#include "occupancy.h"

namespace hexopt {
namespace {

constexpr int kOctBlock = 256;

struct OctBuf {
  void* p = nullptr;
  size_t bytes = 0;
  OctBuf() = default;
  OctBuf(const OctBuf&) = delete;
  OctBuf& operator=(const OctBuf&) = delete;
  ~OctBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t nb) {
    if (nb <= bytes && p) return hipSuccess;
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, std::max<size_t>(nb, 16));
    if (e == hipSuccess) bytes = std::max<size_t>(nb, 16);
    return e;
  }
  // grows keeping the first `keep` bytes
  hipError_t grow(size_t nb, size_t keep, hipStream_t st) {
    if (nb <= bytes && p) return hipSuccess;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return e;
    if (keep) {
      e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) {
        (void)hipFree(q);
        return e;
      }
    }
    if (p) (void)hipFree(p);
    p = q;
    bytes = nb;
    return hipSuccess;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

struct Dims {
  int x, y, z;
  long long count() const { return (long long)x * y * z; }
};

__device__ __forceinline__ long long lin3(int x, int y, int z, Dims d) { return ((long long)z * d.y + y) * d.x + x; }

__device__ __forceinline__ void unlin3(long long i, Dims d, int& x, int& y, int& z) {
  x = (int)(i % d.x);
  const long long t = i / d.x;
  y = (int)(t % d.y);
  z = (int)(t / d.y);
}

// corner k of a cell at (0/1 per axis), in the reference's CORNER_X?_Y?_Z? order (octree.h)
__device__ __forceinline__ void corner_offset(int k, int& dx, int& dy, int& dz) {
  dx = ((k + 1) >> 1) & 1;  // 0 1 1 0 0 1 1 0
  dy = (k >> 1) & 1;        // 0 0 1 1 0 0 1 1
  dz = k >> 2;              // 0 0 0 0 1 1 1 1
}

// predicate flags of a candidate list: drop extent-1 cells (the reference cannot split them),
// mark the rest in the level's split grid
__global__ __launch_bounds__(kOctBlock) void mark_split_kernel(const int4* __restrict__ cand, uint8_t* __restrict__ flag,
                                                                int n, Dims d, uint8_t* __restrict__ S) {
  const int i = blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n) return;
  const int4 c = cand[i];
  const bool f = flag[i] != 0 && c.w > 1;
  flag[i] = f ? 1 : 0;
  if (f) S[lin3(c.x / c.w, c.y / c.w, c.z / c.w, d)] = 1;
}

__global__ __launch_bounds__(kOctBlock) void children_kernel(const int4* __restrict__ sel, int n8,
                                                              int4* __restrict__ out) {
  const int i = blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n8) return;
  const int4 c = sel[i >> 3];
  const int h = c.w >> 1;
  int dx, dy, dz;
  corner_offset(i & 7, dx, dy, dz);
  out[i] = make_int4(c.x + h * dx, c.y + h * dy, c.z + h * dz, h);
}

// pairing: a parent (dims dp) with one split child gets all eight children split
__global__ __launch_bounds__(kOctBlock) void pair_kernel(uint8_t* __restrict__ S, Dims dc, Dims dp, long long np) {
  const long long i = (long long)blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= np) return;
  int px, py, pz;
  unlin3(i, dp, px, py, pz);
  bool any = false;
  for (int k = 0; k < 8; ++k) {
    int dx, dy, dz;
    corner_offset(k, dx, dy, dz);
    any |= S[lin3(2 * px + dx, 2 * py + dy, 2 * pz + dz, dc)] != 0;
  }
  if (!any) return;
  for (int k = 0; k < 8; ++k) {
    int dx, dy, dz;
    corner_offset(k, dx, dy, dz);
    S[lin3(2 * px + dx, 2 * py + dy, 2 * pz + dz, dc)] = 1;
  }
}

// a split cell splits its parent and, graded, the parents of its face and edge neighbours
__global__ __launch_bounds__(kOctBlock) void up_kernel(const uint8_t* __restrict__ S, Dims d, long long n,
                                                        uint8_t* __restrict__ Sp, Dims dp, int graded) {
  const long long i = (long long)blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n || !S[i]) return;
  int x, y, z;
  unlin3(i, d, x, y, z);
  Sp[lin3(x >> 1, y >> 1, z >> 1, dp)] = 1;
  if (!graded) return;
  for (int oz = -1; oz <= 1; ++oz)
    for (int oy = -1; oy <= 1; ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        const int m = (ox != 0) + (oy != 0) + (oz != 0);
        if (m == 0 || m == 3) continue;
        const int a = x + ox, b = y + oy, c = z + oz;
        if (a < 0 || b < 0 || c < 0 || a >= d.x || b >= d.y || c >= d.z) continue;
        Sp[lin3(a >> 1, b >> 1, c >> 1, dp)] = 1;
      }
}

// leaf flags of one level: the cell exists (root level, or its parent is split) and is not split
__global__ __launch_bounds__(kOctBlock) void leaf_flag_kernel(const uint8_t* __restrict__ Sp, Dims dp,
                                                               const uint8_t* __restrict__ S, Dims d, long long n,
                                                               uint8_t* __restrict__ flag) {
  const long long i = (long long)blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  unlin3(i, d, x, y, z);
  const bool exists = Sp == nullptr || Sp[lin3(x >> 1, y >> 1, z >> 1, dp)] != 0;
  const bool split = S != nullptr && S[i] != 0;
  flag[i] = exists && !split ? 1 : 0;
}

__global__ __launch_bounds__(kOctBlock) void leaf_emit_kernel(const int* __restrict__ idx, int n, Dims d, int e,
                                                               int4* __restrict__ out) {
  const int i = blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  unlin3(idx[i], d, x, y, z);
  out[i] = make_int4(x * e, y * e, z * e, e);
}

__global__ __launch_bounds__(kOctBlock) void node_mark_kernel(const int4* __restrict__ leaves, int n, int ef, Dims m,
                                                               uint8_t* __restrict__ nflag) {
  const int i = blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n) return;
  const int4 c = leaves[i];
  for (int k = 0; k < 8; ++k) {
    int dx, dy, dz;
    corner_offset(k, dx, dy, dz);
    nflag[lin3((c.x + dx * c.w) / ef, (c.y + dy * c.w) / ef, (c.z + dz * c.w) / ef, m)] = 1;
  }
}

__global__ __launch_bounds__(kOctBlock) void node_emit_kernel(const uint8_t* __restrict__ nflag,
                                                               const int* __restrict__ nid, Dims m, long long n, int ef,
                                                               double b0, double b1, double b2, double vs,
                                                               int* __restrict__ nodes, double* __restrict__ V) {
  const long long i = (long long)blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n || !nflag[i]) return;
  int x, y, z;
  unlin3(i, m, x, y, z);
  const long long j = nid[i];
  const int p[3] = {x * ef, y * ef, z * ef};
  nodes[3 * j] = p[0];
  nodes[3 * j + 1] = p[1];
  nodes[3 * j + 2] = p[2];
  // mesh_transform + origin + position * voxel_size, in the reference's order
  V[3 * j] = b0 + (double)p[0] * vs;
  V[3 * j + 1] = b1 + (double)p[1] * vs;
  V[3 * j + 2] = b2 + (double)p[2] * vs;
}

__global__ __launch_bounds__(kOctBlock) void hex_kernel(const int4* __restrict__ leaves, int n,
                                                         const int* __restrict__ nid, Dims m, int ef,
                                                         int* __restrict__ H) {
  const int i = blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n) return;
  const int4 c = leaves[i];
  for (int k = 0; k < 8; ++k) {
    int dx, dy, dz;
    corner_offset(k, dx, dy, dz);
    H[8 * (long long)i + k] = nid[lin3((c.x + dx * c.w) / ef, (c.y + dy * c.w) / ef, (c.z + dz * c.w) / ef, m)];
  }
}

// listed cells of one level for a targeted subdivision: a cell is split when it is a leaf now
// (its parent is split, or it is a root, and it is not split itself) and the predicate held
__global__ __launch_bounds__(kOctBlock) void refine_mark_kernel(const int4* __restrict__ cells,
                                                                 const uint8_t* __restrict__ pred, int n,
                                                                 const uint8_t* __restrict__ Sp, Dims dp,
                                                                 uint8_t* __restrict__ S, Dims d) {
  const int i = blockIdx.x * kOctBlock + threadIdx.x;
  if (i >= n) return;
  const int4 c = cells[i];
  const int x = c.x / c.w, y = c.y / c.w, z = c.z / c.w;
  const bool exists = Sp == nullptr || Sp[lin3(x >> 1, y >> 1, z >> 1, dp)] != 0;
  if (exists && pred[i] != 0 && c.w > 1) S[lin3(x, y, z, d)] = 1;
}

struct ToInt {
  __host__ __device__ int operator()(uint8_t v) const { return v; }
};

inline unsigned blocks(long long n) { return (unsigned)((n + kOctBlock - 1) / kOctBlock); }

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace

#define OCT_CK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      err = std::string("octree: ") + #expr + ": " + hipGetErrorString(e_);   \
      return false;                                                           \
    }                                                                         \
  } while (0)

struct OctreeDev {
  int grid[3];
  double base[3];
  double vs;
  bool graded, paired;
  int R = 0, D = 0;  // root extent, level of extent-1 cells
  Dims root{0, 0, 0};
  std::vector<std::unique_ptr<OctBuf>> S;  // split flags per level, allocated and zeroed on demand
  int Lmax = -1;                            // deepest level whose grid may hold a split
  OctBuf cand, flag, sel, tmp, num, idx;
  int* h_num = nullptr;
  ~OctreeDev() {
    if (h_num) (void)hipHostFree(h_num);
  }
  Dims dims(int L) const { return Dims{root.x << L, root.y << L, root.z << L}; }
  int extent(int L) const { return R >> L; }
};

namespace {

constexpr long long kMaxGrid = (1LL << 31) - 1;

bool level_grid(OctreeDev& o, int L, hipStream_t st, std::string& err) {
  while ((int)o.S.size() <= L) o.S.emplace_back(new OctBuf());
  if (o.S[L]->p) return true;
  const Dims d = o.dims(L);
  if (d.count() > kMaxGrid) {
    err = "octree: level " + std::to_string(L) + " exceeds the dense grid limit of 2^31 cells";
    return false;
  }
  OCT_CK(o.S[L]->ensure((size_t)d.count()));
  OCT_CK(hipMemsetAsync(o.S[L]->p, 0, (size_t)d.count(), st));
  return true;
}

bool read_count(OctreeDev& o, hipStream_t st, int& n, std::string& err) {
  OCT_CK(hipMemcpyAsync(o.h_num, o.num.p, sizeof(int), hipMemcpyDeviceToHost, st));
  OCT_CK(hipStreamSynchronize(st));
  n = *o.h_num;
  return true;
}

// pairing and grading closure, bottom-up over every allocated level
bool close_splits(OctreeDev& o, hipStream_t st, std::string& err) {
  for (int L = o.Lmax; L >= 1; --L) {
    const Dims d = o.dims(L), dp = o.dims(L - 1);
    if (o.paired)
      hipLaunchKernelGGL(pair_kernel, dim3(blocks(dp.count())), dim3(kOctBlock), 0, st, o.S[L]->as<uint8_t>(), d, dp,
                         dp.count());
    hipLaunchKernelGGL(up_kernel, dim3(blocks(d.count())), dim3(kOctBlock), 0, st, o.S[L]->as<uint8_t>(), d,
                       d.count(), o.S[L - 1]->as<uint8_t>(), dp, o.graded ? 1 : 0);
    OCT_CK(hipGetLastError());
  }
  if (o.Lmax >= 0 && o.paired) {  // one split root splits every root
    std::vector<uint8_t> r((size_t)o.root.count());
    OCT_CK(hipMemcpyAsync(r.data(), o.S[0]->p, r.size(), hipMemcpyDeviceToHost, st));
    OCT_CK(hipStreamSynchronize(st));
    if (std::find(r.begin(), r.end(), 1) != r.end()) OCT_CK(hipMemsetAsync(o.S[0]->p, 1, r.size(), st));
  }
  return true;
}

}  // namespace

OctreeDev* octree_create(const int grid[3], const double base[3], double voxel_size, bool graded, bool paired,
                         std::string& err) {
  for (int a = 0; a < 3; ++a)
    if (!is_pow2(grid[a]) || grid[a] > (1 << 30)) {
      err = "octree: grid sizes must be powers of two up to 2^30";
      return nullptr;
    }
  auto o = std::make_unique<OctreeDev>();
  for (int a = 0; a < 3; ++a) {
    o->grid[a] = grid[a];
    o->base[a] = base[a];
  }
  o->vs = voxel_size;
  o->graded = graded;
  o->paired = paired;
  o->R = std::min(grid[0], std::min(grid[1], grid[2]));
  while ((1 << o->D) < o->R) ++o->D;
  o->root = Dims{grid[0] / o->R, grid[1] / o->R, grid[2] / o->R};
  if (hipHostMalloc((void**)&o->h_num, sizeof(int)) != hipSuccess || o->num.ensure(sizeof(int)) != hipSuccess) {
    err = "octree: device allocation failed";
    return nullptr;
  }
  return o.release();
}

void octree_destroy(OctreeDev* o) { delete o; }

bool octree_subdivide(OctreeDev& o, const SurfaceDev& s, int stop_extent, hipStream_t st, long& evaluated,
                      std::string& err) {
  // ---- predicate splits, top-down from the root cells ----
  std::vector<int> roots;
  const int R = o.R;
  for (int z = 0; z < o.root.z; ++z)
    for (int y = 0; y < o.root.y; ++y)
      for (int x = 0; x < o.root.x; ++x) roots.insert(roots.end(), {x * R, y * R, z * R, R});
  long long ncand = (long long)roots.size() / 4;
  OCT_CK(o.cand.ensure(roots.size() * sizeof(int)));
  OCT_CK(hipMemcpyAsync(o.cand.p, roots.data(), roots.size() * sizeof(int), hipMemcpyHostToDevice, st));
  for (int L = 0; L <= o.D && ncand > 0; ++L) {
    if (ncand > kMaxGrid / 8) {
      err = "octree: more than 2^28 candidate cells at level " + std::to_string(L);
      return false;
    }
    if (!level_grid(o, L, st, err)) return false;
    const Dims d = o.dims(L);
    evaluated += ncand;
    OCT_CK(o.flag.ensure((size_t)ncand));
    launch_occ_classify(s, (int)ncand, o.cand.as<int>(), o.base, o.vs, stop_extent, o.flag.as<uint8_t>(), st);
    OCT_CK(hipGetLastError());
    hipLaunchKernelGGL(mark_split_kernel, dim3(blocks(ncand)), dim3(kOctBlock), 0, st, o.cand.as<int4>(),
                       o.flag.as<uint8_t>(), (int)ncand, d, o.S[L]->as<uint8_t>());
    OCT_CK(hipGetLastError());
    OCT_CK(o.sel.ensure((size_t)ncand * sizeof(int4)));
    size_t tb = 0;
    OCT_CK(hipcub::DeviceSelect::Flagged(nullptr, tb, o.cand.as<int4>(), o.flag.as<uint8_t>(), o.sel.as<int4>(),
                                         o.num.as<int>(), (int)ncand, st));
    OCT_CK(o.tmp.ensure(tb));
    OCT_CK(hipcub::DeviceSelect::Flagged(o.tmp.p, tb, o.cand.as<int4>(), o.flag.as<uint8_t>(), o.sel.as<int4>(),
                                         o.num.as<int>(), (int)ncand, st));
    int nsel = 0;
    if (!read_count(o, st, nsel, err)) return false;
    if (nsel == 0) break;
    o.Lmax = std::max(o.Lmax, L);
    ncand = 8LL * nsel;
    OCT_CK(o.cand.ensure((size_t)ncand * sizeof(int4)));
    hipLaunchKernelGGL(children_kernel, dim3(blocks(ncand)), dim3(kOctBlock), 0, st, o.sel.as<int4>(), (int)ncand,
                       o.cand.as<int4>());
    OCT_CK(hipGetLastError());
  }
  // ---- cascades, bottom-up ----
  return close_splits(o, st, err);
}

bool octree_refine(OctreeDev& o, const SurfaceDev& s, int stop_extent, int n, const int* cells, hipStream_t st,
                   long& evaluated, std::string& err) {
  // the listed cells, level by level (a cell's level follows from its extent)
  std::vector<std::vector<int>> by_level(o.D + 1);
  for (int i = 0; i < n; ++i) {
    const int* c = cells + 4 * (size_t)i;
    const int e = c[3];
    int L = 0;
    while (L <= o.D && o.extent(L) != e) ++L;
    bool ok = L <= o.D;
    for (int a = 0; a < 3 && ok; ++a) ok = c[a] >= 0 && c[a] % e == 0 && c[a] + e <= o.grid[a];
    if (!ok) {
      err = "octree: refine cell " + std::to_string(i) + " is not a cell of this octree's grid";
      return false;
    }
    by_level[L].insert(by_level[L].end(), c, c + 4);
  }
  for (int L = 0; L <= o.D; ++L) {
    const long long m = (long long)by_level[L].size() / 4;
    if (m == 0) continue;
    if (!level_grid(o, L, st, err)) return false;
    if (L > 0 && !level_grid(o, L - 1, st, err)) return false;
    evaluated += m;
    OCT_CK(o.cand.ensure((size_t)m * sizeof(int4)));
    OCT_CK(hipMemcpyAsync(o.cand.p, by_level[L].data(), (size_t)m * sizeof(int4), hipMemcpyHostToDevice, st));
    OCT_CK(o.flag.ensure((size_t)m));
    launch_occ_classify(s, (int)m, o.cand.as<int>(), o.base, o.vs, stop_extent, o.flag.as<uint8_t>(), st);
    OCT_CK(hipGetLastError());
    hipLaunchKernelGGL(refine_mark_kernel, dim3(blocks(m)), dim3(kOctBlock), 0, st, o.cand.as<int4>(),
                       o.flag.as<uint8_t>(), (int)m, L > 0 ? o.S[L - 1]->as<uint8_t>() : nullptr,
                       L > 0 ? o.dims(L - 1) : o.dims(0), o.S[L]->as<uint8_t>(), o.dims(L));
    OCT_CK(hipGetLastError());
    // the copy above reuses `cand`: wait before the next level overwrites it
    OCT_CK(hipStreamSynchronize(st));
    o.Lmax = std::max(o.Lmax, L);
  }
  return close_splits(o, st, err);
}

bool octree_export(OctreeDev& o, hipStream_t st, OctreeOut& out, std::string& err) {
  const long evaluated = out.evaluated;
  out = OctreeOut();
  out.evaluated = evaluated;
  const int Lleaf = std::min(o.Lmax + 1, o.D);  // deepest level that can hold leaves
  for (int L = 0; L <= Lleaf; ++L)
    if (!level_grid(o, L, st, err)) return false;
  OctBuf leaves;
  long long nleaves = 0;
  int Lf = 0;  // deepest level that holds leaves
  for (int L = 0; L <= Lleaf; ++L) {
    const Dims d = o.dims(L);
    const uint8_t* Sp = L > 0 ? o.S[L - 1]->as<uint8_t>() : nullptr;
    const uint8_t* Sl = L <= o.Lmax ? o.S[L]->as<uint8_t>() : nullptr;
    OCT_CK(o.flag.ensure((size_t)d.count()));
    OCT_CK(o.idx.ensure((size_t)d.count() * sizeof(int)));
    hipLaunchKernelGGL(leaf_flag_kernel, dim3(blocks(d.count())), dim3(kOctBlock), 0, st, Sp,
                       L > 0 ? o.dims(L - 1) : d, Sl, d, d.count(), o.flag.as<uint8_t>());
    OCT_CK(hipGetLastError());
    hipcub::CountingInputIterator<int> it(0);
    size_t tb = 0;
    OCT_CK(hipcub::DeviceSelect::Flagged(nullptr, tb, it, o.flag.as<uint8_t>(), o.idx.as<int>(), o.num.as<int>(),
                                         (int)d.count(), st));
    OCT_CK(o.tmp.ensure(tb));
    OCT_CK(hipcub::DeviceSelect::Flagged(o.tmp.p, tb, it, o.flag.as<uint8_t>(), o.idx.as<int>(), o.num.as<int>(),
                                         (int)d.count(), st));
    int nl = 0;
    if (!read_count(o, st, nl, err)) return false;
    if (nl == 0) continue;
    if (nleaves + nl > kMaxGrid / 8) {
      err = "octree: more than 2^28 leaves";
      return false;
    }
    OCT_CK(leaves.grow((size_t)(nleaves + nl) * sizeof(int4) * 2, (size_t)nleaves * sizeof(int4), st));
    hipLaunchKernelGGL(leaf_emit_kernel, dim3(blocks(nl)), dim3(kOctBlock), 0, st, o.idx.as<int>(), nl, d,
                       o.extent(L), leaves.as<int4>() + nleaves);
    OCT_CK(hipGetLastError());
    nleaves += nl;
    Lf = L;
  }
  out.levels = Lf + 1;
  const int ef = o.extent(Lf);
  const Dims df = o.dims(Lf);
  const Dims m{df.x + 1, df.y + 1, df.z + 1};
  if (m.count() > kMaxGrid) {
    err = "octree: the node lattice at level " + std::to_string(Lf) + " exceeds 2^31 points";
    return false;
  }
  OctBuf nflag, nid, dnodes, dV, dH;
  OCT_CK(nflag.ensure((size_t)m.count()));
  OCT_CK(nid.ensure((size_t)m.count() * sizeof(int)));
  OCT_CK(hipMemsetAsync(nflag.p, 0, (size_t)m.count(), st));
  hipLaunchKernelGGL(node_mark_kernel, dim3(blocks(nleaves)), dim3(kOctBlock), 0, st, leaves.as<int4>(),
                     (int)nleaves, ef, m, nflag.as<uint8_t>());
  OCT_CK(hipGetLastError());
  {
    hipcub::TransformInputIterator<int, ToInt, const uint8_t*> in(nflag.as<uint8_t>(), ToInt());
    size_t tb = 0;
    OCT_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, nid.as<int>(), (int)m.count(), st));
    OCT_CK(o.tmp.ensure(tb));
    OCT_CK(hipcub::DeviceScan::ExclusiveSum(o.tmp.p, tb, in, nid.as<int>(), (int)m.count(), st));
  }
  int last = 0;
  uint8_t lastf = 0;
  OCT_CK(hipMemcpyAsync(&last, nid.as<int>() + (m.count() - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  OCT_CK(hipMemcpyAsync(&lastf, nflag.as<uint8_t>() + (m.count() - 1), 1, hipMemcpyDeviceToHost, st));
  OCT_CK(hipStreamSynchronize(st));
  const long long nnodes = (long long)last + lastf;
  OCT_CK(dnodes.ensure((size_t)nnodes * 3 * sizeof(int)));
  OCT_CK(dV.ensure((size_t)nnodes * 3 * sizeof(double)));
  OCT_CK(dH.ensure((size_t)nleaves * 8 * sizeof(int)));
  hipLaunchKernelGGL(node_emit_kernel, dim3(blocks(m.count())), dim3(kOctBlock), 0, st, nflag.as<uint8_t>(),
                     nid.as<int>(), m, m.count(), ef, o.base[0], o.base[1], o.base[2], o.vs, dnodes.as<int>(),
                     dV.as<double>());
  hipLaunchKernelGGL(hex_kernel, dim3(blocks(nleaves)), dim3(kOctBlock), 0, st, leaves.as<int4>(), (int)nleaves,
                     nid.as<int>(), m, ef, dH.as<int>());
  OCT_CK(hipGetLastError());
  out.leaves.resize((size_t)nleaves * 4);
  out.nodes.resize((size_t)nnodes * 3);
  out.V.resize((size_t)nnodes * 3);
  out.hexes.resize((size_t)nleaves * 8);
  OCT_CK(hipMemcpyAsync(out.leaves.data(), leaves.p, out.leaves.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  OCT_CK(hipMemcpyAsync(out.nodes.data(), dnodes.p, out.nodes.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  OCT_CK(hipMemcpyAsync(out.V.data(), dV.p, out.V.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  OCT_CK(hipMemcpyAsync(out.hexes.data(), dH.p, out.hexes.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  OCT_CK(hipStreamSynchronize(st));
  return true;
}

// one timed subdivide + export (device time between events on the stream, host copies of the
// results excluded)
bool octree_timed(OctreeDev& o, const SurfaceDev* s, int stop_extent, int n, const int* cells, hipStream_t st,
                  OctreeOut& out, std::string& err) {
  hipEvent_t ev0, ev1;
  OCT_CK(hipEventCreate(&ev0));
  if (hipEventCreate(&ev1) != hipSuccess) {
    (void)hipEventDestroy(ev0);
    err = "octree: hipEventCreate failed";
    return false;
  }
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{ev0, ev1};
  long evaluated = 0;
  OCT_CK(hipEventRecord(ev0, st));
  const bool ok = cells ? octree_refine(o, *s, stop_extent, n, cells, st, evaluated, err)
                        : octree_subdivide(o, *s, stop_extent, st, evaluated, err);
  if (!ok) return false;
  OCT_CK(hipEventRecord(ev1, st));
  out.evaluated = evaluated;
  if (!octree_export(o, st, out, err)) return false;
  float ms = 0;
  OCT_CK(hipEventElapsedTime(&ms, ev0, ev1));
  out.ms = ms;
  return true;
}

}  // namespace hexopt
