// bvh.hip -- closest point on a triangle surface with the pseudonormal of
// igl::signed_distance_pseudonormal (signed_distance.cpp:186-257), FP64.
//
// One lane per query. The traversal replays igl::AABB::squared_distance's recursion
// (AABB.cpp:364-441) exactly -- visit a child whose box contains the point, then the nearer
// child (by exterior box distance) and then the other, each only while it can still beat the
// current best, leaves compared with a strict '<' (set_min, AABB.cpp:735-779) -- so the face
// chosen on ties is the reference's. The recursion becomes an explicit stack of packed
// frames (node | visit state | visited-children bits) that lives in LDS, one column per lane.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "bvh.h"
#include "kernels.h"

namespace hexopt {
namespace {

__device__ __forceinline__ bool box_contains(const double* b, const double p[3]) {
  return b[0] <= p[0] && p[0] <= b[3] && b[1] <= p[1] && p[1] <= b[4] && b[2] <= p[2] && p[2] <= b[5];
}

// AlignedBox::squaredExteriorDistance
__device__ __forceinline__ double box_ext_dist(const double* b, const double p[3]) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (b[k] > p[k]) { double a = b[k] - p[k]; s += a * a; }
    else if (p[k] > b[3 + k]) { double a = p[k] - b[3 + k]; s += a * a; }
  }
  return s;
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// igl's Dot of 1x3 rows as Eigen 3.2.9 unrolls it: the first product plus the sum of the other two
__device__ __forceinline__ double dotE(const double a[3], const double b[3]) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }

// Ericson's closest point on a triangle (point_simplex_squared_distance.cpp:44-109)
__device__ __forceinline__ void closest_on_tri(const double p[3], const double a[3], const double b[3], const double c[3], double r[3]) {
  double ab[3], ac[3], ap[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
  const double d1 = dotE(ab, ap), d2 = dotE(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) { r[0] = a[0]; r[1] = a[1]; r[2] = a[2]; return; }
  double bp[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
  const double d3 = dotE(ab, bp), d4 = dotE(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) { r[0] = b[0]; r[1] = b[1]; r[2] = b[2]; return; }
  const double vc = d1 * d4 - d3 * d2;
  if (!(a[0] == b[0] && a[1] == b[1] && a[2] == b[2])) {
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
      const double v = d1 / (d1 - d3);
#pragma unroll
      for (int k = 0; k < 3; ++k) r[k] = a[k] + v * ab[k];
      return;
    }
  }
  double cp[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
  const double d5 = dotE(ab, cp), d6 = dotE(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) { r[0] = c[0]; r[1] = c[1]; r[2] = c[2]; return; }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = a[k] + w * ac[k];
    return;
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = b[k] + w * (c[k] - b[k]);
    return;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
#pragma unroll
  for (int k = 0; k < 3; ++k) r[k] = a[k] + ab[k] * v + ac[k] * w;
}

__device__ __forceinline__ void load3(const double* V, int i, double o[3]) {
  o[0] = V[(size_t)i * 3]; o[1] = V[(size_t)i * 3 + 1]; o[2] = V[(size_t)i * 3 + 2];
}

// frame packing: node (26 bits) | state (3 bits) << 26 | visited (2 bits) << 29
__device__ __forceinline__ int fr_node(unsigned f) { return (int)(f & 0x3ffffffu); }
__device__ __forceinline__ unsigned fr_state(unsigned f) { return (f >> 26) & 7u; }
__device__ __forceinline__ unsigned fr_seen(unsigned f) { return (f >> 29) & 3u; }
__device__ __forceinline__ unsigned fr_make(int n, unsigned st, unsigned seen) { return (unsigned)n | (st << 26) | (seen << 29); }

// six waves per SIMD (80 registers, 2 values spilled) instead of the five that 82 registers
// leave: more walks in flight to cover the dependent node fetches (the four patch walks of the
// 1.1M-hex torus refresh: 275-279 against 294-295 us; eight waves with 33 spills, 284-287)
__global__ __launch_bounds__(kBvhBlock) __attribute__((amdgpu_waves_per_eu(6))) void closest_kernel(SurfaceDev s0, const PatchSeg* __restrict__ segs, int nseg,
                                                            int nq, const double* P, int soa, int stride,
                                                            const int* ids, const int* order, double* S, int* I,
                                                            double* Cout, double* Nout, double* sqrDout, int* status) {
  // one frame per tree level: (depth + 1) x kBvhBlock entries of dynamic LDS (see
  // launch_closest), so shallow trees leave room for more resident blocks
  extern __shared__ unsigned stack[];
  // with segments, each block belongs to one (the last segment starting at or before it) and
  // walks that segment's tree for that segment's queries; the choice is the same in every lane
  SurfaceDev s = s0;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (nseg > 0) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segs[mid].block0 <= (int)blockIdx.x) lo = mid;
      else hi = mid - 1;
    }
    s = segs[lo].s;
    const int local = ((int)blockIdx.x - segs[lo].block0) * blockDim.x + threadIdx.x;
    if (local >= segs[lo].n) return;
    k = segs[lo].q0 + local;
  } else if (k >= nq) {
    return;
  }
  const int cap = stack_frames(s.depth);
  const int q = order ? order[k] : k;
  const int pi = ids ? ids[q] : q;
  double p[3];
  if (soa) {
    p[0] = P[pi]; p[1] = P[(size_t)stride + pi]; p[2] = P[2 * (size_t)stride + pi];
  } else {
    load3(P, pi, p);
  }
  unsigned* stk = stack + threadIdx.x;  // column per lane: entry k at stk[k*kBvhBlock]
  int top = 0;
  double best = INFINITY;
  int bf = -1;
  double bc[3] = {0, 0, 0};
  stk[0] = fr_make(0, 0, 0);
  top = 1;
  bool overflow = false;
  while (top > 0) {
    unsigned f = stk[(top - 1) * kBvhBlock];
    const int n = fr_node(f);
    // the node's record in one go: children's boxes (inner) or the face's corners (leaf)
    const double2* r2 = reinterpret_cast<const double2*>(s.rec + (size_t)n * kNodeRec);
    double x[12];
#pragma unroll
    for (int j = 0; j < 6; ++j) { const double2 v = r2[j]; x[2 * j] = v.x; x[2 * j + 1] = v.y; }
    const int4 w = reinterpret_cast<const int4*>(r2)[6];
    const int pr = w.z;
    if (pr >= 0) {  // leaf_squared_distance + set_min
      const double a[3] = {x[0], x[1], x[2]}, b[3] = {x[3], x[4], x[5]}, c[3] = {x[6], x[7], x[8]};
      double r[3];
      closest_on_tri(p, a, b, c, r);
      // (p - c).squaredNorm() in Eigen's order
      const double d = (p[0] - r[0]) * (p[0] - r[0]) + ((p[1] - r[1]) * (p[1] - r[1]) + (p[2] - r[2]) * (p[2] - r[2]));
      if (d < best) { best = d; bf = pr; bc[0] = r[0]; bc[1] = r[1]; bc[2] = r[2]; }
      --top;
      continue;
    }
    const int L = w.x, R = w.y;
    const double* bl = x;
    const double* br = x + 6;
    unsigned state = fr_state(f), seen = fr_seen(f);
    int child = -1;
    while (child < 0 && state < 4) {
      if (state == 0) {                         // must look left if in box
        state = 1;
        if (box_contains(bl, p)) { seen |= 1; child = L; }
      } else if (state == 1) {                  // must look right if in box
        state = 2;
        if (box_contains(br, p)) { seen |= 2; child = R; }
      } else {                                  // then nearer first, each only if it can win
        const double lm = box_ext_dist(bl, p), rm = box_ext_dist(br, p);
        const bool left_first = lm < rm;
        const bool first = state == 2;
        state += 1;
        const bool take_left = left_first == first;
        if (take_left) {
          if (!(seen & 1) && lm < best) { seen |= 1; child = L; }
        } else {
          if (!(seen & 2) && rm < best) { seen |= 2; child = R; }
        }
      }
    }
    if (child < 0) { --top; continue; }
    stk[(top - 1) * kBvhBlock] = fr_make(n, state, seen);
    if (top >= cap) { overflow = true; break; }
    stk[top * kBvhBlock] = fr_make(child, 0, 0);
    ++top;
  }
  if (overflow) {
    atomicOr(status, 1);
    return;
  }
  if (bf < 0) {  // no finite distance (non-finite query): report NaN
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    if (S) S[q] = nan;
    if (I) I[q] = -1;
    if (sqrDout) sqrDout[q] = nan;
    if (Cout) { Cout[(size_t)q * 3] = nan; Cout[(size_t)q * 3 + 1] = nan; Cout[(size_t)q * 3 + 2] = nan; }
    if (Nout) { Nout[(size_t)q * 3] = nan; Nout[(size_t)q * 3 + 1] = nan; Nout[(size_t)q * 3 + 2] = nan; }
    return;
  }
  // pseudonormal_test (pseudonormal_test.cpp:14-119)
  double A[3], B[3], Cc[3];
  load3(s.V, s.F[(size_t)bf * 3], A);
  load3(s.V, s.F[(size_t)bf * 3 + 1], B);
  load3(s.V, s.F[(size_t)bf * 3 + 2], Cc);
  double area;
  {  // doublearea: Kahan's Heron formula on the decreasing edge lengths
    double l0 = sqrt((B[0] - Cc[0]) * (B[0] - Cc[0]) + (B[1] - Cc[1]) * (B[1] - Cc[1]) + (B[2] - Cc[2]) * (B[2] - Cc[2]));
    double l1 = sqrt((Cc[0] - A[0]) * (Cc[0] - A[0]) + (Cc[1] - A[1]) * (Cc[1] - A[1]) + (Cc[2] - A[2]) * (Cc[2] - A[2]));
    double l2 = sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]));
    double t;
    if (l1 > l0) { t = l0; l0 = l1; l1 = t; }
    if (l2 > l1) { t = l1; l1 = l2; l2 = t; }
    if (l1 > l0) { t = l0; l0 = l1; l1 = t; }
    double arg = (l0 + (l1 + l2)) * (l2 - (l0 - l1)) * (l2 + (l0 - l1)) * (l0 + (l1 - l2));
    area = 2.0 * 0.25 * sqrt(arg);
  }
  const double eps = 1e-12;
  double nn[3];
  if (area > 1e-4) {
    double v0[3], v1[3], v2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { v0[k] = B[k] - A[k]; v1[k] = Cc[k] - A[k]; v2[k] = bc[k] - A[k]; }
    const double d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    const double denom = d00 * d11 - d01 * d01;
    double bb[3];
    bb[1] = (d11 * d20 - d01 * d21) / denom;
    bb[2] = (d00 * d21 - d01 * d20) / denom;
    bb[0] = 1.0f - (bb[1] + bb[2]);
    const int type = (bb[0] <= eps) + (bb[1] <= eps) + (bb[2] <= eps);
    if (type == 2) {
      int x = bb[0] > eps ? 0 : (bb[1] > eps ? 1 : 2);
      load3(s.VN, s.F[(size_t)bf * 3 + x], nn);
    } else if (type == 1) {
      int x = bb[0] <= eps ? 0 : (bb[1] <= eps ? 1 : 2);
      nn[0] = s.FEN[(size_t)bf * 9 + x * 3]; nn[1] = s.FEN[(size_t)bf * 9 + x * 3 + 1]; nn[2] = s.FEN[(size_t)bf * 9 + x * 3 + 2];
    } else {
      load3(s.FN, bf, nn);
    }
  } else {
    bool found = false;
    const double* cv[3] = {A, B, Cc};
    // loops unrolled so that the corner picks are compile-time (no scratch for A, B, Cc)
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      if (found) continue;
      const double* w = cv[v];
      double dd = sqrt((bc[0] - w[0]) * (bc[0] - w[0]) + (bc[1] - w[1]) * (bc[1] - w[1]) + (bc[2] - w[2]) * (bc[2] - w[2]));
      if (dd < eps) { found = true; load3(s.VN, s.F[(size_t)bf * 3 + v], nn); }
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      if (found) continue;
      const double* s0 = cv[(e + 1) % 3];
      const double* s1 = cv[(e + 2) % 3];
      double dv[3] = {s1[0] - s0[0], s1[1] - s0[1], s1[2] - s0[2]};
      double pv[3] = {bc[0] - s0[0], bc[1] - s0[1], bc[2] - s0[2]};
      double t = dot3(dv, pv) / dot3(dv, dv);
      double q2;
      if (t < 0) {
        q2 = dot3(pv, pv);
      } else if (t > 1) {
        double w0 = bc[0] - s1[0], w1 = bc[1] - s1[1], w2 = bc[2] - s1[2];
        q2 = w0 * w0 + w1 * w1 + w2 * w2;
      } else {
        double w0 = s0[0] + t * dv[0] - bc[0], w1 = s0[1] + t * dv[1] - bc[1], w2 = s0[2] + t * dv[2] - bc[2];
        q2 = w0 * w0 + w1 * w1 + w2 * w2;
      }
      if (sqrt(q2) < eps) {
        nn[0] = s.FEN[(size_t)bf * 9 + e * 3]; nn[1] = s.FEN[(size_t)bf * 9 + e * 3 + 1]; nn[2] = s.FEN[(size_t)bf * 9 + e * 3 + 2];
        found = true;
      }
    }
    if (!found) load3(s.FN, bf, nn);
  }
  const double sgn = ((p[0] - bc[0]) * nn[0] + (p[1] - bc[1]) * nn[1] + (p[2] - bc[2]) * nn[2]) >= 0 ? 1. : -1.;
  if (S) S[q] = sgn * sqrt(best);
  if (I) I[q] = bf;
  if (sqrDout) sqrDout[q] = best;
  if (Cout) { Cout[(size_t)q * 3] = bc[0]; Cout[(size_t)q * 3 + 1] = bc[1]; Cout[(size_t)q * 3 + 2] = bc[2]; }
  if (Nout) { Nout[(size_t)q * 3] = nn[0]; Nout[(size_t)q * 3 + 1] = nn[1]; Nout[(size_t)q * 3 + 2] = nn[2]; }
}

// S = S0 + lamT n n^T, b = b0 + lamT n d with d = n . C, for the plane rows (one row per
// plane vertex; rows of other constraint kinds stay S0/b0). One thread per row, in row
// (vertex) order: trow[r] is the plane index of row r, or -1.
__global__ void refold_planes_kernel(const int* trow, int nrows, const double* Cp, const double* t_n, double* t_d,
                                     const double* S0, const double* b0, double* S, double* b, double lam) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int i = trow[r];
  if (i < 0) return;
  const double n0 = t_n[(size_t)i * 3], n1 = t_n[(size_t)i * 3 + 1], n2 = t_n[(size_t)i * 3 + 2];
  const double d = n0 * Cp[(size_t)i * 3] + n1 * Cp[(size_t)i * 3 + 1] + n2 * Cp[(size_t)i * 3 + 2];
  t_d[i] = d;
  const double* s0 = S0 + (size_t)r * 6;
  const double* c0 = b0 + (size_t)r * 3;
  double* s1 = S + (size_t)r * 6;
  double* c1 = b + (size_t)r * 3;
  s1[0] = s0[0] + lam * n0 * n0; s1[1] = s0[1] + lam * n1 * n1; s1[2] = s0[2] + lam * n2 * n2;
  s1[3] = s0[3] + lam * n0 * n1; s1[4] = s0[4] + lam * n0 * n2; s1[5] = s0[5] + lam * n1 * n2;
  c1[0] = c0[0] + lam * n0 * d; c1[1] = c0[1] + lam * n1 * d; c1[2] = c0[2] + lam * n2 * d;
}

}  // namespace

void launch_refold_planes(const SoftDev& sd, const int* trow, int nrows, const double* Cp, double* t_d,
                          const double* S0, const double* b0, double* S, double* b, double lam, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(refold_planes_kernel, dim3((nrows + 255) / 256), dim3(256), 0, st, trow, nrows, Cp, sd.t_n, t_d,
                     S0, b0, S, b, lam);
}

// Vertex-sampled metro statistics of per-query squared distances: a query counts when its
// distance is below `upper`; per block (max, sum, count), then one block folds the blocks
// in order.
// vcg PSDist (vcg/space/deprecated_point3.h:553-565): distance from p to the segment v1 v2
__device__ __forceinline__ double ps_dist(const double p[3], const double v1[3], const double v2[3]) {
  const double e[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  double t = ((p[0] - v1[0]) * e[0] + (p[1] - v1[1]) * e[1] + (p[2] - v1[2]) * e[2]) /
             (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  if (t < 0) t = 0;
  else if (t > 1) t = 1;
  const double q[3] = {v1[0] + e[0] * t, v1[1] + e[1] * t, v1[2] + e[2] * t};
  const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// vcg face::PointDistanceEP (vcg/simplex/face/distance.h:48-180) for the face with corners
// a, b, c and edge-plane record r (build_ep_faces): lowers dist to the face's distance when
// that is smaller. The corner differences and the inside tests keep the reference's operand
// order; the three axis cases differ only in which two coordinates the edge tests use.
__device__ __forceinline__ void point_distance_ep(const double* __restrict__ r, const double a[3], const double b[3],
                                                  const double c[3], const double q[3], double& dist) {
  const double* E0 = r;
  const double* E1 = r + 3;
  const double* E2 = r + 6;
  const double* n = r + 9;
  const double d = (n[0] * q[0] + n[1] * q[1] + n[2] * q[2]) - r[12];
  if (d > dist || d < -dist) return;
  const double p[3] = {q[0] - n[0] * d, q[1] - n[1] * d, q[2] - n[2] * d};
  const int axis = (int)r[13];
  // (u, v): the two coordinates of the projection, as in the NORMX / NORMY / NORMZ cases
  const int u = axis == 0 ? 1 : axis == 1 ? 2 : 0;
  const int v = axis == 0 ? 2 : axis == 1 ? 0 : 1;
  const double b0 = E1[u] * (p[v] - b[v]) - E1[v] * (p[u] - b[u]);
  if (b0 <= 0) {
    const double bt = ps_dist(q, b, c);
    if (dist > bt) dist = bt;
    return;
  }
  const double b1 = E2[u] * (p[v] - c[v]) - E2[v] * (p[u] - c[u]);
  if (b1 <= 0) {
    const double bt = ps_dist(q, c, a);
    if (dist > bt) dist = bt;
    return;
  }
  const double b2 = E0[u] * (p[v] - a[v]) - E0[v] * (p[u] - a[u]);
  if (b2 <= 0) {
    const double bt = ps_dist(q, a, b);
    if (dist > bt) dist = bt;
    return;
  }
  const double bm = b0 < b1 ? (b0 < b2 ? b0 : b2) : (b1 < b2 ? b1 : b2);
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double x[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  const double area2 = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (bm < 0.000001 * area2) {
    const double bt = bm == b0 ? ps_dist(q, b, c) : bm == b1 ? ps_dist(q, c, a) : ps_dist(q, a, b);
    if (dist > bt) dist = bt;
    return;
  }
  dist = fabs(d);
}

// min(upper, min over faces of point_distance_ep) through the tree: nearer child first, a
// child skipped when its box lies further than the current distance plus `margin` (a bound
// far above the rounding of the face distances, so no face that could lower it is skipped)
__global__ __launch_bounds__(kBvhBlock) void metro_ep_kernel(SurfaceDev s, const double* __restrict__ ep, int nq,
                                                             const double* __restrict__ P, double upper,
                                                             double margin, double* __restrict__ dist, int* status) {
  extern __shared__ int mstack[];
  const int cap = stack_frames(s.depth);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  double q[3];
  load3(P, k, q);
  double best = upper;
  int* stk = mstack + threadIdx.x;
  int top = 0;
  stk[0] = 0;
  top = 1;
  bool overflow = false;
  while (top > 0) {
    const int n = stk[(--top) * kBvhBlock];
    const double2* r2 = reinterpret_cast<const double2*>(s.rec + (size_t)n * kNodeRec);
    double x[12];
#pragma unroll
    for (int j = 0; j < 6; ++j) { const double2 w = r2[j]; x[2 * j] = w.x; x[2 * j + 1] = w.y; }
    const int4 w = reinterpret_cast<const int4*>(r2)[6];
    if (w.z >= 0) {
      point_distance_ep(ep + (size_t)w.z * kEPRec, x, x + 3, x + 6, q, best);
      continue;
    }
    const double lim = (best + margin) * (best + margin);
    const double dl = box_ext_dist(x, q), dr = box_ext_dist(x + 6, q);
    const int nearc = dl <= dr ? w.x : w.y, farc = dl <= dr ? w.y : w.x;
    const double dn = dl <= dr ? dl : dr, df = dl <= dr ? dr : dl;
    if (top + 2 > cap) { overflow = true; break; }
    if (farc >= 0 && df <= lim) stk[(top++) * kBvhBlock] = farc;
    if (nearc >= 0 && dn <= lim) stk[(top++) * kBvhBlock] = nearc;
  }
  if (overflow) atomicOr(status, 1);
  dist[k] = best;
}

void launch_metro_ep(const SurfaceDev& s, const double* ep, int nq, const double* P, double upper, double* dist,
                     int* status, hipStream_t st) {
  if (nq <= 0) return;
  const size_t lds = (size_t)stack_frames(s.depth) * kBvhBlock * sizeof(int);
  hipLaunchKernelGGL(metro_ep_kernel, dim3((nq + kBvhBlock - 1) / kBvhBlock), dim3(kBvhBlock), lds, st, s, ep, nq, P,
                     upper, 1e-9 * upper, dist, status);
}


void launch_closest(const SurfaceDev& s, int nq, const double* P, int soa, int stride, const int* ids, double* S, int* I,
                    double* C, double* N, double* sqrD, int* status, hipStream_t st, const int* order) {
  if (nq <= 0) return;
  const size_t lds = (size_t)stack_frames(s.depth) * kBvhBlock * sizeof(unsigned);
  hipLaunchKernelGGL(closest_kernel, dim3((nq + kBvhBlock - 1) / kBvhBlock), dim3(kBvhBlock), lds, st, s,
                     (const PatchSeg*)nullptr, 0, nq, P, soa, stride, ids, order, S, I, C, N, sqrD, status);
}

void launch_closest_segs(const PatchSeg* segs, int nseg, int nblocks, int max_depth, const double* P, int soa,
                         int stride, const int* ids, const int* order, double* C, double* N, int* status,
                         hipStream_t st) {
  if (nseg <= 0 || nblocks <= 0) return;
  const size_t lds = (size_t)stack_frames(max_depth) * kBvhBlock * sizeof(unsigned);
  SurfaceDev none{};
  hipLaunchKernelGGL(closest_kernel, dim3(nblocks), dim3(kBvhBlock), lds, st, none, segs, nseg, 0, P, soa, stride,
                     ids, order, (double*)nullptr, (int*)nullptr, C, N, (double*)nullptr, status);
}

}  // namespace hexopt
