// lines.hip -- feature-curve projection of LINE vertices, FP64: the LINE branch of
// dirty_graph_projection (grid_hex_meshing.cpp:3961-4026). One wave per vertex projects it
// onto the segments of its curve (point_line_projection, global_functions.cpp:3454-3465) and
// keeps the nearest, the lowest segment index on equal distances, which is the first entry
// after the reference's sort of (distance, index) pairs. Built without fused multiply-add;
// sums of three products add x + (y + z), the order of Eigen 3.2.9's unrolled reduction of a
// 3-vector (the first coefficient, then the sum of the other two), which the reference's
// dot() and squaredNorm() compile to.
#include <hip/hip_runtime.h>
#include "lines.h"
#include "kernels.h"

namespace hexopt {
namespace {

__device__ __forceinline__ double norm3(double x, double y, double z) { return sqrt(x * x + (y * y + z * z)); }

// (a - b).normalized(): Eigen 3.2.9 divides by the norm without a zero test
__device__ __forceinline__ void normalized_diff(const double* a, const double* b, double o[3]) {
  const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  const double n = norm3(d0, d1, d2);
  o[0] = d0 / n;
  o[1] = d1 / n;
  o[2] = d2 / n;
}

// point_line_projection of v onto segment p0 p1: the closest point and its distance
__device__ __forceinline__ double segment_point(const double v[3], const double* p0, const double* p1, double q3[3]) {
  const double a0 = v[0] - p0[0], a1 = v[1] - p0[1], a2 = v[2] - p0[2];
  const double e0 = p1[0] - p0[0], e1 = p1[1] - p0[1], e2 = p1[2] - p0[2];
  const double l2 = e0 * e0 + (e1 * e1 + e2 * e2);
  double t = 0;
  if (l2 >= 1.e-7) t = (a0 * e0 + (a1 * e1 + a2 * e2)) / l2;
  if (t >= 0.0 && t <= 1.0) {
    for (int d = 0; d < 3; ++d) q3[d] = p0[d] + t * (p1[d] - p0[d]);
  } else if (t < 0.0) {
    for (int d = 0; d < 3; ++d) q3[d] = p0[d];
  } else {
    for (int d = 0; d < 3; ++d) q3[d] = p1[d];
  }
  return norm3(v[0] - q3[0], v[1] - q3[1], v[2] - q3[2]);
}

// One wave per LINE vertex: lanes take every 64th segment, then the wave keeps the nearest
// (ties to the lower segment index, the first entry of the reference's sorted
// (distance, index) list; a NaN distance at segment 0 wins as in a sequential strict-< scan).
constexpr int kLineWaves = 4;
__global__ __launch_bounds__(64 * kLineWaves) void project_lines_kernel(CurvesDev cv, const double* __restrict__ X,
                                                                        int nV, int n, const int* __restrict__ rows,
                                                                        const int* __restrict__ curve_of,
                                                                        double* __restrict__ origin,
                                                                        double* __restrict__ axis) {
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.x * kLineWaves + (threadIdx.x >> 6);
  if (q >= n) return;
  const int iv = rows[q];
  const double v[3] = {X[iv], X[(size_t)nV + iv], X[2 * (size_t)nV + iv]};
  const int c = curve_of[q];
  const int* curve = cv.vs + cv.off[c];
  const int csize = cv.off[c + 1] - cv.off[c];
  const bool circle = cv.circles[c] != 0;
  const int curve_len = circle ? csize : csize - 1;
  double pv[3] = {0, 0, 0}, tangent[3] = {1, 0, 0};
  if (curve_len > 0) {
    // sequential scan (best = d[0], replaced when d[j] < best): a NaN at segment 0 is kept;
    // otherwise the result is the first minimum over the segments with a non-NaN distance
    double best = INFINITY;
    int bj = -1, nan0 = 0;
    for (int j = lane; j < curve_len; j += 64) {
      double q3[3];
      const double dist = segment_point(v, cv.V + (size_t)curve[j] * 3, cv.V + (size_t)curve[(j + 1) % csize] * 3, q3);
      if (dist != dist) {
        nan0 |= j == 0;
        continue;
      }
      if (bj < 0 || dist < best) { best = dist; bj = j; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oj = __shfl_xor(bj, o, 64);
      nan0 |= __shfl_xor(nan0, o, 64);
      if (oj >= 0 && (bj < 0 || ob < best || (ob == best && oj < bj))) { best = ob; bj = oj; }
    }
    if (lane != 0) return;
    if (nan0) bj = 0;
    const double* p0 = cv.V + (size_t)curve[bj] * 3;
    const double* p1 = cv.V + (size_t)curve[(bj + 1) % csize] * 3;
    segment_point(v, p0, p1, pv);
    normalized_diff(p1, p0, tangent);
  } else {
    // no segment (an open curve of one vertex): nearest curve vertex, tangent from its
    // neighbours added to (1, 0, 0), then normalised
    if (lane != 0) return;
    int pos = 0;
    double best = 0;
    for (int j = 0; j < csize; ++j) {
      const double* w = cv.V + (size_t)curve[j] * 3;
      const double dist = norm3(w[0] - v[0], w[1] - v[1], w[2] - v[2]);
      if (j == 0 || dist < best) { best = dist; pos = j; }
    }
    const double* pp = cv.V + (size_t)curve[pos] * 3;
    for (int d = 0; d < 3; ++d) pv[d] = pp[d];
    double e[3];
    if (circle || (pos != 0 && pos != csize - 1)) {
      normalized_diff(pp, cv.V + (size_t)curve[(pos - 1 + csize) % csize] * 3, e);
      for (int d = 0; d < 3; ++d) tangent[d] += e[d];
      normalized_diff(cv.V + (size_t)curve[(pos + 1) % csize] * 3, pp, e);
      for (int d = 0; d < 3; ++d) tangent[d] += e[d];
    } else if (pos == 0) {
      normalized_diff(cv.V + (size_t)curve[(pos + 1) % csize] * 3, pp, e);
      for (int d = 0; d < 3; ++d) tangent[d] += e[d];
    } else {
      normalized_diff(pp, cv.V + (size_t)curve[(pos - 1 + csize) % csize] * 3, e);
      for (int d = 0; d < 3; ++d) tangent[d] += e[d];
    }
    const double nt = norm3(tangent[0], tangent[1], tangent[2]);
    for (int d = 0; d < 3; ++d) tangent[d] /= nt;
  }
  for (int d = 0; d < 3; ++d) {
    origin[(size_t)q * 3 + d] = pv[d];
    axis[(size_t)q * 3 + d] = tangent[d];
  }
}

// Line rows of the folded constraint form after a projection (set_constraints' folding with
// the new origin o, axis a and weight lamL; a line vertex carries no corner or plane term):
// S = S0 = lamL I, b = b0 = lamL o, and the a_L row has rhs lamL sum_d(-a_d o_d) and
// inverse diagonal 1 / (lamL |a|^2).
__global__ void refold_lines_kernel(const int* __restrict__ lv, const int* __restrict__ cidx, int nL,
                                    const double* __restrict__ origin, const double* __restrict__ axis, double lamL,
                                    double* lt, double* lo, double* S0, double* S, double* b0, double* b,
                                    double* rhs_a, double* invd_a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nL) return;
  const int r = cidx[lv[i]];
  for (int k = 0; k < 6; ++k) {
    const double v = k < 3 ? 0.0 + lamL : 0.0;
    S0[(size_t)r * 6 + k] = v;
    S[(size_t)r * 6 + k] = v;
  }
  double s = 0, tt = 0;
  for (int d = 0; d < 3; ++d) {
    const double o = origin[(size_t)i * 3 + d], a = axis[(size_t)i * 3 + d];
    lt[(size_t)i * 3 + d] = a;
    lo[(size_t)i * 3 + d] = o;
    const double bo = 0.0 + lamL * o;
    b0[(size_t)r * 3 + d] = bo;
    b[(size_t)r * 3 + d] = bo;
    s += -a * o;
    tt += a * a;
  }
  rhs_a[i] = lamL * s;
  const double dg = lamL * tt;
  invd_a[i] = dg != 0 ? 1.0 / dg : 1.0;
}

}  // namespace

void launch_refold_lines(const SoftDev& sd, const double* origin, const double* axis, double* S0, double* b0,
                         double* rhs_a, double* invd_a, hipStream_t st) {
  if (sd.nL <= 0) return;
  hipLaunchKernelGGL(refold_lines_kernel, dim3((sd.nL + 127) / 128), dim3(128), 0, st, sd.lv, sd.cidx, sd.nL, origin,
                     axis, sd.lamL, const_cast<double*>(sd.lt), const_cast<double*>(sd.lo), S0,
                     const_cast<double*>(sd.S), b0, const_cast<double*>(sd.b), rhs_a, invd_a);
}

void launch_project_lines(const CurvesDev& cv, const double* X, int nV, int n, const int* rows, const int* curve_of,
                          double* origin, double* axis, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(project_lines_kernel, dim3((n + kLineWaves - 1) / kLineWaves), dim3(64 * kLineWaves), 0, st, cv, X,
                     nV, n, rows, curve_of, origin, axis);
}

}  // namespace hexopt
