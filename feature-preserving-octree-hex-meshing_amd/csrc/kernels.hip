// kernels.hip -- HIP kernels of the hex-mesh SLIM optimiser for gfx950 (MI355X).
//
// One workgroup of 256 lanes (4 wave64s) owns one tile of up to 256 Morton-ordered hexes:
// lane i computes the 8 corner tets of hex i in registers, the per-corner results are
// accumulated into an LDS vertex buffer colour by colour (no atomics), then every tile
// vertex is either written straight to HBM (all its hexes are in this tile) or parked in a
// halo slot that the fixup kernel folds in a fixed order. Reductions are two-level and in a
// fixed order, so every kernel is bitwise deterministic.
//
// Data layout (structure of arrays, all arrays indexed in the Morton numbering):
//   vertex fields   x[d*nV + v]           d = 0,1,2 (the reference's flattened unknowns)
//   connectivity    conn[k*nH + h]         k = hex corner in the reference's Hybrid::vs order
//   per-tet scalars f[(c*F + j)*nH + h]    c = corner tet, j = field component
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <type_traits>
#include "device_math.h"
#include "kernels.h"

namespace hexopt {

// ---------------------------------------------------------------------------------------
// helpers

// Maps a workgroup id to a tile so that consecutive tiles (neighbours along the Morton
// curve, which share halo vertices) run on the same XCD and hit the same L2: hardware
// hands workgroup b to XCD b % 8.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n / 8, r = n % 8;
  const int x = b % 8, k = b / 8;
  return x * q + (x < r ? x : r) + k;
}

template <class T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  // wave64 butterfly, then the 4 wave sums in fixed order
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// N sums at once, each as block_sum computes it, behind one pair of barriers; red holds
// N * 16 values
template <int N>
__device__ __forceinline__ void block_sum_n(double v[N], double* red) {
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) red[k * 16 + w] = v[k];
  }
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double t = 0;
    for (int i = 0; i < nw; ++i) t += red[k * 16 + i];
    v[k] = t;
  }
}

__device__ __forceinline__ double block_min(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 1; i < nw; ++i) s = fmin(s, red[i]);
  return s;
}

// Reference operator of corner tet c: J = [x_n1-x_c, x_n2-x_c, x_n3-x_c] * B.
template <class T>
__device__ __forceinline__ void tet_B(const RefOp& ro, int nH, int h, int c, T B[3][3]) {
  if (ro.mode == REF_CUBE) {
    const T s = (T)ro.invs[(size_t)c * nH + h];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) B[i][j] = s * (T)kPi[c][i][j];
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) B[i][j] = (T)ro.B[((size_t)c * 9 + i * 3 + j) * nH + h];
  }
}

// Rows of the per-hex corner arrays are selected, not indexed: a corner-table entry is the
// same in every lane, but a run-time index into a register array would move the array to
// scratch memory. With a constant index (unrolled corner loops) the selects fold away.
template <class T, int N>
__device__ __forceinline__ void pick_row(const T a[8][N], int i, T r[N]) {
#pragma unroll
  for (int d = 0; d < N; ++d) r[d] = a[0][d];
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    const bool m = i == k;
#pragma unroll
    for (int d = 0; d < N; ++d) r[d] = m ? a[k][d] : r[d];
  }
}

// a[i][OFF + d] += v[d] for d < M, by selects (adds 0 to the other rows)
template <int OFF, int M, class T, int N>
__device__ __forceinline__ void add_row(T a[8][N], int i, const T v[M]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool m = i == k;
#pragma unroll
    for (int d = 0; d < M; ++d) a[k][OFF + d] += m ? v[d] : T(0);
  }
}

template <class T>
__device__ __forceinline__ void tet_edges(const T p[8][3], int c, T E[3][3]) {
  T a[3];
  pick_row<T, 3>(p, kHexTet[c][0], a);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T b[3];
    pick_row<T, 3>(p, kHexTet[c][k + 1], b);
#pragma unroll
    for (int d = 0; d < 3; ++d) E[d][k] = b[d] - a[d];
  }
}

template <class T>
__device__ __forceinline__ void matmul3(const T A[3][3], const T B[3][3], T C[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

// Adds the vertex forces of one tet from F = dE/dJ-like 3x3 (columns = reference axes):
// force on edge k = (F B^T)[:,k]; + to n_k, - to the corner.
template <class T>
__device__ __forceinline__ void scatter_tet(const T F[3][3], const T B[3][3], int c, T out[8][3]) {
  const int v0 = kHexTet[c][0];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T f[3], mf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      f[d] = F[d][0] * B[k][0] + F[d][1] * B[k][1] + F[d][2] * B[k][2];
      mf[d] = -f[d];
    }
    add_row<0, 3>(out, kHexTet[c][k + 1], f);
    add_row<0, 3>(out, v0, mf);
  }
}

template <class P>
__device__ __forceinline__ void gather8(const P* x, int nV, const int* conn, int nH, int h, int vid[8], double p[8][3]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) vid[k] = conn[(size_t)k * nH + h];
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) p[k][d] = (double)x[(size_t)d * nV + vid[k]];
}

// Operator blocks K: 48 values per hex (the 6 symmetric components of each corner tet,
// tet-major), grouped in 16-byte vectors so that one load brings 4 floats or 2 doubles:
// value m of hex h sits at element (m / W * nH + h) * W + m % W, W = 16 / sizeof(T).
template <class T>
__device__ __forceinline__ size_t k_at(int m, int h, int nH) {
  constexpr int W = 16 / (int)sizeof(T);
  return ((size_t)(m / W) * nH + h) * W + m % W;
}

// the 12 values of corner tets 2q and 2q+1
template <class CT>
__device__ __forceinline__ void k_pair(const float* K, int nH, int h, int q, CT k[12]) {
  const float4* K4 = reinterpret_cast<const float4*>(K);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float4 v = K4[(size_t)(3 * q + i) * nH + h];
    k[4 * i] = (CT)v.x; k[4 * i + 1] = (CT)v.y; k[4 * i + 2] = (CT)v.z; k[4 * i + 3] = (CT)v.w;
  }
}
template <class CT>
__device__ __forceinline__ void k_pair(const double* K, int nH, int h, int q, CT k[12]) {
  const double2* K2 = reinterpret_cast<const double2*>(K);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double2 v = K2[(size_t)(6 * q + i) * nH + h];
    k[2 * i] = (CT)v.x; k[2 * i + 1] = (CT)v.y;
  }
}

// Adds the tile's per-hex corner values c[8][NC] into its LDS accumulator (NC planes of
// kMaxLV), without atomics and in a fixed order per vertex, so the sums are deterministic.
// Tiles where no vertex sits at the same corner index of two hexes (tile_corner, true for
// every tile of a consistently oriented grid) add corner k of all hexes in phase k: 8
// phases with every thread busy. Other tiles add hex colour by colour. The caller has made
// the accumulator's initial values visible (a barrier); the last phase ends in one.
template <class T, int NC>
__device__ __forceinline__ void scatter_tile(const TileMesh& tm, int tile, int nh, bool active, const int li[8],
                                             const T c[8][NC], T* acc) {
  const int tid = threadIdx.x;
  if (tm.tile_corner[tile]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (active) {
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[j * kMaxLV + li[k]] += c[k][j];
      }
      __syncthreads();
    }
    return;
  }
  const uint16_t* off = tm.tile_col + (size_t)tile * (kMaxCol + 1);
  for (int col = 0; col < kMaxCol; ++col) {
    const int a = off[col], b = off[col + 1];
    if (a >= nh) break;
    if (active && tid >= a && tid < b) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[j * kMaxLV + li[k]] += c[k][j];
    }
    __syncthreads();
  }
}

// scatter_tile for 3 values per vertex kept as float4 in LDS (one 16-byte read-modify-write
// per corner)
__device__ __forceinline__ void scatter_tile4(const TileMesh& tm, int tile, int nh, bool active, const int li[8],
                                              const float c[8][3], float4* acc) {
  const int tid = threadIdx.x;
  if (tm.tile_corner[tile]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (active) {
        float4 v = acc[li[k]];
        v.x += c[k][0]; v.y += c[k][1]; v.z += c[k][2];
        acc[li[k]] = v;
      }
      __syncthreads();
    }
    return;
  }
  const uint16_t* off = tm.tile_col + (size_t)tile * (kMaxCol + 1);
  for (int col = 0; col < kMaxCol; ++col) {
    const int a = off[col], b = off[col + 1];
    if (a >= nh) break;
    if (active && tid >= a && tid < b) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        float4 v = acc[li[k]];
        v.x += c[k][0]; v.y += c[k][1]; v.z += c[k][2];
        acc[li[k]] = v;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// Tiled scatter skeleton. Op provides:
//   __device__ double hex(int h, T out[8][NC])        per-hex corner contributions; returns
//                                                       the hex's share of an optional sum
//   __device__ double finish(int v, T acc[NC])        writes vertex v; returns its share of
//                                                       the kernel's optional vertex sum
// minimum waves per SIMD the register allocation of a tile pass must leave room for (1: no
// constraint)
template <class Op>
struct TileWaves {
  static constexpr int value = 1;
};

template <class T, int NC, class Op>
__global__ __launch_bounds__(kTileHex) __attribute__((amdgpu_waves_per_eu(TileWaves<Op>::value)))
void tile_scatter_kernel(TileMesh tm, Op op, T* halo, double* partial, double* hex_partial) {
  __shared__ T acc[NC * kMaxLV];
  __shared__ double red[16];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int h0 = tm.tile_hex[tile], nh = tm.tile_hex[tile + 1] - h0;
  const int lv0 = tm.tile_lv[tile], nlv = tm.tile_lv[tile + 1] - lv0;
  for (int i = tid; i < NC * kMaxLV; i += kTileHex) acc[i] = T(0);
  T c[8][NC];
  int li[8];
  const bool active = tid < nh;
  double hv = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int j = 0; j < NC; ++j) c[k][j] = T(0);
  if (active) {
    hv = op.hex(h0 + tid, c);
#pragma unroll
    for (int k = 0; k < 8; ++k) li[k] = tm.lidx[(size_t)k * tm.nH + h0 + tid];
  }
  if (hex_partial) {
    double sh = block_sum(hv, red);
    if (tid == 0) hex_partial[tile] = sh;
  }
  __syncthreads();
  scatter_tile<T, NC>(tm, tile, nh, active, li, c, acc);
  double part = 0.0;
  for (int lv = tid; lv < nlv; lv += kTileHex) {
    const int g = tm.lv_gid[lv0 + lv], s = tm.lv_slot[lv0 + lv];
    T v[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) v[j] = acc[j * kMaxLV + lv];
    if (s < 0) {
      part += op.finish(g, v);
    } else {
#pragma unroll
      for (int j = 0; j < NC; ++j) halo[(size_t)j * tm.nSlots + s] = v[j];
    }
  }
  if (partial) {
    double s = block_sum(part, red);
    if (tid == 0) partial[tile] = s;
  }
}

template <class T, int NC, class Op>
__global__ __launch_bounds__(256) void fixup_kernel(TileMesh tm, Op op, const T* halo, double* partial) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double part = 0.0;
  if (i < tm.nShared) {
    const int g = tm.sv_gid[i];
    T v[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) v[j] = T(0);
    for (int k = tm.sv_off[i]; k < tm.sv_off[i + 1]; ++k) {
      const int s = tm.sv_slots[k];
#pragma unroll
      for (int j = 0; j < NC; ++j) v[j] += halo[(size_t)j * tm.nSlots + s];
    }
    part = op.finish(g, v);
  }
  if (partial) {
    double s = block_sum(part, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

template <class T, int NC, class Op>
void launch_scatter(const TileMesh& tm, const Op& op, T* halo, double* part_tile, double* part_fix, hipStream_t st,
                    double* hex_part = nullptr) {
  if (tm.nTiles > 0)
    hipLaunchKernelGGL((tile_scatter_kernel<T, NC, Op>), dim3(tm.nTiles), dim3(kTileHex), 0, st, tm, op, halo, part_tile,
                       hex_part);
  if (tm.nShared > 0)
    hipLaunchKernelGGL((fixup_kernel<T, NC, Op>), dim3((tm.nShared + 255) / 256), dim3(256), 0, st, tm, op,
                       (const T*)halo, part_fix);
}

// ---------------------------------------------------------------------------------------
// Reference operator: compute_referenceMesh + grad_tet_ref (uniform) per corner tet.

// hex2tet24 (global_functions.cpp:3994-4022)
__device__ double hex_volume24(const double p[8][3]) {
  const int F[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {0, 3, 7, 4}, {3, 2, 6, 7}, {1, 5, 6, 2}};
  double hc[3] = {0, 0, 0};
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) hc[d] += p[i][d];
  for (int d = 0; d < 3; ++d) hc[d] /= 8;
  double vol = 0;
  for (int i = 0; i < 6; ++i) {
    double fc[3] = {0, 0, 0};
    for (int j = 0; j < 4; ++j)
      for (int d = 0; d < 3; ++d) fc[d] += p[F[i][j]][d];
    for (int d = 0; d < 3; ++d) fc[d] /= 4;
    for (int j = 0; j < 4; ++j) {
      double M[3][3];
      for (int d = 0; d < 3; ++d) {
        M[d][0] = p[F[i][j]][d] - hc[d];
        M[d][1] = p[F[i][(j + 1) % 4]][d] - hc[d];
        M[d][2] = fc[d] - hc[d];
      }
      vol += fabs(det3(M) / 6);
    }
  }
  return vol;
}

// signed tet volume (igl::volume convention) of corner tet c
__device__ __forceinline__ double corner_volume(const double p[8][3], int c) {
  double a[3], b[3], cc[3], d[3];
  pick_row<double, 3>(p, kHexTet[c][0], a);
  pick_row<double, 3>(p, kHexTet[c][1], b);
  pick_row<double, 3>(p, kHexTet[c][2], cc);
  pick_row<double, 3>(p, kHexTet[c][3], d);
  double ad[3] = {a[0] - d[0], a[1] - d[1], a[2] - d[2]};
  double bd[3] = {b[0] - d[0], b[1] - d[1], b[2] - d[2]};
  double cd[3] = {cc[0] - d[0], cc[1] - d[1], cc[2] - d[2]};
  double cr[3] = {bd[1] * cd[2] - bd[2] * cd[1], bd[2] * cd[0] - bd[0] * cd[2], bd[0] * cd[1] - bd[1] * cd[0]};
  return -(ad[0] * cr[0] + ad[1] * cr[1] + ad[2] * cr[2]) / 6.0;
}

// grad_tet_ref (grad.cpp:372-478) for a general reference tet r: returns B (J = E B).
__device__ void general_B(const double x[4][3], const double r[4][3], double vol, double B[3][3]) {
  (void)x;
  double e01[3], e02[3], e03[3];
  for (int d = 0; d < 3; ++d) { e01[d] = r[0][d] - r[1][d]; e02[d] = r[0][d] - r[2][d]; e03[d] = r[0][d] - r[3][d]; }
  double cr[3] = {e01[1] * e02[2] - e01[2] * e02[1], e01[2] * e02[0] - e01[0] * e02[2], e01[0] * e02[1] - e01[1] * e02[0]};
  double vref = fabs(cr[0] * e03[0] + cr[1] * e03[1] + cr[2] * e03[2]) / 6;
  double scale = cbrt(vol / vref);
  const int faces[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
  const int opp[4] = {3, 1, 2, 0};
  double G[3][4];
  for (int k = 0; k < 4; ++k) {
    double a[3], b[3];
    for (int d = 0; d < 3; ++d) {
      a[d] = (r[faces[k][0]][d] - r[faces[k][1]][d]) * scale;
      b[d] = (r[faces[k][0]][d] - r[faces[k][2]][d]) * scale;
    }
    double c1[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double area = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]) / 2.0;
    double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]), nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    for (int d = 0; d < 3; ++d) { a[d] /= na; b[d] /= nb; }
    double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double w = area / (3.0 * vol);
    for (int d = 0; d < 3; ++d) G[d][opp[k]] = w * (n[d] / nn);
  }
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < 3; ++c) B[k][c] = G[c][k + 1];
}

// Pass 1: per hex, cube-reference operator if the hex qualifies (compute_referenceMesh with
// cube=true, |vol| > 1e-20, H_flag true); counts the hexes that need the general operator.
__global__ __launch_bounds__(256) void ref_cube_kernel(const double* X, int nV, const int* conn, int nH, const uint8_t* hflag, int cube,
                                double* invs, int* n_general) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nH) return;
  int vid[8];
  double p[8][3];
  gather8(X, nV, conn, nH, h, vid, p);
  bool ok = cube && (hflag == nullptr || hflag[h] != 0);
  if (ok) {
    double vol = hex_volume24(p);
    ok = fabs(vol) > 1e-20;
  }
  if (!ok) { atomicAdd(n_general, 1); return; }
  for (int c = 0; c < 8; ++c) invs[(size_t)c * nH + h] = 1.0 / cbrt(6.0 * corner_volume(p, c));
}

// Pass 2 (only when some hex needs it): the general operator B for every tet.
// The reference shape comes from X (via conn); the corner-tet volumes that scale it come from
// Xv (via connv) when given, else from X (grad_tet_ref takes them from the mesh being
// optimised, which for the glue is the exploded mesh, not the one the shape comes from).
__global__ __launch_bounds__(256) void ref_general_kernel(const double* X, int nV, const int* conn, int nH,
                                                          const uint8_t* hflag, int cube, double* Bout,
                                                          const double* Xv, int nVv, const int* connv) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nH) return;
  int vid[8];
  double p[8][3], pv[8][3];
  gather8(X, nV, conn, nH, h, vid, p);
  if (Xv) {
    gather8(Xv, nVv, connv, nH, h, vid, pv);
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) pv[i][d] = p[i][d];
  }
  double r8[8][3];
  const bool hf = hflag == nullptr || hflag[h] != 0;
  if (!hf) {
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d) r8[i][d] = p[i][d];
  } else {
    double vol = hex_volume24(p);
    double e0 = 0, e1 = 0, e2 = 0;
    auto L = [&](int a, int b) {
      double q0 = p[a][0] - p[b][0], q1 = p[a][1] - p[b][1], q2 = p[a][2] - p[b][2];
      return sqrt(q0 * q0 + q1 * q1 + q2 * q2);
    };
    if (cube && fabs(vol) > 1e-20) {
      e0 = e1 = e2 = cbrt(vol);
    } else {
      e0 += L(0, 1); e0 += L(3, 2); e0 += L(4, 5); e0 += L(7, 6); e0 /= 4;
      e1 += L(0, 3); e1 += L(1, 2); e1 += L(4, 7); e1 += L(5, 6); e1 /= 4;
      e2 += L(0, 4); e2 += L(1, 5); e2 += L(2, 6); e2 += L(3, 7); e2 /= 4;
      double ratio = cbrt(vol / (e0 * e1 * e2));
      e0 *= ratio; e1 *= ratio; e2 *= ratio;
    }
    const double cc[8][3] = {{0, 0, 0}, {e0, 0, 0}, {e0, e1, 0}, {0, e1, 0},
                             {0, 0, e2}, {e0, 0, e2}, {e0, e1, e2}, {0, e1, e2}};
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d) r8[i][d] = cc[i][d];
  }
  for (int c = 0; c < 8; ++c) {
    double x[4][3], r[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      pick_row<double, 3>(p, kHexTet[c][j], x[j]);
      pick_row<double, 3>(r8, kHexTet[c][j], r[j]);
    }
    double B[3][3];
    general_B(x, r, corner_volume(pv, c), B);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Bout[((size_t)c * 9 + i * 3 + j) * nH + h] = B[i][j];
  }
}

// ---------------------------------------------------------------------------------------
// Soft-constraint epilogue data (per vertex): S (sym 3x3, xx yy zz xy xz yz), rhs b, line.
__device__ __forceinline__ void soft_apply(const SoftDev& sd, int g, const double p[3], int nV, const double* pa,
                                           double out[3]) {
  const int ci = sd.cidx ? sd.cidx[g] : -1;
  if (ci < 0) return;
  const double* S = sd.S + (size_t)ci * 6;
  out[0] += S[0] * p[0] + S[3] * p[1] + S[4] * p[2];
  out[1] += S[3] * p[0] + S[1] * p[1] + S[5] * p[2];
  out[2] += S[4] * p[0] + S[5] * p[1] + S[2] * p[2];
  const int li = sd.cline[ci];
  if (li >= 0) {
    const double a = pa[li];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] -= sd.lamL * sd.lt[(size_t)li * 3 + d] * a;
  }
  (void)nV;
}

// Glue rows of vertex g (add_soft_constraints' equality block, slim_m.cpp:611-645): with k
// copies in its group and S their sum (of the vector being applied), the rows of
// lamG A^T A are lamG k / (k - 1)^2 (k p_g - S); none for a vertex with no other copy.
__device__ __forceinline__ void glue_rows(const SoftDev& sd, int g, const double pv[3], double r[3]) {
  if (!sd.g_of) return;
  const int gr = sd.g_of[g];
  if (gr < 0) return;
  const int k = sd.g_off[gr + 1] - sd.g_off[gr];
  if (k < 2) return;
  const double c = sd.lamG * (double)k / ((double)(k - 1) * (double)(k - 1));
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] += c * ((double)k * pv[d] - sd.g_sum[(size_t)gr * 3 + d]);
}

// Per-hex corner arrays, each ring (corners 0-3, 4-7) turned one step: row i takes the row
// that followed it. The corner tets of a ring are one pattern of rows turned round the ring
// (kHexTet: tets 0-3 are {c, c-1, c+4, c+1}, tets 4-7 {c, c-1, c+1, c-4}, within each ring),
// so a loop over a ring's tets that turns the arrays after each tet reads and writes the same
// compile-time rows every time, with no selects over the 8 rows; four turns bring them back.
template <class T, int N>
__device__ __forceinline__ void rotate_rings(T a[8][N]) {
#pragma unroll
  for (int d = 0; d < N; ++d) {
    const T b0 = a[0][d], t0 = a[4][d];
    a[0][d] = a[1][d]; a[1][d] = a[2][d]; a[2][d] = a[3][d]; a[3][d] = b0;
    a[4][d] = a[5][d]; a[5][d] = a[6][d]; a[6][d] = a[7][d]; a[7][d] = t0;
  }
}

// ---------------------------------------------------------------------------------------
// Local step: per tet SVD -> W, R (slim_m.cpp:236-377); stores K = M W^T W (scaled by
// invs^2 for the cube operator) and scatters the rhs A^T M (W R) (buildRhs, :1044-1092) and
// the Jacobi diagonal of A^T M A (the DiagonalPreconditioner of the CG).
struct LocalStepOp {
  RefOp ro; int nV, nH; const int* conn; const double* X; double* K;
  int energy; double M, exp_f, prox;
  SoftDev sd; double* rhs; double* invd;
  // corner tet c on rows T0..T3 of p / out (compile-time indices, see rotate_rings); the sums
  // are those of scatter_tet into a zeroed per-tet array added to out, in the same order
  template <int T0, int T1, int T2, int T3>
  __device__ __forceinline__ void tet(int h, int c, const double p[8][3], double out[8][6]) const {
    constexpr int tc[4] = {T0, T1, T2, T3};
    double B[3][3], E[3][3], J[3][3], W[3][3], R[3][3];
    tet_B(ro, nH, h, c, B);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) E[d][k] = p[tc[k + 1]][d] - p[T0][d];
    matmul3(E, B, J);
    slim_weights(J, energy, exp_f, W, R);
    double Kt[3][3];  // M W^T W
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Kt[i][j] = M * (W[0][i] * W[0][j] + W[1][i] * W[1][j] + W[2][i] * W[2][j]);
    double s2 = 1.0;
    if (ro.mode == REF_CUBE) { double s = ro.invs[(size_t)c * nH + h]; s2 = s * s; }
    const double ks[6] = {Kt[0][0], Kt[1][1], Kt[2][2], Kt[0][1], Kt[0][2], Kt[1][2]};
#pragma unroll
    for (int j = 0; j < 6; ++j) K[k_at<double>(c * 6 + j, h, nH)] = ks[j] * s2;
    // rhs: (W^T M W R) scattered through B
    double Z[3][3];
    {
      double F[3][3];
      matmul3(W, R, F);
#pragma unroll
      for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) Z[b2][cc] = M * (W[0][b2] * F[0][cc] + W[1][b2] * F[1][cc] + W[2][b2] * F[2][cc]);
    }
    // row k + 1 gets Z B_k, row 0 the sum of their negatives (first from 0), each then added
    // to out
    double t0[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double f = Z[d][0] * B[k][0] + Z[d][1] * B[k][1] + Z[d][2] * B[k][2];
        out[tc[k + 1]][d] += f;
        t0[d] += -f;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) out[T0][d] += t0[d];
    // diagonal: Kt_bb * |G[:,j]|^2; G[:,k] = B[k-1,:], G[:,0] = -sum_k B[k,:]
    double g0[3] = {-(B[0][0] + B[1][0] + B[2][0]), -(B[0][1] + B[1][1] + B[2][1]), -(B[0][2] + B[1][2] + B[2][2])};
    double n0 = g0[0] * g0[0] + g0[1] * g0[1] + g0[2] * g0[2];
    double nk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) nk[k] = B[k][0] * B[k][0] + B[k][1] * B[k][1] + B[k][2] * B[k][2];
    {
      // each product rounded before its sum (no fused multiply-add)
#pragma clang fp contract(off)
#pragma unroll
      for (int d = 0; d < 3; ++d) out[T0][3 + d] += Kt[d][d] * n0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) out[tc[k + 1]][3 + d] += Kt[d][d] * nk[k];
    }
  }
  __device__ double hex(int h, double out[8][6]) const {
    int vid[8];
    double p[8][3];
    gather8(X, nV, conn, nH, h, vid, p);
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      tet<0, 3, 4, 1>(h, r, p, out);
      rotate_rings(p);
      rotate_rings(out);
    }
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      tet<4, 7, 5, 0>(h, 4 + r, p, out);
      rotate_rings(p);
      rotate_rings(out);
    }
    return 0.0;
  }
  __device__ double finish(int g, double acc[6]) const {
    double dg[3] = {acc[3] + prox, acc[4] + prox, acc[5] + prox};
    double r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = acc[d] + prox * X[(size_t)d * nV + g];
    const int ci = sd.cidx ? sd.cidx[g] : -1;
    if (ci >= 0) {
      const double* S = sd.S + (size_t)ci * 6;
      const double* b = sd.b + (size_t)ci * 3;
#pragma unroll
      for (int d = 0; d < 3; ++d) { dg[d] += S[d]; r[d] += b[d]; }
    }
    if (sd.g_of && sd.g_of[g] >= 0) {  // glue diagonal lamG k / (k - 1); its rhs is 0
      const int k = sd.g_off[sd.g_of[g] + 1] - sd.g_off[sd.g_of[g]];
      if (k > 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) dg[d] += sd.lamG * (double)k / (double)(k - 1);
    }
    // a fixed row leaves the solve: invd 0 marks it for the CG kernels
    const bool pinned = sd.fixed && sd.fixed[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rhs[(size_t)d * nV + g] = r[d];
      invd[(size_t)d * nV + g] = pinned ? 0.0 : (dg[d] != 0.0 ? 1.0 / dg[d] : 1.0);
    }
    return 0.0;
  }
};

// the FP64 local step needs all 256 registers on its own (one wave per SIMD); two waves per
// SIMD hide its FP64 latency better than the registers it then spills cost (125 spilled
// values: 2.13 ms against 2.63 ms per pass on the 1.1M-hex octree torus)
template <>
struct TileWaves<LocalStepOp> {
  static constexpr int value = 2;
};

// CG operator L p = A^T M A p + prox p + soft terms; finish() also returns p.(Lp).
// KT: stored K precision, PT: vector precision, CT: arithmetic, OT: output precision.
// Cube reference: invs^2 is folded into K, so the forces on the edges of corner tet c are
// K E (columns = edges), added at the edge ends and taken off the tet's corner.
template <class CT>
__device__ __forceinline__ void tet_forces_cube(const CT E[3][3], const CT kk[6], int c, CT o[8][3]) {
  const CT k00 = kk[0], k11 = kk[1], k22 = kk[2], k01 = kk[3], k02 = kk[4], k12 = kk[5];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const CT f0 = k00 * E[0][k] + k01 * E[1][k] + k02 * E[2][k];
    const CT f1 = k01 * E[0][k] + k11 * E[1][k] + k12 * E[2][k];
    const CT f2 = k02 * E[0][k] + k12 * E[1][k] + k22 * E[2][k];
    const int vk = kHexTet[c][k + 1], v0 = kHexTet[c][0];
    o[vk][0] += f0; o[vk][1] += f1; o[vk][2] += f2;
    o[v0][0] -= f0; o[v0][1] -= f1; o[v0][2] -= f2;
  }
}

// Element part of the CG operator for hex h: corner forces o (added) from the corner
// values q of the direction, with the per-tet K = invs^2 M W W^T (6 symmetric components).
template <class KT, class CT>
__device__ __forceinline__ void hex_forces(const RefOp& ro, const KT* K, int nH, int h, const CT q[8][3], CT o[8][3]) {
  CT kp[12];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    CT E[3][3], B[3][3], Y[3][3];
    tet_edges(q, c, E);
    if ((c & 1) == 0) k_pair(K, nH, h, c >> 1, kp);
    const CT* kk = kp + 6 * (c & 1);
    const CT k00 = kk[0], k11 = kk[1], k22 = kk[2], k01 = kk[3], k02 = kk[4], k12 = kk[5];
    if (ro.mode == REF_CUBE) {
      tet_forces_cube(E, kk, c, o);
    } else {
      tet_B(ro, nH, h, c, B);
      CT J[3][3];
      matmul3(E, B, J);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Y[0][j] = k00 * J[0][j] + k01 * J[1][j] + k02 * J[2][j];
        Y[1][j] = k01 * J[0][j] + k11 * J[1][j] + k12 * J[2][j];
        Y[2][j] = k02 * J[0][j] + k12 * J[1][j] + k22 * J[2][j];
      }
      scatter_tet(Y, B, c, o);
    }
  }
}

template <class KT, class PT, class CT, class OT>
struct ApplyOp {
  RefOp ro; int nV, nH; const int* conn; const KT* K; const PT* p; OT* out; double prox;
  SoftDev sd; const PT* pa; int nL;
  __device__ double hex(int h, CT o[8][3]) const {
    CT q[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int v = conn[(size_t)k * nH + h];
#pragma unroll
      for (int d = 0; d < 3; ++d) q[k][d] = (CT)p[(size_t)d * nV + v];
    }
    hex_forces<KT, CT>(ro, K, nH, h, q, o);
    return 0.0;
  }
  __device__ double finish(int g, CT acc[3]) const {
    double pv[3] = {(double)p[g], (double)p[(size_t)nV + g], (double)p[2 * (size_t)nV + g]};
    double r[3] = {(double)acc[0] + prox * pv[0], (double)acc[1] + prox * pv[1], (double)acc[2] + prox * pv[2]};
    const int ci = sd.cidx ? sd.cidx[g] : -1;
    if (ci >= 0) {
      const double* S = sd.S + (size_t)ci * 6;
      r[0] += S[0] * pv[0] + S[3] * pv[1] + S[4] * pv[2];
      r[1] += S[3] * pv[0] + S[1] * pv[1] + S[5] * pv[2];
      r[2] += S[4] * pv[0] + S[5] * pv[1] + S[2] * pv[2];
      const int li = sd.cline[ci];
      if (li >= 0) {
        const double a = (double)pa[li];
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] -= sd.lamL * sd.lt[(size_t)li * 3 + d] * a;
      }
    }
    glue_rows(sd, g, pv, r);
#pragma unroll
    for (int d = 0; d < 3; ++d) out[(size_t)d * nV + g] = (OT)r[d];
    return pv[0] * r[0] + pv[1] * r[1] + pv[2] * r[2];
  }
};

// ---------------------------------------------------------------------------------------
// FP32 adjoint + local step of the symmetric-Dirichlet energy (the FP32 optimiser's only
// pass over the elements per iteration): per corner tet in FP32 arithmetic the Jacobian,
// its SVD, the SLIM weights K = M W^T W (stored FP32), the energy, and dE/dJ; scatters the
// direction rhs -dE/dx / 2 and the Jacobi diagonal. The corner-determinant sign is tested
// in FP64 (inverted-tet count); the energy is summed in FP64.
struct AdjLocalOp32;
// three waves per SIMD: the FP64 corners and the hex's 8 x 6 sums stay live through the
// corner-tet loops; 37 values spill, and it still runs faster than two waves without spills
// (415 against 420 us on the 1.1M-hex octree torus)
template <>
struct TileWaves<AdjLocalOp32> {
  static constexpr int value = 3;
};

struct AdjLocalOp32 {
  RefOp ro; int nV, nH; const int* conn; const double* X; float* K; double M, prox;
  SoftDev sd; float* rhs; float* invd; int* n_inverted;

  // Corner tet c whose corners are rows T0, T1, T2, T3 of pd / out (compile-time indices: the
  // caller rotates the rows so that every corner tet of a ring takes the same four rows).
  // Per tet, FP64 first: the Jacobian J = E B from the FP64 corners, its orientation (sign of
  // det J; for a cube reference B = invs * (a permutation with determinant +1 for every
  // corner) and invs > 0, so the sign is that of det E), and its deviation from isometry
  // D = J J^T - I. Then FP32 for the rest: the weights (the reference's m = 1 branch,
  // |s_i - 1| < 1e-8 (slim_m.cpp:351-353), is taken for the tet when |D|_F < 2e-8 puts every
  // singular value inside it, a threshold FP32 cannot resolve), and the energy and dE/dJ in
  // their deviation form, which keeps FP32's relative precision where J is near a rotation.
  template <int T0, int T1, int T2, int T3>
  __device__ __forceinline__ void tet(int h, int c, const double pd[8][3], float out[8][6], double& e, int& inv) const {
    constexpr int tc[4] = {T0, T1, T2, T3};
    double E64[3][3], B64[3][3], J64[3][3], D64[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) E64[d][k] = pd[tc[k + 1]][d] - pd[T0][d];
    tet_B(ro, nH, h, c, B64);
    matmul3(E64, B64, J64);
    inv += (ro.mode == REF_CUBE ? det3(E64) : det3(J64)) < 0.0;
    const double dev2 = isometry_deviation(J64, D64);
    const bool iso = dev2 < 4e-16;
    float B[3][3], J[3][3], D[3][3], W[3][3], P[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j2 = 0; j2 < 3; ++j2) {
        B[i][j2] = (float)B64[i][j2];
        J[i][j2] = (float)J64[i][j2];
        D[i][j2] = (float)D64[i][j2];
      }
    if (iso) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j2 = 0; j2 < 3; ++j2) W[i][j2] = i == j2 ? 1.0f : 0.0f;
    } else {
      slim_weights_sd_stable(J, W);
    }
    // the deviation form while I + D is well conditioned (|D|_F < 0.5); the direct form
    // else, where inverting J J^T would square J's condition number
    if (dev2 < 0.25) e += M * (6.0 + (double)sd_energy_grad_dev(J, D, P));
    else e += M * (double)sd_energy_grad(J, P);
    float Kt[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Kt[i][j] = (float)M * (W[0][i] * W[0][j] + W[1][i] * W[1][j] + W[2][i] * W[2][j]);
    float s2 = 1.0f;
    if (ro.mode == REF_CUBE) { float sc = (float)ro.invs[(size_t)c * nH + h]; s2 = sc * sc; }
    const float ks[6] = {Kt[0][0], Kt[1][1], Kt[2][2], Kt[0][1], Kt[0][2], Kt[1][2]};
#pragma unroll
    for (int j = 0; j < 6; ++j) K[k_at<float>(c * 6 + j, h, nH)] = ks[j] * s2;
    // -dE/dx / 2 of the tet's corners: row k + 1 gets P B_k, row 0 minus their sum
    const float hm = -0.5f * (float)M;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float f = hm * (P[d][0] * B[k][0] + P[d][1] * B[k][1] + P[d][2] * B[k][2]);
        out[tc[k + 1]][d] += f;
        out[T0][d] -= f;
      }
    }
    // Jacobi diagonal
    const float g0[3] = {-(B[0][0] + B[1][0] + B[2][0]), -(B[0][1] + B[1][1] + B[2][1]),
                         -(B[0][2] + B[1][2] + B[2][2])};
    const float n0 = g0[0] * g0[0] + g0[1] * g0[1] + g0[2] * g0[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[T0][3 + d] += Kt[d][d] * n0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float nk = B[k][0] * B[k][0] + B[k][1] * B[k][1] + B[k][2] * B[k][2];
#pragma unroll
      for (int d = 0; d < 3; ++d) out[tc[k + 1]][3 + d] += Kt[d][d] * nk;
    }
  }

  __device__ double hex(int h, float out[8][6]) const {
    int vid[8];
    double pd[8][3];
    gather8(X, nV, conn, nH, h, vid, pd);
    double e = 0.0;
    int inv = 0;
    // the ring's first tet pattern on rows turned r steps is tet r (rotate_rings)
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      tet<0, 3, 4, 1>(h, r, pd, out, e, inv);
      rotate_rings(pd);
      rotate_rings(out);
    }
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      tet<4, 7, 5, 0>(h, 4 + r, pd, out, e, inv);
      rotate_rings(pd);
      rotate_rings(out);
    }
    if (inv) atomicAdd(n_inverted, inv);
    return e;
  }
  __device__ double finish(int g, float acc[6]) const {
    double dg[3] = {(double)acc[3] + prox, (double)acc[4] + prox, (double)acc[5] + prox};
    double r[3] = {(double)acc[0], (double)acc[1], (double)acc[2]};
    const int ci = sd.cidx ? sd.cidx[g] : -1;
    if (ci >= 0) {
      const double* S = sd.S + (size_t)ci * 6;
      const double* b = sd.b + (size_t)ci * 3;
      const double x[3] = {X[g], X[(size_t)nV + g], X[2 * (size_t)nV + g]};
      r[0] -= S[0] * x[0] + S[3] * x[1] + S[4] * x[2] - b[0];
      r[1] -= S[3] * x[0] + S[1] * x[1] + S[5] * x[2] - b[1];
      r[2] -= S[4] * x[0] + S[5] * x[1] + S[2] * x[2] - b[2];
#pragma unroll
      for (int d = 0; d < 3; ++d) dg[d] += S[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rhs[(size_t)d * nV + g] = (float)r[d];
      invd[(size_t)d * nV + g] = dg[d] != 0.0 ? (float)(1.0 / dg[d]) : 1.0f;
    }
    return 0.0;
  }
};

// rhs of the a_L rows of the direction system at a_L = 0: lamL t . (x_v - o)
__global__ void line_rhs32_kernel(SoftDev sd, int nV, const double* X, float* rhs_a, float* invd_a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sd.nL) return;
  const int v = sd.lv[i];
  const double* t = sd.lt + (size_t)i * 3;
  const double* o = sd.lo + (size_t)i * 3;
  double s = 0, tt = 0;
  for (int d = 0; d < 3; ++d) { s += t[d] * (X[(size_t)d * nV + v] - o[d]); tt += t[d] * t[d]; }
  rhs_a[i] = (float)(sd.lamL * s);
  double dg = sd.lamL * tt;
  invd_a[i] = dg != 0 ? (float)(1.0 / dg) : 1.0f;
}

// Rows of the feature-line unknowns a_L: (L p)_a = lamL (|t|^2 pa - t . p_v).
template <class PT, class OT>
__global__ void line_apply_kernel(SoftDev sd, int nV, const PT* p, const PT* pa, OT* outa, int nL, double* partial) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double part = 0;
  if (i < nL) {
    const int v = sd.lv[i];
    const double* t = sd.lt + (size_t)i * 3;
    double tp = t[0] * (double)p[v] + t[1] * (double)p[(size_t)nV + v] + t[2] * (double)p[2 * (size_t)nV + v];
    double tt = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    double r = sd.lamL * (tt * (double)pa[i] - tp);
    outa[i] = (OT)r;
    part = (double)pa[i] * r;
  }
  double s = block_sum(part, red);
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------------------
// CG vector kernels (ConjugateGradient.h:27-90). The flattened unknown vector has
// n = 3 nV + nL entries; T is the vector precision, every dot product is summed in FP64.
struct CGState {
  double rhsNorm2, threshold, absNew, alpha, beta, rn2, tol_error;
  int done, iters, zero_rhs, pad;
  double qz, bzq, qq;  // FP32 CG: z.R z, z.R q, q.R q of the next direction z + beta q
  double hist[2];      // FP32 CG: r.z after update k-1 in hist[k & 1] (the beta step's absOld)
};

// The update pass's five partial arrays (|r|^2, r.z, z.R z, z.R q, q.R q) of CG iteration k,
// n per array; n == 0 when the operator pass does not form the beta step itself.
struct UpdParts {
  const double* p[5];
  int n, k;
};

// r = rhs - tmp, partial |rhs|^2, |r|^2
// (rows with invd == 0 are fixed and left out: r = 0 there)
template <class T>
__global__ void cg_init_kernel(int n, const T* rhs, const T* tmp, T* r, const T* invd, double* part_b,
                               double* part_r) {
  __shared__ double red[16];
  double sb = 0, sr = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    T v = invd[i] == T(0) ? T(0) : rhs[i] - tmp[i];
    r[i] = v;
    sb += (double)rhs[i] * (double)rhs[i];
    sr += (double)v * (double)v;
  }
  double a = block_sum(sb, red);
  double b = block_sum(sr, red);
  if (threadIdx.x == 0) { part_b[blockIdx.x] = a; part_r[blockIdx.x] = b; }
}

__device__ double sum_parts(const double* p, int n, double* red) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  return block_sum(s, red);
}

// sum_parts of N arrays of n partials in one pass; red holds N * 16 values
template <int N>
__device__ void sum_parts_n(const double* const (&p)[N], int n, double* red, double out[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) out[k] = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] += p[k][i];
  }
  block_sum_n<N>(out, red);
}

// sum_parts_n with the first U * blockDim partials of every array loaded before any is added
// (one round trip instead of U); the additions keep sum_parts_n's order, so the sums are the
// same
template <int N, int U>
__device__ void sum_parts_n_wide(const double* const (&p)[N], int n, double* red, double out[N]) {
  double v[U][N];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const int i = threadIdx.x + j * (int)blockDim.x;
#pragma unroll
    for (int k = 0; k < N; ++k) v[j][k] = i < n ? p[k][i] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < N; ++k) out[k] = 0;
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const int i = threadIdx.x + j * (int)blockDim.x;
#pragma unroll
    for (int k = 0; k < N; ++k)
      if (i < n) out[k] += v[j][k];
  }
  for (int i = threadIdx.x + U * (int)blockDim.x; i < n; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] += p[k][i];
  }
  block_sum_n<N>(out, red);
}

// Eigen ConjugateGradient.h:44-62 (zero rhs, early exit) and the first direction.
__global__ void cg_init_fin_kernel(CGState* st, const double* part_b, int nb, const double* part_r, double tol) {
  __shared__ double red[16];
  double b2 = sum_parts(part_b, nb, red);
  double r2 = sum_parts(part_r, nb, red);
  if (threadIdx.x == 0) {
    st->rhsNorm2 = b2;
    st->iters = 0;
    st->done = 0;
    st->zero_rhs = 0;
    if (b2 == 0) { st->done = 1; st->zero_rhs = 1; st->tol_error = 0; return; }
    st->threshold = tol * tol * b2;
    st->rn2 = r2;
    if (r2 < st->threshold) { st->done = 1; st->tol_error = sqrt(r2 / b2); }
  }
}

// p = invd r, partial r.p
template <class T>
__global__ void cg_first_dir_kernel(int n, const CGState* st, const T* invd, const T* r, T* p, double* part) {
  __shared__ double red[16];
  double s = 0;
  if (!st->done)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      T v = invd[i] * r[i];
      p[i] = v;
      s += (double)r[i] * (double)v;
    }
  double a = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void cg_absnew_fin_kernel(CGState* st, const double* part, int nb) {
  __shared__ double red[16];
  double a = sum_parts(part, nb, red);
  if (threadIdx.x == 0 && !st->done) st->absNew = a;
}

// alpha = absNew / p.(Lp) ; the dot product is split over three partial arrays
__global__ void cg_alpha_kernel(CGState* st, const double* p1, int n1, const double* p2, int n2, const double* p3, int n3) {
  __shared__ double red[16];
  double a = sum_parts(p1, n1, red);
  double b = sum_parts(p2, n2, red);
  double c = sum_parts(p3, n3, red);
  if (threadIdx.x == 0 && !st->done) st->alpha = st->absNew / (a + b + c);
}

// x += alpha p ; r -= alpha tmp ; z = invd r ; partial |r|^2, r.z
template <class T>
__global__ void cg_update_kernel(int n, const CGState* st, T* x, T* r, T* z, const T* p, const T* tmp, const T* invd,
                                 double* part_rr, double* part_rz) {
  __shared__ double red[16];
  double srr = 0, srz = 0;
  if (!st->done) {
    const T alpha = (T)st->alpha;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      x[i] += alpha * p[i];
      T rv = invd[i] == T(0) ? T(0) : r[i] - alpha * tmp[i];  // fixed rows stay out
      r[i] = rv;
      T zv = invd[i] * rv;
      z[i] = zv;
      srr += (double)rv * (double)rv;
      srz += (double)rv * (double)zv;
    }
  }
  double a = block_sum(srr, red);
  double b = block_sum(srz, red);
  if (threadIdx.x == 0) { part_rr[blockIdx.x] = a; part_rz[blockIdx.x] = b; }
}

// convergence test and beta (ConjugateGradient.h:73-85)
__global__ void cg_beta_kernel(CGState* st, const double* part_rr, const double* part_rz, int nb) {
  __shared__ double red[16];
  double rr = sum_parts(part_rr, nb, red);
  double rz = sum_parts(part_rz, nb, red);
  if (threadIdx.x == 0 && !st->done) {
    st->rn2 = rr;
    if (rr < st->threshold) {
      st->done = 1;
    } else {
      double absOld = st->absNew;
      st->absNew = rz;
      st->beta = rz / absOld;
      st->iters += 1;
    }
    st->tol_error = sqrt(rr / st->rhsNorm2);
  }
}

template <class T>
__global__ void cg_pdir_kernel(int n, const CGState* st, T* p, const T* z) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

template <class T>
__global__ void cg_zero_if_kernel(int n, const CGState* st, T* x) {
  if (!st->zero_rhs) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0;
}

// ---------------------------------------------------------------------------------------
// FP32 CG with the vector work fused into the operator pass (the FP32 optimiser's global
// step, Eigen's algorithm with FP64 dot products). L = E + R: E the element operator, R the
// vertex-local rows (prox, the folded constraint block S, and each line unknown's coupling
// to its vertex). Per CG iteration:
//   cg32_apply   per tile: p = z + beta p_prev staged in LDS (p double-buffered, each vertex
//                written once: by its own tile or by the first tile of its halo slots),
//                E p of the tile's elements, and a per-tile partial of p.E p from the element
//                quadratic forms p_e.(K_e p_e)
//   cg32_lines   p and the a_L rows of L p for the line unknowns (feature lines only)
//   cg32_alpha   one block: p of the vertices without hexes, then
//                alpha = r.z / (p.E p + p.R p). p.R p needs no pass of its own: with
//                p = z + beta q (q the previous direction) it is
//                z.R z + 2 beta z.R q + beta^2 q.R q, three sums the previous update left
//   cg32_update  per vertex: E p (its tile's value or the sum of its halo slots) + R p,
//                x += alpha p, r -= alpha L p, z = r / diag for the vertex and its line
//                unknown; FP64 partials of |r|^2, r.z and of z.R z, z.R p, p.R p
//   cg32_beta    one block: convergence test, beta, and the three R sums
// Every reduction is two-level in a fixed order; kernel boundaries order the levels (an
// in-kernel last-block reduction would need a device-scope fence per block, which on this
// multi-XCD part writes back L2 each time).

// rows of R p at vertex g: prox p + S p - lamL t a
__device__ __forceinline__ void vertex_rows(const SoftDev& sd, double prox, int ci, int li, const double pv[3], double pa,
                                            double r[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] = prox * pv[d];
  if (ci >= 0) {
    const double* S = sd.S + (size_t)ci * 6;
    r[0] += S[0] * pv[0] + S[3] * pv[1] + S[4] * pv[2];
    r[1] += S[3] * pv[0] + S[1] * pv[1] + S[5] * pv[2];
    r[2] += S[4] * pv[0] + S[5] * pv[1] + S[2] * pv[2];
    if (li >= 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) r[d] -= sd.lamL * sd.lt[(size_t)li * 3 + d] * pa;
    }
  }
}

// R as a symmetric bilinear form over vertex g and its line unknown a (the a_L row
// lamL (|t|^2 p_a - t.p_v) included): x.R y = prox x.y + x^T S y
// + lamL (|t|^2 x_a y_a - x_a t.y - y_a t.x)
__device__ __forceinline__ double vertex_form(const SoftDev& sd, double prox, int ci, int li, const double x[3],
                                              double xa, const double y[3], double ya) {
  double f = prox * (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]);
  if (ci >= 0) {
    const double* S = sd.S + (size_t)ci * 6;
    f += x[0] * (S[0] * y[0] + S[3] * y[1] + S[4] * y[2]) + x[1] * (S[3] * y[0] + S[1] * y[1] + S[5] * y[2]) +
         x[2] * (S[4] * y[0] + S[5] * y[1] + S[2] * y[2]);
    if (li >= 0) {
      const double* t = sd.lt + (size_t)li * 3;
      const double tt = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
      const double tx = t[0] * x[0] + t[1] * x[1] + t[2] * x[2];
      const double ty = t[0] * y[0] + t[1] * y[1] + t[2] * y[2];
      f += sd.lamL * (tt * xa * ya - xa * ty - ya * tx);
    }
  }
  return f;
}

// x = 0, r = b, z = b / diag, p_prev = 0; partials of |b|^2 and r.z
__global__ __launch_bounds__(256) void cg32_init_kernel(int n, const float* __restrict__ b, const float* __restrict__ invd,
                                                        float* __restrict__ x, float* __restrict__ r,
                                                        float* __restrict__ z, float* __restrict__ pprev,
                                                        double* part_b, double* part_rz) {
  __shared__ double red[16];
  double sb = 0, srz = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float bv = b[i];
    const float zv = invd[i] * bv;
    x[i] = 0.0f;
    r[i] = bv;
    z[i] = zv;
    pprev[i] = 0.0f;
    sb += (double)bv * (double)bv;
    srz += (double)bv * (double)zv;
  }
  const double a = block_sum(sb, red);
  const double c = block_sum(srz, red);
  if (threadIdx.x == 0) { part_b[blockIdx.x] = a; part_rz[blockIdx.x] = c; }
}

// partials of z.R z for the first direction (p_prev = 0, so z.R q = q.R q = 0)
__global__ __launch_bounds__(256) void cg32_form0_kernel(SoftDev sd, double prox, int nV, const float* __restrict__ z,
                                                         double* part) {
  __shared__ double red[16];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  double f = 0.0;
  if (g < nV) {
    const size_t nA = 3 * (size_t)nV;
    const double zv[3] = {(double)z[g], (double)z[(size_t)nV + g], (double)z[2 * (size_t)nV + g]};
    const int ci = sd.cidx ? sd.cidx[g] : -1;
    const int ln = ci >= 0 ? sd.cline[ci] : -1;
    const double za = ln >= 0 ? (double)z[nA + ln] : 0.0;
    f = vertex_form(sd, prox, ci, ln, zv, za, zv, za);
  }
  const double a = block_sum(f, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// ConjugateGradient.h:44-62 with x0 = 0 (so r0 = b and |r0|^2 = |b|^2)
__global__ void cg32_init_fin_kernel(CGState* st, const double* part_b, const double* part_rz, int nb,
                                     const double* part_q, int nq, double tol) {
  __shared__ double red[16];
  const double b2 = sum_parts(part_b, nb, red);
  const double rz = sum_parts(part_rz, nb, red);
  const double qz = sum_parts(part_q, nq, red);
  if (threadIdx.x == 0) {
    st->rhsNorm2 = b2;
    st->iters = 0;
    st->done = 0;
    st->zero_rhs = 0;
    st->beta = 0;
    st->tol_error = 0;
    st->qz = qz;
    st->bzq = 0;
    st->qq = 0;
    if (b2 == 0) { st->done = 1; st->zero_rhs = 1; return; }
    st->threshold = tol * tol * b2;
    st->rn2 = b2;
    if (b2 < st->threshold) { st->done = 1; st->tol_error = 1; return; }
    st->absNew = rz;
    st->hist[0] = rz;
    st->hist[1] = 0;
  }
}

template <bool CUBE>
__global__ __launch_bounds__(kTileHex) void cg32_apply_kernel(TileMesh tm, RefOp ro, const float* __restrict__ K,
                                                              const float* __restrict__ z,
                                                              const float* __restrict__ pprev, float* __restrict__ pnew,
                                                              float* __restrict__ out, float4* __restrict__ halo,
                                                              const CGState* st, double* partial) {
  __shared__ float sp[3 * kMaxLV];
  __shared__ float acc[3 * kMaxLV];
  __shared__ double red[16];
  if (st->done) return;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int nV = tm.nV, nH = tm.nH;
  const int h0 = tm.tile_hex[tile], nh = tm.tile_hex[tile + 1] - h0;
  const int lv0 = tm.tile_lv[tile], nlv = tm.tile_lv[tile + 1] - lv0;
  const float beta = (float)st->beta;
  // loads that need nothing from the staging go out first: the hex's corner ids and, for
  // the cube reference, its 48 operator values
  const bool active = tid < nh;
  int li[8];
  float kr[CUBE ? 48 : 1];
  if (active) {
#pragma unroll
    for (int k = 0; k < 8; ++k) li[k] = tm.lidx[(size_t)k * nH + h0 + tid];
    if constexpr (CUBE) {
#pragma unroll
      for (int q = 0; q < 4; ++q) k_pair(K, nH, h0 + tid, q, kr + 12 * q);
    }
  }
  // stage p of the tile's vertices
  for (int lv = tid; lv < nlv; lv += kTileHex) {
    const int g = tm.lv_gid[lv0 + lv];
    float pv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pv[d] = z[(size_t)d * nV + g] + beta * pprev[(size_t)d * nV + g];
      sp[d * kMaxLV + lv] = pv[d];
      acc[d * kMaxLV + lv] = 0.0f;
    }
    if (tm.lv_first[lv0 + lv]) {
#pragma unroll
      for (int d = 0; d < 3; ++d) pnew[(size_t)d * nV + g] = pv[d];
    }
  }
  __syncthreads();
  float c[8][3];
  double part = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) c[k][d] = 0.0f;
  if (active) {
    const int h = h0 + tid;
    float q[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) q[k][d] = sp[d * kMaxLV + li[k]];
    if constexpr (CUBE) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        float E[3][3];
        tet_edges(q, t, E);
        tet_forces_cube(E, kr + 6 * t, t, c);
      }
    } else {
      hex_forces<float, float>(ro, K, nH, h, q, c);
    }
    float e = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) e += q[k][d] * c[k][d];
    part = (double)e;
  }
  scatter_tile<float, 3>(tm, tile, nh, active, li, c, acc);
  // E p of each local vertex to its place; the update sums each vertex's run
  for (int lv = tid; lv < nlv; lv += kTileHex)
    halo[tm.lv_dst[lv0 + lv]] = make_float4(acc[lv], acc[kMaxLV + lv], acc[2 * kMaxLV + lv], 0.0f);
  const double bs = block_sum(part, red);
  if (tid == 0) partial[tile] = bs;
}

// line unknown i: p_a = z_a + beta p_prev_a and the a_L row (L p)_a = lamL (|t|^2 p_a - t.p_v)
__device__ __forceinline__ void line_row(const SoftDev& sd, int nV, const float* __restrict__ z,
                                         const float* __restrict__ pprev, float* __restrict__ pnew,
                                         float* __restrict__ out, float beta, int i) {
  const size_t nA = 3 * (size_t)nV;
  const int v = sd.lv[i];
  const float pa = z[nA + i] + beta * pprev[nA + i];
  pnew[nA + i] = pa;
  const double* t = sd.lt + (size_t)i * 3;
  double tp = 0, tt = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double pv = (double)(z[(size_t)d * nV + v] + beta * pprev[(size_t)d * nV + v]);
    tp += t[d] * pv;
    tt += t[d] * t[d];
  }
  out[nA + i] = (float)(sd.lamL * (tt * (double)pa - tp));
}

// Everything a tile pass loads before its first use, fetched one tile ahead by the
// persistent operator pass below: the tile's ranges, the thread's local vertices (id, place
// of its E p, first-writer flag; at most kMaxLV / kTileHex = 4 per thread), its hex's corner
// ids and 48 operator values.
struct PipeTile {
  int h0, nh, lv0, nlv;
  int g[kMaxLV / kTileHex], s[kMaxLV / kTileHex];
  unsigned first;  // bit j: local vertex tid + j * kTileHex writes p
  int li[8];
  float kr[48];
};

__device__ __forceinline__ void pipe_fetch(const TileMesh& tm, const float* __restrict__ K, int t, PipeTile& P) {
  const int tid = threadIdx.x;
  P.h0 = tm.tile_hex[t];
  P.nh = tm.tile_hex[t + 1] - P.h0;
  P.lv0 = tm.tile_lv[t];
  P.nlv = tm.tile_lv[t + 1] - P.lv0;
  P.first = 0;
#pragma unroll
  for (int j = 0; j < kMaxLV / kTileHex; ++j) {
    const int lv = tid + j * kTileHex;
    if (lv < P.nlv) {
      P.g[j] = tm.lv_gid[P.lv0 + lv];
      P.s[j] = tm.lv_dst[P.lv0 + lv];
      P.first |= tm.lv_first[P.lv0 + lv] ? (1u << j) : 0u;
    }
  }
  if (tid < P.nh) {
#pragma unroll
    for (int k = 0; k < 8; ++k) P.li[k] = tm.lidx[(size_t)k * tm.nH + P.h0 + tid];
#pragma unroll
    for (int q = 0; q < 4; ++q) k_pair(K, tm.nH, P.h0 + tid, q, P.kr + 12 * q);
  }
}

// The cube-reference operator pass of cg32_apply_kernel as a persistent kernel: the 8 XCDs
// take contiguous eighths of the tiles (as xcd_tile deals them), and within an XCD block r
// takes tiles r, r + B, r + 2B, ... of that eighth (B blocks on the XCD). While a block
// scatters and writes one tile, the loads of its next tile are in flight.
__global__ __launch_bounds__(kTileHex) void cg32_apply_pipe_kernel(TileMesh tm, const float* __restrict__ K,
                                                                   const float* __restrict__ z,
                                                                   const float* __restrict__ pprev,
                                                                   float* __restrict__ pnew, float* __restrict__ out,
                                                                   float4* __restrict__ halo, SoftDev sd,
                                                                   const int* __restrict__ iso, int niso,
                                                                   CGState* st, double* partial, UpdParts up) {
  __shared__ float4 sp[kMaxLV];
  __shared__ float4 acc[kMaxLV];
  __shared__ double red[5 * 16];
  if (st->done) return;
  const int tid = threadIdx.x;
  const int nV = tm.nV, nT = tm.nTiles;
  const int q8 = nT / 8, r8 = nT % 8;
  const int x = blockIdx.x % 8, rank = blockIdx.x / 8;
  const int nb = ((int)gridDim.x + 7 - x) / 8;  // blocks on XCD x
  const int tend = (x + 1) * q8 + (x + 1 < r8 ? x + 1 : r8);
  int tile = x * q8 + (x < r8 ? x : r8) + rank;
  PipeTile P;
  if (tile < tend) pipe_fetch(tm, K, tile, P);
  // (the first tile's loads are in flight while the beta step is formed)
  float beta;
  if (up.n > 0) {
    // the beta step of the previous CG iteration (cg32_beta_kernel's), formed by every block
    // from the update pass's partials; block 0 records it in the state
    const double* const parts[5] = {up.p[0], up.p[1], up.p[2], up.p[3], up.p[4]};
    double v[5];
    sum_parts_n_wide<5, 6>(parts, up.n, red, v);
    const double rr = v[0], rz = v[1];
    const bool conv = rr < st->threshold;
    const double absOld = st->hist[up.k & 1];
    if (blockIdx.x == 0 && tid == 0) {
      st->rn2 = rr;
      if (conv) {
        st->done = 1;
      } else {
        st->absNew = rz;
        st->hist[(up.k + 1) & 1] = rz;
        st->beta = rz / absOld;
        st->iters += 1;
        st->qz = v[2];
        st->bzq = v[3];
        st->qq = v[4];
      }
      st->tol_error = sqrt(rr / st->rhsNorm2);
    }
    if (conv) return;
    beta = (float)(rz / absOld);
  } else {
    beta = (float)st->beta;
  }
  // the work outside the tiles, spread over all blocks: line unknowns, and p of the
  // vertices without hexes
  for (int i = blockIdx.x * kTileHex + tid; i < sd.nL; i += gridDim.x * kTileHex)
    line_row(sd, nV, z, pprev, pnew, out, beta, i);
  for (int i = blockIdx.x * kTileHex + tid; i < niso; i += gridDim.x * kTileHex) {
    const int g = iso[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) pnew[(size_t)d * nV + g] = z[(size_t)d * nV + g] + beta * pprev[(size_t)d * nV + g];
  }
  double part = 0.0;  // this thread's element quadratic forms, over the block's tiles in order
  while (tile < tend) {
    // stage p of the tile's vertices
#pragma unroll
    for (int j = 0; j < kMaxLV / kTileHex; ++j) {
      const int lv = tid + j * kTileHex;
      if (lv < P.nlv) {
        const int g = P.g[j];
        float pv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) pv[d] = z[(size_t)d * nV + g] + beta * pprev[(size_t)d * nV + g];
        sp[lv] = make_float4(pv[0], pv[1], pv[2], 0.0f);
        acc[lv] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (P.first & (1u << j)) {
#pragma unroll
          for (int d = 0; d < 3; ++d) pnew[(size_t)d * nV + g] = pv[d];
        }
      }
    }
    __syncthreads();
    const bool active = tid < P.nh;
    float c[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) c[k][d] = 0.0f;
    if (active) {
      float q[8][3];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float4 v = sp[P.li[k]];
        q[k][0] = v.x; q[k][1] = v.y; q[k][2] = v.z;
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        float E[3][3];
        tet_edges(q, t, E);
        tet_forces_cube(E, P.kr + 6 * t, t, c);
      }
      float e = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) e += q[k][d] * c[k][d];
      part += (double)e;
    }
    // this tile's indices, then the next tile's loads
    const int cur = tile, nh = P.nh, nlv = P.nlv;
    int li[8], g[kMaxLV / kTileHex], sl[kMaxLV / kTileHex];
#pragma unroll
    for (int k = 0; k < 8; ++k) li[k] = P.li[k];
#pragma unroll
    for (int j = 0; j < kMaxLV / kTileHex; ++j) { g[j] = P.g[j]; sl[j] = P.s[j]; }
    tile += nb;
    if (tile < tend) pipe_fetch(tm, K, tile, P);
    scatter_tile4(tm, cur, nh, active, li, c, acc);
#pragma unroll
    for (int j = 0; j < kMaxLV / kTileHex; ++j) {
      const int lv = tid + j * kTileHex;
      if (lv < nlv) {
        halo[sl[j]] = acc[lv];
      }
    }
    // the next tile's staging overwrites sp and acc
    __syncthreads();
  }
  const double bs = block_sum(part, red);
  if (tid == 0) partial[blockIdx.x] = bs;
}

// p of the vertices without hexes when the operator pass left them (niso > 0), then
// alpha = absNew / (p.E p + z.R z + 2 beta z.R q + beta^2 q.R q)
__global__ __launch_bounds__(1024) void cg32_alpha_kernel(CGState* st, const double* part_tile, int ntiles, int nV,
                                                          const int* __restrict__ iso, int niso,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ pprev, float* __restrict__ pnew) {
  __shared__ double red[16];
  if (st->done) return;
  const float beta = (float)st->beta;
  for (int k = threadIdx.x; k < niso; k += blockDim.x) {
    const int g = iso[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) pnew[(size_t)d * nV + g] = z[(size_t)d * nV + g] + beta * pprev[(size_t)d * nV + g];
  }
  const double e = sum_parts(part_tile, ntiles, red);
  if (threadIdx.x == 0) {
    const double b = (double)beta;
    st->alpha = st->absNew / (e + st->qz + 2.0 * b * st->bzq + b * b * st->qq);
  }
}

// line unknowns (the operator pass of general references leaves them to this kernel)
__global__ __launch_bounds__(256) void cg32_lines_kernel(SoftDev sd, int nV, const float* __restrict__ z,
                                                         const float* __restrict__ pprev, float* __restrict__ pnew,
                                                         float* __restrict__ out, const CGState* st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < sd.nL) line_row(sd, nV, z, pprev, pnew, out, (float)st->beta, i);
}

// a grid-stride pass over the vertices (each thread also updates its vertices' line unknowns:
// each line unknown belongs to exactly one vertex). With ntiles > 0 every block first forms
// alpha itself from the operator pass's ntiles partials, as cg32_alpha_kernel would (the
// same sum in every block, so every block holds the same alpha).
__global__ __launch_bounds__(256) void cg32_update_kernel(TileMesh tm, SoftDev sd, double prox,
                                                          const float4* __restrict__ halo,
                                                          const float* __restrict__ p, const float* __restrict__ out,
                                                          const float* __restrict__ invd, float* __restrict__ x,
                                                          float* __restrict__ r, float* __restrict__ z,
                                                          const CGState* st, const double* part_tile, int ntiles,
                                                          double* part_rr, double* part_rz,
                                                          double* part_qz, double* part_bzq, double* part_qq) {
  __shared__ double red[5 * 16];
  if (st->done) return;
  const int nV = tm.nV;
  const size_t nA = 3 * (size_t)nV;
  float alpha;
  if (ntiles > 0) {
    const double* const pt[1] = {part_tile};
    double ev[1];
    sum_parts_n_wide<1, 4>(pt, ntiles, red, ev);
    const double e = ev[0];
    const double b = (double)(float)st->beta;
    alpha = (float)(st->absNew / (e + st->qz + 2.0 * b * st->bzq + b * b * st->qq));
  } else {
    alpha = (float)st->alpha;
  }
  double srr = 0, srz = 0, fz = 0, fzq = 0, fq = 0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nV; g += gridDim.x * blockDim.x) {
    // every load that does not wait on another goes out first: the vertex's run bounds, its
    // p, r, x, 1/diag and constraint row; then the R rows (which wait only on the constraint
    // row) and the run of E p places (one per tile holding the vertex, in tile order)
    const int k0 = tm.v_dst[g], k1 = tm.v_dst[g + 1];
    float pf[3], rf[3], xf[3], df[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t j = (size_t)d * nV + g;
      pf[d] = p[j];
      rf[d] = r[j];
      xf[d] = x[j];
      df[d] = invd[j];
    }
    const int2 cl = sd.cl ? reinterpret_cast<const int2*>(sd.cl)[g] : make_int2(-1, -1);
    const int ci = cl.x, ln = cl.y;
    double pv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) pv[d] = (double)pf[d];
    const double pa = ln >= 0 ? (double)p[nA + ln] : 0.0;
    double vr[3];
    vertex_rows(sd, prox, ci, ln, pv, pa, vr);
    const float4 h0 = k1 > k0 ? halo[k0] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 h1 = k1 > k0 + 1 ? halo[k0 + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    double a[3] = {0.0, 0.0, 0.0};
    if (k1 > k0) { a[0] += (double)h0.x; a[1] += (double)h0.y; a[2] += (double)h0.z; }
    if (k1 > k0 + 1) { a[0] += (double)h1.x; a[1] += (double)h1.y; a[2] += (double)h1.z; }
    for (int k = k0 + 2; k < k1; ++k) {
      const float4 hv = halo[k];
      a[0] += (double)hv.x; a[1] += (double)hv.y; a[2] += (double)hv.z;
    }
    double zn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t j = (size_t)d * nV + g;
      x[j] = xf[d] + alpha * pf[d];
      const float rv = rf[d] - alpha * (float)(a[d] + vr[d]);
      r[j] = rv;
      const float zv = df[d] * rv;
      z[j] = zv;
      zn[d] = (double)zv;
      srr += (double)rv * (double)rv;
      srz += (double)rv * (double)zv;
    }
    double za = 0.0;
    if (ln >= 0) {
      const size_t j = nA + ln;
      x[j] += alpha * p[j];
      const float rv = r[j] - alpha * out[j];
      r[j] = rv;
      const float zv = invd[j] * rv;
      z[j] = zv;
      za = (double)zv;
      srr += (double)rv * (double)rv;
      srz += (double)rv * (double)zv;
    }
    // the next direction is z + beta p: its R sums
    fz += vertex_form(sd, prox, ci, ln, zn, za, zn, za);
    fzq += vertex_form(sd, prox, ci, ln, zn, za, pv, pa);
    fq += vertex_form(sd, prox, ci, ln, pv, pa, pv, pa);
  }
  double v[5] = {srr, srz, fz, fzq, fq};
  block_sum_n<5>(v, red);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = v[0]; part_rz[blockIdx.x] = v[1];
    part_qz[blockIdx.x] = v[2]; part_bzq[blockIdx.x] = v[3]; part_qq[blockIdx.x] = v[4];
  }
}

// cg_beta_kernel plus the three R sums of the next direction
__global__ void cg32_beta_kernel(CGState* st, const double* part_rr, const double* part_rz, const double* part_qz,
                                 const double* part_bzq, const double* part_qq, int nb) {
  __shared__ double red[5 * 16];
  const double* const parts[5] = {part_rr, part_rz, part_qz, part_bzq, part_qq};
  double v[5];
  sum_parts_n<5>(parts, nb, red, v);
  const double rr = v[0], rz = v[1], qz = v[2], bzq = v[3], qq = v[4];
  if (threadIdx.x == 0 && !st->done) {
    st->rn2 = rr;
    if (rr < st->threshold) {
      st->done = 1;
    } else {
      const double absOld = st->absNew;
      st->absNew = rz;
      st->beta = rz / absOld;
      st->iters += 1;
      st->qz = qz;
      st->bzq = bzq;
      st->qq = qq;
    }
    st->tol_error = sqrt(rr / st->rhsNorm2);
  }
}

// residual of the direction system in FP64: r64 = b - L d (L d given in FP64), partial |r64|^2
__global__ void residual64_kernel(int n, const float* b, const double* Ld, float* r32, double* part) {
  __shared__ double red[16];
  double s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double r = (double)b[i] - Ld[i];
    r32[i] = (float)r;
    s += r * r;
  }
  double a = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// FP64 right-hand side of the direction system from the FP64 adjoint gradient g (in place):
// b = -g / 2 on the 3 nV vertex rows, and the a_L rows at a_L = 0, lamL t . (x_v - o).
__global__ void rhs64_from_grad_kernel(int n3, double* g, SoftDev sd, int nV, const double* X) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n3 + sd.nL; i += gridDim.x * blockDim.x) {
    if (i < n3) {
      g[i] *= -0.5;
    } else {
      const int l = i - n3, v = sd.lv[l];
      const double* t = sd.lt + (size_t)l * 3;
      const double* o = sd.lo + (size_t)l * 3;
      double s = 0;
      for (int d = 0; d < 3; ++d) s += t[d] * (X[(size_t)d * nV + v] - o[d]);
      g[i] = sd.lamL * s;
    }
  }
}

// r = b - L d in FP64 with an FP64 right-hand side; FP32 copy of r; partial sums of r^2 and b^2
__global__ void residual64b_kernel(int n, const double* b, const double* Ld, float* r32, double* part_r,
                                   double* part_b) {
  __shared__ double red[16];
  double s = 0, sb = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double r = b[i] - Ld[i];
    if (r32) r32[i] = (float)r;
    s += r * r;
    sb += b[i] * b[i];
  }
  double a = block_sum(s, red);
  double c = block_sum(sb, red);
  if (threadIdx.x == 0) { part_r[blockIdx.x] = a; part_b[blockIdx.x] = c; }
}

__global__ void d2f_kernel(int n, const double* a, float* b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (float)a[i];
}

__global__ void f2d_kernel(int n, const float* a, double* b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (double)a[i];
}

__global__ void add_inplace_kernel(int n, float* d, const float* e) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] += e[i];
}

__global__ void sum_fin_kernel(const double* part, int n, double* out) {
  __shared__ double red[16];
  double a = sum_parts(part, n, red);
  if (threadIdx.x == 0) *out = a;
}

// ---------------------------------------------------------------------------------------
// Flip-avoiding step bound (flip_avoiding_line_search.cpp:24-298), FP64.
__device__ __forceinline__ int solve_p3(double x[3], double a, double b, double c) {
  double a2 = a * a;
  double q = (a2 - 3 * b) / 9;
  double r = (a * (2 * a2 - 9 * b) + 27 * c) / 54;
  double r2 = r * r, q3 = q * q * q;
  if (r2 < q3) {
    double t = r / sqrt(q3);
    if (t < -1) t = -1;
    if (t > 1) t = 1;
    t = acos(t);
    a /= 3;
    q = -2 * sqrt(q);
    x[0] = q * cos(t / 3) - a;
    x[1] = q * cos((t + (2 * M_PI)) / 3) - a;
    x[2] = q * cos((t - (2 * M_PI)) / 3) - a;
    return 3;
  }
  double A = -pow(fabs(r) + sqrt(r2 - q3), 1. / 3);
  if (r < 0) A = -A;
  double B = (A == 0) ? 0 : q / A;
  a /= 3;
  x[0] = (A + B) - a;
  x[1] = -0.5 * (A + B) - a;
  x[2] = 0.5 * sqrt(3.) * (A - B);
  if (fabs(x[2]) < 1e-14) { x[2] = x[1]; return 2; }
  return 1;
}

__device__ __forceinline__ double smallest_pos_quad_zero(double a, double b, double c) {
  double t1, t2;
  if (fabs(a) > 1.0e-10) {
    double delta_in = b * b - 4 * a * c;
    if (delta_in <= 0) return INFINITY;
    double delta = sqrt(delta_in);
    if (b >= 0) {
      double bd = -b - delta;
      t1 = 2 * c / bd;
      t2 = bd / (2 * a);
    } else {
      double bd = -b + delta;
      t1 = bd / (2 * a);
      t2 = (2 * c) / bd;
    }
    if (a < 0) { double t = t1; t1 = t2; t2 = t; }
    if (t1 > 0) return t2 > 0 ? t2 : t1;
    return INFINITY;
  }
  if (b == 0) return INFINITY;
  t1 = -c / b;
  return t1 > 0 ? t1 : INFINITY;
}

__device__ __forceinline__ double dd3(const double a[3], const double b[3], const double c[3]) {
  return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

// smallest positive root of the corner tet's volume along x + t d; the coefficients are
// those of det(E + tD) = det E + t(...) + t^2(...) + t^3 det D, negated to match the
// reference's 4x4 volume determinant (the sign does not move the roots but enters the
// |a| <= 1e-10 test only through its magnitude).
//
// Only roots below 1.25 can matter: the line search uses min(1, 0.8 t_min). A tet whose
// cubic p(t) = dd + c t + b t^2 + a t^3 provably keeps |p(t)| >= |dd|/2 > 0 on [0, kRootCut]
// (|c| T + |b| T^2 + |a| T^3 < |dd|/2, which also covers the quadratic branch) has no root
// there; it reports INFINITY without solving, which leaves min(1, 0.8 t_min) unchanged since
// 0.8 kRootCut > 1. NaN coefficients fail the test and take the exact path.
constexpr double kRootCut = 1.3;
__device__ __forceinline__ double tet_min_root(const double x[4][3], const double d[4][3]) {
  double e0[3], e1[3], e2[3], f0[3], f1[3], f2[3];
  for (int k = 0; k < 3; ++k) {
    e0[k] = x[1][k] - x[0][k]; e1[k] = x[2][k] - x[0][k]; e2[k] = x[3][k] - x[0][k];
    f0[k] = d[1][k] - d[0][k]; f1[k] = d[2][k] - d[0][k]; f2[k] = d[3][k] - d[0][k];
  }
  double c3 = dd3(f0, f1, f2);
  double c2 = dd3(e0, f1, f2) + dd3(f0, e1, f2) + dd3(f0, f1, e2);
  double c1 = dd3(f0, e1, e2) + dd3(e0, f1, e2) + dd3(e0, e1, f2);
  double c0 = dd3(e0, e1, e2);
  double a = -c3, b = -c2, c = -c1, dd = -c0;
  const double T = kRootCut;
  if (fabs(c) * T + fabs(b) * (T * T) + fabs(a) * (T * T * T) < 0.5 * fabs(dd)) return INFINITY;
  if (fabs(a) <= 1.e-10) return smallest_pos_quad_zero(b, c, dd);
  b /= a; c /= a; dd /= a;
  double res[3];
  int n = solve_p3(res, b, c, dd);
  if (n == 1) return (res[0] >= 0) ? res[0] : INFINITY;
  if (n == 2) {
    double mx = fmax(res[0], res[1]), mn = fmin(res[0], res[1]);
    if (mn > 0) return mn;
    if (mx > 0) return mx;
    return INFINITY;
  }
  // sort 3
  double r0 = res[0], r1 = res[1], r2 = res[2], t;
  if (r1 < r0) { t = r0; r0 = r1; r1 = t; }
  if (r2 < r1) { t = r1; r1 = r2; r2 = t; }
  if (r1 < r0) { t = r0; r0 = r1; r1 = t; }
  if (r0 > 0) return r0;
  if (r1 > 0) return r1;
  if (r2 > 0) return r2;
  return INFINITY;
}

// Point on the search line. DIFF: dst is the global-step solution, the direction is
// dst - cur (flip_avoiding_line_search.cpp:307); otherwise dst already is the direction.
template <bool DIFF, class DT>
__device__ __forceinline__ double line_point(double a, DT q, double t) {
  if (DIFF) return a + t * ((double)q - a);
  return a + t * (double)q;
}
template <bool DIFF, class DT>
__device__ __forceinline__ double line_dir(double a, DT q) {
  if (DIFF) return (double)q - a;
  return (double)q;
}

// direction d = dst - cur, taken per vertex on the fly
template <bool DIFF, class DT>
__global__ __launch_bounds__(256) void max_step_kernel(const double* cur, const DT* dst, int nV, const int* conn,
                                                       int nH, double* part) {
  __shared__ double red[16];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double m = INFINITY;
  if (h < nH) {
    int vid[8];
    double p[8][3], q[8][3];
    gather8(cur, nV, conn, nH, h, vid, p);
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) q[k][d] = line_dir<DIFF>(p[k][d], dst[(size_t)d * nV + vid[k]]);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      double x[4][3], dd[4][3];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        pick_row<double, 3>(p, kHexTet[c][j], x[j]);
        pick_row<double, 3>(q, kHexTet[c][j], dd[j]);
      }
      m = fmin(m, tet_min_root(x, dd));
    }
  }
  double s = block_min(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void min_fin_kernel(const double* part, int n, double* out) {
  __shared__ double red[16];
  double m = INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, part[i]);
  m = block_min(m, red);
  if (threadIdx.x == 0) *out = m;
}

// ---------------------------------------------------------------------------------------
// Energy of X = cur + step (dst - cur) (compute_energy_with_jacobians); per-block partials.
template <class T, bool DIFF, class DT>
__global__ __launch_bounds__(256) void energy_kernel(RefOp ro, const double* cur, const DT* dst, double step, int nV, const int* conn,
                              int nH, int energy, double M, double exp_f, double* part) {
  __shared__ double red[16];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (h < nH) {
    int vid[8];
    double p[8][3];
    gather8(cur, nV, conn, nH, h, vid, p);
    if (dst) {
      for (int k = 0; k < 8; ++k)
        for (int d = 0; d < 3; ++d) p[k][d] = line_point<DIFF>(p[k][d], dst[(size_t)d * nV + vid[k]], step);
    }
    T pt[8][3];
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) pt[k][d] = (T)p[k][d];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      T B[3][3], E[3][3], J[3][3];
      tet_B(ro, nH, h, c, B);
      tet_edges(pt, c, E);
      matmul3(E, B, J);
      e += M * (double)tet_energy(J, energy, (T)exp_f);
    }
  }
  double s = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// compute_soft_const_energy (slim_m.cpp:735-790) at X = cur + step (dst - cur)
template <bool DIFF, class DT>
__global__ void soft_energy_kernel(SoftDev sd, const double* cur, const DT* dst, double step, int nV,
                                   const double* aL, double* part) {
  __shared__ double red[16];
  const int tot = sd.nb + sd.nC + sd.nT + sd.nL;
  double e = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
    int j = i, v;
    int kind;
    if (j < sd.nb) { kind = 0; v = sd.b_ids[j]; }
    else if ((j -= sd.nb) < sd.nC) { kind = 1; v = sd.c_ids[j]; }
    else if ((j -= sd.nC) < sd.nT) { kind = 2; v = sd.t_ids[j]; }
    else { j -= sd.nT; kind = 3; v = sd.lv[j]; }
    double x[3];
    for (int d = 0; d < 3; ++d) {
      double a = cur[(size_t)d * nV + v];
      x[d] = dst ? line_point<DIFF>(a, dst[(size_t)d * nV + v], step) : a;
    }
    if (kind == 0) {
      double q0 = sd.b_pos[j * 3] - x[0], q1 = sd.b_pos[j * 3 + 1] - x[1], q2 = sd.b_pos[j * 3 + 2] - x[2];
      e += sd.lamP * (q0 * q0 + q1 * q1 + q2 * q2);
    } else if (kind == 1) {
      double q0 = sd.c_pos[j * 3] - x[0], q1 = sd.c_pos[j * 3 + 1] - x[1], q2 = sd.c_pos[j * 3 + 2] - x[2];
      e += sd.lamC * (q0 * q0 + q1 * q1 + q2 * q2);
    } else if (kind == 2) {
      double a = sd.t_n[j * 3] * x[0] + sd.t_n[j * 3 + 1] * x[1] + sd.t_n[j * 3 + 2] * x[2] - sd.t_d[j];
      e += sd.lamT * a * a;
    } else {
      double s = 0;
      for (int d = 0; d < 3; ++d) {
        double q = x[d] - aL[j] * sd.lt[(size_t)j * 3 + d] - sd.lo[(size_t)j * 3 + d];
        s += q * q;
      }
      e += sd.lamL * s;
    }
  }
  double s = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// two partial arrays -> (quality, soft) on device
__global__ void energy_fin_kernel(const double* pq, int nq, const double* ps, int ns, double* out2) {
  __shared__ double red[16];
  double a = sum_parts(pq, nq, red);
  double b = sum_parts(ps, ns, red);
  if (threadIdx.x == 0) { out2[0] = a; out2[1] = b; }
}

// cur += step (dst - cur)  (or cur += step d)
template <bool DIFF, class DT>
__global__ void step_kernel(int n, double* cur, const DT* dst, double step) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    cur[i] = line_point<DIFF>(cur[i], dst[i], step);
}

// ---------------------------------------------------------------------------------------
// Adjoint: symmetric-Dirichlet energy, dE/dx (all tets + soft terms) and Hessian-vector
// products; corner Jacobians are recomputed from x, nothing is stored between calls.
template <class T, class P>
struct GradOp {
  RefOp ro; int nV, nH; const int* conn; const P* X; P* G; double M;
  SoftDev sd; const double* aL; double* epart_unused;
  // corner tet c on rows T0..T3 of pd / o (compile-time indices, see rotate_rings)
  template <int T0, int T1, int T2, int T3>
  __device__ __forceinline__ void tet(int h, int c, const double pd[8][3], T o[8][3], double& e) const {
    constexpr int tc[4] = {T0, T1, T2, T3};
    T B[3][3], J[3][3], Pm[3][3];
    if constexpr (sizeof(T) == 4) {
      // FP32: J and D = J J^T - I formed in FP64 from the corners, the energy and dE/dJ in
      // the deviation form (FP32 keeps its relative precision near a rotation)
      double B64[3][3], E64[3][3], J64[3][3], D64[3][3];
      tet_B(ro, nH, h, c, B64);
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) E64[d][k] = pd[tc[k + 1]][d] - pd[T0][d];
      matmul3(E64, B64, J64);
      const double dev2 = isometry_deviation(J64, D64);
      T D[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { B[i][j] = (T)B64[i][j]; J[i][j] = (T)J64[i][j]; D[i][j] = (T)D64[i][j]; }
      if (dev2 < 0.25) e += M * (6.0 + (double)sd_energy_grad_dev(J, D, Pm));   // see AdjLocalOp32
      else e += M * (double)sd_energy_grad(J, Pm);
    } else {
      T E[3][3];
      tet_B(ro, nH, h, c, B);
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) E[d][k] = (T)pd[tc[k + 1]][d] - (T)pd[T0][d];
      matmul3(E, B, J);
      e += M * (double)sd_energy_grad(J, Pm);
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Pm[i][j] *= (T)M;
    // scatter_tet's sums in its order: row k + 1 gets F B_k, row 0 minus it
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T f[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] = Pm[d][0] * B[k][0] + Pm[d][1] * B[k][1] + Pm[d][2] * B[k][2];
#pragma unroll
      for (int d = 0; d < 3; ++d) o[tc[k + 1]][d] += f[d];
#pragma unroll
      for (int d = 0; d < 3; ++d) o[T0][d] += -f[d];
    }
  }
  __device__ double hex(int h, T o[8][3]) const {
    int vid[8];
    double pd[8][3];
    gather8(X, nV, conn, nH, h, vid, pd);
    double e = 0.0;
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      tet<0, 3, 4, 1>(h, r, pd, o, e);
      rotate_rings(pd);
      rotate_rings(o);
    }
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      tet<4, 7, 5, 0>(h, 4 + r, pd, o, e);
      rotate_rings(pd);
      rotate_rings(o);
    }
    return e;
  }
  // soft gradient 2 (S x - b - lamL a t) for constrained vertices
  __device__ double finish(int g, T acc[3]) const {
    double x[3] = {(double)X[g], (double)X[(size_t)nV + g], (double)X[2 * (size_t)nV + g]};
    double r[3] = {(double)acc[0], (double)acc[1], (double)acc[2]};
    const int ci = sd.cidx ? sd.cidx[g] : -1;
    if (ci >= 0) {
      const double* S = sd.S + (size_t)ci * 6;
      const double* b = sd.b + (size_t)ci * 3;
      double sx[3] = {S[0] * x[0] + S[3] * x[1] + S[4] * x[2] - b[0], S[3] * x[0] + S[1] * x[1] + S[5] * x[2] - b[1],
                      S[4] * x[0] + S[5] * x[1] + S[2] * x[2] - b[2]};
      const int li = sd.cline[ci];
      if (li >= 0)
        for (int d = 0; d < 3; ++d) sx[d] -= sd.lamL * aL[li] * sd.lt[(size_t)li * 3 + d];
      for (int d = 0; d < 3; ++d) r[d] += 2.0 * sx[d];
    }
    for (int d = 0; d < 3; ++d) G[(size_t)d * nV + g] = (P)r[d];
    return 0.0;
  }
};

template <class T, class P>
struct HvpOp {
  RefOp ro; int nV, nH; const int* conn; const P* X; const P* Vd; P* Hv; double M; SoftDev sd;
  __device__ double hex(int h, T o[8][3]) const {
    int vid[8];
    double pd[8][3], vd[8][3];
    gather8(X, nV, conn, nH, h, vid, pd);
    gather8(Vd, nV, conn, nH, h, vid, vd);
    T p[8][3], v[8][3];
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) { p[k][d] = (T)pd[k][d]; v[k][d] = (T)vd[k][d]; }
#pragma unroll 1
    for (int c = 0; c < 8; ++c) {
      T B[3][3], E[3][3], J[3][3], Ed[3][3], dJ[3][3], Hd[3][3];
      tet_B(ro, nH, h, c, B);
      tet_edges(p, c, E);
      matmul3(E, B, J);
      tet_edges(v, c, Ed);
      matmul3(Ed, B, dJ);
      sd_hess_dir(J, dJ, Hd);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Hd[i][j] *= (T)M;
      scatter_tet(Hd, B, c, o);
    }
    return 0.0;
  }
  __device__ double finish(int g, T acc[3]) const {
    double vv[3] = {(double)Vd[g], (double)Vd[(size_t)nV + g], (double)Vd[2 * (size_t)nV + g]};
    double r[3] = {(double)acc[0], (double)acc[1], (double)acc[2]};
    const int ci = sd.cidx ? sd.cidx[g] : -1;
    if (ci >= 0) {
      const double* S = sd.S + (size_t)ci * 6;
      r[0] += 2.0 * (S[0] * vv[0] + S[3] * vv[1] + S[4] * vv[2]);
      r[1] += 2.0 * (S[3] * vv[0] + S[1] * vv[1] + S[5] * vv[2]);
      r[2] += 2.0 * (S[4] * vv[0] + S[5] * vv[1] + S[2] * vv[2]);
    }
    for (int d = 0; d < 3; ++d) Hv[(size_t)d * nV + g] = (P)r[d];
    return 0.0;
  }
};

// ---------------------------------------------------------------------------------------
// scaled_jacobian (global_functions.cpp:2309-2358, a_jacobian :2422-2442): per-hex min of
// the 8 corner scaled Jacobians; partial min and sum.
__global__ __launch_bounds__(256) void scaled_jacobian_kernel(const double* X, int nV, const int* conn, int nH, double* per_hex,
                                       double* part_min, double* part_sum) {
  __shared__ double red[16];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double hmin = 1, hs = 0;
  if (h < nH) {
    int vid[8];
    double p[8][3];
    gather8(X, nV, conn, nH, h, vid, p);
    for (int c = 0; c < 8; ++c) {
      const double* a = p[kHexTet[c][0]];
      double M[3][3];
      for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) M[d][k] = (p[kHexTet[c][k + 1]][d] - a[d]) * .5;
      double n1 = sqrt(M[0][0] * M[0][0] + M[1][0] * M[1][0] + M[2][0] * M[2][0]);
      double n2 = sqrt(M[0][1] * M[0][1] + M[1][1] * M[1][1] + M[2][1] * M[2][1]);
      double n3 = sqrt(M[0][2] * M[0][2] + M[1][2] * M[1][2] + M[2][2] * M[2][2]);
      double sj = det3(M);
      const double Precision = 1.e-7;
      if (!(fabs(n1) < Precision || fabs(n2) < Precision || fabs(n3) < Precision)) sj /= n1 * n2 * n3;
      if (hmin > sj) hmin = sj;
    }
    if (per_hex) per_hex[h] = hmin;
    hs = hmin;
  }
  double m = block_min(h < nH ? hmin : INFINITY, red);
  double s = block_sum(hs, red);
  if (threadIdx.x == 0) { part_min[blockIdx.x] = m; part_sum[blockIdx.x] = s; }
}

// ---------------------------------------------------------------------------------------
// host launch wrappers (kernels.h)

// ---------------------------------------------------------------------------------------
// Glue (padding untangle). One thread per group of copies, summing in the caller's order.
__global__ __launch_bounds__(256) void group_sum_kernel(SoftDev sd, int nV, const double* __restrict__ x,
                                                        double* __restrict__ gs) {
  const int gr = blockIdx.x * blockDim.x + threadIdx.x;
  if (gr >= sd.nG) return;
  double s[3] = {0, 0, 0};
  for (int k = sd.g_off[gr]; k < sd.g_off[gr + 1]; ++k) {
    const int v = sd.g_ids[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] += x[(size_t)d * nV + v];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) gs[(size_t)gr * 3 + d] = s[d];
}

__global__ __launch_bounds__(256) void group_mean_kernel(SoftDev sd, int nV, const double* __restrict__ x,
                                                         double* __restrict__ mean) {
  const int gr = blockIdx.x * blockDim.x + threadIdx.x;
  if (gr >= sd.nG) return;
  const int a = sd.g_off[gr], b = sd.g_off[gr + 1];
  double s[3] = {0, 0, 0};
  for (int k = a; k < b; ++k) {
    const int v = sd.g_ids[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] += x[(size_t)d * nV + v];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) mean[(size_t)d * sd.nG + gr] = s[d] / (double)(b - a);
}

// compute_soft_const_energy's glue term (slim_m.cpp:748-766) at y = cur + t (dst - cur): for
// each copy j, lamG |y_j - (sum of the other copies, in order) / (k - 1)|^2
__global__ __launch_bounds__(256) void glue_energy_kernel(SoftDev sd, int nV, const double* __restrict__ cur,
                                                          const double* __restrict__ dst, double step,
                                                          double* part) {
  __shared__ double red[16];
  const int gr = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (gr < sd.nG) {
    const int a = sd.g_off[gr], b = sd.g_off[gr + 1];
    if (b - a > 1) {
      const double coef = 1.0 / (double)(b - a - 1);
      for (int j = a; j < b; ++j) {
        const int vj = sd.g_ids[j];
        double cv[3] = {0, 0, 0}, yj[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double c0 = cur[(size_t)d * nV + vj];
          yj[d] = dst ? c0 + step * (dst[(size_t)d * nV + vj] - c0) : c0;
        }
        for (int k = a; k < b; ++k) {
          if (k == j) continue;
          const int vk = sd.g_ids[k];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double c0 = cur[(size_t)d * nV + vk];
            cv[d] += dst ? c0 + step * (dst[(size_t)d * nV + vk] - c0) : c0;
          }
        }
        double q = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double t = yj[d] - cv[d] * coef;
          q += t * t;
        }
        e += sd.lamG * q;
      }
    }
  }
  const double s = block_sum(e, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void glue_split_kernel(int n, const double* invd, const double* cur, double* xfix, double* x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const bool f = invd[i] == 0.0;
    xfix[i] = f ? cur[i] : 0.0;
    x[i] = f ? 0.0 : cur[i];
  }
}

__global__ void glue_rhs_kernel(int n, const double* invd, double* rhs, const double* tmp) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    rhs[i] = invd[i] == 0.0 ? 0.0 : rhs[i] - tmp[i];
}

__global__ void glue_restore_kernel(int n, const double* invd, const double* cur, double* x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (invd[i] == 0.0) x[i] = cur[i];
}

void launch_group_sum(const SoftDev& sd, int nV, const double* x, double* gs, hipStream_t st) {
  if (sd.nG <= 0) return;
  hipLaunchKernelGGL(group_sum_kernel, dim3((sd.nG + 255) / 256), dim3(256), 0, st, sd, nV, x, gs);
}

void launch_group_mean(const SoftDev& sd, int nV, const double* x, double* mean, hipStream_t st) {
  if (sd.nG <= 0) return;
  hipLaunchKernelGGL(group_mean_kernel, dim3((sd.nG + 255) / 256), dim3(256), 0, st, sd, nV, x, mean);
}

void launch_glue_energy(const SoftDev& sd, int nV, const double* cur, const double* dst, double step, double* part,
                        double* out, hipStream_t st) {
  const int nb = std::max(1, (sd.nG + 255) / 256);
  hipLaunchKernelGGL(glue_energy_kernel, dim3(nb), dim3(256), 0, st, sd, nV, cur, dst, step, part);
  hipLaunchKernelGGL(sum_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part, nb, out);
}

namespace {
int flat_blocks(int n) { return std::max(1, std::min((n + 255) / 256, 4096)); }
}  // namespace

void launch_glue_split(int nV, const double* invd, const double* cur, double* xfix, double* x, hipStream_t st) {
  hipLaunchKernelGGL(glue_split_kernel, dim3(flat_blocks(3 * nV)), dim3(256), 0, st, 3 * nV, invd, cur, xfix, x);
}

void launch_glue_rhs(int nV, const double* invd, double* rhs, const double* tmp, hipStream_t st) {
  hipLaunchKernelGGL(glue_rhs_kernel, dim3(flat_blocks(3 * nV)), dim3(256), 0, st, 3 * nV, invd, rhs, tmp);
}

void launch_glue_restore(int nV, const double* invd, const double* cur, double* x, hipStream_t st) {
  hipLaunchKernelGGL(glue_restore_kernel, dim3(flat_blocks(3 * nV)), dim3(256), 0, st, 3 * nV, invd, cur, x);
}

void launch_ref_cube(const double* X, int nV, const int* conn, int nH, const uint8_t* hflag, int cube, double* invs,
                     int* n_general, hipStream_t st) {
  if (nH == 0) return;
  hipLaunchKernelGGL(ref_cube_kernel, dim3((nH + 255) / 256), dim3(256), 0, st, X, nV, conn, nH, hflag, cube, invs,
                     n_general);
}

void launch_ref_general(const double* X, int nV, const int* conn, int nH, const uint8_t* hflag, int cube, double* B,
                        hipStream_t st, const double* Xvol, int nVvol, const int* connvol) {
  if (nH == 0) return;
  hipLaunchKernelGGL(ref_general_kernel, dim3((nH + 255) / 256), dim3(256), 0, st, X, nV, conn, nH, hflag, cube, B,
                     Xvol, nVvol, connvol);
}

void launch_local_step(const TileMesh& tm, const RefOp& ro, const double* X, double* K, int energy, double M,
                       double exp_f, double prox, const SoftDev& sd, double* rhs, double* invd, double* halo,
                       hipStream_t st) {
  LocalStepOp op{ro, tm.nV, tm.nH, tm.conn, X, K, energy, M, exp_f, prox, sd, rhs, invd};
  launch_scatter<double, 6>(tm, op, halo, nullptr, nullptr, st);
}

void launch_adj_local32(const TileMesh& tm, const RefOp& ro, const double* X, float* K, double M, double prox,
                        const SoftDev& sd, float* rhs, float* invd, float* halo, double* hex_part, int* n_inverted,
                        hipStream_t st) {
  AdjLocalOp32 op{ro, tm.nV, tm.nH, tm.conn, X, K, M, prox, sd, rhs, invd, n_inverted};
  launch_scatter<float, 6>(tm, op, halo, nullptr, nullptr, st, hex_part);
  if (sd.nL > 0)
    hipLaunchKernelGGL(line_rhs32_kernel, dim3((sd.nL + 255) / 256), dim3(256), 0, st, sd, tm.nV, X,
                       rhs + 3 * (size_t)tm.nV, invd + 3 * (size_t)tm.nV);
}

template <class KT, class PT, class CT, class OT>
void launch_apply_t(const TileMesh& tm, const RefOp& ro, const KT* K, const PT* p, OT* out, double prox,
                    const SoftDev& sd, CT* halo, double* part_tile, double* part_fix, double* part_line, hipStream_t st) {
  const PT* pa = p + 3 * (size_t)tm.nV;
  ApplyOp<KT, PT, CT, OT> op{ro, tm.nV, tm.nH, tm.conn, K, p, out, prox, sd, pa, sd.nL};
  launch_scatter<CT, 3>(tm, op, halo, part_tile, part_fix, st);
  if (sd.nL > 0)
    hipLaunchKernelGGL((line_apply_kernel<PT, OT>), dim3((sd.nL + 255) / 256), dim3(256), 0, st, sd, tm.nV, p, pa,
                       out + 3 * (size_t)tm.nV, sd.nL, part_line);
}

void launch_apply(const TileMesh& tm, const RefOp& ro, const double* K, const double* p, double* out, double prox,
                  const SoftDev& sd, double* halo, double* part_tile, double* part_fix, double* part_line,
                  hipStream_t st) {
  launch_apply_t<double, double, double, double>(tm, ro, K, p, out, prox, sd, halo, part_tile, part_fix, part_line, st);
}

void launch_apply32(const TileMesh& tm, const RefOp& ro, const float* K, const float* p, float* out, double prox,
                    const SoftDev& sd, float* halo, double* part_tile, double* part_fix, double* part_line,
                    hipStream_t st) {
  launch_apply_t<float, float, float, float>(tm, ro, K, p, out, prox, sd, halo, part_tile, part_fix, part_line, st);
}

void launch_apply32_f64(const TileMesh& tm, const RefOp& ro, const float* K, const float* p, double* out, double prox,
                        const SoftDev& sd, double* halo, hipStream_t st) {
  launch_apply_t<float, float, double, double>(tm, ro, K, p, out, prox, sd, halo, nullptr, nullptr, nullptr, st);
}

template <class T>
void cg_init_t(int n, const T* rhs, const T* tmp, T* r, double* pb, double* pr, int nb, void* state, const T* invd, T* p,
               double* pabs, double tol, hipStream_t st) {
  CGState* S = (CGState*)state;
  hipLaunchKernelGGL(cg_init_kernel<T>, dim3(nb), dim3(256), 0, st, n, rhs, tmp, r, invd, pb, pr);
  hipLaunchKernelGGL(cg_init_fin_kernel, dim3(1), dim3(1024), 0, st, S, (const double*)pb, nb, (const double*)pr, tol);
  hipLaunchKernelGGL(cg_first_dir_kernel<T>, dim3(nb), dim3(256), 0, st, n, (const CGState*)S, invd, (const T*)r, p, pabs);
  hipLaunchKernelGGL(cg_absnew_fin_kernel, dim3(1), dim3(1024), 0, st, S, (const double*)pabs, nb);
}

void launch_cg_init(int n, const double* rhs, const double* tmp, double* r, double* pb, double* pr, int nb,
                    void* state, const double* invd, double* p, double* pabs, double tol, hipStream_t st) {
  cg_init_t<double>(n, rhs, tmp, r, pb, pr, nb, state, invd, p, pabs, tol, st);
}
void launch_cg_init32(int n, const float* rhs, const float* tmp, float* r, double* pb, double* pr, int nb, void* state,
                      const float* invd, float* p, double* pabs, double tol, hipStream_t st) {
  cg_init_t<float>(n, rhs, tmp, r, pb, pr, nb, state, invd, p, pabs, tol, st);
}

void launch_cg_alpha(void* state, const double* p1, int n1, const double* p2, int n2, const double* p3, int n3,
                     hipStream_t st) {
  hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(1024), 0, st, (CGState*)state, p1, n1, p2, n2, p3, n3);
}

template <class T>
void cg_update_t(int n, void* state, T* x, T* r, T* z, const T* p, const T* tmp, const T* invd, double* prr,
                 double* prz, int nb, T* p_out, hipStream_t st) {
  CGState* S = (CGState*)state;
  hipLaunchKernelGGL(cg_update_kernel<T>, dim3(nb), dim3(256), 0, st, n, (const CGState*)S, x, r, z, p, tmp, invd, prr,
                     prz);
  hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(1024), 0, st, S, (const double*)prr, (const double*)prz, nb);
  hipLaunchKernelGGL(cg_pdir_kernel<T>, dim3(nb), dim3(256), 0, st, n, (const CGState*)S, p_out, (const T*)z);
}

void launch_cg_update(int n, void* state, double* x, double* r, double* z, const double* p, const double* tmp,
                      const double* invd, double* prr, double* prz, int nb, double* p_out, hipStream_t st) {
  cg_update_t<double>(n, state, x, r, z, p, tmp, invd, prr, prz, nb, p_out, st);
}
void launch_cg_update32(int n, void* state, float* x, float* r, float* z, const float* p, const float* tmp,
                        const float* invd, double* prr, double* prz, int nb, float* p_out, hipStream_t st) {
  cg_update_t<float>(n, state, x, r, z, p, tmp, invd, prr, prz, nb, p_out, st);
}

void launch_cg_zero_if(int n, void* state, double* x, int nb, hipStream_t st) {
  hipLaunchKernelGGL(cg_zero_if_kernel<double>, dim3(nb), dim3(256), 0, st, n, (const CGState*)state, x);
}
void launch_cg_zero_if32(int n, void* state, float* x, int nb, hipStream_t st) {
  hipLaunchKernelGGL(cg_zero_if_kernel<float>, dim3(nb), dim3(256), 0, st, n, (const CGState*)state, x);
}

void launch_residual64(int n, const float* b, const double* Ld, float* r32, double* part, int nb, double* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(residual64_kernel, dim3(nb), dim3(256), 0, st, n, b, Ld, r32, part);
  hipLaunchKernelGGL(sum_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part, nb, out);
}

void launch_rhs64_from_grad(int nV, double* g, const SoftDev& sd, const double* X, hipStream_t st) {
  const int n = 3 * nV + sd.nL;
  hipLaunchKernelGGL(rhs64_from_grad_kernel, dim3(std::max(1, std::min((n + 255) / 256, 4096))), dim3(256), 0, st,
                     3 * nV, g, sd, nV, X);
}

void launch_residual64b(int n, const double* b, const double* Ld, float* r32, double* part_r, double* part_b, int nb,
                        double* out2, hipStream_t st) {
  hipLaunchKernelGGL(residual64b_kernel, dim3(nb), dim3(256), 0, st, n, b, Ld, r32, part_r, part_b);
  hipLaunchKernelGGL(energy_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part_r, nb, (const double*)part_b,
                     nb, out2);
}

void launch_add_inplace(int n, float* d, const float* e, int nb, hipStream_t st) {
  hipLaunchKernelGGL(add_inplace_kernel, dim3(nb), dim3(256), 0, st, n, d, e);
}

void launch_sum_partials(const double* part, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(sum_fin_kernel, dim3(1), dim3(1024), 0, st, part, n, out);
}

size_t cg_state_bytes() { return sizeof(CGState); }

void read_cg_state(const void* host_state, int* iters, double* tol_error, int* done) {
  const CGState* S = (const CGState*)host_state;
  *iters = S->iters;
  *tol_error = S->tol_error;
  *done = S->done;
}

void launch_d2f(int n, const double* a, float* b, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(d2f_kernel, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, st, n, a, b);
}

void launch_f2d(int n, const float* a, double* b, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(f2d_kernel, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, st, n, a, b);
}

void launch_cg32_init(int nV, int nL, const SoftDev& sd, double prox, const float* b, const float* invd, float* x,
                      float* r, float* z, float* pprev, void* state, double* part_b, double* part_rz, double* part_q,
                      double tol, hipStream_t st) {
  const int n = 3 * nV + nL;
  const int nb = std::max(1, std::min((n + 255) / 256, 2048));
  const int nq = std::max(1, (nV + 255) / 256);
  hipLaunchKernelGGL(cg32_init_kernel, dim3(nb), dim3(256), 0, st, n, b, invd, x, r, z, pprev, part_b, part_rz);
  hipLaunchKernelGGL(cg32_form0_kernel, dim3(nq), dim3(256), 0, st, sd, prox, nV, (const float*)z, part_q);
  hipLaunchKernelGGL(cg32_init_fin_kernel, dim3(1), dim3(1024), 0, st, (CGState*)state, (const double*)part_b,
                     (const double*)part_rz, nb, (const double*)part_q, nq, tol);
}

namespace {
// blocks of the persistent operator pass: as many as fit on the device at once
int pipe_grid(int ntiles) {
  static int fit[64] = {0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) return std::min(ntiles, 8);
  if (fit[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cg32_apply_pipe_kernel, kTileHex, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      per_cu = cus = 0;
    fit[dev] = std::max(8, per_cu * cus);
  }
  return std::min(ntiles, fit[dev]);
}
}  // namespace

int launch_cg32_apply(const TileMesh& tm, const RefOp& ro, const float* K, const SoftDev& sd, const int* iso, int niso,
                      const float* z, const float* pprev, float* pnew, float* out, float* halo, void* state,
                      double* part, hipStream_t st, const UpdParts* fold) {
  float4* h4 = reinterpret_cast<float4*>(halo);
  if (ro.mode == REF_CUBE) {
    const int grid = pipe_grid(tm.nTiles);
    UpdParts up{};
    if (fold) up = *fold;
    hipLaunchKernelGGL(cg32_apply_pipe_kernel, dim3(grid), dim3(kTileHex), 0, st, tm, K, z, pprev, pnew, out, h4, sd,
                       iso, niso, (CGState*)state, part, up);
    return grid;
  }
  hipLaunchKernelGGL(cg32_apply_kernel<false>, dim3(tm.nTiles), dim3(kTileHex), 0, st, tm, ro, K, z, pprev, pnew, out,
                     h4, (const CGState*)state, part);
  if (sd.nL > 0)
    hipLaunchKernelGGL(cg32_lines_kernel, dim3((sd.nL + 255) / 256), dim3(256), 0, st, sd, tm.nV, z, pprev, pnew, out,
                       (const CGState*)state);
  return -tm.nTiles;
}

// blocks of the grid-stride update pass: as many as are resident at once on the device (so
// no block waits for another to finish all its vertices), which keeps the following one-block
// reduction of their partials short
static int update_grid(int nV) {
  static int fit[64] = {0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int want = std::max(1, (nV + 255) / 256);
  if (dev < 0 || dev >= 64) return std::min(want, 1024);
  if (fit[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cg32_update_kernel, 256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      per_cu = cus = 0;
    fit[dev] = std::max(64, per_cu * cus);
  }
  return std::min(want, fit[dev]);
}

int launch_cg32_apply_folded(const TileMesh& tm, const RefOp& ro, const float* K, const SoftDev& sd, const int* iso,
                             int niso, const float* z, const float* pprev, float* pnew, float* out, float* halo,
                             void* state, double* part, const double* part_rr, const double* part_rz,
                             const double* part_qz, const double* part_bzq, const double* part_qq, int k,
                             hipStream_t st) {
  UpdParts up{{part_rr, part_rz, part_qz, part_bzq, part_qq}, update_grid(tm.nV), k};
  return launch_cg32_apply(tm, ro, K, sd, iso, niso, z, pprev, pnew, out, halo, state, part, st, &up);
}

void launch_cg32_iteration(const TileMesh& tm, const RefOp& ro, const float* K, double prox, const SoftDev& sd,
                           const int* iso, int niso, const float* z, const float* pprev, float* pnew,
                           float* out, float* halo, const float* invd, float* x, float* r, float* zout, void* state,
                           double* part_tile, double* part_rr, double* part_rz, double* part_qz, double* part_bzq,
                           double* part_qq, int k, int last, hipStream_t st) {
  CGState* cs = (CGState*)state;
  // the operator pass of cube references forms the previous iteration's beta step itself; the
  // one-block beta pass then runs only after the last update
  const bool fold_beta = ro.mode == REF_CUBE;
  UpdParts up{{part_rr, part_rz, part_qz, part_bzq, part_qq}, update_grid(tm.nV), k - 1};
  // a positive count: the operator pass did the lines and the vertices without hexes
  const int np = launch_cg32_apply(tm, ro, K, sd, iso, niso, z, pprev, pnew, out, halo, cs, part_tile, st,
                                   fold_beta && k > 0 ? &up : nullptr);
  const bool done_iso = np > 0;
  // alpha folds into the update pass unless the vertices without hexes still need p
  const bool fold = done_iso || niso == 0;
  if (!fold)
    hipLaunchKernelGGL(cg32_alpha_kernel, dim3(1), dim3(1024), 0, st, cs, (const double*)part_tile, -np, tm.nV, iso,
                       niso, z, pprev, pnew);
  const int nbu = update_grid(tm.nV);
  hipLaunchKernelGGL(cg32_update_kernel, dim3(nbu), dim3(256), 0, st, tm, sd, prox, (const float4*)halo,
                     (const float*)pnew, (const float*)out, invd, x, r, zout, (const CGState*)cs,
                     (const double*)part_tile, fold ? std::abs(np) : 0, part_rr, part_rz, part_qz, part_bzq, part_qq);
  if (!fold_beta || last)
    hipLaunchKernelGGL(cg32_beta_kernel, dim3(1), dim3(1024), 0, st, cs, (const double*)part_rr,
                       (const double*)part_rz, (const double*)part_qz, (const double*)part_bzq,
                       (const double*)part_qq, nbu);
}

void launch_max_step(const double* cur, const double* dst, int nV, const int* conn, int nH, double* part,
                     double* out, hipStream_t st) {
  const int nb = (nH + 255) / 256;
  if (nb == 0) return;
  hipLaunchKernelGGL((max_step_kernel<true, double>), dim3(nb), dim3(256), 0, st, cur, dst, nV, conn, nH, part);
  hipLaunchKernelGGL(min_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part, nb, out);
}

void launch_max_step_dir32(const double* cur, const float* d, int nV, const int* conn, int nH, double* part,
                           double* out, hipStream_t st) {
  const int nb = (nH + 255) / 256;
  if (nb == 0) return;
  hipLaunchKernelGGL((max_step_kernel<false, float>), dim3(nb), dim3(256), 0, st, cur, d, nV, conn, nH, part);
  hipLaunchKernelGGL(min_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part, nb, out);
}

void launch_energy(const RefOp& ro, const double* cur, const double* dst, double step, int nV, const int* conn, int nH,
                   int energy, double M, double exp_f, const SoftDev& sd, const double* aL, double* part_q,
                   double* part_s, int ns_blocks, double* out2, hipStream_t st) {
  const int nb = (nH + 255) / 256;
  if (nb > 0)
    hipLaunchKernelGGL((energy_kernel<double, true, double>), dim3(nb), dim3(256), 0, st, ro, cur, dst, step, nV, conn,
                       nH, energy, M, exp_f, part_q);
  hipLaunchKernelGGL((soft_energy_kernel<true, double>), dim3(ns_blocks), dim3(256), 0, st, sd, cur, dst, step, nV, aL,
                     part_s);
  hipLaunchKernelGGL(energy_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part_q, nb, (const double*)part_s,
                     ns_blocks, out2);
}

// The line search of the FP32 path compares energies whose difference is ~1e-8 of the total
// near isometry (e = 6 + O(|strain|^2) per tet), below what per-tet FP32 values of e resolve:
// summed from FP32, the total drifts by ~1e-8 relative as soon as the corners leave the
// lattice, which rejects good steps. The energy is therefore evaluated in FP64 along the FP32
// direction; it is a streaming pass, so FP64 costs little more than FP32 here.
void launch_energy_dir32(const RefOp& ro, const double* cur, const float* d, double step, int nV, const int* conn,
                         int nH, double M, const SoftDev& sd, const double* aL, double* part_q, double* part_s,
                         int ns_blocks, double* out2, hipStream_t st) {
  const int nb = (nH + 255) / 256;
  if (nb > 0)
    hipLaunchKernelGGL((energy_kernel<double, false, float>), dim3(nb), dim3(256), 0, st, ro, cur, d, step, nV, conn, nH,
                       (int)E_SYMMETRIC_DIRICHLET, M, 1.0, part_q);
  hipLaunchKernelGGL((soft_energy_kernel<false, float>), dim3(ns_blocks), dim3(256), 0, st, sd, cur, d, step, nV, aL,
                     part_s);
  hipLaunchKernelGGL(energy_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)part_q, nb, (const double*)part_s,
                     ns_blocks, out2);
}

void launch_step(int n, double* cur, const double* dst, double step, hipStream_t st) {
  hipLaunchKernelGGL((step_kernel<true, double>), dim3((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096), dim3(256), 0,
                     st, n, cur, dst, step);
}
void launch_step_dir32(int n, double* cur, const float* d, double step, hipStream_t st) {
  hipLaunchKernelGGL((step_kernel<false, float>), dim3((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096), dim3(256), 0,
                     st, n, cur, d, step);
}

template <class T, class P>
void launch_grad_t(const TileMesh& tm, const RefOp& ro, const P* X, P* G, double M, const SoftDev& sd, const double* aL,
                   T* halo, hipStream_t st) {
  GradOp<T, P> op{ro, tm.nV, tm.nH, tm.conn, X, G, M, sd, aL, nullptr};
  launch_scatter<T, 3>(tm, op, halo, nullptr, nullptr, st);
}

template <class T, class P>
void launch_hvp_t(const TileMesh& tm, const RefOp& ro, const P* X, const P* Vd, P* Hv, double M, const SoftDev& sd,
                  T* halo, hipStream_t st) {
  HvpOp<T, P> op{ro, tm.nV, tm.nH, tm.conn, X, Vd, Hv, M, sd};
  launch_scatter<T, 3>(tm, op, halo, nullptr, nullptr, st);
}

void launch_grad(int prec, const TileMesh& tm, const RefOp& ro, const void* X, void* G, double M, const SoftDev& sd,
                 const double* aL, void* halo, hipStream_t st) {
  if (prec == 64)
    launch_grad_t<double, double>(tm, ro, (const double*)X, (double*)G, M, sd, aL, (double*)halo, st);
  else
    launch_grad_t<float, float>(tm, ro, (const float*)X, (float*)G, M, sd, aL, (float*)halo, st);
}

void launch_hvp(int prec, const TileMesh& tm, const RefOp& ro, const void* X, const void* Vd, void* Hv, double M,
                const SoftDev& sd, void* halo, hipStream_t st) {
  if (prec == 64)
    launch_hvp_t<double, double>(tm, ro, (const double*)X, (const double*)Vd, (double*)Hv, M, sd, (double*)halo, st);
  else
    launch_hvp_t<float, float>(tm, ro, (const float*)X, (const float*)Vd, (float*)Hv, M, sd, (float*)halo, st);
}

void launch_scaled_jacobian(const double* X, int nV, const int* conn, int nH, double* per_hex, double* pmin,
                            double* psum, double* out2, hipStream_t st) {
  const int nb = (nH + 255) / 256;
  if (nb == 0) return;
  hipLaunchKernelGGL(scaled_jacobian_kernel, dim3(nb), dim3(256), 0, st, X, nV, conn, nH, per_hex, pmin, psum);
  hipLaunchKernelGGL(min_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)pmin, nb, out2);
  hipLaunchKernelGGL(energy_fin_kernel, dim3(1), dim3(1024), 0, st, (const double*)psum, nb, (const double*)psum, 0,
                     out2 + 1);
}

}  // namespace hexopt
