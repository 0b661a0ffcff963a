// device_math.h -- per-corner-tet algebra for the hex optimisation kernels (gfx950).
//
// Everything here is register-resident 3x3 work on one corner tet: the Jacobian from the
// tet's three edge vectors, a one-sided Jacobi SVD, the SLIM weights of every energy of
// slim_m.cpp:236-377, the energies of slim_m.cpp:855-913 and the symmetric-Dirichlet
// derivatives used by the adjoint kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hexopt {

// global_types.h:163-173: corner tet c = {c, n1, n2, n3}
__constant__ static const int kHexTet[8][4] = {
    {0, 3, 4, 1}, {1, 0, 5, 2}, {2, 1, 6, 3}, {3, 2, 7, 0},
    {4, 7, 5, 0}, {5, 4, 6, 1}, {6, 5, 7, 2}, {7, 6, 4, 3}};

// For the cube reference of compute_referenceMesh(cube=true), the reference gradient
// operator of corner tet c is invs_c * Pi_c with Pi_c the signed permutation below:
// J = [x_n1-x_c, x_n2-x_c, x_n3-x_c] * Pi_c * invs_c, invs_c = 1/cbrt(6 vol_c).
// (Each Jacobian column is the hex edge through corner c along one reference axis.)
__constant__ static const int kPi[8][3][3] = {
    {{0, 1, 0}, {0, 0, 1}, {1, 0, 0}},    {{-1, 0, 0}, {0, 0, 1}, {0, 1, 0}},
    {{0, -1, 0}, {0, 0, 1}, {-1, 0, 0}},  {{1, 0, 0}, {0, 0, 1}, {0, -1, 0}},
    {{0, 1, 0}, {1, 0, 0}, {0, 0, -1}},   {{-1, 0, 0}, {0, 1, 0}, {0, 0, -1}},
    {{0, -1, 0}, {-1, 0, 0}, {0, 0, -1}}, {{1, 0, 0}, {0, -1, 0}, {0, 0, -1}}};

enum EnergyKind { E_ARAP = 0, E_LOG_ARAP, E_SYMMETRIC_DIRICHLET, E_CONFORMAL, E_EXP_CONFORMAL, E_EXP_SD };

template <class T> struct Mat3 { T m[3][3]; };

template <class T> __device__ __forceinline__ T det3(const T a[3][3]) {
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

template <class T> __device__ __forceinline__ T dsqrt(T x) { return sqrt(x); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return sqrtf(x); }

// Reciprocal and reciprocal square root for the Jacobi rotations: exact division and sqrt
// in FP64; in FP32 the hardware's 1-ulp approximations, well inside the FP32 SVD's own
// rounding (its orthogonality test stops at 1e-6).
template <class T> __device__ __forceinline__ T rot_rcp(T x) { return T(1) / x; }
template <> __device__ __forceinline__ float rot_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <class T> __device__ __forceinline__ T rot_rsqrt(T x) { return T(1) / sqrt(x); }
template <> __device__ __forceinline__ float rot_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// One-sided Jacobi SVD of a 3x3 matrix A = U diag(s) V^T, s >= 0 descending, U, V
// orthogonal (the contract of Eigen::JacobiSVD that igl::polar_svd relies on). Column pairs
// of B = A V are rotated until orthogonal; the column norms are the singular values.
// Convergence: a sweep that finds every column pair orthogonal to `tol` (relative) ends the
// iteration. In FP32 the measure settles at rounding level (~1e-7), so its tolerance sits
// just above that; with quadratic convergence the last rotating sweep already reached it.
template <class T>
__device__ void svd3(const T A[3][3], T U[3][3], T s[3], T V[3][3]) {
  const T tol = sizeof(T) == 8 ? T(1e-17) : T(1e-6);
  const int max_sweeps = sizeof(T) == 8 ? 12 : 8;
  T B[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) { B[i][j] = A[i][j]; V[i][j] = (i == j) ? T(1) : T(0); }
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    T off = 0;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0;
      const int q = pq == 0 ? 1 : 2;
      T alpha = B[0][p] * B[0][p] + B[1][p] * B[1][p] + B[2][p] * B[2][p];
      T beta = B[0][q] * B[0][q] + B[1][q] * B[1][q] + B[2][q] * B[2][q];
      T gamma = B[0][p] * B[0][q] + B[1][p] * B[1][q] + B[2][p] * B[2][q];
      if (gamma == T(0)) continue;
      T conv, zeta, t, c;
      if (sizeof(T) == 8) {
        conv = fabs(gamma) / dsqrt(alpha * beta);
        zeta = (beta - alpha) / (T(2) * gamma);
        t = (zeta >= T(0) ? T(1) : T(-1)) / (fabs(zeta) + dsqrt(T(1) + zeta * zeta));
        c = T(1) / dsqrt(T(1) + t * t);
      } else {
        conv = fabs(gamma) * rot_rsqrt(alpha * beta);
        zeta = (beta - alpha) * rot_rcp(T(2) * gamma);
        t = (zeta >= T(0) ? T(1) : T(-1)) * rot_rcp(fabs(zeta) + dsqrt(T(1) + zeta * zeta));
        c = rot_rsqrt(T(1) + t * t);
      }
      if (!(conv > tol)) continue;
      off = fmax(off, conv);
      const T sn = c * t;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        T bp = B[i][p], bq = B[i][q];
        B[i][p] = c * bp - sn * bq;
        B[i][q] = sn * bp + c * bq;
        T vp = V[i][p], vq = V[i][q];
        V[i][p] = c * vp - sn * vq;
        V[i][q] = sn * vp + c * vq;
      }
    }
    if (off < tol) break;
  }
  T n[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) n[j] = dsqrt(B[0][j] * B[0][j] + B[1][j] * B[1][j] + B[2][j] * B[2][j]);
  // sort descending by a 3-element compare-and-swap network on whole columns (ties keep the
  // original order); columns move by selects, so every array stays in registers
  T Vs[3][3], Bs[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s[k] = n[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) { Vs[i][k] = V[i][k]; Bs[i][k] = B[i][k]; }
  }
  auto cswap = [&](int a, int b) {  // a < b are constants after unrolling
    const bool sw = s[b] > s[a];
    T t = s[a]; s[a] = sw ? s[b] : t; s[b] = sw ? t : s[b];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T v = Vs[i][a]; Vs[i][a] = sw ? Vs[i][b] : v; Vs[i][b] = sw ? v : Vs[i][b];
      T w = Bs[i][a]; Bs[i][a] = sw ? Bs[i][b] : w; Bs[i][b] = sw ? w : Bs[i][b];
    }
  };
  cswap(0, 1);
  cswap(1, 2);
  cswap(0, 1);
  const T tiny = sizeof(T) == 8 ? T(1e-300) : T(1e-37);
  T u[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T inv = s[k] > tiny ? T(1) / s[k] : T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) u[k][i] = Bs[i][k] * inv;
  }
  if (!(s[1] > tiny)) {
    if (!(s[0] > tiny)) { u[0][0] = 1; u[0][1] = 0; u[0][2] = 0; }
    T tv[3] = {fabs(u[0][0]) < T(0.9) ? T(1) : T(0), fabs(u[0][0]) < T(0.9) ? T(0) : T(1), T(0)};
    T bx = u[0][1] * tv[2] - u[0][2] * tv[1], by = u[0][2] * tv[0] - u[0][0] * tv[2], bz = u[0][0] * tv[1] - u[0][1] * tv[0];
    T bn = T(1) / dsqrt(bx * bx + by * by + bz * bz);
    u[1][0] = bx * bn; u[1][1] = by * bn; u[1][2] = bz * bn;
  }
  if (!(s[2] > tiny)) {
    u[2][0] = u[0][1] * u[1][2] - u[0][2] * u[1][1];
    u[2][1] = u[0][2] * u[1][0] - u[0][0] * u[1][2];
    u[2][2] = u[0][0] * u[1][1] - u[0][1] * u[1][0];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) { U[i][k] = u[k][i]; V[i][k] = Vs[i][k]; }
}

// SLIM local step for one corner tet (slim_m.cpp:236-377): from J, the weight matrix W
// (symmetric, W = U diag(m) U^T) and the target R (closest rotation, or U diag(c) V^T for
// the conformal energies). exp_f is the exponential-energy factor.
template <class T>
__device__ void slim_weights(const T J[3][3], int energy, T exp_f, T W[3][3], T R[3][3]) {
  T U[3][3], s[3], V[3][3];
  svd3(J, U, s, V);
  // polar_svd: R = U V^T; a reflection is removed by negating V's last column in R only
  T Rt[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Rt[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] + U[i][2] * V[j][2];
  if (det3(Rt) < T(0)) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Rt[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] - U[i][2] * V[j][2];
  }
  const T s1 = s[0], s2 = s[1], s3 = s[2];
  T m[3] = {1, 1, 1};
  switch (energy) {
    case E_ARAP: break;
    case E_LOG_ARAP: {
      T g1 = 2 * (log(s1) / s1), g2 = 2 * (log(s2) / s2), g3 = 2 * (log(s3) / s3);
      m[0] = dsqrt(g1 / (2 * (s1 - 1))); m[1] = dsqrt(g2 / (2 * (s2 - 1))); m[2] = dsqrt(g3 / (2 * (s3 - 1)));
      break;
    }
    case E_SYMMETRIC_DIRICHLET: {
      T g1 = 2 * (s1 - pow(s1, T(-3))), g2 = 2 * (s2 - pow(s2, T(-3))), g3 = 2 * (s3 - pow(s3, T(-3)));
      m[0] = dsqrt(g1 / (2 * (s1 - 1))); m[1] = dsqrt(g2 / (2 * (s2 - 1))); m[2] = dsqrt(g3 / (2 * (s3 - 1)));
      break;
    }
    case E_EXP_SD: {
      T g1 = 2 * (s1 - pow(s1, T(-3))), g2 = 2 * (s2 - pow(s2, T(-3))), g3 = 2 * (s3 - pow(s3, T(-3)));
      T in_exp = exp_f * (s1 * s1 + pow(s1, T(-2)) + s2 * s2 + pow(s2, T(-2)) + s3 * s3 + pow(s3, T(-2)));
      T ex = exp(in_exp);
      g1 *= ex * exp_f; g2 *= ex * exp_f; g3 *= ex * exp_f;
      m[0] = dsqrt(g1 / (2 * (s1 - 1))); m[1] = dsqrt(g2 / (2 * (s2 - 1))); m[2] = dsqrt(g3 / (2 * (s3 - 1)));
      break;
    }
    case E_CONFORMAL:
    case E_EXP_CONFORMAL: {
      T cd = 9 * pow(s1 * s2 * s3, T(5. / 3.));
      T g1 = (-2 * s2 * s3 * (s2 * s2 + s3 * s3 - 2 * s1 * s1)) / cd;
      T g2 = (-2 * s1 * s3 * (s1 * s1 + s3 * s3 - 2 * s2 * s2)) / cd;
      T g3 = (-2 * s1 * s2 * (s1 * s1 + s2 * s2 - 2 * s3 * s3)) / cd;
      T cs = dsqrt(s1 * s1 + s3 * s3) / dsqrt(T(2));
      if (energy == E_EXP_CONFORMAL) {
        T ex = exp(exp_f * ((s1 * s1 + s2 * s2 + s3 * s3) / (3 * pow(s1 * s2 * s3, T(2. / 3)))));
        g1 *= ex * exp_f; g2 *= ex * exp_f; g3 *= ex * exp_f;
      }
      m[0] = dsqrt(g1 / (2 * (s1 - cs))); m[1] = dsqrt(g2 / (2 * (s2 - cs))); m[2] = dsqrt(g3 / (2 * (s3 - cs)));
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Rt[i][j] = cs * (U[i][0] * V[j][0] + U[i][1] * V[j][1] + U[i][2] * V[j][2]);
      break;
    }
  }
  const T eps = T(1e-8);
  if (fabs(s1 - 1) < eps) m[0] = 1;
  if (fabs(s2 - 1) < eps) m[1] = 1;
  if (fabs(s3 - 1) < eps) m[2] = 1;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      W[i][j] = U[i][0] * m[0] * U[j][0] + U[i][1] * m[1] * U[j][1] + U[i][2] * m[2] * U[j][2];
      R[i][j] = Rt[i][j];
    }
}

// Symmetric-Dirichlet SLIM weight matrix for the FP32 path: W = U diag(m) U^T with
// m^2 = (s - s^-3)/(s - 1) evaluated as (s + 1)(s^2 + 1)/s^3 (the same function without the
// cancellation at s -> 1 that FP32 cannot afford), and m = 1 where |s - 1| < 1e-8 as in
// slim_m.cpp:351-353.
template <class T>
__device__ void slim_weights_sd_stable(const T J[3][3], T W[3][3]) {
  T U[3][3], s[3], V[3][3];
  svd3(J, U, s, V);
  T m[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T si = s[i];
    const T is = T(1) / si;
    m[i] = fabs(si - T(1)) < T(1e-8) ? T(1) : dsqrt((si + T(1)) * (si * si + T(1)) * is * is * is);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) W[i][j] = U[i][0] * m[0] * U[j][0] + U[i][1] * m[1] * U[j][1] + U[i][2] * m[2] * U[j][2];
}

// Per-tet energy density (slim_m.cpp:855-913). The symmetric Dirichlet and conformal
// energies are evaluated in closed form from invariants (sum s_i^2 = |J|_F^2,
// sum s_i^-2 = |adj J|_F^2 / det^2, s1 s2 s3 = |det J|), the rest from the singular values.
template <class T>
__device__ T tet_energy(const T J[3][3], int energy, T exp_f) {
  if (energy == E_SYMMETRIC_DIRICHLET || energy == E_CONFORMAL || energy == E_EXP_CONFORMAL || energy == E_EXP_SD) {
    T f2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) f2 += J[i][j] * J[i][j];
    T d = det3(J);
    if (energy == E_CONFORMAL) return f2 / (3 * pow(fabs(d), T(2. / 3.)));
    if (energy == E_EXP_CONFORMAL) return exp(f2 / (3 * pow(fabs(d), T(2. / 3.))));
    T a[3][3];  // adjugate
    a[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1]; a[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    a[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1]; a[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    a[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0]; a[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    a[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0]; a[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    a[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    T a2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) a2 += a[i][j] * a[i][j];
    T e = f2 + a2 / (d * d);
    if (energy == E_EXP_SD) return exp(exp_f * e);
    return e;
  }
  T U[3][3], s[3], V[3][3];
  svd3(J, U, s, V);
  if (energy == E_ARAP) return (s[0] - 1) * (s[0] - 1) + (s[1] - 1) * (s[1] - 1) + (s[2] - 1) * (s[2] - 1);
  // LOG_ARAP
  T l1 = log(s[0]), l2 = log(fabs(s[1])), l3 = log(fabs(s[2]));
  return l1 * l1 + l2 * l2 + l3 * l3;
}

template <class T>
__device__ __forceinline__ void inverse3(const T J[3][3], T A[3][3], T& d) {
  A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1]; A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1]; A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0]; A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0]; A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  d = J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
  T id = T(1) / d;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] *= id;
}

// Symmetric Dirichlet e(J) = |J|^2 + |J^-1|^2 and dE/dJ = 2 (J - A^T A A^T), A = J^-1.
template <class T>
__device__ __forceinline__ T sd_energy_grad(const T J[3][3], T P[3][3]) {
  T A[3][3], d;
  inverse3(J, A, d);
  T AAt[3][3];  // A A^T
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) AAt[i][j] = A[i][0] * A[j][0] + A[i][1] * A[j][1] + A[i][2] * A[j][2];
  T e = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) e += J[i][j] * J[i][j] + A[i][j] * A[i][j];
  // A^T A A^T = A^T (A A^T)
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T q = A[0][i] * AAt[0][j] + A[1][i] * AAt[1][j] + A[2][i] * AAt[2][j];
      P[i][j] = T(2) * (J[i][j] - q);
    }
  return e;
}

// Symmetric Dirichlet density from the deviation from isometry D = J J^T - I, given to its own
// relative precision (formed from an FP64 Jacobian):
//   e = |J|^2 + |J^-1|^2 = tr(S) + tr(S^-1) = 6 + tr((I + D)^-1 D^2),            S = I + D
//   dE/dJ = 2 (J - J^-T J^-1 J^-T) = 2 (S - S^-1) J^-T = 2 (I + D)^-1 (2 D + D^2) J^-T.
// Near a rotation both are small and come out to FP32's relative precision, where the
// direct forms lose it to cancellation. Returns e - 6; P = dE/dJ.
template <class T>
__device__ __forceinline__ T sd_energy_grad_dev(const T J[3][3], const T D[3][3], T P[3][3]) {
  T A[3][3], dJ;
  inverse3(J, A, dJ);  // A = J^-1
  T S[3][3], Si[3][3], dS;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] = D[i][j] + (i == j ? T(1) : T(0));
  inverse3(S, Si, dS);
  T DD[3][3], N[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      DD[i][j] = D[i][0] * D[0][j] + D[i][1] * D[1][j] + D[i][2] * D[2][j];
      N[i][j] = T(2) * D[i][j] + DD[i][j];
    }
  T ex = 0, Q[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ex += Si[i][j] * DD[j][i];
      Q[i][j] = Si[i][0] * N[0][j] + Si[i][1] * N[1][j] + Si[i][2] * N[2][j];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = T(2) * (Q[i][0] * A[j][0] + Q[i][1] * A[j][1] + Q[i][2] * A[j][2]);
  return ex;
}

// D = J J^T - I in FP64 and its squared Frobenius norm
__device__ __forceinline__ double isometry_deviation(const double J[3][3], double D[3][3]) {
  double f2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      D[i][j] = J[i][0] * J[j][0] + J[i][1] * J[j][1] + J[i][2] * J[j][2] - (i == j ? 1.0 : 0.0);
      f2 += D[i][j] * D[i][j];
    }
  return f2;
}

// Second derivative of the symmetric Dirichlet density along dJ:
// d/de [2 (J - A^T A A^T)] = 2 (dJ - (dA^T A A^T + A^T dA A^T + A^T A dA^T)), dA = -A dJ A.
template <class T>
__device__ __forceinline__ void sd_hess_dir(const T J[3][3], const T dJ[3][3], T Hd[3][3]) {
  T A[3][3], d;
  inverse3(J, A, d);
  T t1[3][3], dA[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t1[i][j] = dJ[i][0] * A[0][j] + dJ[i][1] * A[1][j] + dJ[i][2] * A[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) dA[i][j] = -(A[i][0] * t1[0][j] + A[i][1] * t1[1][j] + A[i][2] * t1[2][j]);
  T AAt[3][3], AtA[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      AAt[i][j] = A[i][0] * A[j][0] + A[i][1] * A[j][1] + A[i][2] * A[j][2];
      AtA[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T a = dA[0][i] * AAt[0][j] + dA[1][i] * AAt[1][j] + dA[2][i] * AAt[2][j];        // dA^T (A A^T)
      T b = 0;                                                                          // A^T dA A^T
#pragma unroll
      for (int k = 0; k < 3; ++k) b += A[k][i] * (dA[k][0] * A[j][0] + dA[k][1] * A[j][1] + dA[k][2] * A[j][2]);
      T c = AtA[i][0] * dA[j][0] + AtA[i][1] * dA[j][1] + AtA[i][2] * dA[j][2];          // (A^T A) dA^T
      Hd[i][j] = T(2) * (dJ[i][j] - (a + b + c));
    }
}

}  // namespace hexopt
