// hexopt_oracle.cpp -- CPU restatement of the reference's hex-mesh optimisation loop,
// closest-point projection and occupancy predicates.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP product path: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark may load it. Nothing in
// the product links it or falls back to it.
//
// It restates, in plain serial FP64 C++ with no Eigen, the algorithm of:
//   slim_m.cpp                       (SLIM local/global/line-search loop, soft constraints)
//   optimization.cpp:883-900         (slim_m_opt)
//   global_functions.cpp:2309-2442   (scaled_jacobian, a_jacobian)
//   global_functions.cpp:3735-3806,3994-4022 (compute_referenceMesh, hex2cuboid_, hex2tet24)
//   extern/libigl/include/igl/grad.cpp:372-478          (grad_tet_ref, uniform=true)
//   extern/libigl/include/igl/polar_svd.cpp:36-64       (polar_svd semantics)
//   extern/libigl/include/igl/flip_avoiding_line_search.cpp:24-317
//   extern/libigl/include/igl/line_search.cpp:10-56
//   extern/libigl/external/nanogui/ext/eigen/Eigen/src/IterativeLinearSolvers/ConjugateGradient.h:27-90
//   extern/libigl/include/igl/AABB.cpp:93-200,364-441,735-779 (tree build and closest point)
//   extern/libigl/include/igl/point_simplex_squared_distance.cpp:44-109
//   extern/libigl/include/igl/pseudonormal_test.cpp:14-119
//   grid_hex_meshing.cpp:502-517 + geogram mesh_AABB.h:214-237 (cell box vs facet bbox)
//   grid_meshing/voxelization.cpp:57-93, voxelization.h:169-272 (+z ray parity)
//
// Parity of each function is pinned by the tests under tests/ (golden vectors written by
// the compiled reference, closed forms, brute force); see the table in the design notes.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <array>
#include <functional>
#include <limits>
#include <numeric>

namespace {

// global_types.h:163-173
const int kHexTet[8][4] = {
    {0, 3, 4, 1}, {1, 0, 5, 2}, {2, 1, 6, 3}, {3, 2, 7, 0},
    {4, 7, 5, 0}, {5, 4, 6, 1}, {6, 5, 7, 2}, {7, 6, 4, 3}};
// global_types.h:154-161
const int kHexFace[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1},
                            {0, 3, 7, 4}, {3, 2, 6, 7}, {1, 5, 6, 2}};

enum Energy { ARAP = 0, LOG_ARAP, SYMMETRIC_DIRICHLET, CONFORMAL, EXP_CONFORMAL, EXP_SYMMETRIC_DIRICHLET };

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// a dot product of fixed-size Eigen 3-vectors as Eigen 3.2.9 unrolls it: the first product plus
// the sum of the other two
inline double dotE(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline double det3(V3 c0, V3 c1, V3 c2) { return dot(c0, cross(c1, c2)); }

typedef double M3[3][3];

// ---------------------------------------------------------------------------------------
// 3x3 SVD by one-sided (Hestenes) Jacobi: rotate column pairs of A until they are mutually
// orthogonal; V accumulates the rotations, the column norms are the singular values.
// Output follows Eigen::JacobiSVD's contract used by igl::polar_svd: A = U diag(s) V^T,
// s >= 0 sorted descending, U and V orthogonal.
void svd3(const M3 A, M3 U, double s[3], M3 V) {
  double B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { B[i][j] = A[i][j]; V[i][j] = (i == j) ? 1.0 : 0.0; }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < 3; ++i) {
          alpha += B[i][p] * B[i][p];
          beta += B[i][q] * B[i][q];
          gamma += B[i][p] * B[i][q];
        }
        if (gamma == 0.0) continue;
        double conv = std::fabs(gamma) / std::sqrt(alpha * beta);
        if (!(conv > 1e-17)) continue;
        off = std::max(off, conv);
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < 3; ++i) {
          double bp = B[i][p], bq = B[i][q];
          B[i][p] = c * bp - sn * bq;
          B[i][q] = sn * bp + c * bq;
          double vp = V[i][p], vq = V[i][q];
          V[i][p] = c * vp - sn * vq;
          V[i][q] = sn * vp + c * vq;
        }
      }
    if (off < 1e-17) break;
  }
  double nrm[3];
  for (int j = 0; j < 3; ++j) nrm[j] = std::sqrt(B[0][j] * B[0][j] + B[1][j] * B[1][j] + B[2][j] * B[2][j]);
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + 3, [&](int a, int b) { return nrm[a] > nrm[b]; });
  double Vs[3][3], Bs[3][3];
  for (int k = 0; k < 3; ++k) {
    s[k] = nrm[ord[k]];
    for (int i = 0; i < 3; ++i) { Vs[i][k] = V[i][ord[k]]; Bs[i][k] = B[i][ord[k]]; }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = Vs[i][j];
  // U columns: normalised B columns; complete an orthonormal basis where s is tiny.
  V3 u[3];
  for (int k = 0; k < 3; ++k) {
    V3 c = {Bs[0][k], Bs[1][k], Bs[2][k]};
    u[k] = (s[k] > 1e-300) ? mul(c, 1.0 / s[k]) : V3{0, 0, 0};
  }
  if (!(s[1] > 1e-300)) {
    V3 a = u[0];
    if (!(s[0] > 1e-300)) a = {1, 0, 0};
    u[0] = a;
    V3 t = std::fabs(a.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    V3 b = cross(a, t);
    u[1] = mul(b, 1.0 / norm(b));
  }
  if (!(s[2] > 1e-300)) u[2] = cross(u[0], u[1]);
  for (int k = 0; k < 3; ++k) { U[0][k] = u[k].x; U[1][k] = u[k].y; U[2][k] = u[k].z; }
}

double det33(const M3 a) {
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

// igl::polar_svd (polar_svd.cpp:36-64): R = U V^T, a reflection removed by negating the last
// column of V (the one belonging to the smallest singular value) in R only.
void polar_svd3(const M3 A, M3 R, M3 U, double s[3], M3 V) {
  svd3(A, U, s, V);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] + U[i][2] * V[j][2];
  if (det33(R) < 0) {
    // V itself is returned unchanged; only R uses the reflected copy (polar_svd.cpp:52-58)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] - U[i][2] * V[j][2];
  }
}

// ---------------------------------------------------------------------------------------
// Per-tet gradient operator (grad.cpp:372-478 with uniform=true). G is 3x4: G[c][j] is the
// weight of tet vertex j in the derivative along reference axis c.
void grad_tet_ref(const V3 x[4], const V3 r[4], double G[3][4]) {
  // signed tet volume, igl::volume convention: -(a-d).((b-d)x(c-d))/6
  double vol = -dot(sub(x[0], x[3]), cross(sub(x[1], x[3]), sub(x[2], x[3]))) / 6.0;
  V3 e01 = sub(r[0], r[1]), e02 = sub(r[0], r[2]), e03 = sub(r[0], r[3]);
  double vref = std::fabs(dot(cross(e01, e02), e03)) / 6.0;
  double scale = std::cbrt(vol / vref);
  static const int faces[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
  static const int opposite[4] = {3, 1, 2, 0};
  for (int k = 0; k < 4; ++k) {
    V3 a = mul(sub(r[faces[k][0]], r[faces[k][1]]), scale);
    V3 b = mul(sub(r[faces[k][0]], r[faces[k][2]]), scale);
    double area = norm(cross(a, b)) / 2.0;
    a = mul(a, 1.0 / norm(a));
    b = mul(b, 1.0 / norm(b));
    V3 n = cross(a, b);
    n = mul(n, 1.0 / norm(n));
    double w = area / (3.0 * vol);
    int j = opposite[k];
    G[0][j] = w * n.x;
    G[1][j] = w * n.y;
    G[2][j] = w * n.z;
  }
}

// hex2tet24 (global_functions.cpp:3994-4022): hex volume from 24 face-centre tets.
double hex_volume24(const V3 p[8]) {
  V3 fc[6];
  for (int i = 0; i < 6; ++i) {
    V3 s = {0, 0, 0};
    for (int j = 0; j < 4; ++j) s = add(s, p[kHexFace[i][j]]);
    fc[i] = mul(s, 1.0 / 4);
  }
  V3 hc = {0, 0, 0};
  for (int i = 0; i < 8; ++i) hc = add(hc, p[i]);
  hc = mul(hc, 1.0 / 8);
  double vol = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 4; ++j) {
      V3 a = sub(p[kHexFace[i][j]], hc), b = sub(p[kHexFace[i][(j + 1) % 4]], hc), c = sub(fc[i], hc);
      vol += std::fabs(det3(a, b, c) / 6);
    }
  return vol;
}

// hex2cuboid_ (global_functions.cpp:3746-3806): the 8 reference corners of one hex.
void hex_reference_corners(const V3 p[8], bool cube, bool h_flag, V3 out[8]) {
  if (!h_flag) {
    for (int i = 0; i < 8; ++i) out[i] = p[i];
    return;
  }
  double vol = hex_volume24(p);
  double e0 = 0, e1 = 0, e2 = 0;
  if (cube && std::fabs(vol) > 1e-20) {
    e0 = e1 = e2 = std::cbrt(vol);
  } else {
    e0 += norm(sub(p[0], p[1])); e0 += norm(sub(p[3], p[2])); e0 += norm(sub(p[4], p[5])); e0 += norm(sub(p[7], p[6]));
    e0 /= 4;
    e1 += norm(sub(p[0], p[3])); e1 += norm(sub(p[1], p[2])); e1 += norm(sub(p[4], p[7])); e1 += norm(sub(p[5], p[6]));
    e1 /= 4;
    e2 += norm(sub(p[0], p[4])); e2 += norm(sub(p[1], p[5])); e2 += norm(sub(p[2], p[6])); e2 += norm(sub(p[3], p[7]));
    e2 /= 4;
    double ratio = std::cbrt(vol / (e0 * e1 * e2));
    e0 *= ratio; e1 *= ratio; e2 *= ratio;
  }
  const double c[8][3] = {{0, 0, 0}, {e0, 0, 0}, {e0, e1, 0}, {0, e1, 0},
                          {0, 0, e2}, {e0, 0, e2}, {e0, e1, e2}, {0, e1, e2}};
  for (int i = 0; i < 8; ++i) out[i] = {c[i][0], c[i][1], c[i][2]};
}

// ---------------------------------------------------------------------------------------
// SLIM state (slim_m.h:10-118 restated without Eigen). Unknowns are flattened as in
// slim_m.cpp: [x(0..nV) | y(0..nV) | z(0..nV) | a_L(0..nL)].
struct Slim {
  int nV = 0, nT = 0, energy = SYMMETRIC_DIRICHLET;
  std::vector<int> T;          // nT*4
  std::vector<double> G;       // nT*12, G[t][c][j]
  std::vector<double> Vo;      // nV*3 row-major current positions
  double M = 1.0, prox = 1e-5, mesh_area = 0;
  double exp_f = 1.0;
  // soft constraints (Tetralize_Set / Feature_Constraints)
  bool projection = false, glue = false;
  double lam_proj = 1e5, lam_glue = 10, lam_C = 0, lam_T = 0, lam_L = 0;
  std::vector<int> b; std::vector<double> bc;                 // projection
  std::vector<int> ranges;                                     // glue V_ranges
  std::vector<int> idsC; std::vector<double> C;                // corners
  std::vector<int> idsT; std::vector<double> nT_, dT;          // tangent planes
  std::vector<int> idsL; std::vector<double> axa, org;         // feature lines
  std::vector<double> aL;
  // local step outputs
  std::vector<double> W, R;    // nT*9 each, row-major 3x3
  double energy_val = 0, e_quality = 0, e_soft = 0;
  int last_cg_iters = 0;
  double last_cg_err = 0, last_step = 0, last_tmax = 0;

  int nvar() const { return 3 * nV + (glue ? 0 : (int)idsL.size()); }

  void jacobian(const double* X, int t, M3 J) const {
    const double* g = &G[(size_t)t * 12];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int j = 0; j < 4; ++j) s += g[c * 4 + j] * X[(size_t)T[t * 4 + j] * 3 + r];
        J[r][c] = s;
      }
  }

  // compute_energy_with_jacobians, 3D branch (slim_m.cpp:855-913)
  double tet_energy(const M3 J) const {
    M3 R_, U_, V_;
    double s[3];
    polar_svd3(J, R_, U_, s, V_);
    double s1 = s[0], s2 = s[1], s3 = s[2];
    switch (energy) {
      case ARAP: return std::pow(s1 - 1, 2) + std::pow(s2 - 1, 2) + std::pow(s3 - 1, 2);
      case SYMMETRIC_DIRICHLET:
        return std::pow(s1, 2) + std::pow(s1, -2) + std::pow(s2, 2) + std::pow(s2, -2) + std::pow(s3, 2) + std::pow(s3, -2);
      case EXP_SYMMETRIC_DIRICHLET:
        return std::exp(exp_f * (std::pow(s1, 2) + std::pow(s1, -2) + std::pow(s2, 2) + std::pow(s2, -2) +
                                 std::pow(s3, 2) + std::pow(s3, -2)));
      case LOG_ARAP:
        return std::pow(std::log(s1), 2) + std::pow(std::log(std::fabs(s2)), 2) + std::pow(std::log(std::fabs(s3)), 2);
      case CONFORMAL:
        return (std::pow(s1, 2) + std::pow(s2, 2) + std::pow(s3, 2)) / (3 * std::pow(s1 * s2 * s3, 2. / 3.));
      case EXP_CONFORMAL:
        return std::exp((std::pow(s1, 2) + std::pow(s2, 2) + std::pow(s3, 2)) / (3 * std::pow(s1 * s2 * s3, 2. / 3.)));
    }
    return 0;
  }

  // compute_soft_const_energy (slim_m.cpp:735-790)
  double soft_energy(const double* X) const {
    double e = 0;
    if (projection)
      for (size_t i = 0; i < b.size(); ++i) {
        double d0 = bc[i * 3] - X[(size_t)b[i] * 3], d1 = bc[i * 3 + 1] - X[(size_t)b[i] * 3 + 1],
               d2 = bc[i * 3 + 2] - X[(size_t)b[i] * 3 + 2];
        e += lam_proj * (d0 * d0 + d1 * d1 + d2 * d2);
      }
    if (glue) {
      for (size_t i = 0; i < ranges.size(); ++i) {
        int start = i ? ranges[i - 1] + 1 : 0, end = ranges[i];
        if (end == start) continue;
        double coef = 1.0 / (end - start);
        for (int j = start; j <= end; ++j) {
          double cv[3] = {0, 0, 0};
          for (int k = start; k <= end; ++k)
            if (k != j)
              for (int d = 0; d < 3; ++d) cv[d] += X[(size_t)k * 3 + d];
          double sq = 0;
          for (int d = 0; d < 3; ++d) { cv[d] *= coef; double q = X[(size_t)j * 3 + d] - cv[d]; sq += q * q; }
          e += lam_glue * sq;
        }
      }
    } else {
      for (size_t i = 0; i < idsC.size(); ++i) {
        double sq = 0;
        for (int d = 0; d < 3; ++d) { double q = C[i * 3 + d] - X[(size_t)idsC[i] * 3 + d]; sq += q * q; }
        e += lam_C * sq;
      }
      for (size_t i = 0; i < idsT.size(); ++i) {
        double a = nT_[i * 3] * X[(size_t)idsT[i] * 3] + nT_[i * 3 + 1] * X[(size_t)idsT[i] * 3 + 1] +
                   nT_[i * 3 + 2] * X[(size_t)idsT[i] * 3 + 2] - dT[i];
        e += lam_T * a * a;
      }
      for (size_t i = 0; i < idsL.size(); ++i) {
        double sq = 0;
        for (int d = 0; d < 3; ++d) {
          double q = X[(size_t)idsL[i] * 3 + d] - aL[i] * axa[i * 3 + d] - org[i * 3 + d];
          sq += q * q;
        }
        e += lam_L * sq;
      }
    }
    return e;
  }

  // compute_energy_total (slim_m.cpp:716-733): unnormalised energy; records the split.
  double energy_total(const double* X) {
    double e = 0;
    for (int t = 0; t < nT; ++t) {
      M3 J;
      jacobian(X, t, J);
      e += M * tet_energy(J);
    }
    e_quality = e;
    double es = soft_energy(X);
    e_soft = es;
    return e + es;
  }

  // update_weights_and_closest_rotations, 3D branch (slim_m.cpp:236-377)
  void local_step(const double* X) {
    W.assign((size_t)nT * 9, 0.0);
    R.assign((size_t)nT * 9, 0.0);
    const double eps = 1e-8, sqrt_2 = std::sqrt(2.0);
    for (int t = 0; t < nT; ++t) {
      M3 J, Rt, Ut, Vt;
      double s[3], m[3] = {1, 1, 1};
      jacobian(X, t, J);
      polar_svd3(J, Rt, Ut, s, Vt);
      double s1 = s[0], s2 = s[1], s3 = s[2];
      switch (energy) {
        case ARAP: m[0] = m[1] = m[2] = 1; break;
        case LOG_ARAP: {
          double g1 = 2 * (std::log(s1) / s1), g2 = 2 * (std::log(s2) / s2), g3 = 2 * (std::log(s3) / s3);
          m[0] = std::sqrt(g1 / (2 * (s1 - 1))); m[1] = std::sqrt(g2 / (2 * (s2 - 1))); m[2] = std::sqrt(g3 / (2 * (s3 - 1)));
          break;
        }
        case SYMMETRIC_DIRICHLET: {
          double g1 = 2 * (s1 - std::pow(s1, -3)), g2 = 2 * (s2 - std::pow(s2, -3)), g3 = 2 * (s3 - std::pow(s3, -3));
          m[0] = std::sqrt(g1 / (2 * (s1 - 1))); m[1] = std::sqrt(g2 / (2 * (s2 - 1))); m[2] = std::sqrt(g3 / (2 * (s3 - 1)));
          break;
        }
        case EXP_SYMMETRIC_DIRICHLET: {
          double g1 = 2 * (s1 - std::pow(s1, -3)), g2 = 2 * (s2 - std::pow(s2, -3)), g3 = 2 * (s3 - std::pow(s3, -3));
          double in_exp = exp_f * (std::pow(s1, 2) + std::pow(s1, -2) + std::pow(s2, 2) + std::pow(s2, -2) +
                                   std::pow(s3, 2) + std::pow(s3, -2));
          double ex = std::exp(in_exp);
          g1 *= ex * exp_f; g2 *= ex * exp_f; g3 *= ex * exp_f;
          m[0] = std::sqrt(g1 / (2 * (s1 - 1))); m[1] = std::sqrt(g2 / (2 * (s2 - 1))); m[2] = std::sqrt(g3 / (2 * (s3 - 1)));
          break;
        }
        case CONFORMAL:
        case EXP_CONFORMAL: {
          double cd = 9 * std::pow(s1 * s2 * s3, 5. / 3.);
          double g1 = (-2 * s2 * s3 * (std::pow(s2, 2) + std::pow(s3, 2) - 2 * std::pow(s1, 2))) / cd;
          double g2 = (-2 * s1 * s3 * (std::pow(s1, 2) + std::pow(s3, 2) - 2 * std::pow(s2, 2))) / cd;
          double g3 = (-2 * s1 * s2 * (std::pow(s1, 2) + std::pow(s2, 2) - 2 * std::pow(s3, 2))) / cd;
          double cs = std::sqrt(std::pow(s1, 2) + std::pow(s3, 2)) / sqrt_2;
          if (energy == EXP_CONFORMAL) {
            double in_exp = exp_f * ((std::pow(s1, 2) + std::pow(s2, 2) + std::pow(s3, 2)) / (3 * std::pow(s1 * s2 * s3, 2. / 3)));
            double ex = std::exp(in_exp);
            g1 *= ex * exp_f; g2 *= ex * exp_f; g3 *= ex * exp_f;
          }
          m[0] = std::sqrt(g1 / (2 * (s1 - cs))); m[1] = std::sqrt(g2 / (2 * (s2 - cs))); m[2] = std::sqrt(g3 / (2 * (s3 - cs)));
          // R <- U diag(cs) V^T (the polar-corrected V)
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Rt[i][j] = cs * (Ut[i][0] * Vt[j][0] + Ut[i][1] * Vt[j][1] + Ut[i][2] * Vt[j][2]);
          break;
        }
      }
      if (std::fabs(s1 - 1) < eps) m[0] = 1;
      if (std::fabs(s2 - 1) < eps) m[1] = 1;
      if (std::fabs(s3 - 1) < eps) m[2] = 1;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          W[(size_t)t * 9 + i * 3 + j] = Ut[i][0] * m[0] * Ut[j][0] + Ut[i][1] * m[1] * Ut[j][1] + Ut[i][2] * m[2] * Ut[j][2];
          R[(size_t)t * 9 + i * 3 + j] = Rt[i][j];
        }
    }
  }

  // The linear system of build_linear_system / add_soft_constraints (slim_m.cpp:568-714),
  // applied matrix-free: L = A^T diag(M) A + prox I + soft terms, where (A x)_t = W_t J_t(x).
  void apply(const std::vector<double>& p, std::vector<double>& out) const {
    const int n = nvar();
    out.assign(n, 0.0);
    for (int t = 0; t < nT; ++t) {
      const double* g = &G[(size_t)t * 12];
      const double* w = &W[(size_t)t * 9];
      double J[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double s = 0;
          for (int j = 0; j < 4; ++j) s += g[c * 4 + j] * p[(size_t)r * nV + T[t * 4 + j]];
          J[r][c] = s;
        }
      double Y[3][3], Z[3][3];
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) Y[a][c] = M * (w[a * 3] * J[0][c] + w[a * 3 + 1] * J[1][c] + w[a * 3 + 2] * J[2][c]);
      for (int b2 = 0; b2 < 3; ++b2)
        for (int c = 0; c < 3; ++c) Z[b2][c] = w[0 * 3 + b2] * Y[0][c] + w[1 * 3 + b2] * Y[1][c] + w[2 * 3 + b2] * Y[2][c];
      for (int j = 0; j < 4; ++j)
        for (int b2 = 0; b2 < 3; ++b2) {
          double s = 0;
          for (int c = 0; c < 3; ++c) s += g[c * 4 + j] * Z[b2][c];
          out[(size_t)b2 * nV + T[t * 4 + j]] += s;
        }
    }
    for (int i = 0; i < 3 * nV; ++i) out[i] += prox * p[i];
    if (projection)
      for (size_t i = 0; i < b.size(); ++i)
        for (int d = 0; d < 3; ++d) out[(size_t)d * nV + b[i]] += lam_proj * p[(size_t)d * nV + b[i]];
    if (glue) {
      for (size_t i = 0; i < ranges.size(); ++i) {
        int start = i ? ranges[i - 1] + 1 : 0, end = ranges[i];
        if (end == start) continue;
        double coef = -1.0 / (end - start);
        for (int d = 0; d < 3; ++d)
          for (int j = start; j <= end; ++j) {
            double row = p[(size_t)d * nV + j];
            for (int k = start; k <= end; ++k)
              if (k != j) row += coef * p[(size_t)d * nV + k];
            // A^T row
            out[(size_t)d * nV + j] += lam_glue * row;
            for (int k = start; k <= end; ++k)
              if (k != j) out[(size_t)d * nV + k] += lam_glue * coef * row;
          }
      }
    } else {
      for (size_t i = 0; i < idsC.size(); ++i)
        for (int d = 0; d < 3; ++d) out[(size_t)d * nV + idsC[i]] += lam_C * p[(size_t)d * nV + idsC[i]];
      for (size_t i = 0; i < idsT.size(); ++i) {
        int v = idsT[i];
        double r = 0;
        for (int d = 0; d < 3; ++d) r += nT_[i * 3 + d] * p[(size_t)d * nV + v];
        for (int d = 0; d < 3; ++d) out[(size_t)d * nV + v] += lam_T * nT_[i * 3 + d] * r;
      }
      for (size_t i = 0; i < idsL.size(); ++i) {
        int v = idsL[i];
        double pa = p[3 * (size_t)nV + i];
        double acc_a = 0;
        for (int d = 0; d < 3; ++d) {
          double r = p[(size_t)d * nV + v] - axa[i * 3 + d] * pa;
          out[(size_t)d * nV + v] += lam_L * r;
          acc_a += -axa[i * 3 + d] * r;
        }
        out[3 * (size_t)nV + i] += lam_L * acc_a;
      }
    }
  }

  void diagonal(std::vector<double>& dg) const {
    const int n = nvar();
    dg.assign(n, 0.0);
    for (int t = 0; t < nT; ++t) {
      const double* g = &G[(size_t)t * 12];
      const double* w = &W[(size_t)t * 9];
      for (int j = 0; j < 4; ++j) {
        double gg = g[0 * 4 + j] * g[0 * 4 + j] + g[1 * 4 + j] * g[1 * 4 + j] + g[2 * 4 + j] * g[2 * 4 + j];
        for (int b2 = 0; b2 < 3; ++b2) {
          double ww = w[0 * 3 + b2] * w[0 * 3 + b2] + w[1 * 3 + b2] * w[1 * 3 + b2] + w[2 * 3 + b2] * w[2 * 3 + b2];
          dg[(size_t)b2 * nV + T[t * 4 + j]] += M * ww * gg;
        }
      }
    }
    for (int i = 0; i < 3 * nV; ++i) dg[i] += prox;
    if (projection)
      for (size_t i = 0; i < b.size(); ++i)
        for (int d = 0; d < 3; ++d) dg[(size_t)d * nV + b[i]] += lam_proj;
    if (glue) {
      for (size_t i = 0; i < ranges.size(); ++i) {
        int start = i ? ranges[i - 1] + 1 : 0, end = ranges[i];
        if (end == start) continue;
        double coef = -1.0 / (end - start);
        int cnt = end - start;  // number of "other" entries per row
        for (int d = 0; d < 3; ++d)
          for (int j = start; j <= end; ++j) dg[(size_t)d * nV + j] += lam_glue * (1.0 + cnt * coef * coef);
      }
    } else {
      for (size_t i = 0; i < idsC.size(); ++i)
        for (int d = 0; d < 3; ++d) dg[(size_t)d * nV + idsC[i]] += lam_C;
      for (size_t i = 0; i < idsT.size(); ++i)
        for (int d = 0; d < 3; ++d) dg[(size_t)d * nV + idsT[i]] += lam_T * nT_[i * 3 + d] * nT_[i * 3 + d];
      for (size_t i = 0; i < idsL.size(); ++i) {
        for (int d = 0; d < 3; ++d) dg[(size_t)d * nV + idsL[i]] += lam_L;
        double s = 0;
        for (int d = 0; d < 3; ++d) s += axa[i * 3 + d] * axa[i * 3 + d];
        dg[3 * (size_t)nV + i] += lam_L * s;
      }
    }
  }

  // buildRhs + soft-constraint rhs terms (slim_m.cpp:1044-1092, 597-714)
  void rhs(std::vector<double>& out) const {
    const int n = nvar();
    out.assign(n, 0.0);
    for (int t = 0; t < nT; ++t) {
      const double* g = &G[(size_t)t * 12];
      const double* w = &W[(size_t)t * 9];
      const double* r = &R[(size_t)t * 9];
      double F[3][3], Z[3][3];  // F = W R, Z = W^T M F
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) F[a][c] = w[a * 3] * r[0 * 3 + c] + w[a * 3 + 1] * r[1 * 3 + c] + w[a * 3 + 2] * r[2 * 3 + c];
      for (int b2 = 0; b2 < 3; ++b2)
        for (int c = 0; c < 3; ++c) Z[b2][c] = M * (w[0 * 3 + b2] * F[0][c] + w[1 * 3 + b2] * F[1][c] + w[2 * 3 + b2] * F[2][c]);
      for (int j = 0; j < 4; ++j)
        for (int b2 = 0; b2 < 3; ++b2) {
          double s = 0;
          for (int c = 0; c < 3; ++c) s += g[c * 4 + j] * Z[b2][c];
          out[(size_t)b2 * nV + T[t * 4 + j]] += s;
        }
    }
    for (int d = 0; d < 3; ++d)
      for (int i = 0; i < nV; ++i) out[(size_t)d * nV + i] += prox * Vo[(size_t)i * 3 + d];
    if (projection)
      for (size_t i = 0; i < b.size(); ++i)
        for (int d = 0; d < 3; ++d) out[(size_t)d * nV + b[i]] += lam_proj * bc[i * 3 + d];
    if (!glue) {
      for (size_t i = 0; i < idsC.size(); ++i)
        for (int d = 0; d < 3; ++d) out[(size_t)d * nV + idsC[i]] += lam_C * C[i * 3 + d];
      for (size_t i = 0; i < idsT.size(); ++i)
        for (int d = 0; d < 3; ++d) out[(size_t)d * nV + idsT[i]] += lam_T * nT_[i * 3 + d] * dT[i];
      for (size_t i = 0; i < idsL.size(); ++i) {
        double s = 0;
        for (int d = 0; d < 3; ++d) {
          out[(size_t)d * nV + idsL[i]] += lam_L * org[i * 3 + d];
          s += -axa[i * 3 + d] * org[i * 3 + d];
        }
        out[3 * (size_t)nV + i] += lam_L * s;
      }
    }
  }

  // Eigen's ConjugateGradient with DiagonalPreconditioner (ConjugateGradient.h:27-90),
  // tol 1e-8, at most 20 iterations, warm start (slim_m.cpp:484-497).
  void cg(const std::vector<double>& rhsv, std::vector<double>& x) {
    const int n = nvar();
    std::vector<double> dg, invd(n), tmp, res(n), p(n), z(n);
    diagonal(dg);
    for (int i = 0; i < n; ++i) invd[i] = dg[i] != 0 ? 1.0 / dg[i] : 1.0;
    const double tol = 1e-8;
    const int maxIters = 20;
    apply(x, tmp);
    for (int i = 0; i < n; ++i) res[i] = rhsv[i] - tmp[i];
    double rhsNorm2 = 0;
    for (int i = 0; i < n; ++i) rhsNorm2 += rhsv[i] * rhsv[i];
    last_cg_iters = 0;
    if (rhsNorm2 == 0) { std::fill(x.begin(), x.end(), 0.0); last_cg_err = 0; return; }
    double threshold = tol * tol * rhsNorm2;
    double rn2 = 0;
    for (int i = 0; i < n; ++i) rn2 += res[i] * res[i];
    if (rn2 < threshold) { last_cg_err = std::sqrt(rn2 / rhsNorm2); return; }
    for (int i = 0; i < n; ++i) p[i] = invd[i] * res[i];
    double absNew = 0;
    for (int i = 0; i < n; ++i) absNew += res[i] * p[i];
    int it = 0;
    while (it < maxIters) {
      apply(p, tmp);
      double pt = 0;
      for (int i = 0; i < n; ++i) pt += p[i] * tmp[i];
      double alpha = absNew / pt;
      for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; res[i] -= alpha * tmp[i]; }
      rn2 = 0;
      for (int i = 0; i < n; ++i) rn2 += res[i] * res[i];
      if (rn2 < threshold) break;
      for (int i = 0; i < n; ++i) z[i] = invd[i] * res[i];
      double absOld = absNew;
      absNew = 0;
      for (int i = 0; i < n; ++i) absNew += res[i] * z[i];
      double beta = absNew / absOld;
      for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      ++it;
    }
    last_cg_err = std::sqrt(rn2 / rhsNorm2);
    last_cg_iters = it;
  }
};

// ---------------------------------------------------------------------------------------
// Flip-avoiding line search (flip_avoiding_line_search.cpp:24-317).
int solve_p3(double x[3], double a, double b, double c) {
  // roots of t^3 + a t^2 + b t + c (trigonometric / Cardano form, as SolveP3)
  double a2 = a * a;
  double q = (a2 - 3 * b) / 9;
  double r = (a * (2 * a2 - 9 * b) + 27 * c) / 54;
  double r2 = r * r, q3 = q * q * q;
  if (r2 < q3) {
    double t = r / std::sqrt(q3);
    if (t < -1) t = -1;
    if (t > 1) t = 1;
    t = std::acos(t);
    a /= 3;
    q = -2 * std::sqrt(q);
    x[0] = q * std::cos(t / 3) - a;
    x[1] = q * std::cos((t + (2 * M_PI)) / 3) - a;
    x[2] = q * std::cos((t - (2 * M_PI)) / 3) - a;
    return 3;
  }
  double A = -std::pow(std::fabs(r) + std::sqrt(r2 - q3), 1. / 3);
  if (r < 0) A = -A;
  double B = (A == 0) ? 0 : q / A;
  a /= 3;
  x[0] = (A + B) - a;
  x[1] = -0.5 * (A + B) - a;
  x[2] = 0.5 * std::sqrt(3.) * (A - B);
  if (std::fabs(x[2]) < 1e-14) { x[2] = x[1]; return 2; }
  return 1;
}

double smallest_pos_quad_zero(double a, double b, double c) {
  double t1, t2;
  if (std::fabs(a) > 1.0e-10) {
    double delta_in = b * b - 4 * a * c;
    if (delta_in <= 0) return INFINITY;
    double delta = std::sqrt(delta_in);
    if (b >= 0) {
      double bd = -b - delta;
      t1 = 2 * c / bd;
      t2 = bd / (2 * a);
    } else {
      double bd = -b + delta;
      t1 = bd / (2 * a);
      t2 = (2 * c) / bd;
    }
    if (a < 0) std::swap(t1, t2);
    if (t1 > 0) return t2 > 0 ? t2 : t1;
    return INFINITY;
  }
  if (b == 0) return INFINITY;
  t1 = -c / b;
  return t1 > 0 ? t1 : INFINITY;
}

// Smallest positive t with det[b-a, c-a, d-a](x + t dx) = 0 for one tet. The cubic's
// coefficients are those of the reference's symbolic 4x4 volume determinant (up to the
// overall sign, which does not move the roots), written as det(E + tD) =
// det E + t tr(adj(E) D) + t^2 tr(E adj(D)) + t^3 det D with E, D the edge matrices.
double min_pos_root_tet(const V3 x[4], const V3 d[4]) {
  V3 e0 = sub(x[1], x[0]), e1 = sub(x[2], x[0]), e2 = sub(x[3], x[0]);
  V3 f0 = sub(d[1], d[0]), f1 = sub(d[2], d[0]), f2 = sub(d[3], d[0]);
  double c3 = det3(f0, f1, f2);
  double c2 = det3(e0, f1, f2) + det3(f0, e1, f2) + det3(f0, f1, e2);
  double c1 = det3(f0, e1, e2) + det3(e0, f1, e2) + det3(e0, e1, f2);
  double c0 = det3(e0, e1, e2);
  // the reference's coefficient of t^3 is -det D (4x4 determinant with a column of ones)
  double a = -c3, b = -c2, c = -c1, dd = -c0;
  if (std::fabs(a) <= 1.e-10) return smallest_pos_quad_zero(b, c, dd);
  b /= a; c /= a; dd /= a;
  double res[3];
  int n = solve_p3(res, b, c, dd);
  switch (n) {
    case 1: return (res[0] >= 0) ? res[0] : INFINITY;
    case 2: {
      double mx = std::max(res[0], res[1]), mn = std::min(res[0], res[1]);
      if (mn > 0) return mn;
      if (mx > 0) return mx;
      return INFINITY;
    }
    default: {
      std::sort(res, res + 3);
      if (res[0] > 0) return res[0];
      if (res[1] > 0) return res[1];
      if (res[2] > 0) return res[2];
      return INFINITY;
    }
  }
}

}  // namespace

// =======================================================================================
// C ABI used by the tests (ctypes)
extern "C" {

int orc_version() { return 1; }

// grad_ref over hexes: hexes nH*8 (reference Hybrid::vs order), V nV*3 row-major.
// hflag (nullable) per hex; cube = compute_referenceMesh's cube argument.
// Out: G nH*8*12 (tet t = 8h+corner, [c][j]).
void orc_grad_ref(int nV, const double* V, int nH, const int* H, const uint8_t* hflag, int cube, double* Gout) {
  (void)nV;
  for (int h = 0; h < nH; ++h) {
    V3 p[8], r8[8];
    for (int i = 0; i < 8; ++i) p[i] = {V[(size_t)H[h * 8 + i] * 3], V[(size_t)H[h * 8 + i] * 3 + 1], V[(size_t)H[h * 8 + i] * 3 + 2]};
    hex_reference_corners(p, cube != 0, hflag ? hflag[h] != 0 : true, r8);
    for (int c = 0; c < 8; ++c) {
      V3 x[4], r[4];
      for (int j = 0; j < 4; ++j) { x[j] = p[kHexTet[c][j]]; r[j] = r8[kHexTet[c][j]]; }
      double G[3][4];
      grad_tet_ref(x, r, G);
      std::memcpy(&Gout[((size_t)h * 8 + c) * 12], G, sizeof(G));
    }
  }
}

struct orc_slim_params {
  int energy;
  int projection;
  int glue;
  int iters;
  double weight_opt;
  double lam_proj, lam_glue, lam_C, lam_T, lam_L;
};

struct orc_slim_constraints {
  int nb; const int* b; const double* bc;
  int nranges; const int* ranges;
  int nC; const int* idsC; const double* C;
  int nT; const int* idsT; const double* normalT; const double* disT;
  int nL; const int* idsL; const double* axaL; const double* originL;
};

// slim_m_opt (optimization.cpp:883-900) = slim_precompute + slim_solve2 fixed branch.
// V: rest positions (ts.V), UV: in = ts.UV, out = result. G from orc_grad_ref.
// log (nullable): per iteration [energy, E_q, E_s, t_max, step, cg_iters, cg_err] (7 doubles)
// Returns 0.
int orc_slim_m_opt(int nV, const double* V, int nH, const int* H, const double* G, double* UV,
                   const orc_slim_params* prm, const orc_slim_constraints* cs, double* out_energy,
                   double* out_eq, double* out_es, double* log, double* aL_out) {
  (void)V;
  Slim s;
  s.nV = nV;
  s.nT = nH * 8;
  s.energy = prm->energy;
  s.T.resize((size_t)s.nT * 4);
  for (int h = 0; h < nH; ++h)
    for (int c = 0; c < 8; ++c)
      for (int j = 0; j < 4; ++j) s.T[((size_t)h * 8 + c) * 4 + j] = H[h * 8 + kHexTet[c][j]];
  s.G.assign(G, G + (size_t)s.nT * 12);
  s.Vo.assign(UV, UV + (size_t)nV * 3);
  s.M = prm->weight_opt;
  s.prox = 0.00001 * prm->weight_opt;
  s.mesh_area = s.M * s.nT;
  s.exp_f = 1.0;
  s.projection = prm->projection != 0;
  s.glue = prm->glue != 0;
  s.lam_proj = prm->lam_proj; s.lam_glue = prm->lam_glue;
  s.lam_C = prm->lam_C; s.lam_T = prm->lam_T; s.lam_L = prm->lam_L;
  s.b.assign(cs->b, cs->b + cs->nb); s.bc.assign(cs->bc, cs->bc + (size_t)cs->nb * 3);
  s.ranges.assign(cs->ranges, cs->ranges + cs->nranges);
  s.idsC.assign(cs->idsC, cs->idsC + cs->nC); s.C.assign(cs->C, cs->C + (size_t)cs->nC * 3);
  s.idsT.assign(cs->idsT, cs->idsT + cs->nT); s.nT_.assign(cs->normalT, cs->normalT + (size_t)cs->nT * 3);
  s.dT.assign(cs->disT, cs->disT + cs->nT);
  s.idsL.assign(cs->idsL, cs->idsL + cs->nL); s.axa.assign(cs->axaL, cs->axaL + (size_t)cs->nL * 3);
  s.org.assign(cs->originL, cs->originL + (size_t)cs->nL * 3);
  s.aL.assign(cs->nL, 0.0);
  if (s.projection && !s.glue && cs->nL > 0) return -2;  // reference solve would mismatch sizes

  s.energy_val = s.energy_total(s.Vo.data()) / s.mesh_area;
  for (int it = 0; it < prm->iters; ++it) {
    std::vector<double> dest = s.Vo;
    s.local_step(dest.data());
    std::vector<double> rhsv;
    s.rhs(rhsv);
    const int n = s.nvar();
    std::vector<double> x(n, 0.0);
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < 3; ++d) x[(size_t)d * nV + i] = dest[(size_t)i * 3 + d];
    s.cg(rhsv, x);
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < 3; ++d) dest[(size_t)i * 3 + d] = x[(size_t)d * nV + i];
    if (!s.projection && !s.glue)
      for (size_t i = 0; i < s.idsL.size(); ++i) s.aL[i] = x[3 * (size_t)nV + i];
    // flip_avoiding_line_search
    std::vector<double> dir((size_t)nV * 3);
    for (size_t i = 0; i < dir.size(); ++i) dir[i] = dest[i] - s.Vo[i];
    double tmin = INFINITY;
    for (int t = 0; t < s.nT; ++t) {
      V3 xx[4], dd[4];
      for (int j = 0; j < 4; ++j) {
        size_t v = (size_t)s.T[t * 4 + j] * 3;
        xx[j] = {s.Vo[v], s.Vo[v + 1], s.Vo[v + 2]};
        dd[j] = {dir[v], dir[v + 1], dir[v + 2]};
      }
      tmin = std::min(tmin, min_pos_root_tet(xx, dd));
    }
    double step = std::min(1., tmin * 0.8);
    s.last_tmax = step;
    // igl::line_search (line_search.cpp:10-56)
    double old_energy = s.energy_total(s.Vo.data());
    double new_energy = old_energy;
    int cur_iter = 0;
    std::vector<double> nx(s.Vo.size());
    while (new_energy >= old_energy) {
      if (cur_iter > 20) step = 0;
      for (size_t i = 0; i < nx.size(); ++i) nx[i] = s.Vo[i] + step * dir[i];
      double cur_e = s.energy_total(nx.data());
      if (std::isnan(cur_e) || std::isinf(cur_e) || cur_e > old_energy) {
        step /= 2;
      } else {
        s.Vo = nx;
        new_energy = cur_e;
        break;
      }
      if (step == 0) break;
      cur_iter++;
    }
    s.last_step = step;
    s.energy_val = new_energy / s.mesh_area;
    if (log) {
      double* L = log + (size_t)it * 7;
      L[0] = s.energy_val; L[1] = s.e_quality; L[2] = s.e_soft; L[3] = s.last_tmax; L[4] = s.last_step;
      L[5] = s.last_cg_iters; L[6] = s.last_cg_err;
    }
  }
  std::memcpy(UV, s.Vo.data(), sizeof(double) * (size_t)nV * 3);
  if (out_energy) *out_energy = s.energy_val;
  if (out_eq) *out_eq = s.e_quality;
  if (out_es) *out_es = s.e_soft;
  if (aL_out) for (size_t i = 0; i < s.aL.size(); ++i) aL_out[i] = s.aL[i];
  return 0;
}

// The global system of the first SLIM iteration at UV (build_linear_system,
// slim_m.cpp:568-714) applied column by column: L (n x n dense, small meshes only) and rhs.
int orc_linear_system(int nV, int nH, const int* H, const double* G, const double* UV, const orc_slim_params* prm,
                      const orc_slim_constraints* cs, double* Ldense, double* rhs_out, const double* p_in,
                      double* Lp_out, double* diag_out) {
  Slim s;
  s.nV = nV; s.nT = nH * 8; s.energy = prm->energy;
  s.T.resize((size_t)s.nT * 4);
  for (int h = 0; h < nH; ++h)
    for (int c = 0; c < 8; ++c)
      for (int j = 0; j < 4; ++j) s.T[((size_t)h * 8 + c) * 4 + j] = H[h * 8 + kHexTet[c][j]];
  s.G.assign(G, G + (size_t)s.nT * 12);
  s.Vo.assign(UV, UV + (size_t)nV * 3);
  s.M = prm->weight_opt; s.prox = 0.00001 * prm->weight_opt;
  s.projection = prm->projection != 0; s.glue = prm->glue != 0;
  s.lam_proj = prm->lam_proj; s.lam_glue = prm->lam_glue; s.lam_C = prm->lam_C; s.lam_T = prm->lam_T; s.lam_L = prm->lam_L;
  s.b.assign(cs->b, cs->b + cs->nb); s.bc.assign(cs->bc, cs->bc + (size_t)cs->nb * 3);
  s.ranges.assign(cs->ranges, cs->ranges + cs->nranges);
  s.idsC.assign(cs->idsC, cs->idsC + cs->nC); s.C.assign(cs->C, cs->C + (size_t)cs->nC * 3);
  s.idsT.assign(cs->idsT, cs->idsT + cs->nT); s.nT_.assign(cs->normalT, cs->normalT + (size_t)cs->nT * 3);
  s.dT.assign(cs->disT, cs->disT + cs->nT);
  s.idsL.assign(cs->idsL, cs->idsL + cs->nL); s.axa.assign(cs->axaL, cs->axaL + (size_t)cs->nL * 3);
  s.org.assign(cs->originL, cs->originL + (size_t)cs->nL * 3);
  s.aL.assign(cs->nL, 0.0);
  s.local_step(UV);
  const int n = s.nvar();
  if (p_in) {
    std::vector<double> pv(p_in, p_in + n), out;
    s.apply(pv, out);
    std::memcpy(Lp_out, out.data(), sizeof(double) * n);
    std::vector<double> dg;
    s.diagonal(dg);
    for (int i = 0; i < n; ++i) diag_out[i] = dg[i];
  }
  std::vector<double> e(n, 0.0), col;
  if (Ldense)
    for (int j = 0; j < n; ++j) {
      e[j] = 1.0;
      s.apply(e, col);
      for (int i = 0; i < n; ++i) Ldense[(size_t)i * n + j] = col[i];
      e[j] = 0.0;
    }
  std::vector<double> r;
  s.rhs(r);
  std::memcpy(rhs_out, r.data(), sizeof(double) * n);
  return n;
}

// Per-tet SLIM energy density, local-step W and R at positions X (for kernel-level tests).
void orc_local_step(int nV, int nH, const int* H, const double* G, const double* X, int energy,
                    double* W, double* R, double* E) {
  Slim s;
  s.nV = nV; s.nT = nH * 8; s.energy = energy;
  s.T.resize((size_t)s.nT * 4);
  for (int h = 0; h < nH; ++h)
    for (int c = 0; c < 8; ++c)
      for (int j = 0; j < 4; ++j) s.T[((size_t)h * 8 + c) * 4 + j] = H[h * 8 + kHexTet[c][j]];
  s.G.assign(G, G + (size_t)s.nT * 12);
  s.local_step(X);
  if (W) std::memcpy(W, s.W.data(), sizeof(double) * s.W.size());
  if (R) std::memcpy(R, s.R.data(), sizeof(double) * s.R.size());
  if (E)
    for (int t = 0; t < s.nT; ++t) { M3 J; s.jacobian(X, t, J); E[t] = s.tet_energy(J); }
}

// Gradient of the symmetric-Dirichlet SLIM energy E = M sum_t (sum_i s_i^2 + s_i^-2) + soft
// with respect to the positions, written through the SVD: dE/dJ = U diag(2 s - 2 s^-3) V^T
// (an expression independent of the adjugate form the kernels use). soft: tangent-plane,
// corner and line terms with a_L = aL. X, G: nV*3 row-major.
void orc_sd_grad(int nV, int nH, const int* H, const double* Gop, const double* X, double M,
                 const orc_slim_constraints* cs, double lamC, double lamT, double lamL, const double* aL,
                 double* Gout, double* Eout) {
  std::memset(Gout, 0, sizeof(double) * (size_t)nV * 3);
  double e = 0;
  for (int h = 0; h < nH; ++h)
    for (int c = 0; c < 8; ++c) {
      const double* g = &Gop[((size_t)h * 8 + c) * 12];
      int v[4];
      for (int j = 0; j < 4; ++j) v[j] = H[h * 8 + kHexTet[c][j]];
      M3 J;
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) {
          double sum = 0;
          for (int j = 0; j < 4; ++j) sum += g[cc * 4 + j] * X[(size_t)v[j] * 3 + r];
          J[r][cc] = sum;
        }
      M3 U, V;
      double sv[3];
      svd3(J, U, sv, V);
      // signed singular values keep J = U diag(s) V^T exact for inverted tets as well
      double dv[3];
      for (int k = 0; k < 3; ++k) {
        dv[k] = 2 * sv[k] - 2 / (sv[k] * sv[k] * sv[k]);
        e += M * (sv[k] * sv[k] + 1 / (sv[k] * sv[k]));
      }
      M3 P;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = M * (U[i][0] * dv[0] * V[j][0] + U[i][1] * dv[1] * V[j][1] + U[i][2] * dv[2] * V[j][2]);
      for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 3; ++r) {
          double sum = 0;
          for (int cc = 0; cc < 3; ++cc) sum += P[r][cc] * g[cc * 4 + j];
          Gout[(size_t)v[j] * 3 + r] += sum;
        }
    }
  for (int i = 0; i < cs->nC; ++i)
    for (int d = 0; d < 3; ++d) {
      double q = X[(size_t)cs->idsC[i] * 3 + d] - cs->C[i * 3 + d];
      Gout[(size_t)cs->idsC[i] * 3 + d] += 2 * lamC * q;
      e += lamC * q * q;
    }
  for (int i = 0; i < cs->nT; ++i) {
    const double* n = &cs->normalT[i * 3];
    const double* x = &X[(size_t)cs->idsT[i] * 3];
    double a = n[0] * x[0] + n[1] * x[1] + n[2] * x[2] - cs->disT[i];
    for (int d = 0; d < 3; ++d) Gout[(size_t)cs->idsT[i] * 3 + d] += 2 * lamT * a * n[d];
    e += lamT * a * a;
  }
  for (int i = 0; i < cs->nL; ++i)
    for (int d = 0; d < 3; ++d) {
      double q = X[(size_t)cs->idsL[i] * 3 + d] - aL[i] * cs->axaL[i * 3 + d] - cs->originL[i * 3 + d];
      Gout[(size_t)cs->idsL[i] * 3 + d] += 2 * lamL * q;
      e += lamL * q * q;
    }
  if (Eout) *Eout = e;
}

// 3x3 SVD (for closed-form tests): A row-major; U, V row-major; s descending.
void orc_svd3(const double* A, double* U, double* s, double* V) {
  M3 a, u, v;
  for (int i = 0; i < 9; ++i) a[i / 3][i % 3] = A[i];
  svd3(a, u, s, v);
  for (int i = 0; i < 9; ++i) { U[i] = u[i / 3][i % 3]; V[i] = v[i / 3][i % 3]; }
}

// Largest flip-free step: min over tets of the smallest positive root (INFINITY if none).
double orc_max_step(int nV, int nH, const int* H, const double* X, const double* D) {
  (void)nV;
  double tmin = INFINITY;
  for (int h = 0; h < nH; ++h)
    for (int c = 0; c < 8; ++c) {
      V3 xx[4], dd[4];
      for (int j = 0; j < 4; ++j) {
        size_t v = (size_t)H[h * 8 + kHexTet[c][j]] * 3;
        xx[j] = {X[v], X[v + 1], X[v + 2]};
        dd[j] = {D[v], D[v + 1], D[v + 2]};
      }
      tmin = std::min(tmin, min_pos_root_tet(xx, dd));
    }
  return tmin;
}

// scaled_jacobian (global_functions.cpp:2309-2358) with a_jacobian(Vector3d) (:2422-2442).
void orc_scaled_jacobian(int nV, const double* X, int nH, const int* H, double* minJ, double* avgJ, double* perHex) {
  (void)nV;
  double mn = 1, avg = 0;
  for (int h = 0; h < nH; ++h) {
    double hmin = 1;
    for (int c = 0; c < 8; ++c) {
      V3 p[4];
      for (int j = 0; j < 4; ++j) {
        size_t v = (size_t)H[h * 8 + kHexTet[c][j]] * 3;
        p[j] = {X[v], X[v + 1], X[v + 2]};
      }
      V3 c0 = mul(sub(p[1], p[0]), .5), c1 = mul(sub(p[2], p[0]), .5), c2 = mul(sub(p[3], p[0]), .5);
      double n1 = norm(c0), n2 = norm(c1), n3 = norm(c2);
      double sj = det3(c0, c1, c2);
      const double Precision = 1.e-7;
      if (!(std::fabs(n1) < Precision || std::fabs(n2) < Precision || std::fabs(n3) < Precision)) sj /= n1 * n2 * n3;
      if (hmin > sj) hmin = sj;
    }
    if (perHex) perHex[h] = hmin;
    avg += hmin;
    if (mn > hmin) mn = hmin;
  }
  *minJ = mn;
  *avgJ = nH ? avg / nH : 0;
}

// ---------------------------------------------------------------------------------------
// Closest point with pseudonormal (signed_distance_pseudonormal, signed_distance.cpp:186-257)
// over igl's AABB tree restated (AABB.cpp:93-200 build, 364-441 query, 735-779 set_min).
namespace {
struct Node { double mn[3], mx[3]; int left = -1, right = -1, prim = -1; };
struct Tree {
  std::vector<Node> nodes;
  const double* V; const int* F;
  std::vector<int> SI;  // nf*3 sorted rank of barycentre per dim
  int build(const std::vector<int>& I) {
    int id = (int)nodes.size();
    nodes.push_back(Node());
    Node nd;
    for (int d = 0; d < 3; ++d) { nd.mn[d] = INFINITY; nd.mx[d] = -INFINITY; }
    for (int i : I)
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
          double v = V[(size_t)F[(size_t)i * 3 + c] * 3 + d];
          nd.mn[d] = std::min(nd.mn[d], v);
          nd.mx[d] = std::max(nd.mx[d], v);
        }
    if (I.size() == 1) {
      nd.prim = I[0];
      nodes[id] = nd;
      return id;
    }
    int max_d = 0;
    double best = nd.mx[0] - nd.mn[0];
    for (int d = 1; d < 3; ++d)
      if (nd.mx[d] - nd.mn[d] > best) { best = nd.mx[d] - nd.mn[d]; max_d = d; }
    std::vector<int> S(I.size());
    for (size_t i = 0; i < I.size(); ++i) S[i] = SI[(size_t)I[i] * 3 + max_d];
    std::vector<int> A = S;
    size_t n = A.size() / 2;
    std::nth_element(A.begin(), A.begin() + n, A.end());
    double med;
    if (A.size() % 2 == 1) med = A[n];
    else {
      // both elements are read after the second selection, as AABB.cpp:161-170 does
      std::nth_element(A.begin(), A.begin() + n - 1, A.end());
      med = 0.5 * (A[n] + A[n - 1]);
    }
    std::vector<int> L, R;
    for (size_t i = 0; i < I.size(); ++i) (S[i] <= med ? L : R).push_back(I[i]);
    // the reference sizes its halves as (n+1)/2 and n/2; ranks are unique, so a split of
    // any other shape would be a mismatch of this restatement
    if (L.size() != (I.size() + 1) / 2) return -1;
    nodes[id] = nd;
    if (!L.empty()) { int l = build(L); if (l < 0) return -1; nodes[id].left = l; }
    if (!R.empty()) { int r = build(R); if (r < 0) return -1; nodes[id].right = r; }
    return id;
  }
  bool contains(const Node& n, const double* p) const {
    for (int d = 0; d < 3; ++d) if (!(n.mn[d] <= p[d] && p[d] <= n.mx[d])) return false;
    return true;
  }
  double ext_dist(const Node& n, const double* p) const {
    double s = 0;
    for (int d = 0; d < 3; ++d) {
      if (n.mn[d] > p[d]) { double a = n.mn[d] - p[d]; s += a * a; }
      else if (p[d] > n.mx[d]) { double a = p[d] - n.mx[d]; s += a * a; }
    }
    return s;
  }
  // Ericson closest point on triangle (point_simplex_squared_distance.cpp:44-109); its Dot
  // is Eigen's dot of 1x3 rows
  void closest_tri(const double* pp, int f, double c[3], double bary[3]) const {
    V3 p = {pp[0], pp[1], pp[2]};
    const double* A_ = &V[(size_t)F[(size_t)f * 3] * 3];
    const double* B_ = &V[(size_t)F[(size_t)f * 3 + 1] * 3];
    const double* C_ = &V[(size_t)F[(size_t)f * 3 + 2] * 3];
    V3 a = {A_[0], A_[1], A_[2]}, b = {B_[0], B_[1], B_[2]}, cc = {C_[0], C_[1], C_[2]};
    V3 ab = sub(b, a), ac = sub(cc, a), ap = sub(p, a);
    double d1 = dotE(ab, ap), d2 = dotE(ac, ap);
    V3 r;
    if (d1 <= 0.0 && d2 <= 0.0) { bary[0] = 1; bary[1] = 0; bary[2] = 0; r = a; goto done; }
    {
      V3 bp = sub(p, b);
      double d3 = dotE(ab, bp), d4 = dotE(ac, bp);
      if (d3 >= 0.0 && d4 <= d3) { bary[0] = 0; bary[1] = 1; bary[2] = 0; r = b; goto done; }
      double vc = d1 * d4 - d3 * d2;
      if (!(a.x == b.x && a.y == b.y && a.z == b.z)) {
        if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
          double v = d1 / (d1 - d3);
          bary[0] = 1 - v; bary[1] = v; bary[2] = 0; r = add(a, mul(ab, v)); goto done;
        }
      }
      V3 cp = sub(p, cc);
      double d5 = dotE(ab, cp), d6 = dotE(ac, cp);
      if (d6 >= 0.0 && d5 <= d6) { bary[0] = 0; bary[1] = 0; bary[2] = 1; r = cc; goto done; }
      double vb = d5 * d2 - d1 * d6;
      if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        bary[0] = 1 - w; bary[1] = 0; bary[2] = w; r = add(a, mul(ac, w)); goto done;
      }
      double va = d3 * d6 - d5 * d4;
      if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary[0] = 0; bary[1] = 1 - w; bary[2] = w; r = add(b, mul(sub(cc, b), w)); goto done;
      }
      double denom = 1.0 / (va + vb + vc);
      double v = vb * denom, w = vc * denom;
      bary[0] = 1.0 - v - w; bary[1] = v; bary[2] = w;
      r = add(add(a, mul(ab, v)), mul(ac, w));
    }
  done:
    c[0] = r.x; c[1] = r.y; c[2] = r.z;
  }
  double query(int node, const double* p, double sqr_d, int& i, double c[3]) const {
    const Node& nd = nodes[node];
    if (nd.prim >= 0) {
      double cc[3], bb[3];
      closest_tri(p, nd.prim, cc, bb);
      // (p - c).squaredNorm() in Eigen's order
      double d = (p[0] - cc[0]) * (p[0] - cc[0]) + ((p[1] - cc[1]) * (p[1] - cc[1]) + (p[2] - cc[2]) * (p[2] - cc[2]));
      if (d < sqr_d) { i = nd.prim; c[0] = cc[0]; c[1] = cc[1]; c[2] = cc[2]; sqr_d = d; }
      return sqr_d;
    }
    bool ll = false, lr = false;
    auto look = [&](int child, bool& flag) {
      int ic = i;
      double cc[3] = {c[0], c[1], c[2]};
      double d = query(child, p, sqr_d, ic, cc);
      if (d < sqr_d) { i = ic; c[0] = cc[0]; c[1] = cc[1]; c[2] = cc[2]; sqr_d = d; }
      flag = true;
    };
    if (contains(nodes[nd.left], p)) look(nd.left, ll);
    if (contains(nodes[nd.right], p)) look(nd.right, lr);
    double lm = ext_dist(nodes[nd.left], p), rm = ext_dist(nodes[nd.right], p);
    if (lm < rm) {
      if (!ll && lm < sqr_d) look(nd.left, ll);
      if (!lr && rm < sqr_d) look(nd.right, lr);
    } else {
      if (!lr && rm < sqr_d) look(nd.right, lr);
      if (!ll && lm < sqr_d) look(nd.left, ll);
    }
    return sqr_d;
  }
};
}  // namespace

// AABB build order + closest point + pseudonormal sign/normal selection.
// FN nf*3, VN nv*3, FEN nf*9 (edge normal of the edge opposite vertex x of face f at
// [f*9 + x*3 ..], i.e. EN(EMAP(f + x*nf))). Outputs S, I, C (nq*3), N (nq*3), sqrD.
void orc_signed_distance_pseudonormal(int nq, const double* P, int nv, const double* V, int nf, const int* F,
                                      const double* FN, const double* VN, const double* FEN, double* S, int* I,
                                      double* Cout, double* Nout, double* sqrD) {
  (void)nv;
  Tree tr;
  tr.V = V; tr.F = F;
  // barycentres and sorted ranks per dimension (igl::barycenter + igl::sort)
  std::vector<double> BC((size_t)nf * 3);
  for (int f = 0; f < nf; ++f)
    for (int d = 0; d < 3; ++d) {
      double s = 0;
      for (int c = 0; c < 3; ++c) s += V[(size_t)F[(size_t)f * 3 + c] * 3 + d];
      BC[(size_t)f * 3 + d] = s / 3.0;
    }
  tr.SI.assign((size_t)nf * 3, 0);
  for (int d = 0; d < 3; ++d) {
    std::vector<size_t> idx(nf);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<double> col(nf);
    for (int f = 0; f < nf; ++f) col[f] = BC[(size_t)f * 3 + d];
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return col[a] < col[b]; });
    for (int r = 0; r < nf; ++r) tr.SI[idx[r] * 3 + d] = r;
  }
  std::vector<int> all(nf);
  std::iota(all.begin(), all.end(), 0);
  if (nf > 0 && tr.build(all) < 0) {
    for (int q = 0; q < nq; ++q) if (I) I[q] = -2;
    return;
  }
  for (int q = 0; q < nq; ++q) {
    const double* p = &P[(size_t)q * 3];
    int f = -1;
    double c[3] = {0, 0, 0};
    double d2 = tr.query(0, p, INFINITY, f, c);
    // pseudonormal_test (pseudonormal_test.cpp:14-119)
    const double* A_ = &V[(size_t)F[(size_t)f * 3] * 3];
    const double* B_ = &V[(size_t)F[(size_t)f * 3 + 1] * 3];
    const double* C_ = &V[(size_t)F[(size_t)f * 3 + 2] * 3];
    V3 a = {A_[0], A_[1], A_[2]}, b = {B_[0], B_[1], B_[2]}, cc = {C_[0], C_[1], C_[2]};
    // doublearea(A,B,C) for 3D rows (doublearea.cpp:81-111,131-170): Kahan's Heron formula
    // on the edge lengths sorted in decreasing order
    double area;
    {
      double l[3] = {norm(sub(b, cc)), norm(sub(cc, a)), norm(sub(a, b))};
      std::sort(l, l + 3, [](double x, double y) { return x > y; });
      double arg = (l[0] + (l[1] + l[2])) * (l[2] - (l[0] - l[1])) * (l[2] + (l[0] - l[1])) * (l[0] + (l[1] - l[2]));
      area = 2.0 * 0.25 * std::sqrt(arg);
    }
    double n[3];
    const double MIN_DOUBLE_AREA = 1e-4, epsilon = 1e-12;
    if (area > MIN_DOUBLE_AREA) {
      V3 v0 = sub(b, a), v1 = sub(cc, a), v2 = {c[0] - a.x, c[1] - a.y, c[2] - a.z};
      double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
      double denom = d00 * d11 - d01 * d01;
      double bb[3];
      bb[1] = (d11 * d20 - d01 * d21) / denom;
      bb[2] = (d00 * d21 - d01 * d20) / denom;
      bb[0] = 1.0f - (bb[1] + bb[2]);
      int type = (bb[0] <= epsilon) + (bb[1] <= epsilon) + (bb[2] <= epsilon);
      if (type == 2) {
        for (int x = 0; x < 3; ++x)
          if (bb[x] > epsilon) { for (int d = 0; d < 3; ++d) n[d] = VN[(size_t)F[(size_t)f * 3 + x] * 3 + d]; break; }
      } else if (type == 1) {
        for (int x = 0; x < 3; ++x)
          if (bb[x] <= epsilon) { for (int d = 0; d < 3; ++d) n[d] = FEN[(size_t)f * 9 + x * 3 + d]; break; }
      } else {
        for (int d = 0; d < 3; ++d) n[d] = FN[(size_t)f * 3 + d];
      }
    } else {
      bool found = false;
      for (int v = 0; v < 3 && !found; ++v) {
        const double* vv = &V[(size_t)F[(size_t)f * 3 + v] * 3];
        double dd = std::sqrt((c[0] - vv[0]) * (c[0] - vv[0]) + (c[1] - vv[1]) * (c[1] - vv[1]) + (c[2] - vv[2]) * (c[2] - vv[2]));
        if (dd < epsilon) { found = true; for (int d = 0; d < 3; ++d) n[d] = VN[(size_t)F[(size_t)f * 3 + v] * 3 + d]; }
      }
      for (int e = 0; e < 3 && !found; ++e) {
        const double* s0 = &V[(size_t)F[(size_t)f * 3 + (e + 1) % 3] * 3];
        const double* s1 = &V[(size_t)F[(size_t)f * 3 + (e + 2) % 3] * 3];
        // project_to_line_segment
        double dv[3] = {s1[0] - s0[0], s1[1] - s0[1], s1[2] - s0[2]};
        double pv[3] = {c[0] - s0[0], c[1] - s0[1], c[2] - s0[2]};
        double t = (dv[0] * pv[0] + dv[1] * pv[1] + dv[2] * pv[2]) / (dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        double q2;
        if (t < 0) q2 = pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2];
        else if (t > 1) { double w0 = c[0] - s1[0], w1 = c[1] - s1[1], w2 = c[2] - s1[2]; q2 = w0 * w0 + w1 * w1 + w2 * w2; }
        else { double w0 = s0[0] + t * dv[0] - c[0], w1 = s0[1] + t * dv[1] - c[1], w2 = s0[2] + t * dv[2] - c[2]; q2 = w0 * w0 + w1 * w1 + w2 * w2; }
        if (std::sqrt(q2) < epsilon) { for (int d = 0; d < 3; ++d) n[d] = FEN[(size_t)f * 9 + e * 3 + d]; found = true; }
      }
      if (!found) for (int d = 0; d < 3; ++d) n[d] = FN[(size_t)f * 3 + d];
    }
    double s = ((p[0] - c[0]) * n[0] + (p[1] - c[1]) * n[1] + (p[2] - c[2]) * n[2]) >= 0 ? 1. : -1.;
    if (S) S[q] = s * std::sqrt(d2);
    if (I) I[q] = f;
    if (sqrD) sqrD[q] = d2;
    for (int d = 0; d < 3; ++d) { Cout[(size_t)q * 3 + d] = c[d]; Nout[(size_t)q * 3 + d] = n[d]; }
  }
}

// ---------------------------------------------------------------------------------------
// Occupancy predicate of octree_mesh (grid_hex_meshing.cpp:502-517): the cell box, built in
// double exactly as the reference does, overlaps (closed intervals, geogram bboxes_overlap)
// the bbox of at least one facet. Brute force over all facets.
// cells: n*4 ints (x, y, z, extent). origin/mesh_transform: 3 doubles each.
void orc_cell_facet_overlap(int n, const int* cells, int nv, const double* V, int nf, const int* F,
                            const double* origin, double voxel_size, const double* mesh_transform,
                            int stop_extent, uint8_t* flag) {
  (void)nv;
  std::vector<double> fb((size_t)nf * 6);
  for (int f = 0; f < nf; ++f)
    for (int d = 0; d < 3; ++d) {
      double mn = INFINITY, mx = -INFINITY;
      for (int c = 0; c < 3; ++c) {
        double v = V[(size_t)F[(size_t)f * 3 + c] * 3 + d];
        mn = std::min(mn, v); mx = std::max(mx, v);
      }
      fb[(size_t)f * 6 + d] = mn; fb[(size_t)f * 6 + 3 + d] = mx;
    }
  for (int i = 0; i < n; ++i) {
    int x = cells[i * 4], y = cells[i * 4 + 1], z = cells[i * 4 + 2], ext = cells[i * 4 + 3];
    if (ext <= stop_extent) { flag[i] = 0; continue; }
    double mn[3], mx[3];
    const int xyz[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
      mn[d] = mesh_transform[d] + origin[d] + voxel_size * xyz[d];
      mx[d] = mn[d] + voxel_size * ext;
    }
    uint8_t hit = 0;
    for (int f = 0; f < nf && !hit; ++f) {
      bool ov = true;
      for (int d = 0; d < 3; ++d)
        if (mx[d] < fb[(size_t)f * 6 + d] || fb[(size_t)f * 6 + 3 + d] < mn[d]) { ov = false; break; }
      if (ov) hit = 1;
    }
    flag[i] = hit;
  }
}
}  // extern "C"

// +z ray parity inside tests: compute_sign(VoxelGrid) (voxelization.h:221-272) and
// compute_sign(OctreeGrid) (voxelization.cpp:101-163) with intersect_ray_z
// (voxelization.h:195-217), point_in_triangle_2d and its SOS orientation
// (voxelization.cpp:57-93). Literal restatements: every facet goes through the reference's
// closed-box prefilter (its query box spans the mesh in z, so only x and y can reject),
// crossings are sorted by (z, sign), and the octree variant keeps the reference's reduced
// crossing list and odd-count rule.
namespace {
int sos_orientation(double x1, double y1, double x2, double y2, double& twice_signed_area) {
  twice_signed_area = y1 * x2 - x1 * y2;
  if (twice_signed_area > 0) return 1;
  if (twice_signed_area < 0) return -1;
  if (y2 > y1) return 1;
  if (y2 < y1) return -1;
  if (x1 > x2) return 1;
  if (x1 < x2) return -1;
  return 0;
}
bool point_in_triangle_2d(double x0, double y0, double x1, double y1, double x2, double y2, double x3, double y3,
                          double& a, double& b, double& c) {
  x1 -= x0; x2 -= x0; x3 -= x0;
  y1 -= y0; y2 -= y0; y3 -= y0;
  int sa = sos_orientation(x2, y2, x3, y3, a);
  if (sa == 0) return false;
  int sb = sos_orientation(x3, y3, x1, y1, b);
  if (sb != sa) return false;
  int sc = sos_orientation(x1, y1, x2, y2, c);
  if (sc != sa) return false;
  double sum = a + b + c;
  a /= sum; b /= sum; c /= sum;
  return true;
}
// returns the crossing sign in {-1,0,1}; z receives the crossing height
int intersect_ray_z(const double* p1, const double* p2, const double* p3, double qx, double qy, double& z) {
  double u, v, w;
  if (point_in_triangle_2d(qx, qy, p1[0], p1[1], p2[0], p2[1], p3[0], p3[1], u, v, w)) {
    z = u * p1[2] + v * p2[2] + w * p3[2];
    double a11 = p2[0] - p1[0], a12 = p2[1] - p1[1], a21 = p3[0] - p1[0], a22 = p3[1] - p1[1];
    double delta = a11 * a22 - a12 * a21;
    return (delta > 0) ? 1 : (delta < 0 ? -1 : 0);
  }
  return 0;
}
}  // namespace

namespace {
// crossings of the vertical line at (qx, qy) with the facets whose closed xy box touches
// [bx0, bx1] x [by0, by1] (geogram get_facet_bbox + bboxes_overlap), sorted by (z, sign)
std::vector<std::pair<double, int>> column_crossings(const double* V, int nf, const int* F, double qx, double qy,
                                                     double bx0, double bx1, double by0, double by1) {
  std::vector<std::pair<double, int>> inter;
  for (int f = 0; f < nf; ++f) {
    const double* p[3] = {&V[(size_t)F[(size_t)f * 3] * 3], &V[(size_t)F[(size_t)f * 3 + 1] * 3],
                          &V[(size_t)F[(size_t)f * 3 + 2] * 3]};
    double mn[2], mx[2];
    for (int d = 0; d < 2; ++d) {
      mn[d] = mx[d] = p[0][d];
      for (int c = 1; c < 3; ++c) { mn[d] = std::min(mn[d], p[c][d]); mx[d] = std::max(mx[d], p[c][d]); }
    }
    if (bx1 < mn[0] || bx0 > mx[0] || by1 < mn[1] || by0 > mx[1]) continue;
    double z;
    int s = intersect_ray_z(p[0], p[1], p[2], qx, qy, z);
    if (s) inter.emplace_back(z, s);
  }
  std::sort(inter.begin(), inter.end());
  return inter;
}
}  // namespace

// inside: nx*ny*nz bytes, index (z*ny + y)*nx + x.
extern "C" void orc_voxel_parity(int nv, const double* V, int nf, const int* F, int nx, int ny, int nz,
                                 const double* origin, double spacing, uint8_t* inside) {
  (void)nv;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      double cx = (x + 0.5) * spacing + origin[0];
      double cy = (y + 0.5) * spacing + origin[1];
      std::vector<std::pair<double, int>> inter = column_crossings(V, nf, F, cx, cy, cx, cx, cy, cy);
      size_t i = 0;
      int s = 0;
      for (int z = 0; z < nz; ++z) {
        double cz = (z + 0.5) * spacing + origin[2];
        for (; i < inter.size() && inter[i].first < cz; ++i) s += inter[i].second;
        inside[((size_t)z * ny + y) * nx + x] = (uint8_t)(s < 0 ? 1 : 0);
      }
    }
}

// OctreeGrid variant: cells n x 4 (x, y, z, extent); inside[i] in {0, 1}.
extern "C" void orc_cell_parity(int nv, const double* V, int nf, const int* F, int n, const int* cells,
                                const double* origin, double spacing, float* inside) {
  (void)nv;
  for (int i = 0; i < n; ++i) {
    const int x = cells[4 * i], y = cells[4 * i + 1], z = cells[4 * i + 2], extent = cells[4 * i + 3];
    double bx0 = origin[0] + spacing * x, by0 = origin[1] + spacing * y;
    double bx1 = bx0 + spacing * extent, by1 = by0 + spacing * extent;
    double cx = bx0 + 0.5 * spacing * extent, cy = by0 + 0.5 * spacing * extent;
    double cz = origin[2] + spacing * z + 0.5 * spacing * extent;
    std::vector<std::pair<double, int>> inter = column_crossings(V, nf, F, cx, cy, bx0, bx1, by0, by1);
    std::vector<double> reduced;
    for (int k = 0, s = 0; k < (int)inter.size(); ++k) {
      const int ds = inter[k].second;
      s += ds;
      if ((s == -1 && ds < 0) || (s == 0 && ds > 0)) reduced.push_back(inter[k].first);
    }
    int num_before = 0;
    for (double zz : reduced)
      if (zz < cz) ++num_before;
    inside[i] = (num_before % 2 == 1) ? 1.0f : 0.0f;
  }
}

// ---------------------------------------------------------------------------------------
// Feature-curve projection of LINE vertices: the LINE branch of dirty_graph_projection
// (grid_hex_meshing.cpp:3961-4026) with point_line_projection (global_functions.cpp:
// 3454-3465, Precision = 1e-7 from global_types.h:23). Curves are CSR lists of surface
// vertex ids (off[c]..off[c+1]); a circle also has the closing segment. Vector norms and
// dot products sum x + (y + z): Eigen 3.2.9 unrolls the reduction of a 3-vector as the first
// coefficient plus the sum of the other two, which is what the reference's dot() and
// squaredNorm() compile to (checked bit for bit against the compiled point_line_projection).
namespace {
double n3(const double* a) { return std::sqrt(a[0] * a[0] + (a[1] * a[1] + a[2] * a[2])); }
void point_line_projection(const double* v1, const double* v2, const double* v, double* pv) {
  double vv1[3], v21[3];
  for (int d = 0; d < 3; ++d) { vv1[d] = v[d] - v1[d]; v21[d] = v2[d] - v1[d]; }
  const double nv21_2 = v21[0] * v21[0] + (v21[1] * v21[1] + v21[2] * v21[2]);
  double t = 0;
  if (nv21_2 >= 1.e-7) t = (vv1[0] * v21[0] + (vv1[1] * v21[1] + vv1[2] * v21[2])) / nv21_2;
  if (t >= 0.0 && t <= 1.0) for (int d = 0; d < 3; ++d) pv[d] = v1[d] + t * (v2[d] - v1[d]);
  else if (t < 0.0) for (int d = 0; d < 3; ++d) pv[d] = v1[d];
  else if (t > 1.0) for (int d = 0; d < 3; ++d) pv[d] = v2[d];
}
void normalized_diff(const double* a, const double* b, double* out) {  // (a - b).normalized()
  double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  const double n = n3(d);
  for (int k = 0; k < 3; ++k) out[k] = d[k] / n;
}
}  // namespace

// point_line_projection alone on n (segment, point) triples: the foot of each.
extern "C" void orc_point_line_projection(int n, const double* A, const double* B, const double* P, double* PV) {
  for (int i = 0; i < n; ++i) point_line_projection(A + 3 * i, B + 3 * i, P + 3 * i, PV + 3 * i);
}

extern "C" void orc_line_projection(int nv, const double* V, int ncurves, const int* off, const int* cvs,
                                    const uint8_t* circles, int n, const double* P, const int* curve_of,
                                    double* origin, double* axis) {
  (void)nv; (void)ncurves;
  for (int q = 0; q < n; ++q) {
    const double* v = P + (size_t)q * 3;
    const int c = curve_of[q];
    const int* curve = cvs + off[c];
    const int csize = off[c + 1] - off[c];
    int curve_len = csize;
    if (!circles[c]) curve_len--;
    std::vector<std::array<double, 3>> pvs, tangents;
    std::vector<std::pair<double, uint32_t>> dis_ids;
    double tangent[3] = {1, 0, 0}, pv[3] = {0, 0, 0};
    for (int j = 0; j < curve_len; ++j) {
      const double* p0 = V + (size_t)curve[j] * 3;
      const double* p1 = V + (size_t)curve[(j + 1) % csize] * 3;
      point_line_projection(p0, p1, v, pv);
      normalized_diff(p1, p0, tangent);
      double dv[3] = {v[0] - pv[0], v[1] - pv[1], v[2] - pv[2]};
      dis_ids.emplace_back(n3(dv), (uint32_t)pvs.size());
      pvs.push_back({pv[0], pv[1], pv[2]});
      tangents.push_back({tangent[0], tangent[1], tangent[2]});
    }
    std::sort(dis_ids.begin(), dis_ids.end());
    if (!dis_ids.empty()) {
      const uint32_t k = dis_ids[0].second;
      for (int d = 0; d < 3; ++d) { pv[d] = pvs[k][d]; tangent[d] = tangents[k][d]; }
    } else {
      for (int j = 0; j < csize; ++j) {
        const double* w = V + (size_t)curve[j] * 3;
        double dv[3] = {w[0] - v[0], w[1] - v[1], w[2] - v[2]};
        dis_ids.emplace_back(n3(dv), (uint32_t)j);
      }
      std::sort(dis_ids.begin(), dis_ids.end());
      const int pos = (int)dis_ids[0].second;
      for (int d = 0; d < 3; ++d) pv[d] = V[(size_t)curve[pos] * 3 + d];
      curve_len = csize;
      double e[3];
      if (circles[c] || (!circles[c] && pos != 0 && pos != curve_len - 1)) {
        const int pos_0 = (pos - 1 + curve_len) % curve_len, pos_1 = (pos + 1) % curve_len;
        normalized_diff(V + (size_t)curve[pos] * 3, V + (size_t)curve[pos_0] * 3, e);
        for (int d = 0; d < 3; ++d) tangent[d] += e[d];
        normalized_diff(V + (size_t)curve[pos_1] * 3, V + (size_t)curve[pos] * 3, e);
        for (int d = 0; d < 3; ++d) tangent[d] += e[d];
      } else if (!circles[c] && pos == 0) {
        const int pos_1 = (pos + 1) % curve_len;
        normalized_diff(V + (size_t)curve[pos_1] * 3, V + (size_t)curve[pos] * 3, e);
        for (int d = 0; d < 3; ++d) tangent[d] += e[d];
      } else if (!circles[c] && pos == curve_len - 1) {
        const int pos_0 = (pos - 1 + curve_len) % curve_len;
        normalized_diff(V + (size_t)curve[pos] * 3, V + (size_t)curve[pos_0] * 3, e);
        for (int d = 0; d < 3; ++d) tangent[d] += e[d];
      }
      const double nt = n3(tangent);
      for (int d = 0; d < 3; ++d) tangent[d] /= nt;
    }
    for (int d = 0; d < 3; ++d) { origin[(size_t)q * 3 + d] = pv[d]; axis[(size_t)q * 3 + d] = tangent[d]; }
  }
}

// ---------------------------------------------------------------------------------------
// Metro Hausdorff distance with vertex sampling only (compute(const Mesh&, const Mesh&, ...),
// metro_hausdorff.cpp:358-505; vcg Sampling::Hausdorff / VertexSampling / AddSample,
// extern/vcg/sampling.h:224-300,548-601): the box of all vertices of both meshes is grown by
// 2% of its diagonal on every side; each face-referenced vertex of one mesh is sampled
// against the other mesh's triangles at vcg's edge-plane point-to-face distance
// (face::PointDistanceEP, vcg/simplex/face/distance.h:48-180, on the planes and scaled edges
// of tri::UpdateEdges::ComputeEdgePlane, vcg/complex/algorithms/update/component_ep.h:48-73),
// brute force over the faces here; a sample counts when its distance is below the grown box's
// diagonal, and the counted distances are added in vertex order. out[0] = that diagonal,
// out[1] = the larger of the two directions' maxima, out[2] = the larger of their means.
namespace {
struct EPFace {
  double e[3][3];  // edges P1-P0, P2-P1, P0-P2 scaled by 1 / n[axis]
  double n[3], off;
  int axis;
  const double* P[3];
};

void unit(double n[3]) {
  const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (len > 0.0) { n[0] /= len; n[1] /= len; n[2] /= len; }
}

EPFace ep_face(const double* V, const int* F, int f) {
  EPFace r;
  for (int j = 0; j < 3; ++j) r.P[j] = V + (size_t)F[3 * (size_t)f + j] * 3;
  for (int j = 0; j < 3; ++j)
    for (int d = 0; d < 3; ++d) r.e[j][d] = r.P[(j + 1) % 3][d] - r.P[j][d];
  r.n[0] = r.e[0][1] * r.e[1][2] - r.e[0][2] * r.e[1][1];
  r.n[1] = r.e[0][2] * r.e[1][0] - r.e[0][0] * r.e[1][2];
  r.n[2] = r.e[0][0] * r.e[1][1] - r.e[0][1] * r.e[1][0];
  unit(r.n);                                                    // SetDirection (normalising plane)
  r.off = r.n[0] * r.P[0][0] + r.n[1] * r.P[0][1] + r.n[2] * r.P[0][2];
  unit(r.n);                                                    // Plane().Normalize()
  const double ax = std::fabs(r.n[0]), ay = std::fabs(r.n[1]), az = std::fabs(r.n[2]);
  double s;
  if (ax > ay && ax > az) { r.axis = 0; s = 1 / r.n[0]; }
  else if (ay > az) { r.axis = 1; s = 1 / r.n[1]; }
  else { r.axis = 2; s = 1 / r.n[2]; }
  for (int j = 0; j < 3; ++j)
    for (int d = 0; d < 3; ++d) r.e[j][d] *= s;
  return r;
}

double seg_dist(const double* p, const double* a, const double* b) {  // PSDist
  const double e[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double t = ((p[0] - a[0]) * e[0] + (p[1] - a[1]) * e[1] + (p[2] - a[2]) * e[2]) / (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  t = t < 0 ? 0 : (t > 1 ? 1 : t);
  const double c[3] = {a[0] + e[0] * t, a[1] + e[1] * t, a[2] + e[2] * t};
  const double w[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  return std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
}

// PointDistanceEP: lowers dist when the face is closer
void ep_distance(const EPFace& f, const double* q, double& dist) {
  const double d = (f.n[0] * q[0] + f.n[1] * q[1] + f.n[2] * q[2]) - f.off;
  if (d > dist || d < -dist) return;
  const double p[3] = {q[0] - f.n[0] * d, q[1] - f.n[1] * d, q[2] - f.n[2] * d};
  static const int U[3] = {1, 2, 0}, W[3] = {2, 0, 1};
  const int u = U[f.axis], w = W[f.axis];
  double b[3];
  const int order[3][3] = {{1, 1, 2}, {2, 2, 0}, {0, 0, 1}};  // edge, its start corner, its end corner
  for (int k = 0; k < 3; ++k) {
    const double* e = f.e[order[k][0]];
    const double* a = f.P[order[k][1]];
    b[k] = e[u] * (p[w] - a[w]) - e[w] * (p[u] - a[u]);
    if (b[k] <= 0) {
      const double t = seg_dist(q, f.P[order[k][1]], f.P[(order[k][1] + 1) % 3]);
      if (dist > t) dist = t;
      return;
    }
  }
  const double bm = std::min(b[0], std::min(b[1], b[2]));
  const double* P0 = f.P[0];
  const double ab[3] = {f.P[1][0] - P0[0], f.P[1][1] - P0[1], f.P[1][2] - P0[2]};
  const double ac[3] = {f.P[2][0] - P0[0], f.P[2][1] - P0[1], f.P[2][2] - P0[2]};
  const double x[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  if (bm < 0.000001 * std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2])) {
    const int k = bm == b[0] ? 0 : (bm == b[1] ? 1 : 2);
    const double t = seg_dist(q, f.P[order[k][1]], f.P[(order[k][1] + 1) % 3]);
    if (dist > t) dist = t;
    return;
  }
  dist = std::fabs(d);
}

void metro_direction(int nvA, const double* VA, int nfA, const int* FA, int nfB, const double* VB, const int* FB,
                     double upper, double& mx, double& mean) {
  std::vector<char> ref(nvA, 0);
  for (long i = 0; i < 3L * nfA; ++i) ref[FA[i]] = 1;
  std::vector<EPFace> faces(nfB);
  for (int f = 0; f < nfB; ++f) faces[f] = ep_face(VB, FB, f);
  mx = -HUGE_VAL;
  double sum = 0;
  long n = 0;
  for (int v = 0; v < nvA; ++v) {
    if (!ref[v]) continue;
    const double* p = VA + (size_t)v * 3;
    double d = upper;
    for (int f = 0; f < nfB; ++f) ep_distance(faces[f], p, d);
    if (d == upper) continue;
    if (d > mx) mx = d;
    sum += d;
    ++n;
  }
  mean = sum / (double)n;
}
}  // namespace

extern "C" void orc_metro_hausdorff(int nv0, const double* V0, int nf0, const int* F0, int nv1, const double* V1,
                                    int nf1, const int* F1, double* out) {
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = 0; i < nv0; ++i)
    for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], V0[(size_t)i * 3 + d]); hi[d] = std::max(hi[d], V0[(size_t)i * 3 + d]); }
  for (int i = 0; i < nv1; ++i)
    for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], V1[(size_t)i * 3 + d]); hi[d] = std::max(hi[d], V1[(size_t)i * 3 + d]); }
  auto diag = [&]() {
    const double a = hi[0] - lo[0], b = hi[1] - lo[1], c = hi[2] - lo[2];
    return std::sqrt(a * a + b * b + c * c);
  };
  const double off = diag() * 0.02;
  for (int d = 0; d < 3; ++d) { lo[d] -= off; hi[d] += off; }
  const double D = diag();
  double m1, a1, m2, a2;
  metro_direction(nv0, V0, nf0, F0, nf1, V1, F1, D, m1, a1);
  metro_direction(nv1, V1, nf1, F1, nf0, V0, F0, D, m2, a2);
  out[0] = D;
  out[1] = std::max(m1, m2);
  out[2] = std::max(a2, a1);
}
