// bvh_host.cpp -- input-surface set-up for the closest-point kernel: the AABB tree in the
// exact shape igl::AABB::init builds (AABB.cpp:93-200: median split of barycentre ranks along
// the longest box axis), flattened for the device, and the face / angle-weighted vertex /
// uniform edge normals of build_aabb_tree (grid_hex_meshing.cpp:4231-4248).
#include "bvh.h"
#include "occupancy.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace hexopt {
namespace {

struct Builder {
  const double* V;
  const int* F;
  std::vector<int> SI;  // nf*3 sorted rank of each barycentre coordinate
  SurfaceTree* T;
  bool ok = true;

  int build(const std::vector<int>& I) {
    const int id = (int)T->left.size();
    T->left.push_back(-1);
    T->right.push_back(-1);
    T->prim.push_back(-1);
    for (int d = 0; d < 6; ++d) T->box.push_back(d < 3 ? INFINITY : -INFINITY);
    double* bx = &T->box[(size_t)id * 6];
    for (int i : I)
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
          const double v = V[(size_t)F[(size_t)i * 3 + c] * 3 + d];
          bx[d] = std::min(bx[d], v);
          bx[3 + d] = std::max(bx[3 + d], v);
        }
    if (I.size() == 1) {
      T->prim[id] = I[0];
      return id;
    }
    int max_d = 0;
    double best = bx[3] - bx[0];
    for (int d = 1; d < 3; ++d)
      if (bx[3 + d] - bx[d] > best) { best = bx[3 + d] - bx[d]; max_d = d; }
    std::vector<int> S(I.size());
    for (size_t i = 0; i < I.size(); ++i) S[i] = SI[(size_t)I[i] * 3 + max_d];
    std::vector<int> A = S;
    const size_t n = A.size() / 2;
    std::nth_element(A.begin(), A.begin() + n, A.end());
    double med;
    if (A.size() % 2 == 1) {
      med = A[n];
    } else {
      std::nth_element(A.begin(), A.begin() + n - 1, A.end());
      med = 0.5 * (A[n] + A[n - 1]);
    }
    std::vector<int> Lh, Rh;
    for (size_t i = 0; i < I.size(); ++i) (S[i] <= med ? Lh : Rh).push_back(I[i]);
    if (Lh.size() != (I.size() + 1) / 2) ok = false;
    if (!Lh.empty()) { int l = build(Lh); T->left[id] = l; }
    if (!Rh.empty()) { int r = build(Rh); T->right[id] = r; }
    return id;
  }
};

double vnorm(const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

}  // namespace

bool build_surface_tree(int nv, const double* V, int nf, const int* F, SurfaceTree& T, const char** err) {
  T = SurfaceTree();
  if (nf <= 0 || nv <= 0) { *err = "empty surface"; return false; }
  for (long i = 0; i < 3L * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) { *err = "face index out of range"; return false; }
  Builder b;
  b.V = V;
  b.F = F;
  b.T = &T;
  // barycentres (igl::barycenter) and their ranks per dimension (igl::sort + inverse)
  std::vector<double> BC((size_t)nf * 3);
  for (int f = 0; f < nf; ++f)
    for (int d = 0; d < 3; ++d) {
      double s = 0;
      for (int c = 0; c < 3; ++c) s += V[(size_t)F[(size_t)f * 3 + c] * 3 + d];
      BC[(size_t)f * 3 + d] = s / 3.0;
    }
  b.SI.assign((size_t)nf * 3, 0);
  for (int d = 0; d < 3; ++d) {
    std::vector<size_t> idx(nf);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<double> col(nf);
    for (int f = 0; f < nf; ++f) col[f] = BC[(size_t)f * 3 + d];
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t c) { return col[a] < col[c]; });
    for (int r = 0; r < nf; ++r) b.SI[idx[r] * 3 + d] = r;
  }
  std::vector<int> all(nf);
  std::iota(all.begin(), all.end(), 0);
  b.build(all);
  if (!b.ok) { *err = "unbalanced median split"; return false; }
  // depth
  std::vector<int> depth(T.left.size(), 0);
  int maxd = 0;
  for (size_t i = 0; i < T.left.size(); ++i) {
    maxd = std::max(maxd, depth[i]);
    if (T.left[i] >= 0) depth[T.left[i]] = depth[i] + 1;
    if (T.right[i] >= 0) depth[T.right[i]] = depth[i] + 1;
  }
  T.depth = maxd + 1;
  return true;
}

void pack_node_records(const SurfaceTree& T, const double* V, const int* F, std::vector<double>& rec) {
  const size_t nn = T.prim.size();
  rec.assign(nn * kNodeRec, 0.0);
  for (size_t i = 0; i < nn; ++i) {
    double* r = &rec[i * kNodeRec];
    if (T.prim[i] >= 0) {
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) r[c * 3 + d] = V[(size_t)F[(size_t)T.prim[i] * 3 + c] * 3 + d];
    } else {
      for (int d = 0; d < 6; ++d) r[d] = T.box[(size_t)T.left[i] * 6 + d];
      for (int d = 0; d < 6; ++d) r[6 + d] = T.box[(size_t)T.right[i] * 6 + d];
    }
    const int32_t w[4] = {T.left[i], T.right[i], T.prim[i], 0};
    std::memcpy(r + 12, w, sizeof(w));
  }
}

void surface_normals(int nv, const double* V, int nf, const int* F, std::vector<double>& FN, std::vector<double>& VN,
                     std::vector<double>& FEN) {
  // per_face_normals: normalised cross product, zero for degenerate faces
  FN.assign((size_t)nf * 3, 0.0);
  for (int f = 0; f < nf; ++f) {
    const double* a = &V[(size_t)F[f * 3] * 3];
    const double* b = &V[(size_t)F[f * 3 + 1] * 3];
    const double* c = &V[(size_t)F[f * 3 + 2] * 3];
    double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
    double r = vnorm(n);
    if (r != 0)
      for (int d = 0; d < 3; ++d) FN[(size_t)f * 3 + d] = n[d] / r;
  }
  // per_vertex_normals, angle weighting (internal_angles from squared edge lengths)
  VN.assign((size_t)nv * 3, 0.0);
  for (int f = 0; f < nf; ++f) {
    double l2[3];
    for (int d = 0; d < 3; ++d) {
      const double* p = &V[(size_t)F[f * 3 + (d + 1) % 3] * 3];
      const double* q = &V[(size_t)F[f * 3 + (d + 2) % 3] * 3];
      l2[d] = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
    }
    for (int d = 0; d < 3; ++d) {
      const double s1 = l2[d], s2 = l2[(d + 1) % 3], s3 = l2[(d + 2) % 3];
      const double w = std::acos((s3 + s2 - s1) / (2. * std::sqrt(s3 * s2)));
      for (int k = 0; k < 3; ++k) VN[(size_t)F[f * 3 + d] * 3 + k] += w * FN[(size_t)f * 3 + k];
    }
  }
  for (int v = 0; v < nv; ++v) {
    double r = vnorm(&VN[(size_t)v * 3]);
    if (r != 0)
      for (int k = 0; k < 3; ++k) VN[(size_t)v * 3 + k] /= r;
  }
  // per_edge_normals, uniform weighting: sum of the normals of the faces on the undirected
  // edge (not normalised); FEN[f][x] is the edge opposite corner x of face f
  std::vector<std::pair<std::pair<int, int>, int>> e;
  e.reserve((size_t)nf * 3);
  for (int f = 0; f < nf; ++f)
    for (int x = 0; x < 3; ++x) {
      int a = F[f * 3 + (x + 1) % 3], b = F[f * 3 + (x + 2) % 3];
      e.push_back({{std::min(a, b), std::max(a, b)}, f * 3 + x});
    }
  std::sort(e.begin(), e.end());
  FEN.assign((size_t)nf * 9, 0.0);
  for (size_t i = 0; i < e.size();) {
    size_t j = i;
    double s[3] = {0, 0, 0};
    // faces in increasing (face, corner) order, as the reference's face loop adds them
    for (; j < e.size() && e[j].first == e[i].first; ++j) {
      const int f = e[j].second / 3;
      for (int k = 0; k < 3; ++k) s[k] += FN[(size_t)f * 3 + k];
    }
    for (size_t k = i; k < j; ++k)
      for (int d = 0; d < 3; ++d) FEN[(size_t)e[k].second * 3 + d] = s[d];
    i = j;
  }
}

void build_column_bins(int nv, const double* V, int nf, const int* F, ColumnBins& B) {
  (void)nv;
  B = ColumnBins();
  B.fbox.assign((size_t)nf * 4, 0.0);
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int f = 0; f < nf; ++f) {
    double* fb = &B.fbox[(size_t)f * 4];
    for (int d = 0; d < 2; ++d) {
      double mn = V[(size_t)F[(size_t)f * 3] * 3 + d], mx = mn;
      for (int c = 1; c < 3; ++c) {
        const double v = V[(size_t)F[(size_t)f * 3 + c] * 3 + d];
        mn = std::min(mn, v);
        mx = std::max(mx, v);
      }
      fb[2 * d] = mn;
      fb[2 * d + 1] = mx;
      lo[d] = std::min(lo[d], mn);
      hi[d] = std::max(hi[d], mx);
    }
  }
  // about one bin per facet, shaped like the xy box, at most 4096 per axis
  const double ex = hi[0] - lo[0], ey = hi[1] - lo[1];
  int nbx = 1, nby = 1;
  if (ex > 0 && ey > 0) {
    nbx = (int)std::lround(std::sqrt((double)nf * ex / ey));
    nby = (int)std::lround(std::sqrt((double)nf * ey / ex));
  } else if (ex > 0) {
    nbx = nf;
  } else if (ey > 0) {
    nby = nf;
  }
  nbx = std::min(4096, std::max(1, nbx));
  nby = std::min(4096, std::max(1, nby));
  for (int d = 0; d < 2; ++d) {
    B.lo[d] = lo[d];
    B.hi[d] = hi[d];
  }
  std::vector<int> r((size_t)nf * 4);
  // coarsen while facets spanning many bins would make the lists much longer than 16 per facet
  const long long cap = 16LL * nf + (1LL << 20);
  for (;;) {
    B.nb[0] = nbx;
    B.nb[1] = nby;
    B.inv[0] = ex > 0 ? nbx / ex : 0.0;
    B.inv[1] = ey > 0 ? nby / ey : 0.0;
    long long total = 0;
    for (int f = 0; f < nf; ++f) {
      const double* fb = &B.fbox[(size_t)f * 4];
      int* q = &r[(size_t)f * 4];
      q[0] = column_bin(fb[0], B.lo[0], B.inv[0], nbx);
      q[1] = column_bin(fb[1], B.lo[0], B.inv[0], nbx);
      q[2] = column_bin(fb[2], B.lo[1], B.inv[1], nby);
      q[3] = column_bin(fb[3], B.lo[1], B.inv[1], nby);
      total += (long long)(q[1] - q[0] + 1) * (q[3] - q[2] + 1);
    }
    if (total <= cap || (nbx == 1 && nby == 1)) break;
    nbx = std::max(1, nbx / 2);
    nby = std::max(1, nby / 2);
  }
  B.off.assign((size_t)nbx * nby + 1, 0);
  for (int f = 0; f < nf; ++f) {
    const int* q = &r[(size_t)f * 4];
    for (int by = q[2]; by <= q[3]; ++by)
      for (int bx = q[0]; bx <= q[1]; ++bx) ++B.off[(size_t)by * nbx + bx + 1];
  }
  for (size_t k = 0; k + 1 < B.off.size(); ++k) B.off[k + 1] += B.off[k];
  B.fid.assign((size_t)B.off.back(), 0);
  std::vector<int> fill(B.off.begin(), B.off.end() - 1);
  for (int f = 0; f < nf; ++f) {
    const int* q = &r[(size_t)f * 4];
    for (int by = q[2]; by <= q[3]; ++by)
      for (int bx = q[0]; bx <= q[1]; ++bx) B.fid[fill[(size_t)by * nbx + bx]++] = f;
  }
}

// vcg tri::UpdateEdges<>::ComputeEdgePlane (vcg/complex/algorithms/update/component_ep.h:48-73)
// with the face's Plane3<double, true>: SetDirection normalises the cross product of the first
// two edges, the offset is taken with that direction, then Normalize() runs once more; the
// edges are scaled by the inverse of the normal's dominant component.
void build_ep_faces(int nv, const double* V, int nf, const int* F, std::vector<double>& ep) {
  (void)nv;
  ep.assign((size_t)nf * kEPRec, 0.0);
  auto normalize = [](double n[3]) {
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0.0) {
      n[0] /= len;
      n[1] /= len;
      n[2] /= len;
    }
  };
  for (int f = 0; f < nf; ++f) {
    const double* P[3] = {V + (size_t)F[3 * (size_t)f] * 3, V + (size_t)F[3 * (size_t)f + 1] * 3,
                          V + (size_t)F[3 * (size_t)f + 2] * 3};
    double E[3][3];
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) E[j][d] = P[(j + 1) % 3][d] - P[j][d];
    double n[3] = {E[0][1] * E[1][2] - E[0][2] * E[1][1], E[0][2] * E[1][0] - E[0][0] * E[1][2],
                   E[0][0] * E[1][1] - E[0][1] * E[1][0]};
    normalize(n);
    const double off = n[0] * P[0][0] + n[1] * P[0][1] + n[2] * P[0][2];
    normalize(n);
    const double nx = std::fabs(n[0]), ny = std::fabs(n[1]), nz = std::fabs(n[2]);
    int axis;
    double dinv;
    if (nx > ny && nx > nz) {
      axis = 0;
      dinv = 1 / n[0];
    } else if (ny > nz) {
      axis = 1;
      dinv = 1 / n[1];
    } else {
      axis = 2;
      dinv = 1 / n[2];
    }
    double* r = &ep[(size_t)f * kEPRec];
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) r[3 * j + d] = E[j][d] * dinv;
    for (int d = 0; d < 3; ++d) r[9 + d] = n[d];
    r[12] = off;
    r[13] = (double)axis;
  }
}

}  // namespace hexopt
