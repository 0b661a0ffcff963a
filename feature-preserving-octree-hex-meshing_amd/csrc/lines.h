// lines.h -- feature-curve projection of LINE vertices (dirty_graph_projection's LINE branch).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hexopt {

// Curves as CSR lists of surface vertex ids: curve c is vs[off[c] .. off[c+1]); circles[c]
// adds the closing segment.
struct CurvesDev {
  const double* V;  // surface vertices, rows of 3
  const int* off;
  const int* vs;
  const uint8_t* circles;
};

// For n LINE vertices at rows `rows` of SoA positions X (x[nV], y[nV], z[nV]): closest point
// origin (n x 3) on curve curve_of[i] and that segment's unit tangent axis (n x 3).
void launch_project_lines(const CurvesDev& cv, const double* X, int nV, int n, const int* rows, const int* curve_of,
                          double* origin, double* axis, hipStream_t st);

struct SoftDev;
// Writes projected origins/axes and the weight sd.lamL into the line constraint rows (axis
// lt, origin lo, folded S0/S and b0/b rows, a_L rhs and inverse diagonal) for line vertices
// without other constraints.
void launch_refold_lines(const SoftDev& sd, const double* origin, const double* axis, double* S0, double* b0,
                         double* rhs_a, double* invd_a, hipStream_t st);

}  // namespace hexopt
