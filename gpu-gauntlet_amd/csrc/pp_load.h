// Consistent nodal pressure loads on tensor-product grids, one node per call.
//
// The grid has columns at xc[0..nx] and rows at yc[0..ny] (FP64 copies of
// the float32 node coordinates).  Pressure is p(x, y) = q0 + sum_k A_k
// exp(-|x - c_k|^2 / (2 s_k^2)) with patches packed (A, cx, cy, s).  A node
// gets its share of every adjacent element's quadrature:
//   Q4 cells      2x2 Gauss rule, N_corner * p * |J| (element.ts:29-41)
//   split cells   triangles (n0,n1,n2) and (n0,n2,n3) of the cell with
//                 corners n0 = (i,j), n1 = (i+1,j), n2 = (i+1,j+1),
//                 n3 = (i,j+1); 3-point rule (barycentric points
//                 (1/6,1/6,2/3) and its rotations), weight area/3, share
//                 the node's barycentric coordinate
// which is the element loop of the consistent load vector
// (solver.ts:363-407 with the Gaussian-patch pressure) regrouped by node.
// Sums run in a fixed order: adjacent cells (i-1,j-1), (i,j-1), (i,j),
// (i-1,j), then quadrature points.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace pp {

__device__ __forceinline__ double patch_pressure(double x, double y, double q0, const double* pt,
                                                 int np) {
  double p = q0;
  for (int k = 0; k < np; k++) {
    const double dx = x - pt[4 * k + 1], dy = y - pt[4 * k + 2], s = pt[4 * k + 3];
    p += pt[4 * k] * exp(-(dx * dx + dy * dy) / (2.0 * s * s));
  }
  return p;
}

__device__ inline double grid_node_load_q4(const double* xc, const double* yc, int nx, int ny,
                                           int i, int j, double q0, const double* pt, int np) {
  const int ei[4] = {i - 1, i, i, i - 1}, ej[4] = {j - 1, j - 1, j, j};
  const int corner[4] = {2, 3, 0, 1};
  const double g = 0.5773502691896257;
  const double gxi[4] = {-g, g, g, -g}, geta[4] = {-g, -g, g, g};
  const double cxi[4] = {-1, 1, 1, -1}, ceta[4] = {-1, -1, 1, 1};
  double f = 0;
  for (int k = 0; k < 4; k++) {
    const int gi = ei[k], gj = ej[k];
    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) continue;
    const double x0 = xc[gi], x1 = xc[gi + 1], y0 = yc[gj], y1 = yc[gj + 1];
    const double detJ = 0.25 * (x1 - x0) * (y1 - y0);
    for (int q = 0; q < 4; q++) {
      const double xi = gxi[q], eta = geta[q];
      const double px = 0.5 * (1 - xi) * x0 + 0.5 * (1 + xi) * x1;
      const double py = 0.5 * (1 - eta) * y0 + 0.5 * (1 + eta) * y1;
      const double N = 0.25 * (1 + cxi[corner[k]] * xi) * (1 + ceta[corner[k]] * eta);
      f += N * patch_pressure(px, py, q0, pt, np) * fabs(detJ);
    }
  }
  return f;
}

// One triangle's share for the vertex at position `pos` (0..2).
__device__ __forceinline__ double tri_share(const double (&xy)[6], int pos, double q0,
                                            const double* pt, int np) {
  const double L[3][3] = {{1.0 / 6, 1.0 / 6, 2.0 / 3}, {2.0 / 3, 1.0 / 6, 1.0 / 6},
                          {1.0 / 6, 2.0 / 3, 1.0 / 6}};
  const double twoA = (xy[0] - xy[2]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[1] - xy[3]);
  const double area = fabs(twoA) / 2;
  double f = 0;
  for (int q = 0; q < 3; q++) {
    const double px = L[q][0] * xy[0] + L[q][1] * xy[2] + L[q][2] * xy[4];
    const double py = L[q][0] * xy[1] + L[q][1] * xy[3] + L[q][2] * xy[5];
    f += L[q][pos] * (patch_pressure(px, py, q0, pt, np) * area / 3.0);
  }
  return f;
}

__device__ inline double grid_node_load_split(const double* xc, const double* yc, int nx, int ny,
                                              int i, int j, double q0, const double* pt, int np) {
  const int ei[4] = {i - 1, i, i, i - 1}, ej[4] = {j - 1, j - 1, j, j};
  const int corner[4] = {2, 3, 0, 1};
  // position of each cell corner in triangle A = (n0,n1,n2), B = (n0,n2,n3)
  const int posA[4] = {0, 1, 2, -1}, posB[4] = {0, -1, 1, 2};
  double f = 0;
  for (int k = 0; k < 4; k++) {
    const int gi = ei[k], gj = ej[k];
    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) continue;
    const double x0 = xc[gi], x1 = xc[gi + 1], y0 = yc[gj], y1 = yc[gj + 1];
    const int c = corner[k];
    if (posA[c] >= 0) {
      const double a[6] = {x0, y0, x1, y0, x1, y1};
      f += tri_share(a, posA[c], q0, pt, np);
    }
    if (posB[c] >= 0) {
      const double b[6] = {x0, y0, x1, y1, x0, y1};
      f += tri_share(b, posB[c], q0, pt, np);
    }
  }
  return f;
}

}  // namespace pp
