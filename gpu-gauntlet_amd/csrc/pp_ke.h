// Q4 Mindlin element stiffness, shared by the coloured path's device set-up
// and the multigrid coarse-operator construction.
#pragma once

#include <math.h>

namespace pp {

// Q4 Mindlin 12x12 for a general quadrilateral (element.ts:258-483).
__host__ __device__ inline void ke_q4(const double* xy, double E, double nu, double t, double* Ke) {
  const double D = E * t * t * t / (12.0 * (1.0 - nu * nu));
  const double D11 = D, D12 = D * nu, D33 = D * (1.0 - nu) / 2.0;
  const double Ds = (5.0 / 6.0) * E / (2.0 * (1.0 + nu)) * t;
  for (int k = 0; k < 144; k++) Ke[k] = 0.0;
  auto phys = [&](double xi, double eta, double* dx, double* dy) {
    const double dxi[4] = {-(1 - eta) / 4, (1 - eta) / 4, (1 + eta) / 4, -(1 + eta) / 4};
    const double deta[4] = {-(1 - xi) / 4, -(1 + xi) / 4, (1 + xi) / 4, (1 - xi) / 4};
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
    for (int i = 0; i < 4; i++) {
      J11 += dxi[i] * xy[2 * i]; J12 += dxi[i] * xy[2 * i + 1];
      J21 += deta[i] * xy[2 * i]; J22 += deta[i] * xy[2 * i + 1];
    }
    const double det = J11 * J22 - J12 * J21, inv = 1.0 / det;
    for (int i = 0; i < 4; i++) {
      dx[i] = inv * (J22 * dxi[i] - J12 * deta[i]);
      dy[i] = inv * (-J21 * dxi[i] + J11 * deta[i]);
    }
    return det;
  };
  const double g = 0.5773502691896257;
  const double gp[4][2] = {{-g, -g}, {g, -g}, {g, g}, {-g, g}};
  for (int q = 0; q < 4; q++) {
    double dx[4], dy[4];
    const double f = fabs(phys(gp[q][0], gp[q][1], dx, dy));
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        const int ti = 3 * i + 1, yi = 3 * i + 2, tj = 3 * j + 1, yj = 3 * j + 2;
        Ke[ti * 12 + tj] += f * (dy[i] * D11 * dy[j] + dx[i] * D33 * dx[j]);
        Ke[yi * 12 + yj] += f * (dx[i] * D11 * dx[j] + dy[i] * D33 * dy[j]);
        Ke[ti * 12 + yj] += f * (-dy[i] * D12 * dx[j] - dx[i] * D33 * dy[j]);
        Ke[yi * 12 + tj] += f * (-dx[i] * D12 * dy[j] - dy[i] * D33 * dx[j]);
      }
  }
  double dx[4], dy[4];
  const double f = 4.0 * fabs(phys(0, 0, dx, dy));
  const double N = 0.25;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      const int wi = 3 * i, wj = 3 * j;
      Ke[wi * 12 + wj] += f * Ds * (dx[i] * dx[j] + dy[i] * dy[j]);
      Ke[wi * 12 + wj + 1] += f * Ds * dy[i] * (-N);
      Ke[wi * 12 + wj + 2] += f * Ds * dx[i] * N;
      Ke[(wi + 1) * 12 + wj] += f * Ds * (-N) * dy[j];
      Ke[(wi + 1) * 12 + wj + 1] += f * Ds * N * N;
      Ke[(wi + 2) * 12 + wj] += f * Ds * N * dx[j];
      Ke[(wi + 2) * 12 + wj + 2] += f * Ds * N * N;
    }
}

}  // namespace pp
