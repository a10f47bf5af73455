/*
 * FP64 CPU oracle for the plate-bending PCG solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is never linked into the product.  Only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load the library built
 * from it.
 *
 * It restates, in double precision, the reference's CPU solve path
 * (cseifert512/gpu-gauntlet, src/lib/plate):
 *   - Q4 Mindlin element stiffness        element.ts:258-483 (+ :56-240)
 *   - "DKT" 9x9 element stiffness         element.ts:642-943 (as written)
 *   - constrained DOF identification      solver.ts:92-134
 *   - 3x3 nodal block diagonal + inverse  solver.ts:188-304
 *   - block preconditioner apply          pcg.ts:74-88, solver.ts:313-339
 *   - matrix-free colour-ordered K.x      solver.ts:422-483
 *   - PCG with the OR stopping rule       pcg.ts:98-215
 *   - tridiagonal known-answer system     pcg.ts:221-247
 *   - nodal bending moments               postprocess.ts:40-270
 * and adds what the reference lacks but the workloads need:
 *   - consistent pressure-load integration (uniform + Gaussian patches)
 *     over Q4 (2x2 Gauss) and triangles (3-point rule).  The reference has
 *     only nearest-node point loads (solver.ts:363-392); this part is pinned
 *     by closed forms in the tests (total load = integral of the pressure).
 *   - per-element thickness (Ke scales by t^3 for bending, t for shear,
 *     element.ts:205,236-237).
 *
 * Everything is double: coordinates come in as the float32 node array the
 * reference mesher produces (mesher.ts:381-388) and are widened exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ---------------------------------------------------------------- material */

static void orc_material(double E, double nu, double t, double* D11, double* D12,
                         double* D33, double* Ds) {
  /* element.ts:202-240 and types.ts:26-28 */
  const double D = E * t * t * t / (12.0 * (1.0 - nu * nu));
  *D11 = D;
  *D12 = D * nu;
  *D33 = D * (1.0 - nu) / 2.0;
  const double kappa = 5.0 / 6.0;
  const double G = E / (2.0 * (1.0 + nu));
  *Ds = kappa * G * t;
}

/* ------------------------------------------------------------- Q4 Mindlin */

static void q4_dN(double xi, double eta, double dxi[4], double deta[4]) {
  /* element.ts:80-106 */
  const double xm = 1 - xi, xp = 1 + xi, em = 1 - eta, ep = 1 + eta;
  dxi[0] = -em / 4; dxi[1] = em / 4; dxi[2] = ep / 4; dxi[3] = -ep / 4;
  deta[0] = -xm / 4; deta[1] = -xp / 4; deta[2] = xp / 4; deta[3] = xm / 4;
}

static double q4_phys(const double* xy, double xi, double eta, double dNdx[4],
                      double dNdy[4]) {
  /* element.ts:119-184: Jacobian, determinant and J^-1 applied to dN/dxi */
  double dxi[4], deta[4];
  q4_dN(xi, eta, dxi, deta);
  double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
  for (int i = 0; i < 4; i++) {
    J11 += dxi[i] * xy[2 * i];
    J12 += dxi[i] * xy[2 * i + 1];
    J21 += deta[i] * xy[2 * i];
    J22 += deta[i] * xy[2 * i + 1];
  }
  const double detJ = J11 * J22 - J12 * J21;
  const double inv = 1.0 / detJ;
  for (int i = 0; i < 4; i++) {
    dNdx[i] = inv * (J22 * dxi[i] - J12 * deta[i]);
    dNdy[i] = inv * (-J21 * dxi[i] + J11 * deta[i]);
  }
  return detJ;
}

/* Q4 Mindlin 12x12 stiffness, row-major, DOFs [w,thx,thy] per node.
 * element.ts:258-483: bending 2x2 Gauss (weights 1), curvature
 * {d thy/dx, -d thx/dy, d thy/dy - d thx/dx}; shear 1-point (weight 4),
 * gamma {dw/dx + thy, dw/dy - thx}. */
ORC_API void orc_ke_q4(const double* xy, double E, double nu, double t, double* Ke) {
  static const double g = 0.5773502691896257;
  static const double gp[4][2] = {{-g, -g}, {g, -g}, {g, g}, {-g, g}};
  double D11, D12, D33, Ds;
  orc_material(E, nu, t, &D11, &D12, &D33, &Ds);
  const double D22 = D11;
  memset(Ke, 0, 144 * sizeof(double));
  for (int q = 0; q < 4; q++) {
    double dNdx[4], dNdy[4];
    const double detJ = q4_phys(xy, gp[q][0], gp[q][1], dNdx, dNdy);
    const double f = 1.0 * fabs(detJ);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        const int tx_i = 3 * i + 1, ty_i = 3 * i + 2, tx_j = 3 * j + 1, ty_j = 3 * j + 2;
        const double xi_ = dNdx[i], yi_ = dNdy[i], xj_ = dNdx[j], yj_ = dNdy[j];
        Ke[tx_i * 12 + tx_j] += f * (yi_ * D22 * yj_ + xi_ * D33 * xj_);
        Ke[ty_i * 12 + ty_j] += f * (xi_ * D11 * xj_ + yi_ * D33 * yj_);
        Ke[tx_i * 12 + ty_j] += f * (-yi_ * D12 * xj_ - xi_ * D33 * yj_);
        Ke[ty_i * 12 + tx_j] += f * (-xi_ * D12 * yj_ - yi_ * D33 * xj_);
      }
  }
  {
    double dNdx[4], dNdy[4];
    const double detJ = q4_phys(xy, 0.0, 0.0, dNdx, dNdy);
    const double f = 4.0 * fabs(detJ);
    const double N = 0.25; /* all four shape functions at the centre */
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        const int w_i = 3 * i, tx_i = w_i + 1, ty_i = w_i + 2;
        const int w_j = 3 * j, tx_j = w_j + 1, ty_j = w_j + 2;
        Ke[w_i * 12 + w_j] += f * (dNdx[i] * Ds * dNdx[j] + dNdy[i] * Ds * dNdy[j]);
        Ke[w_i * 12 + tx_j] += f * (dNdy[i] * Ds * (-N));
        Ke[w_i * 12 + ty_j] += f * (dNdx[i] * Ds * N);
        Ke[tx_i * 12 + w_j] += f * ((-N) * Ds * dNdy[j]);
        Ke[tx_i * 12 + tx_j] += f * ((-N) * Ds * (-N));
        Ke[ty_i * 12 + w_j] += f * (N * Ds * dNdx[j]);
        Ke[ty_i * 12 + ty_j] += f * (N * Ds * N);
      }
  }
}

/* --------------------------------------------------------------- "DKT" 9x9 */

/* element.ts:833-887 */
static void dkt_bx(double L1, double L2, const double P[3], const double q[3],
                   const double c[3], double d1[9], double d2[9]) {
  const double P4 = P[0], P5 = P[1], P6 = P[2], q4 = q[0], q5 = q[1], q6 = q[2];
  const double c4 = c[0], c5 = c[1], c6 = c[2];
  d1[0] = P6 * (1 - 2 * L1) + (P5 - P6) * L2;
  d1[1] = q6 * (1 - 2 * L1) - (q5 + q6) * L2;
  d1[2] = -4 + 6 * (L1 + L2) + (c5 - c6) * L2 + c6 * (1 - 2 * L1);
  d2[0] = -P6 * (1 - 2 * L2) + (P4 - P5) * L1 - P5 * (1 - 2 * L2);
  d2[1] = q6 * (1 - 2 * L2) - (q4 + q5) * L1 + q5 * (1 - 2 * L2);
  d2[2] = -4 + 6 * (L1 + L2) + (c4 - c5) * L1 + c5 * (1 - 2 * L2);
  d1[3] = P4 * L2 + P6 * (1 - 2 * L1);
  d1[4] = q4 * L2 - q6 * (1 - 2 * L1);
  d1[5] = c4 * L2 - c6 * (1 - 2 * L1);
  d2[3] = P4 * (1 - 2 * L2) + P5 * L1;
  d2[4] = q4 * (1 - 2 * L2) - q5 * L1;
  d2[5] = c4 * (1 - 2 * L2) - c5 * L1;
  d1[6] = -P5 * L2 - P6 * (1 - 2 * L1);
  d1[7] = q5 * L2 + q6 * (1 - 2 * L1);
  d1[8] = -c5 * L2 + c6 * (1 - 2 * L1);
  d2[6] = -P4 * L1 - P5 * (1 - 2 * L2);
  d2[7] = q4 * L1 + q5 * (1 - 2 * L2);
  d2[8] = -c4 * L1 + c5 * (1 - 2 * L2);
}

/* element.ts:893-943 */
static void dkt_by(double L1, double L2, const double T[3], const double r[3],
                   const double e[3], const double b[3], double d1[9], double d2[9]) {
  const double t4 = T[0], t5 = T[1], t6 = T[2], r4 = r[0], r5 = r[1], r6 = r[2];
  const double e4 = e[0], e5 = e[1], e6 = e[2], b4 = b[0], b5 = b[1], b6 = b[2];
  d1[0] = t6 * (1 - 2 * L1) + (t5 - t6) * L2;
  d1[1] = 1 + r6 * (1 - 2 * L1) - (r5 + r6) * L2;
  d1[2] = -e6 * (1 - 2 * L1) + (b5 - b6) * L2;
  d2[0] = -t6 * (1 - 2 * L2) + (t4 - t5) * L1 - t5 * (1 - 2 * L2);
  d2[1] = 1 + r6 * (1 - 2 * L2) - (r4 + r5) * L1 + r5 * (1 - 2 * L2);
  d2[2] = e6 * (1 - 2 * L2) + (b4 - b5) * L1 - b5 * (1 - 2 * L2);
  d1[3] = t4 * L2 + t6 * (1 - 2 * L1);
  d1[4] = -1 + r4 * L2 - r6 * (1 - 2 * L1);
  d1[5] = -e4 * L2 - b6 * (1 - 2 * L1);
  d2[3] = t4 * (1 - 2 * L2) + t5 * L1;
  d2[4] = -1 + r4 * (1 - 2 * L2) - r5 * L1;
  d2[5] = -e4 * (1 - 2 * L2) + b5 * L1;
  d1[6] = -t5 * L2 - t6 * (1 - 2 * L1);
  d1[7] = r5 * L2 + r6 * (1 - 2 * L1);
  d1[8] = e5 * L2 + b6 * (1 - 2 * L1);
  d2[6] = -t4 * L1 - t5 * (1 - 2 * L2);
  d2[7] = r4 * L1 + r5 * (1 - 2 * L2);
  d2[8] = e4 * L1 - b5 * (1 - 2 * L2);
}

/* 9x9 triangle stiffness exactly as element.ts:642-827 writes it (including
 * its signed 1/(2A) and its non-standard B columns), in double. */
ORC_API void orc_ke_dkt(const double* xy, double E, double nu, double t, double* Ke) {
  memset(Ke, 0, 81 * sizeof(double));
  const double x1 = xy[0], y1 = xy[1], x2 = xy[2], y2 = xy[3], x3 = xy[4], y3 = xy[5];
  const double x12 = x1 - x2, y12 = y1 - y2, x23 = x2 - x3, y23 = y2 - y3;
  const double x31 = x3 - x1, y31 = y3 - y1;
  const double L12 = x12 * x12 + y12 * y12, L23 = x23 * x23 + y23 * y23;
  const double L31 = x31 * x31 + y31 * y31;
  const double twoA = x12 * y31 - x31 * y12;
  const double area = fabs(twoA) / 2;
  if (area < 1e-15) return; /* element.ts:672-676 */
  double D11, D12, D33, Ds;
  orc_material(E, nu, t, &D11, &D12, &D33, &Ds);
  const double D22 = D11;
  const double xs[3] = {x23, x31, x12}, ys[3] = {y23, y31, y12}, Ls[3] = {L23, L31, L12};
  double a[3], b[3], c[3], d[3], e[3];
  for (int k = 0; k < 3; k++) {
    a[k] = -xs[k] / Ls[k];
    b[k] = (3 * xs[k] * ys[k]) / (4 * Ls[k]);
    c[k] = (xs[k] * xs[k] - 2 * ys[k] * ys[k]) / (4 * Ls[k]);
    d[k] = -ys[k] / Ls[k];
    e[k] = (ys[k] * ys[k] - 2 * xs[k] * xs[k]) / (4 * Ls[k]);
  }
  double P[3], T[3], q[3], r[3];
  for (int k = 0; k < 3; k++) {
    P[k] = -6 * a[k];
    T[k] = -6 * d[k];
    q[k] = 3 * a[k];
    r[k] = 3 * d[k];
  }
  static const double gpL[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
  const double wgt = 1.0 / 3;
  for (int g = 0; g < 3; g++) {
    const double L1 = gpL[g][0], L2 = gpL[g][1];
    double bx1[9], bx2[9], by1[9], by2[9];
    dkt_bx(L1, L2, P, q, c, bx1, bx2);
    dkt_by(L1, L2, T, r, e, b, by1, by2);
    const double inv2A = 1 / twoA;
    double B[27], DB[27];
    for (int j = 0; j < 9; j++) {
      B[j] = inv2A * ((-y23) * bx1[j] + (-y31) * bx2[j]);
      B[9 + j] = inv2A * (x23 * by1[j] + x31 * by2[j]);
      B[18 + j] = inv2A * (x23 * bx1[j] + x31 * bx2[j] + (-y23) * by1[j] + (-y31) * by2[j]);
    }
    for (int j = 0; j < 9; j++) {
      DB[j] = D11 * B[j] + D12 * B[9 + j];
      DB[9 + j] = D12 * B[j] + D22 * B[9 + j];
      DB[18 + j] = D33 * B[18 + j];
    }
    const double f = wgt * area;
    for (int i = 0; i < 9; i++)
      for (int j = 0; j < 9; j++) {
        double s = 0;
        for (int k = 0; k < 3; k++) s += B[k * 9 + i] * DB[k * 9 + j];
        Ke[i * 9 + j] += f * s;
      }
  }
}

/* ------------------------------------------------------------------ mesh */

typedef struct {
  const float* nodes;       /* 2 per node, float32 as the reference mesher stores */
  const uint32_t* elems;    /* npe per element */
  int64_t n_nodes, n_elems;
  int npe;                  /* 4 = Q4, 3 = DKT triangle */
  double E, nu, t;
  const double* elem_t;     /* optional per-element thickness (NULL = t) */
} orc_mesh;

static void elem_xy(const orc_mesh* m, int64_t e, double* xy) {
  for (int i = 0; i < m->npe; i++) {
    const uint32_t n = m->elems[e * m->npe + i];
    xy[2 * i] = (double)m->nodes[2 * n];
    xy[2 * i + 1] = (double)m->nodes[2 * n + 1];
  }
}

static void elem_ke(const orc_mesh* m, int64_t e, double* Ke) {
  double xy[8];
  elem_xy(m, e, xy);
  const double t = m->elem_t ? m->elem_t[e] : m->t;
  if (m->npe == 3) orc_ke_dkt(xy, m->E, m->nu, t, Ke);
  else orc_ke_q4(xy, m->E, m->nu, t, Ke);
}

/* Ke for every element, npe*3 squared doubles each (the getOrComputeKe memo
 * of solver.ts:47-76, computed eagerly). */
ORC_API void orc_all_ke(const float* nodes, const uint32_t* elems, int64_t n_nodes,
                        int64_t n_elems, int npe, double E, double nu, double t,
                        const double* elem_t, double* Ke_all) {
  orc_mesh m = {nodes, elems, n_nodes, n_elems, npe, E, nu, t, elem_t};
  const int nd = npe * 3;
  for (int64_t e = 0; e < n_elems; e++) elem_ke(&m, e, Ke_all + e * nd * nd);
}

/* --------------------------------------------------- block diag + inverse */

/* solver.ts:188-225 accumulation, then solver.ts:231-304 inversion with the
 * constrained-DOF identity rows/cols and the singular-block fallback.  cmask
 * is one byte per DOF (nonzero = constrained). */
ORC_API void orc_block_inverse(const double* Ke_all, const uint32_t* elems, int64_t n_nodes,
                               int64_t n_elems, int npe, const uint8_t* cmask,
                               double* blk /* 9 per node */) {
  const int nd = npe * 3;
  memset(blk, 0, (size_t)n_nodes * 9 * sizeof(double));
  for (int64_t e = 0; e < n_elems; e++) {
    const double* Ke = Ke_all + e * nd * nd;
    for (int i = 0; i < npe; i++) {
      const uint32_t n = elems[e * npe + i];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) blk[n * 9 + r * 3 + c] += Ke[(3 * i + r) * nd + 3 * i + c];
    }
  }
  for (int64_t n = 0; n < n_nodes; n++) {
    double* b = blk + n * 9;
    const int cw = cmask[3 * n] != 0, cx = cmask[3 * n + 1] != 0, cy = cmask[3 * n + 2] != 0;
    if (cw && cx && cy) {
      for (int k = 0; k < 9; k++) b[k] = (k % 4 == 0) ? 1.0 : 0.0;
      continue;
    }
    if (cw) { b[0] = 1; b[1] = 0; b[2] = 0; b[3] = 0; b[6] = 0; }
    if (cx) { b[4] = 1; b[1] = 0; b[7] = 0; b[3] = 0; b[5] = 0; }
    if (cy) { b[8] = 1; b[2] = 0; b[5] = 0; b[6] = 0; b[7] = 0; }
    const double a00 = b[0], a01 = b[1], a02 = b[2], a10 = b[3], a11 = b[4], a12 = b[5];
    const double a20 = b[6], a21 = b[7], a22 = b[8];
    const double det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) +
                       a02 * (a10 * a21 - a11 * a20);
    if (fabs(det) < 1e-30) {
      for (int k = 0; k < 9; k++) b[k] = (k % 4 == 0) ? 1.0 : 0.0;
      continue;
    }
    const double id = 1.0 / det;
    b[0] = id * (a11 * a22 - a12 * a21);
    b[1] = id * (a02 * a21 - a01 * a22);
    b[2] = id * (a01 * a12 - a02 * a11);
    b[3] = id * (a12 * a20 - a10 * a22);
    b[4] = id * (a00 * a22 - a02 * a20);
    b[5] = id * (a02 * a10 - a00 * a12);
    b[6] = id * (a10 * a21 - a11 * a20);
    b[7] = id * (a01 * a20 - a00 * a21);
    b[8] = id * (a00 * a11 - a01 * a10);
  }
}

/* pcg.ts:74-88 */
ORC_API void orc_apply_block(const double* blk, int64_t n_nodes, const double* r, double* z) {
  for (int64_t n = 0; n < n_nodes; n++) {
    const double* b = blk + n * 9;
    const double r0 = r[3 * n], r1 = r[3 * n + 1], r2 = r[3 * n + 2];
    z[3 * n] = b[0] * r0 + b[1] * r1 + b[2] * r2;
    z[3 * n + 1] = b[3] * r0 + b[4] * r1 + b[5] * r2;
    z[3 * n + 2] = b[6] * r0 + b[7] * r1 + b[8] * r2;
  }
}

/* --------------------------------------------------------------- K . x */

/* solver.ts:422-483: zero y, loop colours in order, element gather / Ke x /
 * scatter-add, then y[c] = x[c] on constrained DOFs.  With order == NULL the
 * elements are visited in index order (same sum up to rounding). */
ORC_API void orc_apply_k(const double* Ke_all, const uint32_t* elems, int64_t n_nodes,
                         int64_t n_elems, int npe, const uint32_t* order,
                         const uint8_t* cmask, const double* x, double* y) {
  const int nd = npe * 3;
  memset(y, 0, (size_t)n_nodes * 3 * sizeof(double));
  for (int64_t k = 0; k < n_elems; k++) {
    const int64_t e = order ? (int64_t)order[k] : k;
    const double* Ke = Ke_all + e * nd * nd;
    int64_t dof[12];
    double xl[12];
    for (int i = 0; i < npe; i++)
      for (int d = 0; d < 3; d++) dof[3 * i + d] = (int64_t)elems[e * npe + i] * 3 + d;
    for (int j = 0; j < nd; j++) xl[j] = x[dof[j]];
    for (int r = 0; r < nd; r++) {
      double s = 0;
      for (int c = 0; c < nd; c++) s += Ke[r * nd + c] * xl[c];
      y[dof[r]] += s;
    }
  }
  for (int64_t i = 0; i < n_nodes * 3; i++)
    if (cmask[i]) y[i] = x[i];
}

/* ------------------------------------------------------------------- PCG */

static double dotv(const double* a, const double* b, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

typedef struct {
  int64_t iterations;
  double final_residual;
  int converged;
  int breakdown;
} orc_pcg_result;

/* pcg.ts:98-215 with the block preconditioner, in double.  Stops when
 * |r| < tol OR |r|/|b| < tol (pcg.ts:174); |pAp| < 1e-30 is a breakdown
 * (pcg.ts:154-161); |b| < 1e-30 returns x = 0 converged (pcg.ts:118-126). */
ORC_API void orc_pcg(const double* Ke_all, const uint32_t* elems, int64_t n_nodes,
                     int64_t n_elems, int npe, const uint32_t* order, const uint8_t* cmask,
                     const double* blk, const double* b, double tol, int64_t max_iter,
                     double* x, orc_pcg_result* res) {
  const int64_t n = n_nodes * 3;
  double* r = (double*)malloc(n * sizeof(double));
  double* z = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* Ap = (double*)malloc(n * sizeof(double));
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  res->breakdown = 0;
  const double bnorm = sqrt(dotv(b, b, n));
  if (bnorm < 1e-30) {
    res->iterations = 0; res->final_residual = 0; res->converged = 1;
    goto done;
  }
  orc_apply_block(blk, n_nodes, r, z);
  memcpy(p, z, n * sizeof(double));
  double rz = dotv(r, z, n);
  for (int64_t it = 0; it < max_iter; it++) {
    orc_apply_k(Ke_all, elems, n_nodes, n_elems, npe, order, cmask, p, Ap);
    const double pAp = dotv(p, Ap, n);
    if (fabs(pAp) < 1e-30) {
      res->iterations = it + 1; res->final_residual = sqrt(dotv(r, r, n));
      res->converged = 0; res->breakdown = 1;
      goto done;
    }
    const double alpha = rz / pAp;
    for (int64_t i = 0; i < n; i++) x[i] += alpha * p[i];
    for (int64_t i = 0; i < n; i++) r[i] -= alpha * Ap[i];
    const double rn = sqrt(dotv(r, r, n));
    if (rn < tol || rn / bnorm < tol) {
      res->iterations = it + 1; res->final_residual = rn; res->converged = 1;
      goto done;
    }
    orc_apply_block(blk, n_nodes, r, z);
    const double rz_new = dotv(r, z, n);
    const double beta = rz_new / rz;
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rz_new;
  }
  res->iterations = max_iter; res->final_residual = sqrt(dotv(r, r, n)); res->converged = 0;
done:
  free(r); free(z); free(p); free(Ap);
}

/* pcg.ts:221-247: 10x10 tridiag(-1,2,-1), b = e1 + e10, scalar Jacobi M = 2,
 * tol 1e-10, 100 iterations.  The exact answer is x = (1,...,1). */
ORC_API int64_t orc_test_system(double* x /* 10 */) {
  const int n = 10;
  double r[10], z[10], p[10], Ap[10];
  double b[10] = {0};
  b[0] = 1; b[n - 1] = 1;
  memset(x, 0, sizeof(r));
  memcpy(r, b, sizeof(r));
  const double bn = sqrt(dotv(b, b, n));
  for (int i = 0; i < n; i++) z[i] = r[i] / 2.0;
  memcpy(p, z, sizeof(p));
  double rz = dotv(r, z, n);
  for (int it = 0; it < 100; it++) {
    for (int i = 0; i < n; i++) {
      Ap[i] = 2 * p[i];
      if (i > 0) Ap[i] -= p[i - 1];
      if (i < n - 1) Ap[i] -= p[i + 1];
    }
    const double pAp = dotv(p, Ap, n);
    if (fabs(pAp) < 1e-30) return -(it + 1);
    const double a = rz / pAp;
    for (int i = 0; i < n; i++) { x[i] += a * p[i]; r[i] -= a * Ap[i]; }
    const double rn = sqrt(dotv(r, r, n));
    if (rn < 1e-10 || rn / bn < 1e-10) return it + 1;
    for (int i = 0; i < n; i++) z[i] = r[i] / 2.0;
    const double rzn = dotv(r, z, n);
    const double beta = rzn / rz;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rzn;
  }
  return 100;
}

/* ------------------------------------------------------- pressure loads */

/* Pressure field p(x,y) = q0 + sum_k A_k exp(-|x - c_k|^2 / (2 s_k^2)).
 * patches holds 4 doubles per patch: A, cx, cy, sigma. */
static double pressure_at(double x, double y, double q0, const double* patches, int np) {
  double p = q0;
  for (int k = 0; k < np; k++) {
    const double A = patches[4 * k], cx = patches[4 * k + 1], cy = patches[4 * k + 2];
    const double s = patches[4 * k + 3];
    const double dx = x - cx, dy = y - cy;
    p += A * exp(-(dx * dx + dy * dy) / (2 * s * s));
  }
  return p;
}

/* Consistent nodal load F_w(i) = sum_e int_e N_i p dA (theta loads zero),
 * then F zeroed on constrained DOFs (solver.ts:400-407 semantics).
 * Q4: the 2x2 Gauss rule of element.ts:29-41 with bilinear N (:56-71).
 * Triangles: the 3-point rule of element.ts:619-625 with N_i = L_i. */
ORC_API void orc_pressure_load(const float* nodes, const uint32_t* elems, int64_t n_nodes,
                               int64_t n_elems, int npe, double q0, const double* patches,
                               int n_patches, const uint8_t* cmask, double* F) {
  memset(F, 0, (size_t)n_nodes * 3 * sizeof(double));
  orc_mesh m = {nodes, elems, n_nodes, n_elems, npe, 0, 0, 0, NULL};
  for (int64_t e = 0; e < n_elems; e++) {
    double xy[8];
    elem_xy(&m, e, xy);
    if (npe == 4) {
      static const double g = 0.5773502691896257;
      static const double gp[4][2] = {{-g, -g}, {g, -g}, {g, g}, {-g, g}};
      for (int q = 0; q < 4; q++) {
        const double xi = gp[q][0], eta = gp[q][1];
        const double N[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4,
                             (1 + xi) * (1 + eta) / 4, (1 - xi) * (1 + eta) / 4};
        double dNdx[4], dNdy[4];
        const double detJ = q4_phys(xy, xi, eta, dNdx, dNdy);
        double px = 0, py = 0;
        for (int i = 0; i < 4; i++) { px += N[i] * xy[2 * i]; py += N[i] * xy[2 * i + 1]; }
        const double pv = pressure_at(px, py, q0, patches, n_patches) * fabs(detJ);
        for (int i = 0; i < 4; i++) F[3 * (int64_t)elems[e * 4 + i]] += N[i] * pv;
      }
    } else {
      static const double gpL[3][3] = {
          {1.0 / 6, 1.0 / 6, 2.0 / 3}, {2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6}};
      const double twoA = (xy[0] - xy[2]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[1] - xy[3]);
      const double area = fabs(twoA) / 2;
      for (int q = 0; q < 3; q++) {
        const double* L = gpL[q];
        const double px = L[0] * xy[0] + L[1] * xy[2] + L[2] * xy[4];
        const double py = L[0] * xy[1] + L[1] * xy[3] + L[2] * xy[5];
        const double pv = pressure_at(px, py, q0, patches, n_patches) * area / 3.0;
        for (int i = 0; i < 3; i++) F[3 * (int64_t)elems[e * 3 + i]] += L[i] * pv;
      }
    }
  }
  for (int64_t i = 0; i < n_nodes * 3; i++)
    if (cmask[i]) F[i] = 0;
}

/* -------------------------------------------------------- post-processing */

/* Nodal bending moments, postprocess.ts:40-270 (computeMomentsQuad :57-155,
 * computeMomentsTriangle :162-270): curvatures k_x = d thy/dx,
 * k_y = -d thx/dy, k_xy = d thy/dy - d thx/dx at each element's centroid
 * (Q4: element.ts:84-184 at xi = eta = 0; triangles: linear shape functions,
 * |2A| < 1e-15 skipped), M = D (k_x + nu k_y, k_y + nu k_x, (1 - nu) k_xy / 2)
 * with D = E t^3 / (12 (1 - nu^2)) (types.ts:26-28), and each node the mean
 * over its elements, added in element order.
 * ref32 = 1 restates the reference's precision as well: the displacements
 * arrive rounded to float32 (its solution is a Float32Array), the triangle
 * shape derivatives are stored in Float32Arrays, and the sums and counts are
 * Float32Arrays that round after every addition; the rest is double.
 * ref32 = 0 is the same algorithm in double throughout. */
ORC_API void orc_moments(const float* nodes, const uint32_t* elems, int64_t n_nodes,
                         int64_t n_elems, int npe, double E, double nu, double t,
                         const double* elem_t, const double* x, int ref32, double* Mx,
                         double* My, double* Mxy) {
  double* s = (double*)calloc((size_t)n_nodes * 4, sizeof(double));
  for (int64_t e = 0; e < n_elems; e++) {
    double xy[8], thx[4], thy[4], dNdx[4], dNdy[4];
    for (int i = 0; i < npe; i++) {
      const uint32_t n = elems[e * npe + i];
      xy[2 * i] = (double)nodes[2 * n];
      xy[2 * i + 1] = (double)nodes[2 * n + 1];
      thx[i] = ref32 ? (double)(float)x[3 * (int64_t)n + 1] : x[3 * (int64_t)n + 1];
      thy[i] = ref32 ? (double)(float)x[3 * (int64_t)n + 2] : x[3 * (int64_t)n + 2];
    }
    if (npe == 4) {
      q4_phys(xy, 0.0, 0.0, dNdx, dNdy);
    } else {
      const double twoA = (xy[2] - xy[0]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[3] - xy[1]);
      if (fabs(twoA) < 1e-15) continue;
      const double inv = 1 / twoA;
      dNdx[0] = (xy[3] - xy[5]) * inv;
      dNdy[0] = (xy[4] - xy[2]) * inv;
      dNdx[1] = (xy[5] - xy[1]) * inv;
      dNdy[1] = (xy[0] - xy[4]) * inv;
      dNdx[2] = (xy[1] - xy[3]) * inv;
      dNdy[2] = (xy[2] - xy[0]) * inv;
      if (ref32)
        for (int i = 0; i < 3; i++) {
          dNdx[i] = (double)(float)dNdx[i];
          dNdy[i] = (double)(float)dNdy[i];
        }
    }
    double kx = 0, ky = 0, kxy = 0;
    for (int i = 0; i < npe; i++) {
      kx += dNdx[i] * thy[i];
      ky += -dNdy[i] * thx[i];
      kxy += dNdy[i] * thy[i] - dNdx[i] * thx[i];
    }
    const double te = elem_t ? elem_t[e] : t;
    const double D = E * pow(te, 3) / (12 * (1 - nu * nu));
    const double m[3] = {D * (kx + nu * ky), D * (ky + nu * kx), D * (1 - nu) * kxy / 2};
    for (int i = 0; i < npe; i++) {
      double* a = s + 4 * (int64_t)elems[e * npe + i];
      for (int k = 0; k < 3; k++) a[k] = ref32 ? (double)(float)(a[k] + m[k]) : a[k] + m[k];
      a[3] += 1;
    }
  }
  for (int64_t n = 0; n < n_nodes; n++) {
    const double* a = s + 4 * n;
    double* out[3] = {Mx, My, Mxy};
    for (int k = 0; k < 3; k++) {
      const double v = a[3] > 0 ? a[k] / a[3] : 0.0;
      out[k][n] = ref32 ? (double)(float)v : v;
    }
  }
  free(s);
}
