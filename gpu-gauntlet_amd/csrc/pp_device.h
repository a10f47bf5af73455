// Device-side building blocks: wavefront / workgroup / grid reductions, the
// grid barrier of the persistent kernels, and the matrix-free Q4 Mindlin
// element operator on an axis-aligned rectangle.
#pragma once

#include <hip/hip_runtime.h>

#include "pp_internal.h"

namespace pp {

// Spin budget of a grid barrier before a workgroup gives up and raises the
// abort flag (each spin sleeps ~100 cycles and polls one L2 atomic).
constexpr unsigned kSpinLimit = 4u * 1000u * 1000u;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Workgroup sum of NV values; the result is valid in thread 0.  lds holds
// NV * (NT/64) doubles.  Summation order is fixed (deterministic).
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) lds[k * NW + wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0;
      for (int w = 0; w < NW; w++) s += lds[k * NW + w];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Grid-wide all-reduce + barrier.  Thread 0 of every workgroup holds its
// block's NV (<= 2) partial sums in v.  Each workgroup publishes them into
// slot (epoch & 1) of `part` (2 * gridDim.x doubles per slot), arrives on the
// counter and waits until every workgroup has arrived.  Wave 0 then sums the
// partials in a fixed order, so every workgroup gets bit-identical totals and
// takes identical control decisions.  Two slots suffice: a workgroup cannot
// publish epoch e+2 before all have passed epoch e+1, i.e. finished reading e.
// Returns false if the barrier was aborted (spin budget spent somewhere).
template <int NV>
__device__ __forceinline__ bool grid_allreduce(double (&v)[NV], double* part, GridSync* gs,
                                               unsigned& epoch, double* lds) {
  static_assert(NV >= 1 && NV <= 2, "two values per slot");
  const unsigned nwg = gridDim.x;
  double* slot = part + (size_t)(epoch & 1u) * 2 * nwg;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) slot[k * nwg + blockIdx.x] = v[k];
    __hip_atomic_fetch_add(&gs->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (epoch + 1u) * nwg;
    unsigned spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&gs->count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          ++spins > kSpinLimit) {
        __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    ((int*)lds)[2 * NV] = ok;
  }
  epoch++;
  __syncthreads();
  const bool ok = ((int*)lds)[2 * NV] != 0;
  if (ok && threadIdx.x < 64) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0;
      for (unsigned i = threadIdx.x; i < nwg; i += 64) s += slot[k * nwg + i];
      s = wave_sum(s);
      if (threadIdx.x == 0) lds[k] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = lds[k];
  __syncthreads();
  if (!ok) return false;
  // The data written before the barrier by other workgroups is read by all
  // threads after it: make this thread's later loads observe it.
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// y = Ke u for the Q4 Mindlin element of element.ts:258-483 on an
// axis-aligned a x b rectangle, evaluated as B^T D B u without forming Ke.
// On a rectangle J = diag(a/2, b/2), so dN/dx = (2/a) dN/dxi and
// dN/dy = (2/b) dN/deta; bending at the 2x2 Gauss points (weight 1,
// |J| = ab/4), shear at the centre (weight 4).  Curvatures
// {d thy/dx, -d thx/dy, d thy/dy - d thx/dx}, shear strains
// {dw/dx + thy, dw/dy - thx}; corner order SW, SE, NE, NW.
template <typename T>
__device__ __forceinline__ void q4_rect_apply(T a, T b, T ia, T ib, T D11, T D12, T D33, T Ds,
                                              const T (&u)[12], T (&y)[12]) {
  const T g = (T)0.5773502691896257;
  const T fb = (T)0.25 * a * b;
  const T qa = (T)0.5 * ia, qb = (T)0.5 * ib;  // (2/a)/4, (2/b)/4
#pragma unroll
  for (int k = 0; k < 12; k++) y[k] = (T)0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const T xi = (q == 1 || q == 2) ? g : -g;
    const T eta = (q >= 2) ? g : -g;
    const T dx[4] = {-((T)1 - eta) * qa, ((T)1 - eta) * qa, ((T)1 + eta) * qa, -((T)1 + eta) * qa};
    const T dy[4] = {-((T)1 - xi) * qb, -((T)1 + xi) * qb, ((T)1 + xi) * qb, ((T)1 - xi) * qb};
    T kx = 0, ky = 0, kxy = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      kx += dx[i] * u[3 * i + 2];
      ky -= dy[i] * u[3 * i + 1];
      kxy += dy[i] * u[3 * i + 2] - dx[i] * u[3 * i + 1];
    }
    const T mx = fb * (D11 * kx + D12 * ky);
    const T my = fb * (D12 * kx + D11 * ky);
    const T mxy = fb * D33 * kxy;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      y[3 * i + 1] += -dy[i] * my - dx[i] * mxy;
      y[3 * i + 2] += dx[i] * mx + dy[i] * mxy;
    }
  }
  {
    const T dx[4] = {-qa, qa, qa, -qa};
    const T dy[4] = {-qb, -qb, qb, qb};
    T gxz = 0, gyz = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      gxz += dx[i] * u[3 * i] + (T)0.25 * u[3 * i + 2];
      gyz += dy[i] * u[3 * i] - (T)0.25 * u[3 * i + 1];
    }
    const T fs = a * b * Ds;
    const T qx = fs * gxz, qy = fs * gyz;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      y[3 * i] += dx[i] * qx + dy[i] * qy;
      y[3 * i + 1] += (T)(-0.25) * qy;
      y[3 * i + 2] += (T)0.25 * qx;
    }
  }
}

// Symmetric 3x3 inverse of a nodal block with the constrained-DOF rules of
// solver.ts:231-304 (identity rows/cols for constrained DOFs, identity for a
// fully constrained or singular block).  Input full row-major A, output the
// six entries m00 m01 m02 m11 m12 m22.
__device__ __forceinline__ void block_inverse_sym(double (&A)[9], unsigned fix, double* out) {
  const bool cw = fix & PP_FIX_W, cx = fix & PP_FIX_THX, cy = fix & PP_FIX_THY;
  if (cw && cx && cy) {
    out[0] = 1; out[1] = 0; out[2] = 0; out[3] = 1; out[4] = 0; out[5] = 1;
    return;
  }
  if (cw) { A[0] = 1; A[1] = 0; A[2] = 0; A[3] = 0; A[6] = 0; }
  if (cx) { A[4] = 1; A[1] = 0; A[7] = 0; A[3] = 0; A[5] = 0; }
  if (cy) { A[8] = 1; A[2] = 0; A[5] = 0; A[6] = 0; A[7] = 0; }
  const double a00 = A[0], a01 = A[1], a02 = A[2], a11 = A[4], a12 = A[5], a22 = A[8];
  const double a10 = A[3], a20 = A[6], a21 = A[7];
  const double det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) +
                     a02 * (a10 * a21 - a11 * a20);
  if (fabs(det) < 1e-30) {
    out[0] = 1; out[1] = 0; out[2] = 0; out[3] = 1; out[4] = 0; out[5] = 1;
    return;
  }
  const double id = 1.0 / det;
  out[0] = id * (a11 * a22 - a12 * a21);
  out[1] = id * (a02 * a21 - a01 * a22);
  out[2] = id * (a01 * a12 - a02 * a11);
  out[3] = id * (a00 * a22 - a02 * a20);
  out[4] = id * (a02 * a10 - a00 * a12);
  out[5] = id * (a00 * a11 - a01 * a10);
}

template <typename T, typename M>
__device__ __forceinline__ void sym3_apply(const M* m, const T (&r)[3], T (&z)[3]) {
  const T m00 = (T)m[0], m01 = (T)m[1], m02 = (T)m[2], m11 = (T)m[3], m12 = (T)m[4], m22 = (T)m[5];
  z[0] = m00 * r[0] + m01 * r[1] + m02 * r[2];
  z[1] = m01 * r[0] + m11 * r[1] + m12 * r[2];
  z[2] = m02 * r[0] + m12 * r[1] + m22 * r[2];
}

}  // namespace pp
