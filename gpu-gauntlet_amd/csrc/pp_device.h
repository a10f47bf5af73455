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

// The workgroup part of an all-reduce's arrival: the block sum of the
// partials, then every thread waits for its own outstanding stores, then
// the workgroup meets once more, so thread 0's arrival (grid_allreduce)
// still follows every store of the workgroup while those stores drain
// during the block sum.  On the 142 x 142 plate one device took 3.86 us per
// all-reduce waiting before the sum and 3.45 us waiting after it; on others
// both orders measure the same.  publish() is defined below.
__device__ __forceinline__ void publish();
template <int NV, int NT>
__device__ __forceinline__ void block_sum_published(double (&v)[NV], double* lds) {
  block_sum<NV, NT>(v, lds);
  publish();
  __syncthreads();
}

// Cross-workgroup exchange inside a persistent kernel.  Every value that
// another workgroup reads while the kernel runs is written with st_coh and
// read with ld_coh: relaxed atomic accesses at device scope (sc1), served by
// the device-coherent level, so no L2 write-back or invalidate fence is ever
// issued (those fences are most of the cost of a release/acquire barrier).
// Data a workgroup only reads back itself keeps plain loads and stores.
// Measured on MI355X: plain stores with L1-bypassing workgroup-scope loads,
// or a workgroup-scope arrival counter, lose updates between CUs even when
// all workgroups sit on one XCD; device-scope relaxed accesses do not.
__device__ __forceinline__ double ld_coh(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_coh(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(float* p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until this wave's outstanding stores have completed; every thread
// calls it before the workgroup barrier that precedes an arrival.
__device__ __forceinline__ void publish() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Grid-wide all-reduce + barrier of a persistent kernel (all nwg workgroups
// co-resident).  Thread 0 of every workgroup holds its block's NV (<= 4)
// partial sums in v and publishes them into slot (epoch & 1) of `part`
// (4 * nwg words per slot, 8 * nwg in all); wave 0 of every workgroup then
// waits until all nwg partials of this epoch are there and sums them in a
// fixed order, so every workgroup gets bit-identical totals and takes
// identical control decisions.  Two slots suffice: no workgroup can publish
// epoch e+2 before all have passed e+1, i.e. finished reading e.  The
// caller makes every thread's earlier stores complete (publish() before the
// workgroup barrier), so data written before the call is visible to every
// workgroup after it.  Returns false when the barrier gave up (spin budget
// spent here or the abort flag raised elsewhere); the kernel then exits.
//
// No arrival counter: the partial itself is the arrival.  Each partial is
// stored as one 64-bit word whose 8 low mantissa bits carry the epoch tag
// (the value keeps 44 fraction bits, a relative change below 6e-14), so a
// single coherent store publishes both, and readers poll the slots until
// every tag is current.  Atomic-free: device-scope atomics on one address
// serialise across workgroups.  The slot memory must be zeroed before the
// kernel starts (tag 0 never matches; the first epoch uses tag 1).
__device__ __forceinline__ unsigned long long tag_word(double v, unsigned tag) {
  return (__double_as_longlong(v) & ~0xffull) | (unsigned long long)(tag & 0xffu);
}
__device__ __forceinline__ double untag_word(unsigned long long w) {
  return __longlong_as_double((long long)(w & ~0xffull));
}

//
// Polling: lane l reads slots l, l + 64, ... in rounds of U loads that are
// all in flight before any is consumed, so a poll of nwg slots costs
// ceil(nwg / (64 U)) coherent round trips instead of ceil(nwg / 64); each
// lane still adds its slots in increasing order (same bits as one by one).
template <int NV, int U = 4>
__device__ __forceinline__ bool grid_allreduce(double (&v)[NV], double* part, GridSync* gs,
                                               unsigned& epoch, double* lds, unsigned nwg,
                                               unsigned me) {
  static_assert(NV >= 1 && NV <= 4, "four values per slot");
  unsigned long long* slot = (unsigned long long*)part + (size_t)(epoch & 1u) * 4 * nwg;
  const unsigned tag = (epoch + 1u) & 0xffu ? (epoch + 1u) & 0xffu : 0x80u;
  if (threadIdx.x < 64) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NV; k++)
        __hip_atomic_store(&slot[k * nwg + me], tag_word(v[k], tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned spins = 0;
    int ok = 1;
    double s[NV];
    for (;;) {
      int cur = 1;
#pragma unroll
      for (int k = 0; k < NV; k++) s[k] = 0;
      for (unsigned base = 0; base < nwg; base += 64 * U) {
        unsigned long long w[U][NV];
#pragma unroll
        for (int j = 0; j < U; j++) {
          const unsigned i = base + threadIdx.x + 64 * j;
#pragma unroll
          for (int k = 0; k < NV; k++)
            w[j][k] = i < nwg ? __hip_atomic_load(&slot[k * nwg + i], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : (unsigned long long)tag;
        }
#pragma unroll
        for (int j = 0; j < U; j++)
#pragma unroll
          for (int k = 0; k < NV; k++) {
            cur &= (unsigned)(w[j][k] & 0xffu) == tag;
            s[k] += untag_word(w[j][k]);
          }
      }
      if (__all(cur)) break;
      if (++spins > kSpinLimit ||
          __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        if (threadIdx.x == 0)
          __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] = wave_sum(s[k]);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NV; k++) lds[k] = s[k];
      ((int*)lds)[2 * NV] = ok;
    }
  }
  epoch++;
  __syncthreads();
  const bool ok = ((int*)lds)[2 * NV] != 0;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = lds[k];
  __syncthreads();
  return ok;
}

// y = Ke u for the Q4 Mindlin element of element.ts:258-483 on an
// axis-aligned a x b rectangle, evaluated as B^T D B u without forming Ke.
// On a rectangle J = diag(a/2, b/2), so dN/dx = (2/a) dN/dxi and
// dN/dy = (2/b) dN/deta; bending at the 2x2 Gauss points (weight 1,
// |J| = ab/4), shear at the centre (weight 4).  Curvatures
// {d thy/dx, -d thx/dy, d thy/dy - d thx/dx}, shear strains
// {dw/dx + thy, dw/dy - thx}; corner order SW, SE, NE, NW.
template <typename T>
__host__ __device__ __forceinline__ void q4_rect_apply(T a, T b, T ia, T ib, T D11, T D12, T D33, T Ds,
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

// The same y = Ke u as q4_rect_apply with the Gauss sums done in closed
// form.  On a rectangle dN_i/dx = xi_i (1 + eta_i eta) / (2a) and
// dN_i/dy = eta_i (1 + xi_i xi) / (2b) (corner signs xi_i, eta_i = -+1), so
// d/dx of a rotation field depends on eta only and d/dy on xi only.  The
// curvatures at the 2x2 points therefore reduce to four edge differences per
// rotation, and the back-projection B^T m needs only the moment sums
// sum_q m_q, sum_q xi_q m_q, sum_q eta_q m_q (g^2 = 1/3).  About 100 flops
// instead of ~360, and far fewer live registers.
template <typename T>
__host__ __device__ __forceinline__ void q4_rect_apply_fast(T a, T b, T ia, T ib, T D11, T D12,
                                                           T D33, T Ds, const T (&u)[12],
                                                           T (&y)[12]) {
  const T g = (T)0.5773502691896257;
  const T ga = (T)0.5 * ia, gb = (T)0.5 * ib;  // 1/(2a), 1/(2b)
  const T F = a * b;                           // 4 points x |J| (= ab/4)
  // edge differences of thx (index 1) and thy (index 2); corners SW SE NE NW
  const T Sx = u[4] - u[1], Nx = u[7] - u[10], Wx = u[10] - u[1], Ex = u[7] - u[4];
  const T Sy = u[5] - u[2], Ny = u[8] - u[11], Wy = u[11] - u[2], Ey = u[8] - u[5];
  // kx(eta) = ga[(Sy+Ny) + eta(Ny-Sy)], ky(xi) = -gb[(Wx+Ex) + xi(Ex-Wx)],
  // kxy = T(xi) - S(eta), T = gb[(Wy+Ey) + xi(Ey-Wy)], S = ga[(Sx+Nx) + eta(Nx-Sx)].
  // Over the 4 points: sum kx = 4 KXp, sum eta kx = 4 g KXm, sum xi kx = 0, etc.
  const T KXp = ga * (Sy + Ny), KXm = g * ga * (Ny - Sy);
  const T KYp = -gb * (Wx + Ex), KYm = -g * gb * (Ex - Wx);
  const T Tp = gb * (Wy + Ey), Tm = g * gb * (Ey - Wy);
  const T Sp = ga * (Sx + Nx), Sm = g * ga * (Nx - Sx);
  // moment sums over the points (times |J|): m = D kappa
  const T SA = F * (D11 * KXp + D12 * KYp);  // sum mx
  const T EA = g * F * D11 * KXm;            // sum eta mx
  const T SB = F * (D12 * KXp + D11 * KYp);  // sum my
  const T XB = g * F * D11 * KYm;            // sum xi my
  const T SC = F * D33 * (Tp - Sp);          // sum mxy
  const T XC = g * F * D33 * Tm;             // sum xi mxy
  const T EC = -g * F * D33 * Sm;            // sum eta mxy
  // shear at the centre (weight 4, |J| = ab/4)
  const T gxz = ga * (u[3] + u[6] - u[0] - u[9]) + (T)0.25 * (u[2] + u[5] + u[8] + u[11]);
  const T gyz = gb * (u[6] + u[9] - u[0] - u[3]) - (T)0.25 * (u[1] + u[4] + u[7] + u[10]);
  const T fs = a * b * Ds;
  const T qx = fs * gxz, qy = fs * gyz;
  const T xs[4] = {-1, 1, 1, -1}, es[4] = {-1, -1, 1, 1};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const T xi = xs[i], et = es[i];
    y[3 * i] = ga * xi * qx + gb * et * qy;
    y[3 * i + 1] = -gb * et * (SB + xi * XB) - ga * xi * (SC + et * EC) - (T)0.25 * qy;
    y[3 * i + 2] = ga * xi * (SA + et * EA) + gb * et * (SC + xi * XC) + (T)0.25 * qx;
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
