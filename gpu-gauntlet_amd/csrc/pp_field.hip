// Log-normal element thickness of the learned-preconditioner workload,
// synthesised on the device.
//
// g is a unit-variance Gaussian random field on the element centres of an
// nx x ny grid with squared-exponential covariance exp(-d^2 / (2 l^2)),
// made by spectral synthesis on a grid padded to Nx x Ny = 2nx x 2ny (so the
// crop carries no wrap-around correlation):
//   w      standard normals, w[j][i] the (j Nx + i)-th normal of the
//          SplitMix64 stream of `seed` (Box-Muller on consecutive pairs of
//          53-bit uniforms);
//   g      the circular convolution of w with the inverse DFT of the filter
//          exp(-l^2 |k|^2 / 4) (the square root of the spectral density),
//          cropped to ny x nx, minus its mean, over its standard deviation.
// The filter factors as exp(-l^2 kx^2 / 4) exp(-l^2 ky^2 / 4), so the 2-D
// convolution is two 1-D circular convolutions with real, even taps
//   c[m] = (1/N) sum_j exp(-l^2 k_j^2 / 4) cos(2 pi j m / N),
// k_j = 2 pi (j or j - N) / (N h) (the DFT's frequency order).  The taps of
// a smooth field decay like exp(-(m h / l)^2); they are cut where they fall
// below 1e-15 of c[0], the rounding level of their own FP64 evaluation (the
// field then differs from an FFT evaluation by a few 1e-15).
// Thickness t_e = t_median exp(sigma_log g_e), rounded to float32.
//
// Work: pass 1 convolves the Ny padded rows along x (only the nx columns
// the crop keeps), staging a row segment and the taps in LDS; pass 2
// convolves those nx columns along y for the ny kept rows; mean and
// variance are fixed-order two-level FP64 reductions.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pp_internal.h"

namespace pp {

namespace {

constexpr int kFieldNT = 256;
constexpr int kRedParts = 1024;

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// normal i of the stream: uniforms number 2i+1 and 2i+2
__global__ __launch_bounds__(kFieldNT) void k_noise(uint64_t seed, int64_t n, double* w) {
  const int64_t i = (int64_t)blockIdx.x * kFieldNT + threadIdx.x;
  if (i >= n) return;
  const double s = 1.0 / 9007199254740992.0;  // 2^-53
  const double u1 = fmax((double)(splitmix_at(seed, 2 * (uint64_t)i + 1) >> 11) * s, 1e-300);
  const double u2 = (double)(splitmix_at(seed, 2 * (uint64_t)i + 2) >> 11) * s;
  w[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// c[m], m = blockIdx.x in [0, N/2]: one workgroup per tap
__global__ __launch_bounds__(kFieldNT) void k_taps(int N, double h, double corr, double* c) {
  __shared__ double red[kFieldNT];
  const int m = blockIdx.x;
  const double two_pi_over_Nh = 2.0 * M_PI / (N * h);
  double acc = 0;
  for (int j = threadIdx.x; j < N; j += kFieldNT) {
    const int f = j < (N + 1) / 2 ? j : j - N;  // the DFT's frequency order
    const double k = two_pi_over_Nh * f;
    const int64_t r = ((int64_t)j * m) % N;
    acc += exp(-0.25 * corr * corr * k * k) * cospi(2.0 * (double)r / N);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kFieldNT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) c[m] = red[0] / N;
}

// Pass 1: a[j][i] = sum_{d=-M..M} c[|d|] w[j][(i - d) mod Nx] for i < nx.
// One workgroup per (row, 256 outputs); the row segment i0 - M .. i0 + 255 + M
// (circular) and the taps are staged in LDS (dynamic shared memory).
__global__ __launch_bounds__(kFieldNT) void k_conv_rows(const double* w, int Nx, int nx, int M,
                                                       const double* c, double* a) {
  extern __shared__ double sh[];
  double* tap = sh;               // M + 1
  double* seg = sh + M + 1;       // kFieldNT + 2M
  const int j = blockIdx.y, i0 = blockIdx.x * kFieldNT;
  const double* row = w + (int64_t)j * Nx;
  for (int t = threadIdx.x; t <= M; t += kFieldNT) tap[t] = c[t];
  for (int t = threadIdx.x; t < kFieldNT + 2 * M; t += kFieldNT) {
    int x = (i0 - M + t) % Nx;
    if (x < 0) x += Nx;
    seg[t] = row[x];
  }
  __syncthreads();
  const int i = i0 + threadIdx.x;
  if (i >= nx) return;
  // output i reads seg[threadIdx.x + M - d]
  const double* s = seg + threadIdx.x + M;
  double acc = tap[0] * s[0];
  for (int d = 1; d <= M; d++) acc += tap[d] * (s[-d] + s[d]);
  a[(int64_t)j * nx + i] = acc;
}

// Pass 2: g[j][i] = sum_{d=-M..M} c[|d|] a[(j - d) mod Ny][i], j < ny.
// Threads run along i (coalesced rows of a); each computes 4 rows j.
__global__ __launch_bounds__(kFieldNT) void k_conv_cols(const double* a, int Ny, int nx, int ny,
                                                       int M, const double* c, double* g) {
  extern __shared__ double tap[];
  for (int t = threadIdx.x; t <= M; t += kFieldNT) tap[t] = c[t];
  __syncthreads();
  const int i = blockIdx.x * kFieldNT + threadIdx.x;
  const int j0 = blockIdx.y * 4;
  if (i >= nx) return;
  double acc[4] = {0, 0, 0, 0};
  // row r = j - d contributes c[|j - r|] to output j
  for (int r = j0 - M; r <= j0 + 3 + M; r++) {
    int y = r % Ny;
    if (y < 0) y += Ny;
    const double v = a[(int64_t)y * nx + i];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int d = abs(j0 + q - r);
      if (d <= M) acc[q] += tap[d] * v;
    }
  }
  for (int q = 0; q < 4; q++)
    if (j0 + q < ny) g[(int64_t)(j0 + q) * nx + i] = acc[q];
}

// part[b] = sum over the b-th slice of (g - shift)^p (p = 1 or 2), fixed order
__global__ __launch_bounds__(kFieldNT) void k_moment_parts(const double* g, int64_t n,
                                                          const double* shift, int p,
                                                          double* part) {
  __shared__ double red[kFieldNT];
  const double s0 = shift ? *shift : 0.0;
  const int64_t per = (n + kRedParts - 1) / kRedParts;
  const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
  double acc = 0;
  for (int64_t k = lo + threadIdx.x; k < hi; k += kFieldNT) {
    const double v = g[k] - s0;
    acc += p == 1 ? v : v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kFieldNT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = sum(part) / n, one workgroup
__global__ __launch_bounds__(kFieldNT) void k_moment_final(const double* part, int64_t n,
                                                          double* out) {
  __shared__ double red[kFieldNT];
  double acc = 0;
  for (int k = threadIdx.x; k < kRedParts; k += kFieldNT) acc += part[k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kFieldNT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] / (double)n;
}

// g <- (g - mean) / sqrt(var); t = float(t_median exp(sigma g))
__global__ __launch_bounds__(kFieldNT) void k_lognormal(double* g, int64_t n, const double* mv,
                                                       double t_median, double sigma, float* t) {
  const int64_t k = (int64_t)blockIdx.x * kFieldNT + threadIdx.x;
  if (k >= n) return;
  const double z = (g[k] - mv[0]) / sqrt(mv[1]);
  g[k] = z;
  t[k] = (float)(t_median * exp(sigma * z));
}

unsigned blocks(int64_t n) { return (unsigned)((n + kFieldNT - 1) / kFieldNT); }

// taps of one direction, cut where they fall below 1e-15 of c[0]
int make_taps(hipStream_t s, int N, double h, double corr, DevBuf<double>& c, int* M) {
  const int half = N / 2;
  int rc = c.alloc((size_t)half + 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_taps, dim3(half + 1), dim3(kFieldNT), 0, s, N, h, corr, c.p);
  PP_HIP(hipGetLastError());
  std::vector<double> hc((size_t)half + 1);
  PP_HIP(hipMemcpyAsync(hc.data(), c.p, hc.size() * 8, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  int m = half;
  while (m > 0 && fabs(hc[m]) < 1e-15 * fabs(hc[0])) m--;
  // a full circular sum needs every offset once: -(N-1)/2 .. (N-1)/2, plus
  // N/2 once for even N (folded into c[N/2] with weight 1/2 each side)
  if (2 * m + 1 > N) {
    m = half;
    if (N % 2 == 0) {
      hc[half] *= 0.5;
      PP_HIP(hipMemcpyAsync(c.p + half, &hc[half], 8, hipMemcpyHostToDevice, s));
      PP_HIP(hipStreamSynchronize(s));
    }
  }
  *M = m;
  return PP_OK;
}

}  // namespace

}  // namespace pp

using namespace pp;

extern "C" {

int pp_lognormal_thickness(pp_ctx* ctx, int32_t nx, int32_t ny, double width, double height,
                           double corr_len, uint64_t seed, double t_median, double sigma_log,
                           double* field, float* element_t) {
  if (!ctx || (!field && !element_t)) return fail(PP_ERR_INVALID, "NULL argument");
  if (nx < 1 || ny < 1 || nx > 16384 || ny > 16384)
    return fail(PP_ERR_INVALID, "grid must be 1..16384 cells per side");
  if (!(width > 0) || !(height > 0) || !(corr_len > 0) || !(t_median > 0) || !(sigma_log >= 0))
    return fail(PP_ERR_INVALID, "width, height, corr_len, t_median must be > 0, sigma_log >= 0");
  PP_LOCK(ctx);
  PP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int Nx = 2 * nx, Ny = 2 * ny;
  const int64_t nw = (int64_t)Nx * Ny, ng = (int64_t)nx * ny;
  DevBuf<double> w, a, g, cx, cy, part, mv;
  DevBuf<float> t;
  int rc, Mx = 0, My = 0;
  if ((rc = w.alloc(nw)) || (rc = a.alloc((int64_t)Ny * nx)) || (rc = g.alloc(ng)) ||
      (rc = t.alloc(ng)) || (rc = part.alloc(kRedParts)) || (rc = mv.alloc(2)))
    return rc;
  if ((rc = make_taps(s, Nx, width / nx, corr_len, cx, &Mx)) ||
      (rc = make_taps(s, Ny, height / ny, corr_len, cy, &My)))
    return rc;
  const size_t sh1 = ((size_t)Mx + 1 + kFieldNT + 2 * (size_t)Mx) * sizeof(double);
  const size_t sh2 = ((size_t)My + 1) * sizeof(double);
  if (sh1 > 64 * 1024 || sh2 > 64 * 1024)
    return fail(PP_ERR_UNSUPPORTED, "correlation length too long for the grid spacing");
  hipLaunchKernelGGL(k_noise, dim3(blocks(nw)), dim3(kFieldNT), 0, s, seed, nw, w.p);
  hipLaunchKernelGGL(k_conv_rows, dim3(blocks(nx), Ny), dim3(kFieldNT), sh1, s, (const double*)w.p,
                     Nx, nx, Mx, (const double*)cx.p, a.p);
  hipLaunchKernelGGL(k_conv_cols, dim3(blocks(nx), (ny + 3) / 4), dim3(kFieldNT), sh2, s,
                     (const double*)a.p, Ny, nx, ny, My, (const double*)cy.p, g.p);
  hipLaunchKernelGGL(k_moment_parts, dim3(kRedParts), dim3(kFieldNT), 0, s, (const double*)g.p,
                     ng, (const double*)nullptr, 1, part.p);
  hipLaunchKernelGGL(k_moment_final, dim3(1), dim3(kFieldNT), 0, s, (const double*)part.p, ng,
                     mv.p);
  hipLaunchKernelGGL(k_moment_parts, dim3(kRedParts), dim3(kFieldNT), 0, s, (const double*)g.p,
                     ng, (const double*)mv.p, 2, part.p);
  hipLaunchKernelGGL(k_moment_final, dim3(1), dim3(kFieldNT), 0, s, (const double*)part.p, ng,
                     mv.p + 1);
  hipLaunchKernelGGL(k_lognormal, dim3(blocks(ng)), dim3(kFieldNT), 0, s, g.p, ng,
                     (const double*)mv.p, t_median, sigma_log, t.p);
  PP_HIP(hipGetLastError());
  if (field) PP_HIP(hipMemcpyAsync(field, g.p, ng * 8, hipMemcpyDeviceToHost, s));
  if (element_t) PP_HIP(hipMemcpyAsync(element_t, t.p, ng * 4, hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  return PP_OK;
}

}  // extern "C"
