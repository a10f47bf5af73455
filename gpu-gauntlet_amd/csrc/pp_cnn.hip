// Learned preconditioner: a 3-level U-net on the nodal grid (BF16 MFMA,
// FP32 accumulate), trained on the device by this file's own forward and
// backward kernels, and applied inside flexible PCG.
//
// Tensors are channels-last BF16: element (pixel, c) of a view at
// p[pixel * cs + c0 + c], pixel = y * W + x on the (padded) node grid.
// Channel strides and offsets are multiples of 4 so a lane's four channels
// are one 8-byte load.  Level 0 is the node grid padded to 4k+1 on each side;
// levels 1 and 2 are the vertex-centred coarsenings (H_{l+1} = (H_l + 1) / 2)
// with full-weighting restriction R and bilinear prolongation P = 4 R^T.
//
//   x_in(4) -e0a-> a0(16) -e0b-> s0(16) -R-> r1(16) -e1a-> a1(32) -e1b-> s1(32)
//   -R-> r2(32) -ba-> a2(64) -bb-> b(64) -P-> [P(b) | s1](96) -d1a-> c1(32)
//   -d1b-> d1(32) -P-> [P(d1) | s0](48) -d0a-> c0(16) -d0b-> c0b(16) -head-> o(3)
//
// every conv 3x3; with ReLU and biases (all but the head) or, by default,
// without either, so the network is a fixed linear operator.  Inputs: the
// block-Jacobi preconditioned residual z = M^-1 r (3 channels, scaled per
// channel and by its RMS s) and log(t / t0) times the first of them.
// Output: N(r) = z + s * o / cin_scale (the network learns the correction
// beyond block Jacobi).
//
// The preconditioner wraps the network in a smoother S (damped block Jacobi,
// or a Galerkin multigrid V-cycle from pp_mg.hip): z0 = S r, a correction
// from the network on r - K z0 with its energy-optimal step, then S again.
//
// Training pairs: e from a seed (a windowed sum of Gaussian bumps in w with
// Kirchhoff-consistent rotations, zero on constrained DOFs) or from solver
// iterates, r = K e with the solver's own FP64 operator.  Loss: the energy
// norm of the error of the whole preconditioner, (M(r) - e)^T K (M(r) - e) /
// e^T K e.  Adam on FP32 master weights.
//
// Convolution as implicit GEMM on v_mfma_f32_16x16x16bf16_1k: M = 16 output
// channels (the weights are the A operand), N = 16 pixels, K = (tap, input
// channel) in steps of 16, input channels staged through LDS in chunks.  The
// backward-data pass is the same kernel with the flipped, transposed weights;
// the weight gradient is a GEMM over pixels with LDS-staged transposed tiles
// and deterministic partial sums.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pp_cnn.h"

namespace pp {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Eight BF16 values at an 8-byte aligned LDS address, as two 8-byte reads
// (one ds_read2_b64): the K = 32 MFMA operand of one lane.
__device__ __forceinline__ bf8 lds8(const uint16_t* p) {
  const s4 lo = *(const s4*)p, hi = *(const s4*)(p + 4);
  return __builtin_bit_cast(bf8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (NaN kept quiet): one v_cvt_pk_bf16_f32 on gfx950
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

int nblk(int64_t n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// ------------------------------------------------------------ convolution

// four BF16 channels as one 8-byte access
__device__ __forceinline__ void ld4(const uint16_t* p, float (&v)[4]) {
  const uint2 u = *(const uint2*)p;
  v[0] = __uint_as_float(u.x << 16);
  v[1] = __uint_as_float(u.x & 0xffff0000u);
  v[2] = __uint_as_float(u.y << 16);
  v[3] = __uint_as_float(u.y & 0xffff0000u);
}
__device__ __forceinline__ void st4(uint16_t* p, const float (&v)[4]) {
  uint2 u;
  u.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
  u.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
  *(uint2*)p = u;
}


struct ConvArgs {
  const uint16_t* in;
  int in_cs, in_c0, cin;  // cin % 4 == 0
  uint16_t* out;          // bf16 output (or out32)
  int out_cs, out_c0, cout;
  float* out32;
  int out32_cs;
  const uint16_t* w;  // [NB*16][kp] bf16, k = tap * cin + ci
  int kp;             // multiple of 16, >= 9 cin
  int cc;             // input channels per LDS chunk (divides cin, multiple of 4)
  const float* bias;  // [cout] or null
  int H, W, dil, relu;
  const uint16_t* mask;  // out[c] *= (mask[pixel][c] > 0) for c >= mask_from
  int mask_cs, mask_c0, mask_from;
  // two-part input (k_convp): channels [0, p_ch) come from pin (channel stride
  // p_cs, offset p_c0), channels [p_ch, cin) from in at in_c0 + c - p_ch; so a
  // concatenation of two tensors stored apart is read in place
  const uint16_t* pin;
  int p_cs, p_c0, p_ch;
  // k_convp: output channels in groups of NB * 16, each workgroup one group
  // (weights [groups * NB * 16][kp]); 1 = all output blocks per workgroup
  int groups;
};

// ((a + b) + (c + d)) / 4 per BF16 channel, in FP32, rounded to BF16: the
// bilinear value of a fine node from its 1, 2 or 4 coarse neighbours, a
// missing neighbour repeating the one beside it (k_prolong and k_prolong8 sum
// in this order, so both give the same bits).
__device__ __forceinline__ uint32_t avg4_bf2(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  const float lo = ((__uint_as_float(a << 16) + __uint_as_float(b << 16)) +
                    (__uint_as_float(c << 16) + __uint_as_float(d << 16))) * 0.25f;
  const float hi = ((__uint_as_float(a & 0xffff0000u) + __uint_as_float(b & 0xffff0000u)) +
                    (__uint_as_float(c & 0xffff0000u) + __uint_as_float(d & 0xffff0000u))) * 0.25f;
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}
__device__ __forceinline__ uint4 avg4_bf8(const uint4& a, const uint4& b, const uint4& c,
                                          const uint4& d) {
  return uint4{avg4_bf2(a.x, b.x, c.x, d.x), avg4_bf2(a.y, b.y, c.y, d.y),
               avg4_bf2(a.z, b.z, c.z, d.z), avg4_bf2(a.w, b.w, c.w, d.w)};
}

// One workgroup per output tile of MB rows x 64 pixels (MB = 2 or 4).  The input tile
// with its halo, (MB + 2 dil) x (64 + 2 dil) pixels x cin channels, is
// staged in LDS by independent 8-byte loads (pixel stride cin + 4 channels,
// which spreads a wave's 16 pixels over the LDS banks), next to the weights
// [NB * 16][kp].  Wave w owns the 16-pixel blocks w MB .. w MB + MB - 1 of
// the tile (row-major, 4 per row).  The weights are the MFMA's A operand and
// the pixels its B operand, so D[channel 4 kq + i][pixel m] leaves each lane
// four consecutive output channels of one pixel: one 8-byte store.
constexpr int kStage = 8;  // staging loads in flight per thread

template <int NB, int MB, bool U16>
__global__ __launch_bounds__(256) void k_conv(ConvArgs A) {
  extern __shared__ uint16_t smem[];
  // input channels in chunks of CC: per chunk the tile holds CC channels and
  // the weights their K slice [NB * 16][KC], KC = 9 CC rounded up to 16
  const int CC = A.cc, KC = (9 * CC + 15) / 16 * 16;
  const int d = A.dil, TW = 64 + 2 * d, TH = MB + 2 * d, CS = CC + 4;
  uint16_t* w_s = smem;
  uint16_t* in_s = smem + NB * 16 * KC;
  const int tiles_x = (A.W + 63) / 64;
  const int tx0 = (blockIdx.x % tiles_x) * 64, ty0 = (blockIdx.x / tiles_x) * MB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int k9 = 9 * CC;
  int pbase[MB];  // LDS pixel of each owned block's lane-m pixel at tap (0, 0)
#pragma unroll
  for (int j = 0; j < MB; j++) {
    const int mb = wave * MB + j;
    pbase[j] = ((mb >> 2) * TW + (mb & 3) * 16 + m) * CS;
  }
  f4 acc[MB][NB];
#pragma unroll
  for (int j = 0; j < MB; j++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[j][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  // Index splits by float reciprocals, exact here: (q + 1/2) / D is at least
  // 1 / (2 D) >= 1 / 144 from an integer for every divisor D <= 72 used
  // below, and for q < 2^16 its float error is below 1e-2 / 2 (the launch
  // checks both bounds)
  const int cq = CC >> 2, nq = TH * TW * cq, kq4 = KC >> 2, nw4 = NB * 16 * kq4;
  const float rcq = 1.0f / (float)cq, rtw = 1.0f / (float)TW, rkq = 1.0f / (float)kq4;
  for (int c0 = 0; c0 < A.cin; c0 += CC) {
    if (c0) __syncthreads();  // the previous chunk's reads are done
    // staging: kStage independent 8-byte loads in flight per thread, then
    // their LDS stores.  Weights: quad (co, k = 4 i) of the slice is
    // w[co][tap * cin + c0 + ci] for k = tap * CC + ci < 9 CC, else 0.
    for (int i0 = threadIdx.x; i0 < nw4; i0 += 256 * kStage) {
      uint2 v[kStage];
#pragma unroll
      for (int u = 0; u < kStage; u++) {
        const int i = i0 + 256 * u;
        v[u] = uint2{0u, 0u};
        if (i < nw4) {
          const int co = (int)(((float)i + 0.5f) * rkq), k = 4 * (i - co * kq4);
          if (k < k9) {
            const int tap = (int)(((float)k + 0.5f) / (float)CC), ci = k - tap * CC;
            v[u] = *(const uint2*)&A.w[(size_t)co * A.kp + tap * A.cin + c0 + ci];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kStage; u++)
        if (i0 + 256 * u < nw4) *(uint2*)&w_s[4 * (i0 + 256 * u)] = v[u];
    }
    for (int q0 = threadIdx.x; q0 < nq; q0 += 256 * kStage) {
      uint2 v[kStage];
      int dst[kStage];
#pragma unroll
      for (int u = 0; u < kStage; u++) {
        const int q = q0 + 256 * u;
        v[u] = uint2{0u, 0u};
        dst[u] = -1;
        if (q < nq) {
          const int pix = (int)(((float)q + 0.5f) * rcq), c = q - pix * cq;
          const int row = (int)(((float)pix + 0.5f) * rtw), col = pix - row * TW;
          const int gx = tx0 - d + col, gy = ty0 - d + row;
          dst[u] = (row * TW + col) * CS + 4 * c;
          if (gx >= 0 && gx < A.W && gy >= 0 && gy < A.H)
            v[u] = *(const uint2*)&A.in[(gy * A.W + gx) * A.in_cs + A.in_c0 + c0 + 4 * c];
        }
      }
#pragma unroll
      for (int u = 0; u < kStage; u++)
        if (dst[u] >= 0) *(uint2*)&in_s[dst[u]] = v[u];
    }
    __syncthreads();
    if constexpr (U16) {
      // CC % 16 == 0: every 16-wide K step lies within one tap, whose LDS
      // offset is uniform over the wave; the lane adds its pixel and quad
#pragma unroll 1
      for (int tap = 0; tap < 9; tap++) {
        const int ty = tap >= 6 ? 2 : tap >= 3 ? 1 : 0;
        const int toff = (ty * d * TW + (tap - 3 * ty) * d) * CS + 4 * kq;
        for (int c16 = 0; c16 < CC; c16 += 16) {
          const int kb = tap * CC + c16 + 4 * kq;
          s4 b[NB];
#pragma unroll
          for (int nb = 0; nb < NB; nb++) b[nb] = *(const s4*)&w_s[(nb * 16 + m) * KC + kb];
          s4 a[MB];
#pragma unroll
          for (int j = 0; j < MB; j++) a[j] = *(const s4*)&in_s[pbase[j] + toff + c16];
#pragma unroll
          for (int j = 0; j < MB; j++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
              acc[j][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(b[nb], a[j], acc[j][nb], 0, 0, 0);
        }
      }
    } else {
      // CC = 4, 8 or 12: (tap, channel) of this lane's k = k0 + 4 kq, advanced
      // by 16 per step without branches (at most 16 / CC <= 4 wraps).  Steps
      // past 9 CC have zero weights, so their A reads may use any pixel of
      // the tile: they take the lane's own (toff 0) and every read is
      // unconditional.
      int tap = 0, ci = 4 * kq;
      auto wrap = [&]() {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int w = ci >= CC;
          ci -= w * CC;
          tap += w;
        }
      };
      wrap();
      for (int k0 = 0; k0 < KC; k0 += 16) {
        const int kb = k0 + 4 * kq;
        s4 b[NB];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) b[nb] = *(const s4*)&w_s[(nb * 16 + m) * KC + kb];
        const int ty = tap >= 6 ? 2 : tap >= 3 ? 1 : 0;
        const int toff = kb < k9 ? (ty * d * TW + (tap - 3 * ty) * d) * CS + ci : 0;
        s4 a[MB];
#pragma unroll
        for (int j = 0; j < MB; j++) a[j] = *(const s4*)&in_s[pbase[j] + toff];
#pragma unroll
        for (int j = 0; j < MB; j++)
#pragma unroll
          for (int nb = 0; nb < NB; nb++)
            acc[j][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(b[nb], a[j], acc[j][nb], 0, 0, 0);
        ci += 16;
        wrap();
      }
    }
  }
#pragma unroll
  for (int j = 0; j < MB; j++) {
    const int mb = wave * MB + j;
    const int oy = ty0 + (mb >> 2), ox = tx0 + (mb & 3) * 16 + m;
    if (oy >= A.H || ox >= A.W) continue;
    const int op = oy * A.W + ox;  // 32-bit offsets: the launch checks they fit
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const int co = nb * 16 + 4 * kq;
      if (co >= A.cout) continue;
      float v[4];
      for (int i = 0; i < 4; i++) {
        v[i] = acc[j][nb][i] + (A.bias && co + i < A.cout ? A.bias[co + i] : 0.0f);
        if (A.relu) v[i] = fmaxf(v[i], 0.0f);
      }
      if (A.mask && co >= A.mask_from) {
        float q[4];
        ld4(&A.mask[op * A.mask_cs + A.mask_c0 + co], q);
        for (int i = 0; i < 4; i++)
          if (!(q[i] > 0.0f)) v[i] = 0.0f;
      }
      // a float output with room for four channels takes the zero pad channels too
      if (co + 4 <= A.cout || (A.out32 && co + 4 <= A.out32_cs)) {
        if (A.out32) *(float4*)&A.out32[op * A.out32_cs + co] = make_float4(v[0], v[1], v[2], v[3]);
        else st4(&A.out[op * A.out_cs + A.out_c0 + co], v);
      } else {
        for (int i = 0; i < 4 && co + i < A.cout; i++) {
          if (A.out32) A.out32[op * A.out32_cs + co + i] = v[i];
          else A.out[op * A.out_cs + A.out_c0 + co + i] = f2bf(v[i]);
        }
      }
    }
  }
}

// Forward convolution with the shapes fixed at compile time: dilation 1,
// input channels in chunks of CC (16, 32 or 48), NB blocks of 16 output
// channels, an output tile of MB rows x 64 pixels (4 waves, wave w owning the
// 16-pixel blocks w MB .. w MB + MB - 1, row-major, 4 per row).  The same
// implicit GEMM as k_conv, but every staging split is a division by a
// constant, the input and the weights come in 16-byte loads (8 channels),
// all loads of a chunk are issued before the first LDS store, and the tap
// loop is unrolled so every LDS read of the K loop has a constant offset
// from the lane's base.  Weight rows in LDS are padded by 4 channels so the
// 16 rows an MFMA reads fall in distinct banks.  Tiles are dealt to the
// XCDs in row-major bands (workgroup b runs on XCD b % 8), so the halo rows
// two vertically adjacent tiles share are fetched into one L2.  Inputs and
// weights need channel strides and offsets that are multiples of 8 (the
// launch checks).
template <int CC, int NB, int MB>
__global__ __launch_bounds__(256) void k_convf(ConvArgs A, int ntiles) {
  constexpr int TW = 66, TH = MB + 2, CS = CC + 4, KC = 9 * CC, KP = KC + 4;
  constexpr int Q8 = CC / 8;                                // 16-byte units per pixel
  constexpr int NIN = TH * TW * Q8, NW = NB * 16 * KC / 8;  // 16-byte units per chunk
  constexpr int UIN = (NIN + 255) / 256, UW = (NW + 255) / 256;
  static_assert(CC % 16 == 0 && KC % 16 == 0, "16-wide K steps within one tap");
  extern __shared__ uint16_t smem[];
  uint16_t* w_s = smem;
  uint16_t* in_s = smem + NB * 16 * KP;
  const int band = (ntiles + 7) / 8;
  const int tile = (int)(blockIdx.x % 8) * band + (int)(blockIdx.x / 8);
  if (tile >= ntiles) return;
  const int tiles_x = (A.W + 63) / 64;
  const int tx0 = (tile % tiles_x) * 64, ty0 = (tile / tiles_x) * MB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  int pbase[MB];
#pragma unroll
  for (int j = 0; j < MB; j++) {
    const int mb = wave * MB + j;
    pbase[j] = ((mb >> 2) * TW + (mb & 3) * 16 + m) * CS + 4 * kq;
  }
  f4 acc[MB][NB];
#pragma unroll
  for (int j = 0; j < MB; j++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[j][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int c0 = 0; c0 < A.cin; c0 += CC) {
    if (c0) __syncthreads();  // the previous chunk's reads are done
    uint4 vw[UW], vi[UIN];
#pragma unroll
    for (int u = 0; u < UW; u++) {
      const int i = threadIdx.x + 256 * u;
      vw[u] = uint4{0u, 0u, 0u, 0u};
      if (i < NW) {
        const int co = i / (KC / 8), kk = (i % (KC / 8)) * 8;
        const int tap = kk / CC, ci = kk % CC;
        vw[u] = *(const uint4*)&A.w[(size_t)co * A.kp + tap * A.cin + c0 + ci];
      }
    }
#pragma unroll
    for (int u = 0; u < UIN; u++) {
      const int i = threadIdx.x + 256 * u;
      vi[u] = uint4{0u, 0u, 0u, 0u};
      if (i < NIN) {
        const int pix = i / Q8, part = i % Q8;
        const int gx = tx0 - 1 + pix % TW, gy = ty0 - 1 + pix / TW;
        if (gx >= 0 && gx < A.W && gy >= 0 && gy < A.H)
          vi[u] = *(const uint4*)&A.in[(gy * A.W + gx) * A.in_cs + A.in_c0 + c0 + 8 * part];
      }
    }
#pragma unroll
    for (int u = 0; u < UW; u++) {
      const int i = threadIdx.x + 256 * u;
      if (i < NW) {
        uint16_t* d = &w_s[(i / (KC / 8)) * KP + (i % (KC / 8)) * 8];
        *(uint2*)d = uint2{vw[u].x, vw[u].y};
        *(uint2*)(d + 4) = uint2{vw[u].z, vw[u].w};
      }
    }
#pragma unroll
    for (int u = 0; u < UIN; u++) {
      const int i = threadIdx.x + 256 * u;
      if (i < NIN) {
        uint16_t* d = &in_s[(i / Q8) * CS + 8 * (i % Q8)];
        *(uint2*)d = uint2{vi[u].x, vi[u].y};
        *(uint2*)(d + 4) = uint2{vi[u].z, vi[u].w};
      }
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
      const int toff = ((tap / 3) * TW + tap % 3) * CS;
#pragma unroll
      for (int c16 = 0; c16 < CC; c16 += 16) {
        const int kb = tap * CC + c16 + 4 * kq;
        s4 b[NB];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) b[nb] = *(const s4*)&w_s[(nb * 16 + m) * KP + kb];
        s4 a[MB];
#pragma unroll
        for (int j = 0; j < MB; j++) a[j] = *(const s4*)&in_s[pbase[j] + toff + c16];
#pragma unroll
        for (int j = 0; j < MB; j++)
#pragma unroll
          for (int nb = 0; nb < NB; nb++)
            acc[j][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(b[nb], a[j], acc[j][nb], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < MB; j++) {
    const int mb = wave * MB + j;
    const int oy = ty0 + (mb >> 2), ox = tx0 + (mb & 3) * 16 + m;
    if (oy >= A.H || ox >= A.W) continue;
    const int op = oy * A.W + ox;  // 32-bit offsets: the launch checks they fit
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const int co = nb * 16 + 4 * kq;
      if (co >= A.cout) continue;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        v[i] = acc[j][nb][i] + (A.bias && co + i < A.cout ? A.bias[co + i] : 0.0f);
        if (A.relu) v[i] = fmaxf(v[i], 0.0f);
      }
      // a float output with room for four channels takes the zero pad channels too
      if (co + 4 <= A.cout || (A.out32 && co + 4 <= A.out32_cs)) {
        if (A.out32) *(float4*)&A.out32[op * A.out32_cs + co] = make_float4(v[0], v[1], v[2], v[3]);
        else st4(&A.out[op * A.out_cs + A.out_c0 + co], v);
      } else {
        for (int i = 0; i < 4 && co + i < A.cout; i++) {
          if (A.out32) A.out32[op * A.out32_cs + co + i] = v[i];
          else A.out[op * A.out_cs + A.out_c0 + co + i] = f2bf(v[i]);
        }
      }
    }
  }
}

// Persistent form of k_convf: each workgroup stages all the layer's weights
// once ([NB * 16][9 cin + 4], the K order of the packed weights), then walks
// the tiles of its XCD's band (tile = band start + local index, stepping by
// the workgroups on that XCD) chunk by chunk (CC input channels at a time),
// loading the next (tile, chunk) input into registers while the current one
// is multiplied, so the load latency of every stage but the first is hidden
// behind the previous stage's work.  Chunks below A.p_ch come from A.pin (a
// concatenated input whose parts are stored apart).
template <int CC, int NB, int MB>
__global__ __launch_bounds__(256) void k_convp(ConvArgs A, int ntiles) {
  constexpr int TW = 66, TH = MB + 2, CS = CC + 4;
  constexpr int Q8 = CC / 8;
  constexpr int NIN = TH * TW * Q8, UIN = (NIN + 255) / 256;
  static_assert(CC % 16 == 0, "16-wide K steps within one tap");
  extern __shared__ uint16_t smem[];
  const int cin = A.cin, nch = cin / CC, K9 = 9 * cin, KP = K9 + 4;
  uint16_t* w_s = smem;
  uint16_t* in_s = smem + NB * 16 * KP;
  // workgroup b runs on XCD b % 8; the workgroups of an XCD are dealt to
  // the output-channel groups in turn, and each walks the XCD's band of
  // tiles for its group
  const int band = (ntiles + 7) / 8, xcd = (int)(blockIdx.x % 8);
  const int grp = (int)(blockIdx.x / 8) % A.groups;
  const int per_xcd = (int)(gridDim.x / 8) / A.groups, local0 = (int)(blockIdx.x / 8) / A.groups;
  const int tile_lo = xcd * band, tile_hi = min(ntiles, tile_lo + band);
  const int tiles_x = (A.W + 63) / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int co0 = grp * NB * 16;  // the group's first output channel
  const uint16_t* wg = A.w + (size_t)co0 * A.kp;
  // 32-channel chunks multiply on v_mfma_f32_16x16x32_bf16 (lane kq holds K
  // 8 kq .. 8 kq + 7 of A and B), 16-channel chunks on the K = 16 form (4 kq ..)
  constexpr int KQ = CC == 32 ? 8 : 4;
  int pbase[MB];
#pragma unroll
  for (int j = 0; j < MB; j++) {
    const int mb = wave * MB + j;
    pbase[j] = ((mb >> 2) * TW + (mb & 3) * 16 + m) * CS + KQ * kq;
  }
  // weights, once: row co holds k = tap * cin + ci, 9 cin entries; kStage
  // loads in flight per thread before their LDS stores
  {
    const int units = K9 / 8, nw = NB * 16 * units;
    for (int i0 = threadIdx.x; i0 < nw; i0 += 256 * kStage) {
      uint4 v[kStage];
#pragma unroll
      for (int u = 0; u < kStage; u++) {
        const int i = i0 + 256 * u;
        if (i < nw) {
          const int co = i / units, kk = (i - co * units) * 8;
          v[u] = *(const uint4*)&wg[(size_t)co * A.kp + kk];
        }
      }
#pragma unroll
      for (int u = 0; u < kStage; u++) {
        const int i = i0 + 256 * u;
        if (i < nw) {
          const int co = i / units, kk = (i - co * units) * 8;
          uint16_t* d = &w_s[co * KP + kk];
          *(uint2*)d = uint2{v[u].x, v[u].y};
          *(uint2*)(d + 4) = uint2{v[u].z, v[u].w};
        }
      }
    }
  }
  // Staging without branches: every thread loads from an address clamped
  // into the tensor (units past the tile repeat the last one, and LDS then
  // receives the same value twice) and keeps a bit per unit that zeroes it
  // when it is written to LDS.  With no divergent loads the compiler's
  // waits are exact counts, so waiting for one stage's loads does not also
  // wait for the stores issued after them.
  uint4 vi[UIN];
  uint32_t inside = 0;
  auto load_stage = [&](int tile, int c0) {
    const int tx0 = (tile % tiles_x) * 64, ty0 = (tile / tiles_x) * MB;
    const bool first = c0 < A.p_ch;
    const uint16_t* src = first ? A.pin + A.p_c0 + c0 : A.in + A.in_c0 + (c0 - A.p_ch);
    const int cs = first ? A.p_cs : A.in_cs;
    inside = 0;
#pragma unroll
    for (int u = 0; u < UIN; u++) {
      const int i = min((int)threadIdx.x + 256 * u, NIN - 1);
      const int pix = i / Q8, part = i % Q8;
      const int gx = tx0 - 1 + pix % TW, gy = ty0 - 1 + pix / TW;
      const bool ok = gx >= 0 && gx < A.W && gy >= 0 && gy < A.H;
      inside |= (uint32_t)ok << u;
      const int cx = min(max(gx, 0), A.W - 1), cy = min(max(gy, 0), A.H - 1);
      vi[u] = *(const uint4*)&src[(cy * A.W + cx) * cs + 8 * part];
    }
  };
  // the bias, staged in LDS once: a global load in the epilogue would make
  // the epilogue's wait also wait for the next tile's prefetch
  float* bias_s = (float*)(in_s + TH * TW * CS);
  if ((int)threadIdx.x < NB * 16) {
    const int co = co0 + (int)threadIdx.x;
    bias_s[threadIdx.x] = A.bias && co < A.cout ? A.bias[co] : 0.0f;
  }
  // (the halo outside the tensor is zeroed by a second LDS write, not by a
  // select on the loaded value, which the compiler would schedule, with
  // its wait, right after the loads)
  auto store_stage = [&]() {
#pragma unroll
    for (int u = 0; u < UIN; u++) {
      const int i = min((int)threadIdx.x + 256 * u, NIN - 1);
      uint16_t* d = &in_s[(i / Q8) * CS + 8 * (i % Q8)];
      *(uint2*)d = uint2{vi[u].x, vi[u].y};
      *(uint2*)(d + 4) = uint2{vi[u].z, vi[u].w};
      if (!((inside >> u) & 1)) {
        *(uint2*)d = uint2{0u, 0u};
        *(uint2*)(d + 4) = uint2{0u, 0u};
      }
    }
  };
  // Stage order: the next stage's loads are issued before this stage's
  // multiplies and written to LDS after them (behind a barrier), and only
  // then does the epilogue store the tile's outputs.  So a wait for the
  // loads never waits for stores issued after them (the counter of
  // outstanding vector memory operations covers loads and stores in order).
  int tile = tile_lo + local0;
  if (tile < tile_hi) {
    load_stage(tile, 0);
    store_stage();
  }
  __syncthreads();
  for (; tile < tile_hi; tile += per_xcd) {
    f4 acc[MB][NB];
#pragma unroll
    for (int j = 0; j < MB; j++)
#pragma unroll
      for (int nb = 0; nb < NB; nb++) acc[j][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ch = 0; ch < nch; ch++) {
      // LDS holds stage (tile, ch); the next one is loaded into registers.
      // Every stage loads and stores one (the workgroup's last reloads its
      // own first chunk, never read), so no load or wait is conditional.
      const bool more = ch + 1 < nch;
      const int nt = more ? tile : tile + per_xcd < tile_hi ? tile + per_xcd : tile;
      load_stage(nt, more ? (ch + 1) * CC : 0);
      const int kc0 = ch * CC + KQ * kq;
#pragma unroll
      for (int tap = 0; tap < 9; tap++) {
        const int toff = ((tap / 3) * TW + tap % 3) * CS;
        if constexpr (CC == 32) {
          const int kb = tap * cin + kc0;
          bf8 b[NB];
#pragma unroll
          for (int nb = 0; nb < NB; nb++) b[nb] = lds8(&w_s[(nb * 16 + m) * KP + kb]);
          bf8 a[MB];
#pragma unroll
          for (int j = 0; j < MB; j++) a[j] = lds8(&in_s[pbase[j] + toff]);
#pragma unroll
          for (int j = 0; j < MB; j++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
              acc[j][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[nb], a[j], acc[j][nb], 0, 0, 0);
        } else {
#pragma unroll
          for (int c16 = 0; c16 < CC; c16 += 16) {
            const int kb = tap * cin + kc0 + c16;
            s4 b[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) b[nb] = *(const s4*)&w_s[(nb * 16 + m) * KP + kb];
            s4 a[MB];
#pragma unroll
            for (int j = 0; j < MB; j++) a[j] = *(const s4*)&in_s[pbase[j] + toff + c16];
#pragma unroll
            for (int j = 0; j < MB; j++)
#pragma unroll
              for (int nb = 0; nb < NB; nb++)
                acc[j][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(b[nb], a[j], acc[j][nb], 0, 0, 0);
          }
        }
      }
      __syncthreads();  // every wave's reads of this stage are done
      store_stage();
      if (more) __syncthreads();  // the tile's next chunk is in LDS
    }
    const int tx0 = (tile % tiles_x) * 64, ty0 = (tile / tiles_x) * MB;
#pragma unroll
    for (int j = 0; j < MB; j++) {
      const int mb = wave * MB + j;
      const int oy = ty0 + (mb >> 2), ox = tx0 + (mb & 3) * 16 + m;
      if (oy >= A.H || ox >= A.W) continue;
      const int op = oy * A.W + ox;
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        const int co = co0 + nb * 16 + 4 * kq;
        if (co >= A.cout) continue;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          v[i] = acc[j][nb][i] + bias_s[nb * 16 + 4 * kq + i];
          if (A.relu) v[i] = fmaxf(v[i], 0.0f);
        }
        // a float output with room for four channels takes the zero pad channels too
      if (co + 4 <= A.cout || (A.out32 && co + 4 <= A.out32_cs)) {
          if (A.out32) *(float4*)&A.out32[op * A.out32_cs + co] = make_float4(v[0], v[1], v[2], v[3]);
          else st4(&A.out[op * A.out_cs + A.out_c0 + co], v);
        } else {
          for (int i = 0; i < 4 && co + i < A.cout; i++) {
            if (A.out32) A.out32[op * A.out32_cs + co + i] = v[i];
            else A.out[op * A.out_cs + A.out_c0 + co + i] = f2bf(v[i]);
          }
        }
      }
    }
    __syncthreads();  // the next tile's first stage is in LDS
  }
}

// The network's first two convolutions in one kernel: e0a (4 -> 16 channels)
// then e0b (16 -> 16), both 3x3 with dilation 1.  Per output tile of MB rows
// x 64 pixels the 4-channel input region with a two-pixel halo, (MB + 4) x 68
// pixels at 8 bytes each, is staged in LDS; e0a is evaluated on the tile with
// a one-pixel halo, (MB + 2) x 66 pixels taken as a flat list of 16-pixel
// blocks (lane kq of K step s covers tap 4 s + kq, steps past tap 8 having
// zero weights), rounded to BF16 (zero outside the grid, the zero padding e0b
// sees when the two run apart) into a second LDS buffer that e0b reads.  So
// the 16-channel intermediate never goes through memory, and the result has
// the bits of the two separate launches (the same BF16 rounding points and
// MFMA K order).  Persistent as k_convp: both layers' weights staged once per
// workgroup, tiles walked in XCD row bands, the next tile's input loaded
// into registers while the current one is computed.  (Fusing two 16-channel
// layers the same way measured slower than their two launches: the halo
// recomputation and the LDS traffic of the first layer outweigh the saved
// round trip of the intermediate.)
struct Conv2Args {
  const uint16_t* in;  // e0a input, 4 channels at in[pixel * in_cs + in_c0 + c]
  int in_cs, in_c0;
  const uint16_t* wa;  // packed forward weights of e0a [16][kpa], k = tap * 4 + ci
  const uint16_t* wb;  // of e0b [16][kpb], k = tap * 16 + ci
  int kpa, kpb;
  const float* bias_a;  // [16] or null
  const float* bias_b;  // [16] or null
  int relu_a, relu_b;
  uint16_t* out;  // e0b output (BF16)
  int out_cs, out_c0, cout;
  int H, W;
};

template <int MB>
__global__ __launch_bounds__(256) void k_conv2(Conv2Args A, int ntiles) {
  constexpr int IW = 68, IH = MB + 4, MW = 66, MH = MB + 2;
  constexpr int CSA = 4, CSM = 20;  // LDS pixel strides (4 input, 16 + 4 mid channels)
  constexpr int KA = 48, KPA = KA + 4, KPB = 9 * 16 + 4;
  constexpr int NMID = MH * MW, MBLK = (NMID + 15) / 16;
  constexpr int NIN = IH * IW, UIN = (NIN + 255) / 256;  // 8-byte staging units
  extern __shared__ uint16_t smem[];
  uint16_t* wa_s = smem;
  uint16_t* wb_s = wa_s + 16 * KPA;
  uint16_t* in_s = wb_s + 16 * KPB;
  uint16_t* mid_s = in_s + IH * IW * CSA;
  const int band = (ntiles + 7) / 8, xcd = (int)(blockIdx.x % 8);
  const int per_xcd = (int)(gridDim.x / 8), local0 = (int)(blockIdx.x / 8);
  const int tile_lo = xcd * band, tile_hi = min(ntiles, tile_lo + band);
  const int tiles_x = (A.W + 63) / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  // both layers' weights, once (8-byte units; rows padded by 4 entries)
  for (int i = threadIdx.x; i < 16 * (KA / 4); i += 256) {
    const int co = i / (KA / 4), k = 4 * (i - co * (KA / 4));
    *(uint2*)&wa_s[co * KPA + k] = *(const uint2*)&A.wa[(size_t)co * A.kpa + k];
  }
  for (int i = threadIdx.x; i < 16 * 36; i += 256) {
    const int co = i / 36, k = 4 * (i - co * 36);
    *(uint2*)&wb_s[co * KPB + k] = *(const uint2*)&A.wb[(size_t)co * A.kpb + k];
  }
  // layer a: lane m of mid block b is mid pixel min(16 b + m, NMID - 1); its
  // tap (0, 0) input pixel sits at the same (row, col) of the input region
  // staging without branches, as in k_convp: clamped loads, a bit per unit
  // outside the tensor, zeroed by a second LDS write
  uint2 vi[UIN];
  uint32_t inside = 0;
  auto load_stage = [&](int tile) {
    const int tx0 = (tile % tiles_x) * 64 - 2, ty0 = (tile / tiles_x) * MB - 2;
    inside = 0;
#pragma unroll
    for (int u = 0; u < UIN; u++) {
      const int i = min((int)threadIdx.x + 256 * u, NIN - 1);
      const int gx = tx0 + i % IW, gy = ty0 + i / IW;
      inside |= (uint32_t)(gx >= 0 && gx < A.W && gy >= 0 && gy < A.H) << u;
      const int cx = min(max(gx, 0), A.W - 1), cy = min(max(gy, 0), A.H - 1);
      vi[u] = *(const uint2*)&A.in[(cy * A.W + cx) * A.in_cs + A.in_c0];
    }
  };
  auto store_stage = [&]() {
#pragma unroll
    for (int u = 0; u < UIN; u++) {
      uint2* d = (uint2*)&in_s[min((int)threadIdx.x + 256 * u, NIN - 1) * CSA];
      *d = vi[u];
      if (!((inside >> u) & 1)) *d = uint2{0u, 0u};
    }
  };
  // both layers' bias, staged in LDS once (a global load in an epilogue
  // would make its wait also wait for the next tile's prefetch)
  float* bias_s = (float*)(mid_s + NMID * CSM);
  if (threadIdx.x < 16) bias_s[threadIdx.x] = A.bias_a ? A.bias_a[threadIdx.x] : 0.0f;
  else if (threadIdx.x < 32) {
    const int co = (int)threadIdx.x - 16;
    bias_s[16 + co] = A.bias_b && co < A.cout ? A.bias_b[co] : 0.0f;
  }
  // Per tile: the next tile's input is loaded while e0a runs and written to
  // LDS once e0a's reads are done (before e0b and its stores), so waiting
  // for those loads never waits for stores issued after them.  Every tile
  // loads one (the workgroup's last reloads its own, never read).
  int tile = tile_lo + local0;
  if (tile < tile_hi) {
    load_stage(tile);
    store_stage();
  }
  __syncthreads();  // the first tile and the weights are in LDS
  for (; tile < tile_hi; tile += per_xcd) {
    const int tx0 = (tile % tiles_x) * 64, ty0 = (tile / tiles_x) * MB;
    load_stage(tile + per_xcd < tile_hi ? tile + per_xcd : tile);
    // layer a, two mid blocks at a time (independent MFMA chains)
    for (int b0 = wave; b0 < MBLK; b0 += 8) {
      f4 acc[2] = {f4{0.0f, 0.0f, 0.0f, 0.0f}, f4{0.0f, 0.0f, 0.0f, 0.0f}};
      int q[2], base[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        q[h] = min(16 * (b0 + 4 * h) + m, NMID - 1);
        const int row = q[h] / MW, col = q[h] - row * MW;
        base[h] = (row * IW + col) * CSA;
      }
#pragma unroll
      for (int s = 0; s < 3; s++) {
        const int tap = 4 * s + kq;
        const int toff = tap < 9 ? ((tap / 3) * IW + tap % 3) * CSA : 0;
        const s4 w = *(const s4*)&wa_s[m * KPA + 16 * s + 4 * kq];
#pragma unroll
        for (int h = 0; h < 2; h++)
          acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
              w, *(const s4*)&in_s[base[h] + toff], acc[h], 0, 0, 0);
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (b0 + 4 * h >= MBLK || 16 * (b0 + 4 * h) + m >= NMID) continue;
        const int row = q[h] / MW, col = q[h] - row * MW;
        const int gy = ty0 - 1 + row, gx = tx0 - 1 + col;
        const bool inside = gx >= 0 && gx < A.W && gy >= 0 && gy < A.H;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          v[i] = acc[h][i] + bias_s[4 * kq + i];
          if (A.relu_a) v[i] = fmaxf(v[i], 0.0f);
          if (!inside) v[i] = 0.0f;
        }
        st4(&mid_s[q[h] * CSM + 4 * kq], v);
      }
    }
    __syncthreads();  // e0a's reads of the input are done, its output is in LDS
    store_stage();
    // layer b on the tile: wave w owns 16-pixel blocks w MB .. w MB + MB - 1
    f4 acc[MB];
    int mbase[MB];
#pragma unroll
    for (int j = 0; j < MB; j++) {
      const int mb = wave * MB + j;
      mbase[j] = ((mb >> 2) * MW + (mb & 3) * 16 + m) * CSM + 4 * kq;
      acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
      const int toff = ((tap / 3) * MW + tap % 3) * CSM;
      const s4 w = *(const s4*)&wb_s[m * KPB + 16 * tap + 4 * kq];
#pragma unroll
      for (int j = 0; j < MB; j++)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, *(const s4*)&mid_s[mbase[j] + toff],
                                                            acc[j], 0, 0, 0);
    }
    const int co = 4 * kq;
#pragma unroll
    for (int j = 0; j < MB; j++) {
      const int mb = wave * MB + j;
      const int oy = ty0 + (mb >> 2), ox = tx0 + (mb & 3) * 16 + m;
      if (oy >= A.H || ox >= A.W || co >= A.cout) continue;
      const int op = oy * A.W + ox;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        v[i] = acc[j][i] + bias_s[16 + co + i];
        if (A.relu_b) v[i] = fmaxf(v[i], 0.0f);
      }
      st4(&A.out[op * A.out_cs + A.out_c0 + co], v);
    }
    __syncthreads();  // e0b's reads are done, the next tile's input is in LDS
  }
}

// Weight gradient dW[co][tap*cin + ci] = sum_p dout[p][co] in[p + off(tap)][ci]:
// a GEMM over pixels (K), M = output channels, N = (tap, ci).  One row
// segment of 64 pixels at a time is staged transposed in LDS (dout as
// [co][64], the input rows y - dil, y, y + dil as [ci][3][64 + 2 dil]); each
// wave owns 4 (M block, N block) tiles; partial sums per row chunk.
struct WgArgs {
  const uint16_t* in;
  int in_cs, in_c0, cin;
  const uint16_t* dout;
  int dout_cs, dout_c0, cout;
  int H, W, dil;
  int kpad, coutP, rows;
  float* part;   // [nchunks][coutP][kpad]
  float* bpart;  // [nchunks][coutP] bias partials (written by blockIdx.y == 0)
};

__global__ __launch_bounds__(256) void k_wgrad(WgArgs A) {
  extern __shared__ uint16_t sm[];
  const int TW = 64 + 2 * A.dil;
  uint16_t* in_s = sm;                    // [cin][3][TW]
  uint16_t* do_s = sm + A.cin * 3 * TW;   // [coutP][64]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nkb = A.kpad / 16, nmb = A.coutP / 16, npairs = nmb * nkb;
  const int k9 = 9 * A.cin;
  int pm[4], pn[4];
  bool pv[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int q = blockIdx.y * 16 + wave * 4 + j;
    pv[j] = q < npairs;
    pm[j] = pv[j] ? q / nkb : 0;
    pn[j] = pv[j] ? q % nkb : 0;
  }
  f4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;  // bias partial of channel threadIdx.x (y == 0 only)
  const int y0 = blockIdx.x * A.rows, y1 = min(A.H, y0 + A.rows);
  for (int y = y0; y < y1; y++) {
    for (int x0 = 0; x0 < A.W; x0 += 64) {
      __syncthreads();
      for (int i = threadIdx.x; i < A.coutP * 64; i += 256) {
        const int co = i % A.coutP, px = i / A.coutP;
        uint16_t v = 0;
        if (co < A.cout && x0 + px < A.W)
          v = A.dout[((size_t)y * A.W + x0 + px) * A.dout_cs + A.dout_c0 + co];
        do_s[co * 64 + px] = v;
      }
      for (int i = threadIdx.x; i < A.cin * 3 * TW; i += 256) {
        const int ci = i % A.cin, rest = i / A.cin;
        const int xx = rest % TW, r = rest / TW;
        const int gx = x0 - A.dil + xx, gy = y + (r - 1) * A.dil;
        uint16_t v = 0;
        if (gx >= 0 && gx < A.W && gy >= 0 && gy < A.H)
          v = A.in[((size_t)gy * A.W + gx) * A.in_cs + A.in_c0 + ci];
        in_s[(ci * 3 + r) * TW + xx] = v;
      }
      __syncthreads();
      if (blockIdx.y == 0 && threadIdx.x < A.cout)
        for (int px = 0; px < 64; px++) bsum += bf2f(do_s[threadIdx.x * 64 + px]);
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        const int pxb = kb * 16 + 4 * (lane >> 4);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (!pv[j]) continue;  // uniform over the wave
          const int co = pm[j] * 16 + (lane & 15);
          const s4 a = *(const s4*)&do_s[co * 64 + pxb];
          const int n = pn[j] * 16 + (lane & 15);
          s4 b = {0, 0, 0, 0};
          if (n < k9) {
            const int tap = n / A.cin, ci = n - tap * A.cin;
            const int base = (ci * 3 + tap / 3) * TW + pxb + (tap % 3) * A.dil;
            b[0] = (short)in_s[base];
            b[1] = (short)in_s[base + 1];
            b[2] = (short)in_s[base + 2];
            b[3] = (short)in_s[base + 3];
          }
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
      }
    }
  }
  float* P = A.part + (size_t)blockIdx.x * A.coutP * A.kpad;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!pv[j]) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = pm[j] * 16 + 4 * (lane >> 4) + i;
      const int n = pn[j] * 16 + (lane & 15);
      P[(size_t)co * A.kpad + n] = acc[j][i];
    }
  }
  if (blockIdx.y == 0 && threadIdx.x < A.coutP)
    A.bpart[(size_t)blockIdx.x * A.coutP + threadIdx.x] = threadIdx.x < A.cout ? bsum : 0.0f;
}

// grad[co][n] (n < 9 cin) = sum over chunks, fixed order; bias likewise.
__global__ void k_wgrad_reduce(const float* part, const float* bpart, int nchunks, int coutP,
                               int kpad, int cout, int k9, float* gw, float* gb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cout * k9) {
    const int co = i / k9, n = i % k9;
    float s = 0.0f;
    for (int c = 0; c < nchunks; c++) s += part[((size_t)c * coutP + co) * kpad + n];
    gw[i] = s;
  }
  if (i < cout) {
    float s = 0.0f;
    for (int c = 0; c < nchunks; c++) s += bpart[(size_t)c * coutP + i];
    gb[i] = s;
  }
}

// Pack FP32 master weights W[co][tap][ci] into the forward layout
// wf[co][tap*cin + ci] (rows to coutP, columns to kp, zero padded) and the
// backward-data layout wt[ci][tap'*coutA + co] = W[co][8 - tap'][ci].
__global__ void k_pack_w(const float* W, int cout, int cin, int coutP, int kp, int coutA,
                         int cinP, int kpT, uint16_t* wf, uint16_t* wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < coutP * kp) {
    const int co = i / kp, k = i % kp;
    float v = 0.0f;
    if (co < cout && k < 9 * cin) v = W[(size_t)co * 9 * cin + k];
    wf[i] = f2bf(v);
  }
  if (i < cinP * kpT) {
    const int ci = i / kpT, k = i % kpT;
    float v = 0.0f;
    if (ci < cin && k < 9 * coutA) {
      const int tp = k / coutA, co = k % coutA;
      if (co < cout) v = W[((size_t)co * 9 + (8 - tp)) * cin + ci];
    }
    wt[i] = f2bf(v);
  }
}

// ------------------------------------------------------- grid transfer

// Epilogue shared by the resamplers: scale, then += previous out when
// accumulate, then *= (mask > 0); four channels at a time.
__device__ __forceinline__ void resample_store(float (&s)[4], float scale, uint16_t* o,
                                               const uint16_t* mask, int accumulate) {
  float v[4] = {scale * s[0], scale * s[1], scale * s[2], scale * s[3]};
  if (accumulate) {
    float q[4];
    ld4(o, q);
    for (int k = 0; k < 4; k++) v[k] += q[k];
  }
  if (mask) {
    float q[4];
    ld4(mask, q);
    for (int k = 0; k < 4; k++)
      if (!(q[k] > 0.0f)) v[k] = 0.0f;
  }
  st4(o, v);
}

// Full-weighting restriction, four channels per thread (C % 4 == 0):
// out[X,Y][c] = scale * sum w_a w_b in[2X + a, 2Y + b][c] / 16, w = (1, 2, 1).
__global__ void k_restrict(const uint16_t* in, int in_cs, int in_c0, int Hf, int Wf,
                           uint16_t* out, int out_cs, int out_c0, int Hc, int Wc, int C,
                           float scale, const uint16_t* mask, int mask_cs, int mask_c0,
                           int accumulate) {
  const int C4 = C >> 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Hc * Wc * C4) return;
  const int c = (i % C4) * 4, pix = i / C4;
  const int X = pix % Wc, Y = pix / Wc;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int b = -1; b <= 1; b++) {
    const int fy = 2 * Y + b;
    if (fy < 0 || fy >= Hf) continue;
    const float wy = b == 0 ? 2.0f : 1.0f;
    for (int a = -1; a <= 1; a++) {
      const int fx = 2 * X + a;
      if (fx < 0 || fx >= Wf) continue;
      const float w = wy * (a == 0 ? 2.0f : 1.0f);
      float v[4];
      ld4(&in[((size_t)fy * Wf + fx) * in_cs + in_c0 + c], v);
      for (int k = 0; k < 4; k++) s[k] += w * v[k];
    }
  }
  resample_store(s, scale * (1.0f / 16.0f), &out[(size_t)pix * out_cs + out_c0 + c],
                 mask ? &mask[(size_t)pix * mask_cs + mask_c0 + c] : nullptr, accumulate);
}

// Full-weighting restriction of a tensor stored alone (channel stride C, a
// multiple of 8), eight channels per thread in 16-byte accesses, for the
// network's forward pass: the sums of k_restrict with scale 1, in the same
// order, so the same bits.
__global__ void k_restrict8(const uint16_t* in, int Hf, int Wf, uint16_t* out, int Hc, int Wc,
                            int C) {
  const int C8 = C >> 3;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Hc * Wc * C8) return;
  const int c = (i % C8) * 8, pix = i / C8;
  const int X = pix % Wc, Y = pix / Wc;
  float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (int b = -1; b <= 1; b++) {
    const int fy = 2 * Y + b;
    if (fy < 0 || fy >= Hf) continue;
    const float wy = b == 0 ? 2.0f : 1.0f;
    for (int a = -1; a <= 1; a++) {
      const int fx = 2 * X + a;
      if (fx < 0 || fx >= Wf) continue;
      const float w = wy * (a == 0 ? 2.0f : 1.0f);
      const uint4 u = *(const uint4*)&in[((size_t)fy * Wf + fx) * C + c];
      const uint32_t q[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        s[2 * k] += w * __uint_as_float(q[k] << 16);
        s[2 * k + 1] += w * __uint_as_float(q[k] & 0xffff0000u);
      }
    }
  }
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    o[k] = (uint32_t)f2bf(0.0625f * s[2 * k]) | ((uint32_t)f2bf(0.0625f * s[2 * k + 1]) << 16);
  *(uint4*)&out[(size_t)pix * C + c] = uint4{o[0], o[1], o[2], o[3]};
}

// Bilinear prolongation into a tensor stored alone (channel stride C, a
// multiple of 8) for the network's forward pass: eight channels per thread in
// 16-byte accesses, the value k_prolong gives with scale 1.
__global__ void k_prolong8(const uint16_t* in, int Hc, int Wc, uint16_t* out, int Hf, int Wf,
                           int C) {
  const int C8 = C >> 3;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Hf * Wf * C8) return;
  const int c = (i % C8) * 8, pix = i / C8;
  const int x = pix % Wf, y = pix / Wf;
  const int x0 = x >> 1, y0 = y >> 1, x1 = x0 + (x & 1), y1 = y0 + (y & 1);
  const uint4 a = *(const uint4*)&in[((size_t)y0 * Wc + x0) * C + c];
  const uint4 b = *(const uint4*)&in[((size_t)y0 * Wc + x1) * C + c];
  const uint4 d = *(const uint4*)&in[((size_t)y1 * Wc + x0) * C + c];
  const uint4 e = *(const uint4*)&in[((size_t)y1 * Wc + x1) * C + c];
  *(uint4*)&out[(size_t)pix * C + c] = avg4_bf8(a, b, d, e);
}

// out[x,y][c] = scale * bilinear(in) (vertex-centred, Hf = 2 Hc - 1), four
// channels per thread.
__global__ void k_prolong(const uint16_t* in, int in_cs, int in_c0, int Hc, int Wc, uint16_t* out,
                          int out_cs, int out_c0, int Hf, int Wf, int C, float scale,
                          const uint16_t* mask, int mask_cs, int mask_c0, int accumulate) {
  const int C4 = C >> 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Hf * Wf * C4) return;
  const int c = (i % C4) * 4, pix = i / C4;
  const int x = pix % Wf, y = pix / Wf;
  const int x0 = x >> 1, y0 = y >> 1;
  const int x1 = x0 + (x & 1), y1 = y0 + (y & 1);
  // ((v00 + v01) + (v10 + v11)) / 4 with a missing neighbour repeating the
  // one beside it: the bilinear weights 1, 1/2, 1/4, summed as avg4_bf2 does
  float v00[4], v01[4], v10[4], v11[4], s[4];
  ld4(&in[((size_t)y0 * Wc + x0) * in_cs + in_c0 + c], v00);
  ld4(&in[((size_t)y0 * Wc + x1) * in_cs + in_c0 + c], v01);
  ld4(&in[((size_t)y1 * Wc + x0) * in_cs + in_c0 + c], v10);
  ld4(&in[((size_t)y1 * Wc + x1) * in_cs + in_c0 + c], v11);
  for (int k = 0; k < 4; k++) s[k] = ((v00[k] + v01[k]) + (v10[k] + v11[k])) * 0.25f;
  resample_store(s, scale, &out[(size_t)pix * out_cs + out_c0 + c],
                 mask ? &mask[(size_t)pix * mask_cs + mask_c0 + c] : nullptr, accumulate);
}

// ------------------------------------------------------ input / output

// z = M^-1 r (FP64, per node), partial sums of (scale_c z_c)^2 per block.
__global__ void k_bj_norm(int64_t nn, const double* minv, const double* r, double* z, double sw,
                          double st, double* part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nn;
       n += (int64_t)gridDim.x * blockDim.x) {
    double rv[3] = {r[3 * n], r[3 * n + 1], r[3 * n + 2]}, zv[3];
    sym3_apply(minv + 6 * n, rv, zv);
    z[3 * n] = zv[0];
    z[3 * n + 1] = zv[1];
    z[3 * n + 2] = zv[2];
    acc += sw * sw * zv[0] * zv[0] + st * st * (zv[1] * zv[1] + zv[2] * zv[2]);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Fixed-order sum of np (<= 256) partials at stride `stride` by one
// 256-thread block; thread 0 gets the result.
__device__ double block_total(const double* part, int np, int stride) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  sh[t] = t < np ? part[(size_t)t * stride] : 0.0;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) sh[t] += sh[t + k];
    __syncthreads();
  }
  return sh[0];
}

// scal[0] = s = sqrt(sum / count) (1 when zero); fixed order.
__global__ void k_scale(const double* part, int np, double count, double* scal) {
  const double a = block_total(part, np, 1);
  if (threadIdx.x == 0) {
    const double s = sqrt(a / count);
    scal[0] = s > 0.0 ? s : 1.0;
  }
}

// x_in[pixel] = (sw z_w, st z_thx, st z_thy) / s and log(t / t0) times the
// first channel (so the input, and a network without activation and biases,
// stay linear in r); pixels outside the plate are zero.
__global__ void k_pack_in(int nx, int ny, int Wp, int Hp, const double* z, const double* scal,
                          double sw, double st, const float* logt, uint16_t* xin) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)Wp * Hp) return;
  const int x = (int)(p % Wp), y = (int)(p / Wp);
  uint16_t v[4] = {0, 0, 0, 0};
  if (x <= nx && y <= ny) {
    const int64_t n = (int64_t)y * (nx + 1) + x;
    const double is = 1.0 / scal[0];
    const float x0 = (float)(sw * z[3 * n] * is);
    v[0] = f2bf(x0);
    v[1] = f2bf((float)(st * z[3 * n + 1] * is));
    v[2] = f2bf((float)(st * z[3 * n + 2] * is));
    v[3] = f2bf(logt[n] * x0);
  }
  *(uint2*)&xin[4 * p] = make_uint2(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16));
}

// M(r) = z + s * o / scale_c on free DOFs; z on constrained DOFs.
__global__ void k_unpack_out(int nx, int ny, int Wp, const double* z, const float* o,
                             const double* scal, double sw, double st, const uint8_t* cmask,
                             double* out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(nx + 1) * (ny + 1)) return;
  const int x = (int)(n % (nx + 1)), y = (int)(n / (nx + 1));
  const int64_t p = (int64_t)y * Wp + x;
  const unsigned fix = cmask[n];
  const double s = scal[0];
  const double inv[3] = {1.0 / sw, 1.0 / st, 1.0 / st};
  for (int d = 0; d < 3; d++) {
    double v = z[3 * n + d];
    if (!((fix >> d) & 1u)) v += s * inv[d] * (double)o[4 * p + d];
    out[3 * n + d] = v;
  }
}

// dL/do[pixel][c] = factor * s * g[3n + c] / scale_c on free DOFs (bf16, 4 channels).
__global__ void k_pack_dout(int nx, int ny, int Wp, int Hp, const double* g, const double* scal,
                            const double* fac, double sw, double st, const uint8_t* cmask,
                            uint16_t* dout) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)Wp * Hp) return;
  const int x = (int)(p % Wp), y = (int)(p / Wp);
  uint16_t v[4] = {0, 0, 0, 0};
  if (x <= nx && y <= ny) {
    const int64_t n = (int64_t)y * (nx + 1) + x;
    const unsigned fix = cmask[n];
    const double f = fac[0] * scal[0];
    const double inv[3] = {1.0 / sw, 1.0 / st, 1.0 / st};
    for (int d = 0; d < 3; d++)
      if (!((fix >> d) & 1u)) v[d] = f2bf((float)(f * inv[d] * g[3 * n + d]));
  }
  *(uint2*)&dout[4 * p] = make_uint2(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16));
}

// Two dot products a.b and c.d, per-block partials (fixed order).
__global__ void k_dot2(int64_t n, const double* a, const double* b, const double* c,
                       const double* d, double* part) {
  __shared__ double r0[256], r1[256];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    s0 += a[i] * b[i];
    if (c) s1 += c[i] * d[i];
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      r0[threadIdx.x] += r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

// out[0] = sum a.b, out[1] = sum c.d over the partials.
__global__ void k_dot2_finish(const double* part, int np, double* out) {
  const double a = block_total(part, np, 2);
  __syncthreads();
  const double b = block_total(part + 1, np, 2);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

// d = m - e
__global__ void k_sub(int64_t n, const double* m, const double* e, double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = m[i] - e[i];
}

// loss = dKd / eKe, gradient factor 2 / eKe: io = {dKd, eKe} -> out
__global__ void k_loss(const double* io, double* loss, double* fac) {
  if (threadIdx.x == 0) {
    const double eke = io[1] > 0.0 ? io[1] : 1.0;
    loss[0] = io[0] / eke;
    fac[0] = 2.0 / eke;
  }
}

// Adam on the flat FP32 parameters (bias-corrected).
__global__ void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1,
                       float b2, float eps, float c1, float c2, float clip) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * clip;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// sum of squares of the flat gradient, per-block partials
__global__ void k_sumsq(int64_t n, const float* g, double* part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += (double)g[i] * g[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Training error field of one seed: w = phi(x, y) sum_k a_k exp(-|x - c_k|^2 / (2 s_k^2)),
// phi = sin^2(pi u) sin^2(pi v) on the plate's unit square (u, v); thx = dw/dy,
// thy = -dw/dx (the rigid-mode signs of the Q4 element); zero on constrained DOFs.
__global__ void k_gen_error(KGrid G, const double* bumps, int nb, double* e) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(G.nx + 1) * (G.ny + 1)) return;
  const int i = (int)(n % (G.nx + 1)), j = (int)(n / (G.nx + 1));
  const double x = G.xc[i], y = G.yc[j];
  const double x0 = G.xc[0], y0 = G.yc[0];
  const double Lx = G.xc[G.nx] - x0, Ly = G.yc[G.ny] - y0;
  double f = 0.0, fx = 0.0, fy = 0.0;
  for (int k = 0; k < nb; k++) {
    const double a = bumps[4 * k], cx = bumps[4 * k + 1], cy = bumps[4 * k + 2],
                 s = bumps[4 * k + 3];
    const double dx = x - cx, dy = y - cy;
    const double g = a * exp(-(dx * dx + dy * dy) / (2.0 * s * s));
    f += g;
    fx -= g * dx / (s * s);
    fy -= g * dy / (s * s);
  }
  const double pi = 3.141592653589793;
  const double su = sin(pi * (x - x0) / Lx), sv = sin(pi * (y - y0) / Ly);
  const double cu = cos(pi * (x - x0) / Lx), cv = cos(pi * (y - y0) / Ly);
  const double ph = su * su * sv * sv;
  const double phx = 2.0 * su * cu * (pi / Lx) * sv * sv;
  const double phy = su * su * 2.0 * sv * cv * (pi / Ly);
  const double w = ph * f, wx = phx * f + ph * fx, wy = phy * f + ph * fy;
  const unsigned fix = G.cmask[n];
  e[3 * n] = (fix & 1u) ? 0.0 : w;
  e[3 * n + 1] = (fix & 2u) ? 0.0 : wy;
  e[3 * n + 2] = (fix & 4u) ? 0.0 : -wx;
}

// node log(t / t0): mean of the adjacent elements' log thickness
__global__ void k_node_logt(KGrid G, float* logt) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)(G.nx + 1) * (G.ny + 1)) return;
  const int i = (int)(n % (G.nx + 1)), j = (int)(n / (G.nx + 1));
  double s = 0.0;
  int c = 0;
  for (int b = j - 1; b <= j; b++)
    for (int a = i - 1; a <= i; a++)
      if (a >= 0 && a < G.nx && b >= 0 && b < G.ny) {
        const double t = G.elem_t ? (double)G.elem_t[(size_t)b * G.nx + a] : G.t0;
        s += log(t / G.t0);
        c++;
      }
  logt[n] = (float)(c ? s / c : 0.0);
}

// ---------------------------------------------------------- host utils

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uni(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
double gauss(uint64_t& s) {
  const double u1 = std::max(uni(s), 1e-300), u2 = uni(s);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

// out = base + omega B in (B the 3x3 nodal block-Jacobi inverse; base may be null)
__global__ void k_bjw(int64_t nn, const double* minv, const double* in, double omega,
                      const double* base, double* out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double rv[3] = {in[3 * n], in[3 * n + 1], in[3 * n + 2]}, zv[3];
  sym3_apply(minv + 6 * n, rv, zv);
  for (int d = 0; d < 3; d++) out[3 * n + d] = (base ? base[3 * n + d] : 0.0) + omega * zv[d];
}

// out = a + alpha b
__global__ void k_lin(int64_t n, const double* a, double alpha, const double* b, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + alpha * b[i];
}

// out = a + sign * g[0] * b (g on the device)
__global__ void k_axpy_dev(int64_t n, const double* a, const double* g, double sign,
                           const double* b, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + sign * g[0] * b[i];
}

// Partials of b.Kb, b.Ku, u.Ku, b.r, u.r, five per block in a fixed order.
__global__ __launch_bounds__(256) void k_dot5(int64_t n, const double* b, const double* kb,
                                              const double* u, const double* ku, const double* r,
                                              double* part) {
  __shared__ double red[5][256];
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double bi = b[i], ui = u[i], ri = r[i];
    s[0] += bi * kb[i];
    s[1] += bi * ku[i];
    s[2] += ui * ku[i];
    s[3] += bi * ri;
    s[4] += ui * ri;
  }
  for (int k = 0; k < 5; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 5; k++) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) part[5 * (size_t)blockIdx.x + k] = red[k][0];
}

// (g0, g1) minimising the energy error of z0 + g0 b + g1 u (b = B r1, the
// block-Jacobi part of the network's output u): the 2 x 2 normal equations
// [b.Kb b.Ku; b.Ku u.Ku] g = [b.r1; u.r1].  When b and u are parallel in
// the energy inner product (1 - cos^2 < 1e-10, as for a zero network, where
// u = b) the step is along u alone, g1 = u.r1 / u.Ku.
__global__ void k_gamma2(const double* part, int np, double* g) {
  double a[5];
  for (int k = 0; k < 5; k++) {
    a[k] = block_total(part + k, np, 5);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double bb = a[0], bu = a[1], uu = a[2], br = a[3], ur = a[4];
    const double det = bb * uu - bu * bu;
    if (bb > 0.0 && uu > 0.0 && det > 1e-10 * bb * uu) {
      g[0] = (br * uu - ur * bu) / det;
      g[1] = (bb * ur - bu * br) / det;
    } else {
      g[0] = 0.0;
      g[1] = uu > 0.0 ? ur / uu : 0.0;
    }
  }
}

// out = a + sign (g[0] b + g[1] c) (g on the device)
__global__ void k_axpy2_dev(int64_t n, const double* a, const double* g, double sign,
                            const double* b, const double* c, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + sign * (g[0] * b[i] + g[1] * c[i]);
}

// v /= |v| (host-provided norm), for the power iteration
__global__ void k_scale_vec(int64_t n, double* v, double f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= f;
}

constexpr int kDotBlocks = 256;
constexpr int kRowsPerChunk = 8;

}  // namespace

// ------------------------------------------------------------- network

// The persistent kernel's compile-time shapes (input channels per chunk,
// 16-channel output blocks, rows per tile), in the order they are tried.
static const int kConvpShapes[6][3] = {{16, 1, 4}, {16, 2, 4}, {32, 2, 4},
                                        {32, 2, 2}, {32, 1, 4}, {32, 1, 2}};
constexpr int kNConvp = 6;

// The k_convp shape for cin input channels and nb output blocks: the first
// whose output blocks divide nb (the rest in groups over the workgroups), at
// most two of them, and whose weights and tile leave room for two
// workgroups per CU, else one; -1 if none fits.  Measured at 1025^2: the
// 64-channel layers on the 257^2 grid in two groups of 32 channels, ba 13.6
// -> 12.2 us and bb 22.2 -> 19.0 us (their weights take half the LDS, so
// more workgroups stay resident); one block per workgroup was slower for the
// 32-channel layers (d1a 40 -> 46 us).
int Cnn::convp_pick(int cin, int nb) const {
  if (cin % 16) return -1;
  const int cc = cin % 32 == 0 ? 32 : 16;
  const int nb_max = 2;
  for (size_t cap : {lds_max / 2, lds_max})
    for (int k = 0; k < kNConvp; k++) {
      const int N = kConvpShapes[k][1];
      if (cc == kConvpShapes[k][0] && N <= nb_max && nb % N == 0 && convp_lds(cin, k) <= cap)
        return k;
    }
  return -1;
}

// the weights, the input tile and the bias
size_t Cnn::convp_lds(int cin, int k) const {
  const int C = kConvpShapes[k][0], N = kConvpShapes[k][1], R = kConvpShapes[k][2];
  return ((size_t)N * 16 * (9 * cin + 4) + (size_t)(R + 2) * 66 * (C + 4)) * 2 + N * 16 * 4;
}

int Cnn::conv_fwd(const Layer& L, const TV& in, const TV& out, float* out32, int out32_cs,
                  int lev, bool backward, const TV* mask, int mask_from, const TV* first,
                  int first_ch) {
  ConvArgs A;
  memset(&A, 0, sizeof(A));
  const int cin = backward ? L.coutA : L.cin;
  const int cout = backward ? L.cin : L.cout;
  const int nb = backward ? L.cinP / 16 : L.coutP / 16;
  A.in = in.p; A.in_cs = in.cs; A.in_c0 = in.c0; A.cin = cin;
  A.out = out.p; A.out_cs = out.cs; A.out_c0 = out.c0; A.cout = cout;
  A.out32 = out32; A.out32_cs = out32_cs;
  A.w = backward ? wt.p + L.wt_off : wf.p + L.wf_off;
  A.kp = backward ? L.kpT : L.kp;
  A.bias = (backward || linear) ? nullptr : params.p + L.b_off;
  A.H = H[lev]; A.W = W[lev]; A.dil = L.dil;
  A.relu = backward ? 0 : L.relu;
  if (mask) {
    A.mask = mask->p; A.mask_cs = mask->cs; A.mask_c0 = mask->c0; A.mask_from = mask_from;
  }
  int64_t cs_max = std::max({(int64_t)A.in_cs, (int64_t)A.out_cs, (int64_t)A.out32_cs,
                             (int64_t)A.mask_cs});
  if (first) {
    // a two-part input: the persistent kernel only, whole chunks per part
    const int shape = convp_pick(cin, nb);
    if (backward || mask || A.dil != 1 || shape < 0 || first_ch <= 0 || first_ch >= cin ||
        first_ch % kConvpShapes[shape][0] || first->cs % 8 || first->c0 % 8 ||
        A.in_cs % 8 || A.in_c0 % 8)
      return fail(PP_ERR_INVALID, "conv: no two-part input for this layer");
    A.pin = first->p; A.p_cs = first->cs; A.p_c0 = first->c0; A.p_ch = first_ch;
    cs_max = std::max(cs_max, (int64_t)first->cs);
  }
  if ((int64_t)H[lev] * W[lev] * cs_max + cs_max >= (1LL << 31))
    return fail(PP_ERR_INVALID, "conv: tensor too large for 32-bit offsets");
  hipStream_t s = stream;
  // forward passes of the network's own shapes: the compile-time kernels
  if (!backward && !mask && A.dil == 1 && cin % 16 == 0 && A.in_cs % 8 == 0 &&
      A.in_c0 % 8 == 0) {
    // input channels in chunks of 32 or 16; the persistent kernel holds all
    // the weights, the tile-per-workgroup kernel one chunk of them
    const int cc = cin % 32 == 0 ? 32 : 16;
    A.cc = cc;
    const int shape = convp_pick(cin, nb);
    if (shape >= 0) {
      const int R = kConvpShapes[shape][2];
      const int nt = ((W[lev] + 63) / 64) * ((H[lev] + R - 1) / R);
      const size_t lds = convp_lds(cin, shape);
      // workgroups: what stays resident (kernel occupancy x CUs), whole XCD
      // groups, at most one per tile
      // output blocks beyond the shape's in groups, one per workgroup
      A.groups = nb / kConvpShapes[shape][1];
      const int gr = A.groups;
#define PP_CONVP(K, C, N, R)                                                                  \
  if (shape == K) {                                                                             \
    int occ = 0;                                                                                \
    PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_convp<C, N, R>, 256, lds));      \
    const int g = 8 * gr * std::max(1, std::min((nt + 7) / 8, std::max(1, occ) * cu / (8 * gr))); \
    hipLaunchKernelGGL((k_convp<C, N, R>), dim3(g), dim3(256), lds, s, A, nt);                  \
    PP_HIP(hipGetLastError());                                                                  \
    return PP_OK;                                                                               \
  }
      PP_CONVP(0, 16, 1, 4)
      PP_CONVP(1, 16, 2, 4)
      PP_CONVP(2, 32, 2, 4)
      PP_CONVP(3, 32, 2, 2)
      PP_CONVP(4, 32, 1, 4)
      PP_CONVP(5, 32, 1, 2)
#undef PP_CONVP
    }
    auto tiles_of = [&](int mb) { return ((W[lev] + 63) / 64) * ((H[lev] + mb - 1) / mb); };
    auto lds_of = [&](int nbk, int mb) {
      return ((size_t)nbk * 16 * (9 * cc + 4) + (size_t)(mb + 2) * 66 * (cc + 4)) * 2;
    };
#define PP_CONVF(C, N, R)                                                                \
  if (cc == C && nb == N && lds_of(N, R) <= lds_max) {                                   \
    const int nt = tiles_of(R);                                                          \
    hipLaunchKernelGGL((k_convf<C, N, R>), dim3(8 * ((nt + 7) / 8)), dim3(256), lds_of(N, R), \
                       s, A, nt);                                                        \
    PP_HIP(hipGetLastError());                                                           \
    return PP_OK;                                                                        \
  }
    PP_CONVF(16, 1, 8)
    PP_CONVF(16, 2, 4)
    PP_CONVF(32, 2, 4)
    PP_CONVF(32, 4, 2)
#undef PP_CONVF
  }
  // input channels per chunk: all of them up to 32, else 32 or 16
  A.cc = cin <= 32 ? cin : cin % 32 == 0 ? 32 : cin % 16 == 0 ? 16 : 4;
  const int kc = (9 * A.cc + 15) / 16 * 16;
  // rows per tile: 4 when weights and tile fit two workgroups per CU, else 2,
  // else the largest that fits one
  auto lds_of = [&](int mb) {
    return (size_t)nb * 16 * kc * 2 + (size_t)(mb + 2 * A.dil) * (64 + 2 * A.dil) * (A.cc + 4) * 2;
  };
  int MB = 0;
  for (size_t cap : {lds_max / 2, lds_max})
    for (int mb : {4, 2})
      if (!MB && lds_of(mb) <= cap) MB = mb;
  if (!MB || cin % 4) return fail(PP_ERR_INVALID, "conv: layer too wide for LDS");
  if (64 + 2 * A.dil > 72 || (MB + 2 * A.dil) * (64 + 2 * A.dil) * (A.cc / 4) >= 65536 ||
      nb * 16 * (kc / 4) >= 65536 || kc / 4 > 72)
    return fail(PP_ERR_INVALID, "conv: input tile beyond the staging index range");
  const size_t lds = lds_of(MB);
  const int grid = ((W[lev] + 63) / 64) * ((H[lev] + MB - 1) / MB);
  const bool u16 = A.cc % 16 == 0;
  if (!u16 && nb != 1) return fail(PP_ERR_INVALID, "conv: narrow input chunks need one N block");
#define PP_CONV(N, R, U) hipLaunchKernelGGL((k_conv<N, R, U>), dim3(grid), dim3(256), lds, s, A)
#define PP_CONV_CASE(N)                  \
  case N:                                \
    if (MB == 4) PP_CONV(N, 4, true);    \
    else PP_CONV(N, 2, true);            \
    break;
  if (!u16) {
    if (MB == 4) PP_CONV(1, 4, false);
    else PP_CONV(1, 2, false);
  } else {
    switch (nb) {
      PP_CONV_CASE(1)
      PP_CONV_CASE(2)
      PP_CONV_CASE(3)
      PP_CONV_CASE(4)
      PP_CONV_CASE(6)
      default: return fail(PP_ERR_INVALID, "conv: unsupported channel count");
    }
  }
#undef PP_CONV_CASE
#undef PP_CONV
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// LDS bytes of k_conv2<mb>: both layers' weights, the input region and the
// intermediate tile
static size_t conv2_lds(int mb) {
  return ((size_t)16 * (48 + 4) + 16 * (9 * 16 + 4) + (size_t)(mb + 4) * 68 * 4 +
          (size_t)(mb + 2) * 66 * 20) * 2 + 32 * 4;
}

bool Cnn::conv2_ok(const Layer& La, const Layer& Lb) const {
  return La.cin == 4 && La.cout == 16 && Lb.cin == 16 && Lb.cout == 16 && La.dil == 1 &&
         Lb.dil == 1 && La.kp == 48 && Lb.kp == 144 && conv2_lds(kConv2Rows) <= lds_max;
}

int Cnn::conv2_fwd(const Layer& La, const Layer& Lb, const TV& in, const TV& out, int lev) {
  if (!conv2_ok(La, Lb) || in.cs % 4 || in.c0 % 4 || out.cs % 4 || out.c0 % 4)
    return fail(PP_ERR_INVALID, "conv2: layer pair not supported");
  Conv2Args A;
  memset(&A, 0, sizeof(A));
  A.in = in.p; A.in_cs = in.cs; A.in_c0 = in.c0;
  A.wa = wf.p + La.wf_off; A.kpa = La.kp;
  A.wb = wf.p + Lb.wf_off; A.kpb = Lb.kp;
  A.bias_a = linear ? nullptr : params.p + La.b_off;
  A.bias_b = linear ? nullptr : params.p + Lb.b_off;
  A.relu_a = La.relu; A.relu_b = Lb.relu;
  A.out = out.p; A.out_cs = out.cs; A.out_c0 = out.c0; A.cout = Lb.cout;
  A.H = H[lev]; A.W = W[lev];
  const int64_t cs_max = std::max((int64_t)in.cs, (int64_t)out.cs);
  if ((int64_t)H[lev] * W[lev] * cs_max + cs_max >= (1LL << 31))
    return fail(PP_ERR_INVALID, "conv2: tensor too large for 32-bit offsets");
  constexpr int R = kConv2Rows;
  const int nt = ((W[lev] + 63) / 64) * ((H[lev] + R - 1) / R);
  const size_t lds = conv2_lds(R);
  int occ = 0;
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_conv2<R>, 256, lds));
  const int g = 8 * std::max(1, std::min((nt + 7) / 8, std::max(1, occ) * cu / 8));
  hipLaunchKernelGGL(k_conv2<R>, dim3(g), dim3(256), lds, stream, A, nt);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int Cnn::conv_wgrad(const Layer& L, const TV& in, const TV& dout, int lev) {
  WgArgs A;
  A.in = in.p; A.in_cs = in.cs; A.in_c0 = in.c0; A.cin = L.cin;
  A.dout = dout.p; A.dout_cs = dout.cs; A.dout_c0 = dout.c0; A.cout = L.cout;
  A.H = H[lev]; A.W = W[lev]; A.dil = L.dil;
  A.kpad = L.kp; A.coutP = L.coutP; A.rows = kRowsPerChunk;
  A.part = wpart.p; A.bpart = bpart.p;
  const int nchunks = (H[lev] + kRowsPerChunk - 1) / kRowsPerChunk;
  const int npairs = (L.coutP / 16) * (L.kp / 16);
  const size_t lds = ((size_t)L.cin * 3 * (64 + 2 * L.dil) + (size_t)L.coutP * 64) * 2;
  hipLaunchKernelGGL(k_wgrad, dim3(nchunks, (npairs + 15) / 16), dim3(256), lds, stream, A);
  PP_HIP(hipGetLastError());
  const int k9 = 9 * L.cin;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblk((int64_t)L.cout * k9)), dim3(256), 0, stream,
                     (const float*)wpart.p, (const float*)bpart.p, nchunks, L.coutP, L.kp, L.cout,
                     k9, grads.p + L.w_off, grads.p + L.b_off);
  PP_HIP(hipGetLastError());
  if (linear) PP_HIP(hipMemsetAsync(grads.p + L.b_off, 0, L.cout * 4, stream));
  return PP_OK;
}

int Cnn::resample(bool restrict_op, const TV& in, int lin, const TV& out, int lout, int C,
                  float scale, const TV* mask, bool accumulate) {
  if (C % 4 || in.cs % 4 || in.c0 % 4 || out.cs % 4 || out.c0 % 4 ||
      (mask && (mask->cs % 4 || mask->c0 % 4)))
    return fail(PP_ERR_INVALID, "resample: channel counts, strides and offsets must be multiples of 4");
  const int64_t n = (int64_t)H[lout] * W[lout] * (C / 4);
  if (n >= (1LL << 31)) return fail(PP_ERR_INVALID, "resample: grid too large");
  if (restrict_op)
    hipLaunchKernelGGL(k_restrict, dim3(nblk(n)), dim3(256), 0, stream, (const uint16_t*)in.p,
                       in.cs, in.c0, H[lin], W[lin], out.p, out.cs, out.c0, H[lout], W[lout], C,
                       scale, mask ? (const uint16_t*)mask->p : nullptr, mask ? mask->cs : 0,
                       mask ? mask->c0 : 0, accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL(k_prolong, dim3(nblk(n)), dim3(256), 0, stream, (const uint16_t*)in.p,
                       in.cs, in.c0, H[lin], W[lin], out.p, out.cs, out.c0, H[lout], W[lout], C,
                       scale, mask ? (const uint16_t*)mask->p : nullptr, mask ? mask->cs : 0,
                       mask ? mask->c0 : 0, accumulate ? 1 : 0);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

int Cnn::setup(pp_plan* plan, const pp_cnn_config* cfg, uint64_t seed) {
  pl = plan;
  stream = plan->ctx->stream;
  cu = plan->ctx->cu_count;
  nx = plan->nx;
  ny = plan->ny;
  const int c1 = cfg && cfg->channels[0] > 0 ? cfg->channels[0] : 16;
  const int c2 = cfg && cfg->channels[1] > 0 ? cfg->channels[1] : 32;
  const int c3 = cfg && cfg->channels[2] > 0 ? cfg->channels[2] : 64;
  const int bdil = cfg && cfg->bottleneck_dilation > 0 ? cfg->bottleneck_dilation : 1;
  if (c1 % 16 || c2 % 16 || c3 % 16 || c1 > 64 || c2 > 64 || c3 > 64)
    return fail(PP_ERR_INVALID, "channels must be multiples of 16, at most 64");
  sw = 1.0;
  st = cfg && cfg->rotation_scale > 0 ? cfg->rotation_scale : 0.3;
  linear = cfg && cfg->relu ? 0 : 1;
  line_search = cfg && cfg->fixed_step ? 0 : 1;
  fuse_pairs = !(cfg && cfg->unfused_forward);
  // FCG(8) by default; FCG(4) with the multigrid smoother, whose stronger
  // steps need fewer kept directions (1024^2: 217 iterations against 242)
  fcg_window = cfg && cfg->fcg_window > 0 ? cfg->fcg_window : cfg && cfg->multigrid ? 4 : 8;
  if (fcg_window > 16) return fail(PP_ERR_INVALID, "fcg_window must be at most 16");
  // padded level-0 grid: 4k + 1 nodes per side
  W[0] = ((nx + 3) / 4) * 4 + 1;
  H[0] = ((ny + 3) / 4) * 4 + 1;
  for (int l = 1; l < 3; l++) {
    W[l] = (W[l - 1] + 1) / 2;
    H[l] = (H[l - 1] + 1) / 2;
  }
  const int C0 = c1, C1 = c2, C2 = c3;
  auto mk = [&](int cin, int cout, int relu, int dil) {
    Layer L;
    L.cin = cin; L.cout = cout; L.relu = linear ? 0 : relu; L.dil = dil;
    L.coutP = (cout + 15) / 16 * 16;
    L.coutA = (cout + 3) / 4 * 4;
    L.cinP = (cin + 15) / 16 * 16;
    L.kp = (9 * cin + 15) / 16 * 16;
    L.kpT = (9 * L.coutA + 15) / 16 * 16;
    return L;
  };
  layers.clear();
  layers.push_back(mk(4, C0, 1, 1));          // e0a
  layers.push_back(mk(C0, C0, 1, 1));         // e0b
  layers.push_back(mk(C0, C1, 1, 1));         // e1a
  layers.push_back(mk(C1, C1, 1, 1));         // e1b
  layers.push_back(mk(C1, C2, 1, 1));         // ba
  layers.push_back(mk(C2, C2, 1, bdil));      // bb
  layers.push_back(mk(C2 + C1, C1, 1, 1));    // d1a
  layers.push_back(mk(C1, C1, 1, 1));         // d1b
  layers.push_back(mk(C1 + C0, C0, 1, 1));    // d0a
  layers.push_back(mk(C0, C0, 1, 1));         // d0b
  layers.push_back(mk(C0, 3, 0, 1));          // head
  int64_t np = 0, nwf = 0, nwt = 0;
  for (auto& L : layers) {
    L.w_off = np;
    np += (int64_t)L.cout * 9 * L.cin;
    L.b_off = np;
    np += L.cout;
    L.wf_off = nwf;
    nwf += (int64_t)L.coutP * L.kp;
    L.wt_off = nwt;
    nwt += (int64_t)L.cinP * L.kpT;
  }
  n_params = np;
  int rc;
  if ((rc = params.alloc(np)) || (rc = grads.alloc(np)) || (rc = adam_m.alloc(np)) ||
      (rc = adam_v.alloc(np)) || (rc = wf.alloc(nwf)) || (rc = wt.alloc(nwt)))
    return rc;
  // He initialisation, zero biases, and a zero head: the untrained network
  // adds no correction, so training starts from the preconditioner without it
  std::vector<float> h(np, 0.0f);
  uint64_t s = seed;
  for (size_t li = 0; li + 1 < layers.size(); li++) {
    const Layer& L = layers[li];
    const double sd = sqrt(2.0 / (9.0 * L.cin));
    for (int64_t k = 0; k < (int64_t)L.cout * 9 * L.cin; k++) h[L.w_off + k] = (float)(sd * gauss(s));
  }
  PP_HIP(hipMemcpyAsync(params.p, h.data(), np * 4, hipMemcpyHostToDevice, stream));
  PP_HIP(hipMemsetAsync(adam_m.p, 0, np * 4, stream));
  PP_HIP(hipMemsetAsync(adam_v.p, 0, np * 4, stream));
  PP_HIP(hipMemsetAsync(grads.p, 0, np * 4, stream));
  adam_t = 0;
  // activations (forward) and their gradients
  const int64_t P0 = (int64_t)H[0] * W[0], P1 = (int64_t)H[1] * W[1], P2 = (int64_t)H[2] * W[2];
  cat0C = C1 + C0;
  cat1C = C2 + C1;
  struct B { DevBuf<uint16_t>* b; int64_t n; };
  B bufs[] = {{&xin, P0 * 4},     {&a0, P0 * C0},       {&cat0, P0 * cat0C},  {&r1, P1 * C0},
              {&a1, P1 * C1},     {&cat1, P1 * cat1C},  {&r2, P2 * C1},       {&a2, P2 * C2},
              {&bb, P2 * C2},     {&c1b, P1 * C1},      {&d1, P1 * C1},       {&c0, P0 * C0},
              {&c0b, P0 * C0},    {&g_c0b, P0 * C0},    {&g_c0, P0 * C0},     {&g_cat0, P0 * cat0C},
              {&g_d1, P1 * C1},   {&g_c1, P1 * C1},     {&g_cat1, P1 * cat1C}, {&g_bb, P2 * C2},
              {&g_a2, P2 * C2},   {&g_r2, P2 * C1},     {&g_a1, P1 * C1},     {&g_r1, P1 * C0},
              {&g_a0, P0 * C0},   {&g_out, P0 * 4},     {&s0a, P0 * C0},      {&s1a, P1 * C1},
              {&p0a, P0 * C1},    {&p1a, P1 * C2}};
  for (auto& b : bufs)
    if ((rc = b.b->alloc(b.n))) return rc;
  for (auto& b : bufs) PP_HIP(hipMemsetAsync(b.b->p, 0, b.n * 2, stream));
  if ((rc = z0v.alloc(pl->n_dof)) || (rc = z1v.alloc(pl->n_dof)) || (rc = r1v.alloc(pl->n_dof)))
    return rc;
  if ((rc = out32.alloc(P0 * 4)) || (rc = zbj.alloc(pl->n_dof)) || (rc = mout.alloc(pl->n_dof)) ||
      (rc = dvec.alloc(pl->n_dof)) || (rc = kvec.alloc(pl->n_dof)) ||
      (rc = evec.alloc(pl->n_dof)) || (rc = rvec.alloc(pl->n_dof)) ||
      (rc = logt.alloc(pl->n_nodes)) || (rc = scal.alloc(16)) ||
      (rc = dotp.alloc(5 * kDotBlocks)) || (rc = bumps.alloc(4 * 64)) ||
      (rc = kbv.alloc(pl->n_dof)))
    return rc;
  // weight-gradient partials: the largest layer x its chunk count
  int64_t wp = 0, bp = 0;
  for (auto& L : layers)
    for (int l = 0; l < 3; l++) {
      const int64_t ch = (H[l] + kRowsPerChunk - 1) / kRowsPerChunk;
      wp = std::max(wp, ch * L.coutP * L.kp);
      bp = std::max(bp, ch * L.coutP);
    }
  if ((rc = wpart.alloc(wp)) || (rc = bpart.alloc(bp))) return rc;
  KGrid G = make_kgrid(pl);
  hipLaunchKernelGGL(k_node_logt, dim3(nblk(pl->n_nodes)), dim3(256), 0, stream, G, logt.p);
  PP_HIP(hipGetLastError());
  // dynamic LDS above 64 KB for the widest layers
  {
    int v = 0, o = 0;
    PP_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, plan->ctx->device));
    PP_HIP(hipDeviceGetAttribute(&o, hipDeviceAttributeSharedMemPerBlockOptin, plan->ctx->device));
    lds_max = (size_t)std::max(v, o);
  }
  const void* ks[] = {(const void*)k_conv<1, 4, true>,  (const void*)k_conv<2, 4, true>,
                      (const void*)k_conv<3, 4, true>,  (const void*)k_conv<4, 4, true>,
                      (const void*)k_conv<6, 4, true>,  (const void*)k_conv<1, 2, true>,
                      (const void*)k_conv<2, 2, true>,  (const void*)k_conv<3, 2, true>,
                      (const void*)k_conv<4, 2, true>,  (const void*)k_conv<6, 2, true>,
                      (const void*)k_conv<1, 4, false>, (const void*)k_conv<1, 2, false>};
  for (auto k : ks)
    PP_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
  const void* kf[] = {(const void*)k_convf<16, 1, 8>, (const void*)k_convf<16, 2, 4>,
                      (const void*)k_convf<32, 2, 4>, (const void*)k_convf<32, 4, 2>,
                      (const void*)k_convp<16, 1, 4>, (const void*)k_convp<16, 2, 4>,
                      (const void*)k_convp<32, 2, 4>, (const void*)k_convp<32, 2, 2>,
                      (const void*)k_convp<32, 1, 4>,
                      (const void*)k_convp<32, 1, 2>, (const void*)k_conv2<kConv2Rows>};
  for (auto k : kf)
    PP_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
  // the forward pass without backward reads d1a's and d0a's concatenated
  // inputs in two parts when both run on the persistent kernel with whole
  // chunks per part
  split_fwd = C0 % 16 == 0 && C1 % 16 == 0 && C2 % 16 == 0;
  for (int l : {6, 8}) {
    const Layer& Lf = layers[l];
    const int shape = convp_pick(Lf.cin, Lf.coutP / 16), pch = l == 6 ? C2 : C1;
    if (Lf.dil != 1 || shape < 0 || pch % kConvpShapes[shape][0]) split_fwd = false;
  }
  size_t mw = 0;
  for (auto& L : layers)
    mw = std::max(mw, ((size_t)L.cin * 3 * (64 + 2 * L.dil) + (size_t)L.coutP * 64) * 2);
  PP_HIP(hipFuncSetAttribute((const void*)k_wgrad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mw));
  if ((rc = pack_weights())) return rc;
  PP_HIP(hipMemsetAsync(scal.p, 0, 16 * 8, stream));
  {
    const double one = 1.0;  // scal[13] = 1, scal[14] = scal[15] = 0: constants
    PP_HIP(hipMemcpyAsync(scal.p + 13, &one, 8, hipMemcpyHostToDevice, stream));
  }
  if (cfg && cfg->smoother_weight > 0) omega = cfg->smoother_weight;
  else if ((rc = estimate_omega())) return rc;
  if (cfg && cfg->multigrid) {
    mg = new Mg();
    if ((rc = mg->setup(pl, omega, cfg->mg_sweeps, cfg->mg_coarse_sweeps, cfg->mg_jacobi != 0)) ||
        (rc = mgz.alloc(pl->n_dof)) || (rc = mgf_cycle_create(pl, mg, &mgc)))
      return rc;
  }
  PP_HIP(hipStreamSynchronize(stream));
  return PP_OK;
}

int Cnn::pack_weights() {
  for (auto& L : layers) {
    const int64_t n = std::max((int64_t)L.coutP * L.kp, (int64_t)L.cinP * L.kpT);
    hipLaunchKernelGGL(k_pack_w, dim3(nblk(n)), dim3(256), 0, stream, (const float*)params.p + L.w_off,
                       L.cout, L.cin, L.coutP, L.kp, L.coutA, L.cinP, L.kpT, wf.p + L.wf_off,
                       wt.p + L.wt_off);
    PP_HIP(hipGetLastError());
  }
  return PP_OK;
}

// M(r) into out (FP64 DOF vector); r on the device.
int Cnn::apply(const double* r, double* out, bool keep) {
  int rc;
  hipLaunchKernelGGL(k_bj_norm, dim3(kDotBlocks), dim3(256), 0, stream, (int64_t)pl->n_nodes,
                     (const double*)pl->minv.p, r, zbj.p, sw, st, dotp.p);
  hipLaunchKernelGGL(k_scale, dim3(1), dim3(256), 0, stream, (const double*)dotp.p, kDotBlocks,
                     (double)pl->n_nodes, scal.p);
  hipLaunchKernelGGL(k_pack_in, dim3(nblk((int64_t)H[0] * W[0])), dim3(256), 0, stream, nx, ny,
                     W[0], H[0], (const double*)zbj.p, (const double*)scal.p, sw, st,
                     (const float*)logt.p, xin.p);
  PP_HIP(hipGetLastError());
  const Layer* L = layers.data();
  const int C0 = L[0].cout, C1 = L[2].cout, C2 = L[4].cout;
  TV vx{xin.p, 4, 0}, va0{a0.p, C0, 0}, vs0{cat0.p, cat0C, C1}, vr1{r1.p, C0, 0},
      va1{a1.p, C1, 0}, vs1{cat1.p, cat1C, C2}, vr2{r2.p, C1, 0}, va2{a2.p, C2, 0},
      vb{bb.p, C2, 0}, vpb{cat1.p, cat1C, 0}, vcat1{cat1.p, cat1C, 0}, vc1{c1b.p, C1, 0},
      vd1{d1.p, C1, 0}, vpd1{cat0.p, cat0C, 0}, vcat0{cat0.p, cat0C, 0}, vc0{c0.p, C0, 0},
      vc0b{c0b.p, C0, 0};
  // consecutive convolutions at one level fused into one launch (the
  // solver's forward pass; the training pass keeps every activation)
  const bool pair0 = !keep && fuse_pairs && conv2_ok(L[0], L[1]) && split_fwd;
  if (!pair0 && (rc = conv_fwd(L[0], vx, va0, nullptr, 0, 0, false))) return rc;
  if (!keep && split_fwd) {
    // skip and prolonged tensors stored alone (whole cache lines), resampled
    // in 16-byte accesses; the decoder's first layers read the two parts of
    // their concatenated input in place: the concatenations are never written
    const TV vs0a{s0a.p, C0, 0}, vs1a{s1a.p, C1, 0}, vp0a{p0a.p, C1, 0}, vp1a{p1a.p, C2, 0};
    if (pair0) {
      if ((rc = conv2_fwd(L[0], L[1], vx, vs0a, 0))) return rc;
    } else if ((rc = conv_fwd(L[1], va0, vs0a, nullptr, 0, 0, false))) {
      return rc;
    }
    hipLaunchKernelGGL(k_restrict8, dim3(nblk((int64_t)H[1] * W[1] * (C0 / 8))), dim3(256), 0,
                       stream, (const uint16_t*)s0a.p, H[0], W[0], r1.p, H[1], W[1], C0);
    if ((rc = conv_fwd(L[2], vr1, va1, nullptr, 0, 1, false))) return rc;
    if ((rc = conv_fwd(L[3], va1, vs1a, nullptr, 0, 1, false))) return rc;
    hipLaunchKernelGGL(k_restrict8, dim3(nblk((int64_t)H[2] * W[2] * (C1 / 8))), dim3(256), 0,
                       stream, (const uint16_t*)s1a.p, H[1], W[1], r2.p, H[2], W[2], C1);
    PP_HIP(hipGetLastError());
    if ((rc = conv_fwd(L[4], vr2, va2, nullptr, 0, 2, false))) return rc;
    if ((rc = conv_fwd(L[5], va2, vb, nullptr, 0, 2, false))) return rc;
    hipLaunchKernelGGL(k_prolong8, dim3(nblk((int64_t)H[1] * W[1] * (C2 / 8))), dim3(256), 0,
                       stream, (const uint16_t*)bb.p, H[2], W[2], p1a.p, H[1], W[1], C2);
    PP_HIP(hipGetLastError());
    if ((rc = conv_fwd(L[6], vs1a, vc1, nullptr, 0, 1, false, nullptr, 0, &vp1a, C2))) return rc;
    if ((rc = conv_fwd(L[7], vc1, vd1, nullptr, 0, 1, false))) return rc;
    hipLaunchKernelGGL(k_prolong8, dim3(nblk((int64_t)H[0] * W[0] * (C1 / 8))), dim3(256), 0,
                       stream, (const uint16_t*)d1.p, H[1], W[1], p0a.p, H[0], W[0], C1);
    PP_HIP(hipGetLastError());
    if ((rc = conv_fwd(L[8], vs0a, vc0, nullptr, 0, 0, false, nullptr, 0, &vp0a, C1))) return rc;
  } else {
    if ((rc = conv_fwd(L[1], va0, vs0, nullptr, 0, 0, false))) return rc;
    if ((rc = resample(true, vs0, 0, vr1, 1, C0, 1.0f, nullptr, false))) return rc;
    if ((rc = conv_fwd(L[2], vr1, va1, nullptr, 0, 1, false))) return rc;
    if ((rc = conv_fwd(L[3], va1, vs1, nullptr, 0, 1, false))) return rc;
    if ((rc = resample(true, vs1, 1, vr2, 2, C1, 1.0f, nullptr, false))) return rc;
    if ((rc = conv_fwd(L[4], vr2, va2, nullptr, 0, 2, false))) return rc;
    if ((rc = conv_fwd(L[5], va2, vb, nullptr, 0, 2, false))) return rc;
    if ((rc = resample(false, vb, 2, vpb, 1, C2, 1.0f, nullptr, false))) return rc;
    if ((rc = conv_fwd(L[6], vcat1, vc1, nullptr, 0, 1, false))) return rc;
    if ((rc = conv_fwd(L[7], vc1, vd1, nullptr, 0, 1, false))) return rc;
    if ((rc = resample(false, vd1, 1, vpd1, 0, C1, 1.0f, nullptr, false))) return rc;
    if ((rc = conv_fwd(L[8], vcat0, vc0, nullptr, 0, 0, false))) return rc;
  }
  if ((rc = conv_fwd(L[9], vc0, vc0b, nullptr, 0, 0, false))) return rc;
  if ((rc = conv_fwd(L[10], vc0b, TV{nullptr, 0, 0}, out32.p, 4, 0, false))) return rc;
  hipLaunchKernelGGL(k_unpack_out, dim3(nblk(pl->n_nodes)), dim3(256), 0, stream, nx, ny, W[0],
                     (const double*)zbj.p, (const float*)out32.p, (const double*)scal.p, sw, st,
                     (const uint8_t*)pl->cmask.p, out);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// Backward pass from g_out (bf16 dL/do); weight and bias gradients into grads.
int Cnn::backward() {
  int rc;
  const Layer* L = layers.data();
  const int C0 = L[0].cout, C1 = L[2].cout, C2 = L[4].cout;
  TV vx{xin.p, 4, 0}, va0{a0.p, C0, 0}, vs0{cat0.p, cat0C, C1}, vr1{r1.p, C0, 0},
      va1{a1.p, C1, 0}, vs1{cat1.p, cat1C, C2}, vr2{r2.p, C1, 0}, va2{a2.p, C2, 0},
      vb{bb.p, C2, 0}, vcat1{cat1.p, cat1C, 0}, vc1{c1b.p, C1, 0}, vd1{d1.p, C1, 0},
      vcat0{cat0.p, cat0C, 0}, vc0{c0.p, C0, 0}, vc0b{c0b.p, C0, 0};
  TV gout{g_out.p, 4, 0}, gc0b{g_c0b.p, C0, 0}, gc0{g_c0.p, C0, 0}, gcat0{g_cat0.p, cat0C, 0},
      gpd1{g_cat0.p, cat0C, 0}, gs0{g_cat0.p, cat0C, C1}, gd1{g_d1.p, C1, 0}, gc1{g_c1.p, C1, 0},
      gcat1{g_cat1.p, cat1C, 0}, gpb{g_cat1.p, cat1C, 0}, gs1{g_cat1.p, cat1C, C2},
      gbb{g_bb.p, C2, 0}, ga2{g_a2.p, C2, 0}, gr2{g_r2.p, C1, 0}, ga1{g_a1.p, C1, 0},
      gr1{g_r1.p, C0, 0}, ga0{g_a0.p, C0, 0};
  // head (no ReLU): dW; d c0b = W^T g_out masked by c0b > 0
  if ((rc = conv_wgrad(L[10], vc0b, gout, 0))) return rc;
  if ((rc = conv_fwd(L[10], gout, gc0b, nullptr, 0, 0, true, mk(vc0b), 0))) return rc;
  if ((rc = conv_wgrad(L[9], vc0, gc0b, 0))) return rc;
  if ((rc = conv_fwd(L[9], gc0b, gc0, nullptr, 0, 0, true, mk(vc0), 0))) return rc;
  if ((rc = conv_wgrad(L[8], vcat0, gc0, 0))) return rc;
  // d cat0: channels >= C1 are s0 (ReLU output), below are P(d1) (linear)
  if ((rc = conv_fwd(L[8], gc0, gcat0, nullptr, 0, 0, true, mk(vcat0), C1))) return rc;
  // P^T = 4 R, masked by d1 > 0
  if ((rc = resample(true, gpd1, 0, gd1, 1, C1, 4.0f, mk(vd1), false))) return rc;
  if ((rc = conv_wgrad(L[7], vc1, gd1, 1))) return rc;
  if ((rc = conv_fwd(L[7], gd1, gc1, nullptr, 0, 1, true, mk(vc1), 0))) return rc;
  if ((rc = conv_wgrad(L[6], vcat1, gc1, 1))) return rc;
  if ((rc = conv_fwd(L[6], gc1, gcat1, nullptr, 0, 1, true, mk(vcat1), C2))) return rc;
  if ((rc = resample(true, gpb, 1, gbb, 2, C2, 4.0f, mk(vb), false))) return rc;
  if ((rc = conv_wgrad(L[5], va2, gbb, 2))) return rc;
  if ((rc = conv_fwd(L[5], gbb, ga2, nullptr, 0, 2, true, mk(va2), 0))) return rc;
  if ((rc = conv_wgrad(L[4], vr2, ga2, 2))) return rc;
  if ((rc = conv_fwd(L[4], ga2, gr2, nullptr, 0, 2, true, nullptr, 0))) return rc;
  // R^T = P / 4, accumulated into d s1 (skip path), masked by s1 > 0
  if ((rc = resample(false, gr2, 2, gs1, 1, C1, 0.25f, mk(vs1), true))) return rc;
  if ((rc = conv_wgrad(L[3], va1, gs1, 1))) return rc;
  if ((rc = conv_fwd(L[3], gs1, ga1, nullptr, 0, 1, true, mk(va1), 0))) return rc;
  if ((rc = conv_wgrad(L[2], vr1, ga1, 1))) return rc;
  if ((rc = conv_fwd(L[2], ga1, gr1, nullptr, 0, 1, true, nullptr, 0))) return rc;
  if ((rc = resample(false, gr1, 1, gs0, 0, C0, 0.25f, mk(vs0), true))) return rc;
  if ((rc = conv_wgrad(L[1], va0, gs0, 0))) return rc;
  if ((rc = conv_fwd(L[1], gs0, ga0, nullptr, 0, 0, true, mk(va0), 0))) return rc;
  if ((rc = conv_wgrad(L[0], vx, ga0, 0))) return rc;
  return PP_OK;
}

// out = base + S in: S = w B (block Jacobi) or the multigrid V-cycle.
int Cnn::smooth(const double* in, const double* base, double* out) {
  if (!mg) {
    const int64_t nn = pl->n_nodes;
    hipLaunchKernelGGL(k_bjw, dim3(nblk(nn)), dim3(256), 0, stream, nn, (const double*)pl->minv.p,
                       in, omega, base, out);
    PP_HIP(hipGetLastError());
    return PP_OK;
  }
  if (!base) return mgf_cycle_apply(mgc, in, out);
  int rc;
  if ((rc = mgf_cycle_apply(mgc, in, mgz.p))) return rc;
  const int64_t nd = pl->n_dof;
  hipLaunchKernelGGL(k_lin, dim3(nblk(nd)), dim3(256), 0, stream, nd, base, 1.0,
                     (const double*)mgz.p, out);
  PP_HIP(hipGetLastError());
  return PP_OK;
}

// The preconditioner: smoothing S (w B, or a multigrid V-cycle) around the
// network,
//   z0 = S r;  r1 = r - K z0;  u = N(r1);  z1 = z0 + g u;  z = z1 + S (r1 - g K u)
// (the network takes the error the smoother leaves).  With line_search the
// step g = u.r1 / u.K u minimises the energy error of z1 along u, so an
// application is never worse than the smoothing alone; in training g = 1
// (the loss sees the network's own output).
int Cnn::precondition(const double* r, double* out, bool line_search, bool keep) {
  int rc;
  const int64_t nd = pl->n_dof;
  if ((rc = smooth(r, nullptr, z0v.p))) return rc;
  if ((rc = struct_sweep(pl, 2, z0v.p, r, nullptr, 0.0, 0.0, r1v.p))) return rc;  // r - K z0
  if ((rc = apply(r1v.p, mout.p, keep))) return rc;
  if ((rc = struct_apply_k(pl, mout.p, kvec.p))) return rc;  // K u
  double* g = scal.p + 10;
  if (line_search) {
    // two directions, the network's block-Jacobi part b = zbj and its whole
    // output u, each with its own energy-optimal length: an application is
    // then never worse than the same preconditioner with a zero network
    if ((rc = struct_apply_k(pl, zbj.p, kbv.p))) return rc;  // K b
    hipLaunchKernelGGL(k_dot5, dim3(kDotBlocks), dim3(256), 0, stream, nd, (const double*)zbj.p,
                       (const double*)kbv.p, (const double*)mout.p, (const double*)kvec.p,
                       (const double*)r1v.p, dotp.p);
    hipLaunchKernelGGL(k_gamma2, dim3(1), dim3(256), 0, stream, (const double*)dotp.p, kDotBlocks, g);
    hipLaunchKernelGGL(k_axpy2_dev, dim3(nblk(nd)), dim3(256), 0, stream, nd, (const double*)z0v.p,
                       (const double*)g, 1.0, (const double*)zbj.p, (const double*)mout.p, z1v.p);
    hipLaunchKernelGGL(k_axpy2_dev, dim3(nblk(nd)), dim3(256), 0, stream, nd, (const double*)r1v.p,
                       (const double*)g, -1.0, (const double*)kbv.p, (const double*)kvec.p, r1v.p);
    PP_HIP(hipGetLastError());
    return smooth(r1v.p, z1v.p, out);
  } else {
    hipLaunchKernelGGL(k_lin, dim3(1), dim3(1), 0, stream, (int64_t)1, (const double*)scal.p + 14,
                       1.0, (const double*)scal.p + 13, g);  // g = 0 + 1
  }
  hipLaunchKernelGGL(k_axpy_dev, dim3(nblk(nd)), dim3(256), 0, stream, nd, (const double*)z0v.p,
                     (const double*)g, 1.0, (const double*)mout.p, z1v.p);
  hipLaunchKernelGGL(k_axpy_dev, dim3(nblk(nd)), dim3(256), 0, stream, nd, (const double*)r1v.p,
                     (const double*)g, -1.0, (const double*)kvec.p, r1v.p);
  PP_HIP(hipGetLastError());
  return smooth(r1v.p, z1v.p, out);
}

// One pair: e (device, FP64), r = K e, z = precondition(r), e_f = e - z.
// Loss e_f^T K e_f / e^T K e.  With E = I - S K (S the symmetric smoother),
// e_f = E (e - z0 - u) for the network output u, so dL/du = -2 E^T K e_f /
// e^T K e and E^T K e_f = q1 - K S q1 with q1 = K e_f.
int Cnn::loss_grad(const double* e, double* loss_dev) {
  int rc;
  const int64_t nd = pl->n_dof;
  if ((rc = struct_apply_k(pl, e, rvec.p))) return rc;
  if ((rc = precondition(rvec.p, dvec.p, false, true))) return rc;  // z (scal[0] = s of the input)
  hipLaunchKernelGGL(k_sub, dim3(nblk(nd)), dim3(256), 0, stream, nd, e, (const double*)dvec.p,
                     dvec.p);  // e_f
  if ((rc = struct_apply_k(pl, dvec.p, kvec.p))) return rc;  // q1
  hipLaunchKernelGGL(k_dot2, dim3(kDotBlocks), dim3(256), 0, stream, nd, (const double*)dvec.p,
                     (const double*)kvec.p, e, (const double*)rvec.p, dotp.p);
  hipLaunchKernelGGL(k_dot2_finish, dim3(1), dim3(256), 0, stream, (const double*)dotp.p,
                     kDotBlocks, scal.p + 2);
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(64), 0, stream, (const double*)scal.p + 2, loss_dev,
                     scal.p + 4);  // scal[4] = 2 / eKe
  if ((rc = smooth(kvec.p, nullptr, z0v.p))) return rc;     // S q1
  // E^T K e_f = q1 - K S q1
  if ((rc = struct_sweep(pl, 2, z0v.p, kvec.p, nullptr, 0.0, 0.0, r1v.p))) return rc;
  // dL/do = -(2 / eKe) S^T K e_f * s / scale_c
  hipLaunchKernelGGL(k_lin, dim3(1), dim3(1), 0, stream, (int64_t)1, (const double*)scal.p + 15,
                     -1.0, (const double*)scal.p + 4, scal.p + 5);  // scal[5] = 0 - scal[4]
  hipLaunchKernelGGL(k_pack_dout, dim3(nblk((int64_t)H[0] * W[0])), dim3(256), 0, stream, nx, ny,
                     W[0], H[0], (const double*)r1v.p, (const double*)scal.p,
                     (const double*)scal.p + 5, sw, st, (const uint8_t*)pl->cmask.p, g_out.p);
  PP_HIP(hipGetLastError());
  return backward();
}

// omega = 4 / (3 lambda_max(B K)), lambda_max by 40 power iterations on B K
// from a fixed pseudo-random start (x 1.05 against underestimation).
int Cnn::estimate_omega() {
  const int64_t nd = pl->n_dof, nn = pl->n_nodes;
  std::vector<double> h(nd);
  uint64_t s = 12345;
  for (int64_t i = 0; i < nd; i++) h[i] = pl->h_cmask[i / 3] >> (i % 3) & 1 ? 0.0 : uni(s) - 0.5;
  PP_HIP(hipMemcpyAsync(z0v.p, h.data(), nd * 8, hipMemcpyHostToDevice, stream));
  double lam = 0.0;
  for (int it = 0; it < 40; it++) {
    int rc;
    if ((rc = struct_apply_k(pl, z0v.p, kvec.p))) return rc;
    hipLaunchKernelGGL(k_bjw, dim3(nblk(nn)), dim3(256), 0, stream, nn, (const double*)pl->minv.p,
                       (const double*)kvec.p, 1.0, (const double*)nullptr, z1v.p);
    hipLaunchKernelGGL(k_dot2, dim3(kDotBlocks), dim3(256), 0, stream, nd, (const double*)z1v.p,
                       (const double*)z1v.p, (const double*)z0v.p, (const double*)z0v.p, dotp.p);
    double hp[2 * kDotBlocks];
    PP_HIP(hipMemcpyAsync(hp, dotp.p, sizeof(hp), hipMemcpyDeviceToHost, stream));
    PP_HIP(hipStreamSynchronize(stream));
    double a = 0, b = 0;
    for (int k = 0; k < kDotBlocks; k++) {
      a += hp[2 * k];
      b += hp[2 * k + 1];
    }
    if (!(a > 0) || !(b > 0)) break;
    lam = sqrt(a / b);
    hipLaunchKernelGGL(k_scale_vec, dim3(nblk(nd)), dim3(256), 0, stream, nd, z1v.p, 1.0 / sqrt(a));
    std::swap(z0v.p, z1v.p);
  }
  PP_HIP(hipGetLastError());
  omega = lam > 0 ? 4.0 / (3.0 * 1.05 * lam) : 0.5;
  return PP_OK;
}

int Cnn::gen_error(uint64_t seed, double* e) {
  // 32 bumps: amplitude ~ N(0,1) * sigma, sigma log-uniform in [2h, L/2]
  const int nb = 32;
  const double x0 = pl->h_nodes[0], x1 = pl->h_nodes[2 * nx];
  const double y0 = pl->h_nodes[1], y1 = pl->h_nodes[2 * ((size_t)ny * (nx + 1)) + 1];
  const double L = std::max(x1 - x0, y1 - y0), h = std::min((x1 - x0) / nx, (y1 - y0) / ny);
  std::vector<double> b(4 * nb);
  uint64_t s = seed;
  for (int k = 0; k < nb; k++) {
    const double sig = exp(log(2.0 * h) + uni(s) * (log(0.5 * L) - log(2.0 * h)));
    b[4 * k] = gauss(s) * sig;
    b[4 * k + 1] = x0 + uni(s) * (x1 - x0);
    b[4 * k + 2] = y0 + uni(s) * (y1 - y0);
    b[4 * k + 3] = sig;
  }
  PP_HIP(hipMemcpyAsync(bumps.p, b.data(), b.size() * 8, hipMemcpyHostToDevice, stream));
  KGrid G = make_kgrid(pl);
  hipLaunchKernelGGL(k_gen_error, dim3(nblk(pl->n_nodes)), dim3(256), 0, stream, G,
                     (const double*)bumps.p, nb, e);
  PP_HIP(hipGetLastError());
  // the host vector b must outlive the copy
  PP_HIP(hipStreamSynchronize(stream));
  return PP_OK;
}

int Cnn::adam_step(float lr, float max_norm) {
  adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)adam_t), c2 = 1.0f - powf(b2, (float)adam_t);
  float clip = 1.0f;
  if (max_norm > 0) {
    hipLaunchKernelGGL(k_sumsq, dim3(kDotBlocks), dim3(256), 0, stream, n_params,
                       (const float*)grads.p, dotp.p);
    double h[kDotBlocks];
    PP_HIP(hipMemcpyAsync(h, dotp.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    PP_HIP(hipStreamSynchronize(stream));
    double ss = 0;
    for (int i = 0; i < kDotBlocks; i++) ss += h[i];
    const double gn = sqrt(ss);
    last_grad_norm = gn;
    if (gn > max_norm) clip = (float)(max_norm / gn);
  }
  hipLaunchKernelGGL(k_adam, dim3(nblk(n_params)), dim3(256), 0, stream, n_params, params.p,
                     (const float*)grads.p, adam_m.p, adam_v.p, lr, b1, b2, 1e-8f, c1, c2, clip);
  PP_HIP(hipGetLastError());
  return pack_weights();
}

Cnn::~Cnn() {
  mgf_cycle_destroy(mgc);
  delete mg;
  mg = nullptr;
  if (!pl) return;
  if (pl->learned == this) pl->learned = nullptr;
  (void)hipStreamSynchronize(stream);
  pp_plan* p = pl;
  pl = nullptr;
  plan_release(p);
}

}  // namespace pp
