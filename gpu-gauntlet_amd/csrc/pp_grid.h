// Structured tensor-grid description and the LDS tile machinery shared by
// the structured kernels (FP64 PCG, mixed precision, learned preconditioner).
//
// Grid: node (i, j) has index j*(nx+1)+i; element (i, j) has index j*nx+i
// and corners SW (i,j), SE (i+1,j), NE (i+1,j+1), NW (i,j+1) (the element
// order of mesher.ts:381-407).  Every element of a tensor grid built from
// float32 coordinates is an exact axis-aligned rectangle, so an element is
// described by its column width a and row height b alone.
//
// A tile of TX x TY nodes is owned by one workgroup.  The tile plus a
// one-node halo of the input vector is staged in LDS, every element touching
// the tile ((TX+1) x (TY+1)) is applied matrix-free into LDS, and each node
// sums its four elements' contributions in a fixed order: race-free without
// atomics or colour passes, and deterministic.
#pragma once

#include "pp_device.h"

namespace pp {

struct KGrid {
  int nx, ny;
  const double* ax;     // 2 per column: a, 1/a
  const double* by;     // 2 per row: b, 1/b
  const double* xc;     // node x coordinates (nx+1)
  const double* yc;     // node y coordinates (ny+1)
  const float* elem_t;  // nullptr = uniform t0
  double t0, Dfac, Gk, nu;  // D = Dfac t^3, Ds = Gk t
  const uint8_t* cmask;     // PP_FIX_* bits per node
};

template <int TX_, int TY_, int NT_>
struct Tile {
  static constexpr int TX = TX_, TY = TY_, NT = NT_;
  static constexpr int PW = TX + 2, PH = TY + 2, NP = PW * PH;  // staged nodes
  static constexpr int EW = TX + 1, EH = TY + 1, NE = EW * EH;  // elements
  static_assert(TX * TY <= NT, "one node per thread");
};

// Tile of workgroup b in a launch of G workgroups, one per tile.  Workgroup
// b runs on XCD b % 8 (the dispatcher deals workgroups to the eight XCDs in
// turn), so XCD x is given the contiguous row-major tiles
// [x (G / 8) + min(x, G % 8), ...): horizontally adjacent tiles, and most
// vertically adjacent ones, then share one L2 for their halo rows and for the
// partial 128-byte lines at the ends of every staged row.  A bijection of
// [0, G), so per-workgroup partials keep their slots.
__device__ __forceinline__ int xcd_band_tile(int b, int G) {
  const int x = b & 7, q = G >> 3, r = G & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Tiles of workgroup me among nwg strided workgroups.  Workgroup me runs on
// XCD me % 8 (the dispatcher deals workgroups to the eight XCDs in turn), so
// the tiles are cut into eight row-major bands, one per XCD, and the
// workgroups of an XCD take their band's tiles in turn: neighbouring tiles,
// which share halo rows and the partial 128-byte lines at the ends of every
// staged row, are then fetched into one L2.
struct TileBand {
  int first, last, step;
};
__device__ __forceinline__ TileBand xcd_tile_band(unsigned me, unsigned nwg, int ntiles) {
  if (nwg < 8u) return TileBand{(int)me, ntiles, (int)nwg};  // not every XCD has a workgroup
  const int x = (int)(me & 7u), rank = (int)(me >> 3);
  const int on_x = (int)(nwg >> 3) + (x < (int)(nwg & 7u) ? 1 : 0);  // workgroups on XCD x
  const int lo = (int)((int64_t)ntiles * x / 8), hi = (int)((int64_t)ntiles * (x + 1) / 8);
  return TileBand{lo + rank, hi, on_x};
}

template <class C>
__device__ __forceinline__ int tiles_x(const KGrid& G) {
  return (G.nx + 1 + C::TX - 1) / C::TX;
}
template <class C>
__device__ __forceinline__ int tiles_total(const KGrid& G) {
  return tiles_x<C>(G) * ((G.ny + 1 + C::TY - 1) / C::TY);
}

// Load of data other workgroups wrote during a persistent kernel (COH:
// device-coherent ld_coh) or written before the kernel started (plain).
template <bool COH, typename S>
__device__ __forceinline__ S ld_in(const S* p) {
  if constexpr (COH) return ld_coh(p);
  else return *p;
}

// Stage p = src + beta * pold (pold == nullptr: p = src) for the tile plus
// halo into LDS, SoA p_s[c * NP + idx]; outside the grid p = 0.
template <class C, typename T, typename S, bool COH = false>
__device__ __forceinline__ void stage_tile(const KGrid& G, int i0, int j0, const S* src,
                                           const S* pold, T beta, T* p_s) {
  for (int idx = threadIdx.x; idx < C::NP; idx += C::NT) {
    const int lx = idx % C::PW, ly = idx / C::PW;
    const int gi = i0 - 1 + lx, gj = j0 - 1 + ly;
    T v0 = 0, v1 = 0, v2 = 0;
    if (gi >= 0 && gi <= G.nx && gj >= 0 && gj <= G.ny) {
      const size_t n = (size_t)gj * (G.nx + 1) + gi;
      v0 = (T)ld_in<COH>(src + 3 * n);
      v1 = (T)ld_in<COH>(src + 3 * n + 1);
      v2 = (T)ld_in<COH>(src + 3 * n + 2);
      if (pold) {
        v0 += beta * (T)ld_in<COH>(pold + 3 * n);
        v1 += beta * (T)ld_in<COH>(pold + 3 * n + 1);
        v2 += beta * (T)ld_in<COH>(pold + 3 * n + 2);
      }
    }
    p_s[idx] = v0;
    p_s[C::NP + idx] = v1;
    p_s[2 * C::NP + idx] = v2;
  }
}

// Material of element (gi, gj).
__device__ __forceinline__ void elem_material(const KGrid& G, int gi, int gj, double& D,
                                              double& Ds) {
  const double t = G.elem_t ? (double)G.elem_t[(size_t)gj * G.nx + gi] : G.t0;
  D = G.Dfac * t * t * t;
  Ds = G.Gk * t;
}

// Apply every element touching the tile; results to ye_s[k * NE + le].
template <class C, typename T>
__device__ __forceinline__ void tile_elements(const KGrid& G, int i0, int j0, const T* p_s,
                                              T* ye_s) {
  for (int le = threadIdx.x; le < C::NE; le += C::NT) {
    const int ex = le % C::EW, ey = le / C::EW;
    const int gi = i0 - 1 + ex, gj = j0 - 1 + ey;
    T y[12];
    if (gi >= 0 && gi < G.nx && gj >= 0 && gj < G.ny) {
      const int c0 = ey * C::PW + ex, c1 = c0 + 1, c2 = c0 + C::PW + 1, c3 = c0 + C::PW;
      constexpr int NP = C::NP;
      const T u[12] = {p_s[c0], p_s[NP + c0], p_s[2 * NP + c0], p_s[c1], p_s[NP + c1],
                       p_s[2 * NP + c1], p_s[c2], p_s[NP + c2], p_s[2 * NP + c2],
                       p_s[c3], p_s[NP + c3], p_s[2 * NP + c3]};
      double D, Ds;
      elem_material(G, gi, gj, D, Ds);
      q4_rect_apply_fast<T>((T)G.ax[2 * gi], (T)G.by[2 * gj], (T)G.ax[2 * gi + 1],
                            (T)G.by[2 * gj + 1], (T)D, (T)(D * G.nu),
                            (T)(D * (1.0 - G.nu) / 2.0), (T)Ds, u, y);
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) y[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < 12; k++) ye_s[k * C::NE + le] = y[k];
  }
}

// Sum of the four elements around node (tx, ty) of the tile, fixed order.
template <class C, typename T>
__device__ __forceinline__ void node_sum(int tx, int ty, const T* ye_s, T (&y)[3]) {
  constexpr int EW = C::EW, NE = C::NE;
  const int sw = ty * EW + tx, se = sw + 1, ne = sw + EW + 1, nw = sw + EW;
#pragma unroll
  for (int d = 0; d < 3; d++)
    y[d] = ye_s[(6 + d) * NE + sw] + ye_s[(9 + d) * NE + se] + ye_s[d * NE + ne] +
           ye_s[(3 + d) * NE + nw];
}

KGrid make_kgrid(const pp_plan* pl);

}  // namespace pp
