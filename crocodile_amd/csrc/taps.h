// Kernel-window tap walks shared by the tile kernels (grid.hip,
// degrid.hip) and the w-towers job kernels (wtowers.hip).
//
// A visibility's S x S window is walked in tap order t (row i = t / S,
// column t - i*S) over a range [tlo, thi) of whole rows.  Its cells lie in a
// row-major LDS region with row stride RS cells: cell cbase + t + i*(RS - S).
// Rows may carry `pad` extra bytes (a row pitch that is not a multiple of the
// LDS bank width keeps the rows a wavefront touches on different banks);
// `extra` is then the padding before the window's first row.
// Lanes take consecutive taps; blocks of 64*NQ taps issue all NQ kernel
// loads before the first use.  Blocks that lie wholly inside the range run
// with no clamps and no masks; the last, partial block clamps its load
// indices to the last tap (a valid address, value unused) and masks the
// rest.  Kernel tables are 2D [S][S] (k2) or separable (ky[i] * kx[j]).
#pragma once

#include "internal.h"

// Gridding: adds v * k of every tap to an int64 fixed-point region, re and
// im interleaved (cell c at bytes 16c, 16c + 8).  cb = 16 * (cbase + t0 +
// lane), rs16 = 16 * (RS - S), kt = 8 * (t0 + lane); vr, vi = the
// visibility times the fixed-point scale.
template <int NQ, bool FULL, bool SEPARABLE>
__device__ __forceinline__ void spread_block(
    unsigned long long *acc, uint32_t cb, uint32_t rs16, uint32_t kt, int S, int thi,
    uint32_t m, int t0, int lane, const float2 *k2, const float2 *ky, const float2 *kx,
    double vr, double vi) {
  float2 kv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int t = FULL ? t0 + q * 64 + lane : min(t0 + q * 64 + lane, thi - 1);
    if (!SEPARABLE) {
      kv[q] = FULL ? ldk(k2 + q * 64, kt) : ldk(k2, (uint32_t)t * 8u);
    } else {
      const int i = tap_row(t, m);
      kv[q] = cmul(ldk(ky, (uint32_t)i * 8u), ldk(kx, (uint32_t)(t - i * S) * 8u));
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    if (!FULL && t0 + q * 64 >= thi) break;  // wave-uniform
    const int t = t0 + q * 64 + lane;
    if (FULL || t < thi) {
      const uint32_t a = cb + __umul24((uint32_t)tap_row(t, m), rs16) + 1024u * q;
      unsigned long long *cell = (unsigned long long *)((char *)acc + a);
      const double kr = kv[q].x, ki = kv[q].y;
      atomicAdd(cell, fixed2(vr, kr, -vi, ki));
      atomicAdd(cell + 1, fixed2(vr, ki, vi, kr));
    }
  }
}

// Taps [tlo, thi) of one visibility: unmasked blocks of 512 and 256 taps,
// then at most one masked block of the remaining < 256.
template <bool SEPARABLE>
__device__ __forceinline__ void spread_range(
    unsigned long long *acc, int cbase, int RS, int S, int tlo, int thi, int lane,
    const float2 *k2, const float2 *ky, const float2 *kx, double vr, double vi,
    uint32_t pad = 0u, uint32_t extra = 0u) {
  const uint32_t m = __builtin_amdgcn_readfirstlane(tap_div(S));
  const uint32_t rs16 = 16u * (uint32_t)(RS - S) + pad;
  const uint32_t cb0 = 16u * (uint32_t)(cbase + lane) + extra, kt0 = 8u * (uint32_t)lane;
  int t0 = tlo;
  for (; t0 + 512 <= thi; t0 += 512)
    spread_block<8, true, SEPARABLE>(acc, cb0 + 16u * t0, rs16, kt0 + 8u * t0, S, thi, m, t0,
                                     lane, k2, ky, kx, vr, vi);
  if (t0 + 256 <= thi) {
    spread_block<4, true, SEPARABLE>(acc, cb0 + 16u * t0, rs16, kt0 + 8u * t0, S, thi, m, t0,
                                     lane, k2, ky, kx, vr, vi);
    t0 += 256;
  }
  if (t0 < thi)
    spread_block<4, false, SEPARABLE>(acc, cb0 + 16u * t0, rs16, kt0 + 8u * t0, S, thi, m, t0,
                                      lane, k2, ky, kx, vr, vi);
}

// Degridding: per-lane partial sums of g * k over a complex64 region.
// cb = 8 * (cbase + t0 + lane), rs8 = 8 * (RS - S), kt = 8 * (t0 + lane).
template <int NQ, bool FULL, bool SEPARABLE>
__device__ __forceinline__ void gather_block(
    const float2 *sg, uint32_t cb, uint32_t rs8, uint32_t kt, int S, int thi, uint32_t m,
    int t0, int lane, const float2 *k2, const float2 *ky, const float2 *kx, float &ar,
    float &ai) {
  float2 kv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int t = FULL ? t0 + q * 64 + lane : min(t0 + q * 64 + lane, thi - 1);
    if (!SEPARABLE) {
      kv[q] = FULL ? ldk(k2 + q * 64, kt) : ldk(k2, (uint32_t)t * 8u);
    } else {
      const int i = tap_row(t, m);
      kv[q] = cmul(ldk(ky, (uint32_t)i * 8u), ldk(kx, (uint32_t)(t - i * S) * 8u));
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    if (!FULL && t0 + q * 64 >= thi) break;  // wave-uniform
    const int t = t0 + q * 64 + lane;
    if (FULL || t < thi) {
      const uint32_t a = cb + __umul24((uint32_t)tap_row(t, m), rs8) + 512u * q;
      const float2 g = *(const float2 *)((const char *)sg + a);
      ar = fmaf(g.x, kv[q].x, fmaf(-g.y, kv[q].y, ar));
      ai = fmaf(g.x, kv[q].y, fmaf(g.y, kv[q].x, ai));
    }
  }
}

// Taps [tlo, thi) of one visibility: unmasked blocks of 1024 (when
// BIG_BLOCKS), 512 and 256 taps, then at most one masked block of the
// remaining < 256.
template <bool SEPARABLE, bool BIG_BLOCKS = true>
__device__ __forceinline__ void gather_range(
    const float2 *sg, int cbase, int RS, int S, int tlo, int thi, int lane, const float2 *k2,
    const float2 *ky, const float2 *kx, float &ar, float &ai, uint32_t pad = 0u,
    uint32_t extra = 0u) {
  const uint32_t m = __builtin_amdgcn_readfirstlane(tap_div(S));
  const uint32_t rs8 = 8u * (uint32_t)(RS - S) + pad;
  const uint32_t cb0 = 8u * (uint32_t)(cbase + lane) + extra, kt0 = 8u * (uint32_t)lane;
  int t0 = tlo;
  if (BIG_BLOCKS)
    for (; t0 + 1024 <= thi; t0 += 1024)
      gather_block<16, true, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, thi, m, t0,
                                        lane, k2, ky, kx, ar, ai);
  for (; t0 + 512 <= thi; t0 += 512)
    gather_block<8, true, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, thi, m, t0,
                                     lane, k2, ky, kx, ar, ai);
  if (t0 + 256 <= thi) {
    gather_block<4, true, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, thi, m, t0,
                                     lane, k2, ky, kx, ar, ai);
    t0 += 256;
  }
  if (t0 < thi)
    gather_block<4, false, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, thi, m, t0,
                                      lane, k2, ky, kx, ar, ai);
}

// x + (x moved across lanes by the DPP control CTRL); rows outside ROW_MASK
// add 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_add(float x) {
  const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, false);
  return x + __int_as_float(y);
}

// Sum over the 64 lanes of a wavefront, left in lane 63: butterflies within
// rows of 16 lanes (quad permutes, half-row and row mirrors), then the row
// broadcasts of lanes 15 and 31.  VALU data movement only (DPP); no LDS
// round trips on the critical path.
__device__ __forceinline__ float wave_sum_to_63(float x) {
  x = dpp_add<0xb1>(x);        // quad_perm [1, 0, 3, 2]
  x = dpp_add<0x4e>(x);        // quad_perm [2, 3, 0, 1]
  x = dpp_add<0x141>(x);       // row_half_mirror
  x = dpp_add<0x140>(x);       // row_mirror
  x = dpp_add<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  x = dpp_add<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

// Visibility records reach a wavefront 64 at a time: one vector load per
// lane (lane l holds the wave's l-th next record), handed out with
// v_readlane.  A scalar-cache load per visibility would instead make the
// wave wait (lgkmcnt) for its own outstanding LDS operations before every
// visibility.
__device__ __forceinline__ uint4 lane_u4(uint4 x, int l) {
  return make_uint4(__builtin_amdgcn_readlane(x.x, l), __builtin_amdgcn_readlane(x.y, l),
                    __builtin_amdgcn_readlane(x.z, l), __builtin_amdgcn_readlane(x.w, l));
}
__device__ __forceinline__ float2 lane_f2(float2 x, int l) {
  return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(x.x), l)),
                     __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x.y), l)));
}
