// Convolutional degridding: crocodile/synthesis.py:484-503 (convdegrid) on
// CDNA4.
//
// Work item = up to `chunk` visibilities of one tile whose kernel tables lie
// in one range of plane groups.  There is one queue per XCD, each holding a
// contiguous range of the kernel bank, so that an XCD's L2 serves its own
// slice of a bank far larger than one L2.  A workgroup stages the R x R cell
// grid region of its tile in LDS (R = 63 or 95; coalesced, streaming row
// reads), then each wavefront takes one visibility at a time: lanes walk the
// S*S taps, gather grid cells from LDS and kernel taps from the kernel bank,
// accumulate complex products in FP32, and the wavefront sums the 64
// partial sums with DPP row operations.
#include "internal.h"

namespace {

constexpr int DEGRID_NT = 512;

// x + (x moved across lanes by the DPP control `ctrl`); rows outside
// row_mask add 0.
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

// Partial sums of one visibility over the block of 64*NQ taps t0 .. of its
// S x S window (tap t = row i = t / S, column t - i*S; cell base + t +
// i*(R - S)).  Kernel loads are issued before first use.  FULL: every tap
// of the block lies in the window (no clamps, no masks); otherwise load
// indices are clamped to the last tap and the products masked.  cb = 8 *
// (base + t0 + lane) is the lane's LDS byte address, rs8 = 8 * (R - S),
// kt = 8 * (t0 + lane).
template <int NQ, bool FULL, bool SEPARABLE>
__device__ __forceinline__ void gather_block(
    const float2 *sg, uint32_t cb, uint32_t rs8, uint32_t kt, int S, int SS, uint32_t m,
    int t0, int lane, const float2 *k2, const float2 *ky, const float2 *kx, float &ar,
    float &ai) {
  float2 kv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int t = FULL ? t0 + q * 64 + lane : min(t0 + q * 64 + lane, SS - 1);
    if (!SEPARABLE) {
      kv[q] = FULL ? ldk(k2 + q * 64, kt) : ldk(k2, (uint32_t)t * 8u);
    } else {
      const int i = tap_row(t, m);
      kv[q] = cmul(ldk(ky, (uint32_t)i * 8u), ldk(kx, (uint32_t)(t - i * S) * 8u));
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    if (!FULL && t0 + q * 64 >= SS) break;  // wave-uniform
    const int t = t0 + q * 64 + lane;
    if (FULL || t < SS) {
      const uint32_t a = cb + __umul24((uint32_t)tap_row(t, m), rs8) + 512u * q;
      const float2 g = *(const float2 *)((const char *)sg + a);
      ar = fmaf(g.x, kv[q].x, fmaf(-g.y, kv[q].y, ar));
      ai = fmaf(g.x, kv[q].y, fmaf(g.y, kv[q].x, ai));
    }
  }
}

// All taps of one visibility: unmasked blocks of 1024, 512 and 256 taps,
// then at most one masked block of the remaining < 256 taps.
template <bool SEPARABLE, int R>
__device__ __forceinline__ void gather_window(
    const float2 *sg, int base, int S, int lane, const float2 *k2, const float2 *ky,
    const float2 *kx, float &ar, float &ai) {
  const int SS = S * S;
  const uint32_t m = __builtin_amdgcn_readfirstlane(tap_div(S));
  const uint32_t rs8 = 8u * (uint32_t)(R - S);
  const uint32_t cb0 = 8u * (uint32_t)(base + lane), kt0 = 8u * (uint32_t)lane;
  int t0 = 0;
  for (; t0 + 1024 <= SS; t0 += 1024)
    gather_block<16, true, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, SS, m, t0,
                                      lane, k2, ky, kx, ar, ai);
  if (t0 + 512 <= SS) {
    gather_block<8, true, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, SS, m, t0,
                                     lane, k2, ky, kx, ar, ai);
    t0 += 512;
  }
  if (t0 + 256 <= SS) {
    gather_block<4, true, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, SS, m, t0,
                                     lane, k2, ky, kx, ar, ai);
    t0 += 256;
  }
  if (t0 < SS)
    gather_block<4, false, SEPARABLE>(sg, cb0 + 8u * t0, rs8, kt0 + 8u * t0, S, SS, m, t0,
                                      lane, k2, ky, kx, ar, ai);
}

// Items are listed group-major; the groups are cut into nq contiguous
// ranges, queue x holding the items of range x.  A workgroup drains the
// queue of its own XCD first, then helps with the others in turn, so every
// XCD's L2 mostly serves one slice of the kernel bank.
template <int NT, bool SEPARABLE, int R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_degrid_tiles(
    const float2 *__restrict__ grid, int N, int T, int ntx,
    const uint4 *__restrict__ rec, float2 *__restrict__ vis_out,
    const float2 *__restrict__ kern, const int4 *__restrict__ items,
    const uint32_t *__restrict__ gstart, int ngroups, uint32_t *__restrict__ qhead) {
  constexpr int CELLS = R * R;
  constexpr int NW = NT / 64;
  __shared__ float2 sg[CELLS];
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nq = min(XCD_QUEUES, ngroups);
  const int x0 = xcd_id() % nq;

  for (int k = 0; k < nq; k++) {
    const int x = (x0 + k) % nq;
    const int lo = (int)gstart[(int64_t)x * ngroups / nq];
    const int hi = (int)gstart[(int64_t)(x + 1) * ngroups / nq];
    for (;;) {
      if (tid == 0) s_item = lo + (int)atomicAdd(qhead + x, 1u);
      __syncthreads();
      const int it = s_item;
      if (it >= hi) break;
      const int4 item = items[it];
      const int tile = __builtin_amdgcn_readfirstlane(item.x);
      const int tx0 = (tile % ntx) * T, ty0 = (tile / ntx) * T;
      // Stage the region, eight loads in flight per thread.
      constexpr int PER = (CELLS + NT - 1) / NT;
      for (int c0 = 0; c0 < PER; c0 += 8) {
        float2 gv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int c = (c0 + k) * NT + tid;
          const int cy = c / R, cx = c - cy * R;
          const int gy = ty0 + cy, gx = tx0 + cx;
          gv[k] = (c < CELLS && gy < N && gx < N) ? load_stream(grid + (int64_t)gy * N + gx)
                                                  : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int c = (c0 + k) * NT + tid;
          if (c < CELLS) sg[c] = gv[k];
        }
      }
      __syncthreads();

      // Each wave walks its visibilities (j = first + wave, + NW, ...) with
      // the next record already in flight.
      const int jend = __builtin_amdgcn_readfirstlane(item.z);
      int j = __builtin_amdgcn_readfirstlane(item.y) + wave;
      uint4 rn = make_uint4(0, 0, 0, 0);
      if (j < jend) rn = rec[j];
      for (; j < jend; j += NW) {
        const uint4 r = uniform_u4(rn);
        if (j + NW < jend) rn = rec[j + NW];
        const int S = (int)(r.y & 0xffu);
        const int xf = (int)((r.y >> 8) & 0xfffu), yf = (int)(r.y >> 20);
        const int base = ((int)(r.x >> 16) - ty0) * R + ((int)(r.x & 0xffffu) - tx0);
        const float2 *kp = kern + r.z;  // 2D table of (yf, xf) / separable [Q][S]
        const float2 *ky = kp + yf * S, *kx = kp + xf * S;
        float ar = 0.f, ai = 0.f;
        gather_window<SEPARABLE, R>(sg, base, S, lane, kp, ky, kx, ar, ai);
        ar = wave_sum_to_63(ar);
        ai = wave_sum_to_63(ai);
        if (lane == 63) store_stream(vis_out + r.w, make_float2(ar, ai));
      }
      __syncthreads();
    }
    __syncthreads();  // every thread has read the last s_item of this queue
  }
}

}  // namespace

// Degridding of binned visibilities from grid into vis_out.  No status read.
int degrid_binned(croc_ctx *ctx, const BinPlan &plan, const void *grid, int N, void *vis_out,
                  const void *kern, int separable) {
  CROC_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
  const float2 *g = (const float2 *)grid;
  const float2 *kt = (const float2 *)kern;
  float2 *vo = (float2 *)vis_out;
  if (plan.R == TILE_R_SMALL) {
    const int G = persistent_blocks(ctx, 2);  // two fit by registers (16 waves)
    if (separable)
      k_degrid_tiles<DEGRID_NT, true, TILE_R_SMALL><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items_g, plan.gstart, plan.ngroups, plan.qhead);
    else
      k_degrid_tiles<DEGRID_NT, false, TILE_R_SMALL><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items_g, plan.gstart, plan.ngroups, plan.qhead);
  } else {
    const int G = persistent_blocks(ctx, 2);
    if (separable)
      k_degrid_tiles<DEGRID_NT, true, TILE_R><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items_g, plan.gstart, plan.ngroups, plan.qhead);
    else
      k_degrid_tiles<DEGRID_NT, false, TILE_R><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items_g, plan.gstart, plan.ngroups, plan.qhead);
  }
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
  ctx->ev_degrid = true;
  return CROC_OK;
}

extern "C" int croc_convdegrid(croc_ctx *ctx, const void *grid, int N,
                               const double *uvw, int64_t nv, int64_t stride,
                               double lam, void *vis_out, const void *kern,
                               int Q, int separable, int nplanes,
                               const int64_t *plane_off,
                               const int32_t *plane_s, int max_support,
                               int plane_mode, const int32_t *vplane,
                               double wstep, int wbin0) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid && kern && plane_off && plane_s, CROC_EINVAL,
               "null array argument");
  CROC_REQUIRE(nv == 0 || (uvw && vis_out), CROC_EINVAL, "null uvw / vis_out");
  CROC_REQUIRE(stride >= 2, CROC_EINVAL, "uvw stride must be >= 2");
  CROC_REQUIRE(plane_mode != CROC_PLANE_INDEX || vplane, CROC_EINVAL,
               "CROC_PLANE_INDEX needs vplane");
  CROC_REQUIRE(plane_mode != CROC_PLANE_WBIN || (stride >= 3 && wstep > 0),
               CROC_EINVAL, "CROC_PLANE_WBIN needs a w column and wstep > 0");
  CROC_REQUIRE(lam > 0, CROC_EINVAL, "lam must be positive");
  if (nv == 0) return CROC_OK;

  BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_off, separable,
            plane_mode, vplane, wstep, wbin0, nullptr, max_support};
  BinPlan plan;
  rc = bin_visibilities(ctx, ctx->ws, a, max_support, ITEMS_BY_GROUP, &plan);
  if (rc) return rc;
  rc = degrid_binned(ctx, plan, grid, N, vis_out, kern, separable);
  if (rc) return rc;
  rc = read_status(ctx, plan);
  return rc;
}
