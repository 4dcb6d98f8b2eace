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
#include "taps.h"

namespace {

constexpr int DEGRID_NT = 512;

template <int NT, bool SEPARABLE, int R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(R == 63 ? 8 : 4))) void k_degrid_tiles(
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
      __syncthreads();  // every wave has read s_item before the next queue rewrites it
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
        gather_range<SEPARABLE, R != 63>(sg, base, R, S, 0, S * S, lane, kp, ky, kx, ar, ai);
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
  tap_timer_reset(ctx, 1);
  int rc = tap_timer_begin(ctx, 1);
  if (rc) return rc;
  const float2 *g = (const float2 *)grid;
  const float2 *kt = (const float2 *)kern;
  float2 *vo = (float2 *)vis_out;
  if (plan.R == TILE_R_SMALL) {
    const int G = persistent_blocks(ctx, 4);  // four fit by registers (32 waves)
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
  if ((rc = tap_timer_end(ctx, 1))) return rc;
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
            plane_mode, vplane, wstep, wbin0, max_support};
  BinPlan plan;
  rc = bin_visibilities(ctx, ctx->ws, a, max_support, ITEMS_BY_GROUP, &plan);
  if (rc) return rc;
  rc = degrid_binned(ctx, plan, grid, N, vis_out, kern, separable);
  if (rc) return rc;
  rc = read_status(ctx, plan);
  return rc;
}
