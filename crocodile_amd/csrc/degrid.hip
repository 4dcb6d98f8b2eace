// Convolutional degridding: crocodile/synthesis.py:484-503 (convdegrid) on
// CDNA4.
//
// Work item = up to `chunk` visibilities of one (tile, plane-group) bin;
// items are queued group-major, so at any time the whole chip reads kernel
// taps from one L2-sized slice of the kernel bank.  A workgroup stages the
// 95 x 95 cell grid region of its tile in LDS (coalesced row reads), then
// each wavefront takes one visibility at a time: lanes walk the S*S taps,
// gather grid cells from LDS and kernel taps from the kernel bank,
// accumulate complex products in FP32 and the wavefront reduces the 64
// partial sums with cross-lane shuffles.
#include "internal.h"

namespace {

constexpr int DEGRID_NT = 512;

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Partial sums of one visibility over a block of up to 64*NQ taps t0 .. of
// its S x S window (tap t = row i, column t - i*S); loads unpredicated
// (clamped index) and issued before first use.
template <int NQ, bool SEPARABLE, int R>
__device__ __forceinline__ void gather_taps(
    const float2 *sg, int base, int S, int SS, float invS, int t0, int lane,
    const float2 *k2, const float2 *ky, const float2 *kx, float &ar, float &ai) {
  float2 kv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int t = min(t0 + q * 64 + lane, SS - 1);
    if (!SEPARABLE) {
      kv[q] = k2[t];
    } else {
      const int i = (int)(((float)t + 0.5f) * invS);
      kv[q] = cmul(ky[i], kx[t - i * S]);
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int t = t0 + q * 64 + lane;
    if (t < SS) {
      const int i = (int)(((float)t + 0.5f) * invS);
      const float2 g = sg[base + i * (R - S) + t];
      ar += g.x * kv[q].x - g.y * kv[q].y;
      ai += g.x * kv[q].y + g.y * kv[q].x;
    }
  }
}

template <int NT, bool SEPARABLE, int R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_degrid_tiles(
    const float2 *__restrict__ grid, int N, int T, int ntx,
    const uint4 *__restrict__ rec, float2 *__restrict__ vis_out,
    const float2 *__restrict__ kern, const int4 *__restrict__ items,
    const uint32_t *__restrict__ nitems, uint32_t *__restrict__ qhead) {
  constexpr int CELLS = R * R;
  constexpr int NW = NT / 64;
  __shared__ float2 sg[CELLS];
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nit = (int)*nitems;

  for (;;) {
    if (tid == 0) s_item = (int)atomicAdd(qhead, 1u);
    __syncthreads();
    const int it = s_item;
    if (it >= nit) break;
    const int4 item = items[it];
    const int tile = __builtin_amdgcn_readfirstlane(item.x);
    const int tx0 = (tile % ntx) * T, ty0 = (tile / ntx) * T;
    for (int c = tid; c < CELLS; c += NT) {
      const int cy = c / R, cx = c - cy * R;
      const int gy = ty0 + cy, gx = tx0 + cx;
      sg[c] = (gy < N && gx < N) ? grid[(int64_t)gy * N + gx] : make_float2(0.f, 0.f);
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
      const int SS = S * S;
      const float invS = 1.0f / (float)S;
      const float2 *kp = kern + r.z;  // 2D table of (yf, xf) / separable [Q][S]
      const float2 *ky = kp + yf * S, *kx = kp + xf * S;
      float ar = 0.f, ai = 0.f;
      const int nq = (SS + 63) >> 6;
      if (nq <= 4) {
        gather_taps<4, SEPARABLE, R>(sg, base, S, SS, invS, 0, lane, kp, ky, kx, ar, ai);
      } else if (nq <= 8) {
        gather_taps<8, SEPARABLE, R>(sg, base, S, SS, invS, 0, lane, kp, ky, kx, ar, ai);
      } else {
        for (int t0 = 0; t0 < SS; t0 += 64 * 16)
          gather_taps<16, SEPARABLE, R>(sg, base, S, SS, invS, t0, lane, kp, ky, kx, ar, ai);
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) vis_out[r.w] = make_float2(ar, ai);
    }
    __syncthreads();
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
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items, plan.nitems, plan.qhead);
    else
      k_degrid_tiles<DEGRID_NT, false, TILE_R_SMALL><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items, plan.nitems, plan.qhead);
  } else {
    const int G = persistent_blocks(ctx, 2);
    if (separable)
      k_degrid_tiles<DEGRID_NT, true, TILE_R><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items, plan.nitems, plan.qhead);
    else
      k_degrid_tiles<DEGRID_NT, false, TILE_R><<<G, DEGRID_NT, 0, ctx->stream>>>(
          g, N, plan.T, plan.ntx, plan.rec, vo, kt, plan.items, plan.nitems, plan.qhead);
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
