// Plain per-visibility gridding and degridding: crocodile/synthesis.py:464-
// 503 (convgrid / convdegrid) as a direct loop over visibilities on the
// GPU, the baseline the w-towers pipeline is compared with.  One wavefront
// per visibility, its lanes walking the S*S taps: gridding adds every
// product to the caller's complex64 grid with FP32 atomics (so its rounding
// depends on the order of the atomics), degridding gathers the window from
// the grid and sums across the wavefront.  No binning, no LDS
// privatisation.  All coordinates are validated by a first pass, so that an
// error leaves the outputs unmodified.
#include "locate.h"
#include "taps.h"

namespace {

constexpr int PV_NT = 256;

__device__ __forceinline__ uint4 pv_locate(const BinArgs &a, int64_t i) {
  uint32_t st = 0;
  const BinGeom g{a.N, 1, 1, 1};  // one tile, one group
  return locate(a, g, i, &st);
}

__global__ void k_pv_check(BinArgs a, uint32_t *__restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nv) return;
  uint32_t st = 0;
  const BinGeom g{a.N, 1, 1, 1};
  locate(a, g, i, &st);
  if (st) atomicOr(status, st);
}

template <bool SEPARABLE>
__global__ __launch_bounds__(PV_NT) void k_pv_grid(BinArgs a, float2 *__restrict__ grid,
                                                  const float2 *__restrict__ vis,
                                                  const float2 *__restrict__ kern) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * PV_NT) >> 6;
  for (int64_t i = ((int64_t)blockIdx.x * PV_NT + threadIdx.x) >> 6; i < a.nv; i += nw) {
    const uint4 L = pv_locate(a, i);
    const int x0 = (int)(L.x & 0xffffu), y0 = (int)(L.x >> 16);
    const int S = (int)(L.y & 0xffu), xf = (int)((L.y >> 8) & 0xfffu), yf = (int)(L.y >> 20);
    const float2 v = vis[i];
    const float2 *kp = kern + L.z, *ky = kp + yf * S, *kx = kp + xf * S;
    const uint32_t m = tap_div(S);
    for (int t = lane; t < S * S; t += 64) {
      const int r = tap_row(t, m), c = t - r * S;
      const float2 k = SEPARABLE ? cmul(ky[r], kx[c]) : kp[t];
      float2 *cell = grid + (int64_t)(y0 + r) * a.N + (x0 + c);
      atomicAdd(&cell->x, v.x * k.x - v.y * k.y);
      atomicAdd(&cell->y, v.x * k.y + v.y * k.x);
    }
  }
}

template <bool SEPARABLE>
__global__ __launch_bounds__(PV_NT) void k_pv_degrid(BinArgs a, const float2 *__restrict__ grid,
                                                    float2 *__restrict__ vis_out,
                                                    const float2 *__restrict__ kern) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * PV_NT) >> 6;
  for (int64_t i = ((int64_t)blockIdx.x * PV_NT + threadIdx.x) >> 6; i < a.nv; i += nw) {
    const uint4 L = pv_locate(a, i);
    const int x0 = (int)(L.x & 0xffffu), y0 = (int)(L.x >> 16);
    const int S = (int)(L.y & 0xffu), xf = (int)((L.y >> 8) & 0xfffu), yf = (int)(L.y >> 20);
    const float2 *kp = kern + L.z, *ky = kp + yf * S, *kx = kp + xf * S;
    const uint32_t m = tap_div(S);
    float ar = 0.f, ai = 0.f;
    for (int t = lane; t < S * S; t += 64) {
      const int r = tap_row(t, m), c = t - r * S;
      const float2 k = SEPARABLE ? cmul(ky[r], kx[c]) : kp[t];
      const float2 g = grid[(int64_t)(y0 + r) * a.N + (x0 + c)];
      ar = fmaf(g.x, k.x, fmaf(-g.y, k.y, ar));
      ai = fmaf(g.x, k.y, fmaf(g.y, k.x, ai));
    }
    ar = wave_sum_to_63(ar);
    ai = wave_sum_to_63(ai);
    if (lane == 63) vis_out[i] = make_float2(ar, ai);
  }
}

int pv_enter(croc_ctx *ctx, const void *grid, const double *uvw, int64_t nv, int64_t stride,
             const void *vdata, const void *kern, const int64_t *plane_off,
             const int32_t *plane_s, int plane_mode, const int32_t *vplane, double wstep,
             double lam, int max_support) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(max_support >= 1 && max_support <= 64, CROC_EINVAL,
               "kernel support must lie in [1, 64]");
  CROC_REQUIRE(grid && kern && plane_off && plane_s, CROC_EINVAL, "null array argument");
  CROC_REQUIRE(nv == 0 || (uvw && vdata), CROC_EINVAL, "null uvw / visibility array");
  CROC_REQUIRE(stride >= 2, CROC_EINVAL, "uvw stride must be >= 2");
  CROC_REQUIRE(plane_mode != CROC_PLANE_INDEX || vplane, CROC_EINVAL,
               "CROC_PLANE_INDEX needs vplane");
  CROC_REQUIRE(plane_mode != CROC_PLANE_WBIN || (stride >= 3 && wstep > 0), CROC_EINVAL,
               "CROC_PLANE_WBIN needs a w column and wstep > 0");
  CROC_REQUIRE(lam > 0, CROC_EINVAL, "lam must be positive");
  return CROC_OK;
}

int pv_check(croc_ctx *ctx, const BinArgs &a) {
  uint32_t *st = (uint32_t *)ws_get(ctx, WS_SMALL, 256);
  if (!st) return CROC_ENOMEM;
  int rc = zero_u32_async(ctx, st, 1);
  if (rc) return rc;
  k_pv_check<<<(unsigned)((a.nv + PV_NT - 1) / PV_NT), PV_NT, 0, ctx->stream>>>(a, st);
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipMemcpyAsync(ctx->host_status, st, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return croc_status_to_code(*ctx->host_status);
}

}  // namespace

extern "C" int croc_convgrid_pervis(croc_ctx *ctx, void *grid, int N, const double *uvw,
                                    int64_t nv, int64_t stride, double lam, const void *vis,
                                    const void *kern, int Q, int separable, int nplanes,
                                    const int64_t *plane_off, const int32_t *plane_s,
                                    int max_support, int plane_mode, const int32_t *vplane,
                                    double wstep, int wbin0) {
  int rc = pv_enter(ctx, grid, uvw, nv, stride, vis, kern, plane_off, plane_s, plane_mode,
                    vplane, wstep, lam, max_support);
  if (rc) return rc;
  if (nv == 0) return CROC_OK;
  const BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_off, separable,
                  plane_mode, vplane, wstep, wbin0, max_support};
  if ((rc = pv_check(ctx, a))) return rc;
  const int G = persistent_blocks(ctx, 8);
  CROC_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  tap_timer_reset(ctx, 0);
  if ((rc = tap_timer_begin(ctx, 0))) return rc;
  if (separable)
    k_pv_grid<true><<<G, PV_NT, 0, ctx->stream>>>(a, (float2 *)grid, (const float2 *)vis,
                                                  (const float2 *)kern);
  else
    k_pv_grid<false><<<G, PV_NT, 0, ctx->stream>>>(a, (float2 *)grid, (const float2 *)vis,
                                                   (const float2 *)kern);
  CROC_HIP(hipGetLastError());
  if ((rc = tap_timer_end(ctx, 0))) return rc;
  CROC_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  ctx->ev_grid = true;
  return CROC_OK;
}

extern "C" int croc_convdegrid_pervis(croc_ctx *ctx, const void *grid, int N,
                                      const double *uvw, int64_t nv, int64_t stride,
                                      double lam, void *vis_out, const void *kern, int Q,
                                      int separable, int nplanes, const int64_t *plane_off,
                                      const int32_t *plane_s, int max_support, int plane_mode,
                                      const int32_t *vplane, double wstep, int wbin0) {
  int rc = pv_enter(ctx, grid, uvw, nv, stride, vis_out, kern, plane_off, plane_s, plane_mode,
                    vplane, wstep, lam, max_support);
  if (rc) return rc;
  if (nv == 0) return CROC_OK;
  const BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_off, separable,
                  plane_mode, vplane, wstep, wbin0, max_support};
  if ((rc = pv_check(ctx, a))) return rc;
  const int G = persistent_blocks(ctx, 8);
  CROC_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
  tap_timer_reset(ctx, 1);
  if ((rc = tap_timer_begin(ctx, 1))) return rc;
  if (separable)
    k_pv_degrid<true><<<G, PV_NT, 0, ctx->stream>>>(a, (const float2 *)grid, (float2 *)vis_out,
                                                    (const float2 *)kern);
  else
    k_pv_degrid<false><<<G, PV_NT, 0, ctx->stream>>>(a, (const float2 *)grid,
                                                     (float2 *)vis_out, (const float2 *)kern);
  CROC_HIP(hipGetLastError());
  if ((rc = tap_timer_end(ctx, 1))) return rc;
  CROC_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
  ctx->ev_degrid = true;
  return CROC_OK;
}
