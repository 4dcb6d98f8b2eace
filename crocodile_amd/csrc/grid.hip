// Convolutional gridding: crocodile/synthesis.py:464-481 (convgrid) on
// CDNA4.
//
// Work item = up to `chunk` visibilities whose kernel windows start inside
// one T x T grid tile.  A workgroup privatises the 95 x 95 cell region
// covering every such window in LDS (planar re / im FP32 arrays), lets each
// wavefront spread one visibility at a time over its S*S taps with LDS
// float atomics (ds_add_f32), and every `batch` visibilities folds the LDS
// partial sums into per-thread Kahan-compensated register accumulators
// (each thread owns a fixed set of cells, so no two threads touch the same
// accumulator).  At the end of the item the compensated sums are added to
// the HBM grid with one float atomic per touched cell component.
// Workgroups are persistent and pull items from a device queue.
#include "internal.h"

namespace {

constexpr int GRID_NT = 1024;

__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int NT, bool SEPARABLE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_grid_tiles(
    float2 *__restrict__ grid, int N, int T, int ntx,
    const uint4 *__restrict__ rec, const float2 *__restrict__ svis,
    const float2 *__restrict__ kern, int Q,
    const int64_t *__restrict__ plane_off, const int32_t *__restrict__ plane_s,
    const int4 *__restrict__ items, const uint32_t *__restrict__ nitems,
    uint32_t *__restrict__ qhead, int batch) {
  constexpr int R = TILE_R, CELLS = R * R, PER = (CELLS + NT - 1) / NT;
  constexpr int NW = NT / 64;
  __shared__ float sre[CELLS];
  __shared__ float sim[CELLS];
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nit = (int)*nitems;
  float acc_re[PER], acc_im[PER], cmp_re[PER], cmp_im[PER];

  for (;;) {
    if (tid == 0) s_item = (int)atomicAdd(qhead, 1u);
    for (int c = tid; c < CELLS; c += NT) {
      sre[c] = 0.f;
      sim[c] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) acc_re[k] = acc_im[k] = cmp_re[k] = cmp_im[k] = 0.f;
    __syncthreads();
    const int it = s_item;
    if (it >= nit) break;
    const int4 item = items[it];
    const int tx0 = (item.x % ntx) * T, ty0 = (item.x / ntx) * T;

    for (int b0 = item.y; b0 < item.z; b0 += batch) {
      const int b1 = min(item.z, b0 + batch);
      for (int j = b0 + wave; j < b1; j += NW) {
        const uint4 r = rec[j];
        const float2 v = svis[j];
        const int S = plane_s[r.z];
        const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
        const int base = ((int)(r.x >> 16) - ty0) * R + ((int)(r.x & 0xffffu) - tx0);
        const int SS = S * S;
        const float invS = 1.0f / (float)S;
        if (!SEPARABLE) {
          const float2 *k = kern + plane_off[r.z] + (int64_t)(yf * Q + xf) * SS;
          for (int t = lane; t < SS; t += 64) {
            const int i = (int)(((float)t + 0.5f) * invS);
            const int jj = t - i * S;
            const float2 kv = k[t];
            const int c = base + i * R + jj;
            atomicAdd(&sre[c], v.x * kv.x - v.y * kv.y);
            atomicAdd(&sim[c], v.x * kv.y + v.y * kv.x);
          }
        } else {
          const float2 *ky = kern + plane_off[r.z] + (int64_t)yf * S;
          const float2 *kx = kern + plane_off[r.z] + (int64_t)xf * S;
          for (int t = lane; t < SS; t += 64) {
            const int i = (int)(((float)t + 0.5f) * invS);
            const int jj = t - i * S;
            const float2 a = ky[i], b = kx[jj];
            const float kr = a.x * b.x - a.y * b.y, ki = a.x * b.y + a.y * b.x;
            const int c = base + i * R + jj;
            atomicAdd(&sre[c], kr * v.x - ki * v.y);
            atomicAdd(&sim[c], kr * v.y + ki * v.x);
          }
        }
      }
      barrier_after_lds_atomics();
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int c = tid + k * NT;
        if (c < CELLS) {
          kahan_add(acc_re[k], cmp_re[k], sre[c]);
          kahan_add(acc_im[k], cmp_im[k], sim[c]);
          sre[c] = 0.f;
          sim[c] = 0.f;
        }
      }
      __syncthreads();
    }

#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int c = tid + k * NT;
      if (c < CELLS) {
        const int cy = c / R, cx = c - cy * R;
        const int gy = ty0 + cy, gx = tx0 + cx;
        if (gy < N && gx < N) {
          const float re = acc_re[k] - cmp_re[k];
          const float im = acc_im[k] - cmp_im[k];
          float *g = (float *)(grid + (int64_t)gy * N + gx);
          if (re != 0.f) unsafeAtomicAdd(g, re);
          if (im != 0.f) unsafeAtomicAdd(g + 1, im);
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" int croc_convgrid(croc_ctx *ctx, void *grid, int N,
                             const double *uvw, int64_t nv, int64_t stride,
                             double lam, const void *vis, const void *kern,
                             int Q, int separable, int nplanes,
                             const int64_t *plane_off, const int32_t *plane_s,
                             int max_support, int plane_mode,
                             const int32_t *vplane, double wstep, int wbin0) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid && kern && plane_off && plane_s, CROC_EINVAL,
               "null array argument");
  CROC_REQUIRE(nv == 0 || (uvw && vis), CROC_EINVAL, "null uvw / vis");
  CROC_REQUIRE(stride >= 2, CROC_EINVAL, "uvw stride must be >= 2");
  CROC_REQUIRE(plane_mode != CROC_PLANE_INDEX || vplane, CROC_EINVAL,
               "CROC_PLANE_INDEX needs vplane");
  CROC_REQUIRE(plane_mode != CROC_PLANE_WBIN || (stride >= 3 && wstep > 0),
               CROC_EINVAL, "CROC_PLANE_WBIN needs a w column and wstep > 0");
  CROC_REQUIRE(lam > 0, CROC_EINVAL, "lam must be positive");
  if (nv == 0) return CROC_OK;

  BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_mode, vplane,
            wstep, wbin0, (const float2 *)vis, max_support};
  BinPlan plan;
  rc = bin_visibilities(ctx, a, max_support, &plan);
  if (rc) return rc;
  const int G = persistent_blocks(ctx, 1);
  CROC_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  if (separable)
    k_grid_tiles<GRID_NT, true><<<G, GRID_NT, 0, ctx->stream>>>(
        (float2 *)grid, N, plan.T, plan.ntx, plan.rec, plan.svis,
        (const float2 *)kern, Q, plane_off, plane_s, plan.items, plan.nitems,
        plan.qhead, ctx->batch);
  else
    k_grid_tiles<GRID_NT, false><<<G, GRID_NT, 0, ctx->stream>>>(
        (float2 *)grid, N, plan.T, plan.ntx, plan.rec, plan.svis,
        (const float2 *)kern, Q, plane_off, plane_s, plan.items, plan.nitems,
        plan.qhead, ctx->batch);
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  rc = read_status(ctx, plan);
  (void)hipEventElapsedTime(&ctx->grid_ms, ctx->ev[0], ctx->ev[1]);
  return rc;
}
