// Convolutional gridding: crocodile/synthesis.py:464-481 (convgrid) on
// CDNA4.
//
// Work item = up to `chunk` visibilities whose kernel windows start inside
// one T x T grid tile.  A workgroup privatises the R x R cell region (R =
// 63 or 95, T + S - 1 <= R) covering every such window in LDS and lets each
// wavefront spread one visibility at a time over its S*S taps.
//
// Accumulation is exact fixed point: every contribution k*v (the FP64
// product of the FP32 tap and visibility; for separable kernels the tap is
// the FP32 product of its two 1D factors) is scaled by 2^e and rounded to a
// 64-bit integer (fixed2, within 1 unit), then added with integer atomics
// (ds_add_u64 in LDS, global_atomic_add_x2 into an int64 workspace grid).
// e is chosen on the device from max|v| * max|k| and the visibility count so
// that no partial sum can overflow and every single contribution stays below
// 2^50; with 2M visibilities it leaves ~40 fractional bits relative to the
// largest contribution.  Integer addition is
// associative, so the result does not depend on scheduling: gridding is
// bit-reproducible.  (LDS float atomics, ds_add_f32, run about ten times
// slower than the integer ones on gfx950, which is what rules them out.)
// A final pass adds workspace / 2^e into the caller's complex64 grid.
#include "taps.h"

namespace {

constexpr int GRID_NT = 1024;

struct ScaleState {
  unsigned int max_vis_bits;  // max |re|,|im| of the visibilities (float bits)
  unsigned int max_k_bits;    // max |re|,|im| of the kernel bank (float bits)
  double scale;               // 2^e
  double inv_scale;           // 2^-e
  // gridding work-list lengths: dense tiles' range items per XCD queue, then
  // the sparse tiles' items
  unsigned int qcount[XCD_QUEUES + 1];
  unsigned int nonfinite;     // bit 0: a visibility, bit 1: a kernel value is NaN or infinite
};



// Grid-stride maxima of the finite visibility and kernel component
// magnitudes; NaN or infinite values are flagged in st->nonfinite instead.
// The bank is scanned as one extent [0, max_k(plane_off[k] + table size)),
// which the C ABI requires to be a single allocation.  Non-negative floats
// order like their bit patterns, so atomicMax on the bits is a float max.
__global__ void k_scale_maxima(const float2 *__restrict__ vis, int64_t nv,
                               const float2 *__restrict__ kern,
                               const int64_t *__restrict__ plane_off,
                               const int32_t *__restrict__ plane_s, int nplanes,
                               int Q, int separable, ScaleState *st) {
  __shared__ unsigned long long s_nk;
  const int64_t nk = bank_extent(plane_off, plane_s, nplanes, Q, separable, &s_nk);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool bv = false, bk = false;
  block_max_to(array_max(vis, nv, i0, stride, bv), &st->max_vis_bits);
  block_max_to(array_max(kern, nk, i0, stride, bk), &st->max_k_bits);
  if (bv || bk) atomicOr(&st->nonfinite, (bv ? 1u : 0u) | (bk ? 2u : 0u));
}

// 2^e with (nv + 1) * c * 2^e < 2^62 (no sum overflows) and c * 2^e < 2^50
// (the range fixed2 needs), c = 2 * max|v| * max|k|^(1 or 2) bounding one
// contribution.
__global__ void k_scale_pick(ScaleState *st, int64_t nv, int separable) {
  const double mv = (double)__uint_as_float(st->max_vis_bits);
  double mk = (double)__uint_as_float(st->max_k_bits);
  if (separable) mk = mk * mk * 2.0;  // product of two table entries
  const double one = 2.0 * mv * mk;
  const double bound = one * (double)(nv + 1);
  int e = 40;
  if (bound > 0.0 && isfinite(bound)) {
    int ex, ex1;
    frexp(bound, &ex);  // bound < 2^ex
    frexp(one, &ex1);   // one < 2^ex1
    e = min(61 - ex, 50 - ex1);
  }
  e = max(-1000, min(e, 1000));
  st->scale = ldexp(1.0, e);
  st->inv_scale = ldexp(1.0, -e);
}


// One work item: the item's visibilities spread into the LDS region, which
// is then flushed (touched cells only) and cleared for the next item.
template <int NT, bool SEPARABLE, int R>
__device__ __forceinline__ void grid_item(unsigned long long *sacc, const int4 item,
                                          unsigned long long *__restrict__ acc, int N, int T,
                                          int ntx, const uint4 *__restrict__ rec,
                                          const float2 *__restrict__ vis,
                                          const float2 *__restrict__ kern, double scale) {
  constexpr int CELLS = R * R;
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tile = __builtin_amdgcn_readfirstlane(item.x);
  const int tx0 = (tile % ntx) * T, ty0 = (tile / ntx) * T;

  // Each wave walks its visibilities j = first + wave, + NW, ..., records
  // and values fetched 64 at a time (lane l: the l-th of the batch; the
  // value is gathered from the caller's array by the record's index).
  const int jend = __builtin_amdgcn_readfirstlane(item.z);
  const int jfirst = __builtin_amdgcn_readfirstlane(item.y) + wave;
  for (int jb = jfirst; jb < jend; jb += 64 * NW) {
    const int jl = jb + lane * NW;
    uint4 rb = make_uint4(0, 0, 0, 0);
    float2 vb = make_float2(0.f, 0.f);
    if (jl < jend) {
      rb = rec[jl];
      vb = vis[rb.w];
    }
    const int nb = min(64, (jend - jb + NW - 1) / NW);
    for (int it = 0; it < nb; it++) {
      const uint4 r = lane_u4(rb, it);
      const float2 v = lane_f2(vb, it);
      const int S = (int)(r.y & 0xffu);
      const int xf = (int)((r.y >> 8) & 0xfffu), yf = (int)(r.y >> 20);
      const int base = ((int)(r.x >> 16) - ty0) * R + ((int)(r.x & 0xffffu) - tx0);
      const float2 *kp = kern + r.z;  // 2D table of (yf, xf) / separable [Q][S]
      const float2 *ky = kp + yf * S, *kx = kp + xf * S;
      spread_range<SEPARABLE>(sacc, base, R, S, 0, S * S, lane, kp, ky, kx,
                              (double)v.x * scale, (double)v.y * scale);
    }
  }
  barrier_after_lds_atomics();
  // Flush the region (touched cells only) and clear it for the next item.
  for (int c = tid; c < CELLS; c += NT) {
    const unsigned long long re = sacc[2 * c], im = sacc[2 * c + 1];
    if (re | im) {
      const int cy = c / R, cx = c - cy * R;
      const int gy = ty0 + cy, gx = tx0 + cx;
      if (gy < N && gx < N) {
        unsigned long long *g = acc + 2 * ((int64_t)gy * N + gx);
        if (re) atomicAdd(g, re);
        if (im) atomicAdd(g + 1, im);
      }
      sacc[2 * c] = 0ull;
      sacc[2 * c + 1] = 0ull;
    }
  }
  __syncthreads();
}

// Persistent work queues, in two phases.  Phase A (plans, whose per-XCD item
// lists exist, with plane groups): the items of tiles holding at least
// dense_min visibilities, cut per plane-group range, each XCD draining the
// queue of its own range first (its L2 then serves one slice of the kernel
// bank, as in degridding: with uniform w every tile touches every plane, so
// whole-tile items read the whole bank through every XCD's L2).  Dense tiles
// carry most of the visibilities, and their range items are long enough to
// pay for their extra region flushes.  Phase B: every other tile's
// tile-major items, from one shared queue (qhead[XCD_QUEUES]).  Both lists
// are compacted beforehand (k_grid_lists), so a queue holds only its work.
template <int NT, bool SEPARABLE, int R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_grid_tiles(
    unsigned long long *__restrict__ acc, int N, int T, int ntx,
    const uint4 *__restrict__ rec, const float2 *__restrict__ vis,
    const float2 *__restrict__ kern, const int4 *__restrict__ items,
    const uint32_t *__restrict__ nitems, const int4 *__restrict__ dense_g,
    const uint32_t *__restrict__ gstart, int ngroups, uint32_t *__restrict__ qhead,
    const ScaleState *__restrict__ st) {
  constexpr int CELLS = R * R;
  __shared__ unsigned long long sacc[2 * CELLS];  // re, im interleaved
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const double scale = st->scale;

  for (int c = tid; c < 2 * CELLS; c += NT) sacc[c] = 0ull;
  if (ngroups > 1) {
    const int nq = min(XCD_QUEUES, ngroups);
    const int x0 = xcd_id() % nq;
    for (int k = 0; k < nq; k++) {
      const int x = (x0 + k) % nq;
      const int lo = (int)gstart[(int64_t)x * ngroups / nq];
      // Queue x starts x / nq of the way into its tile-major list, so that
      // the XCDs do not flush the same tiles' regions at the same time.
      const int len = (int)st->qcount[x], rot = (int)((int64_t)len * x / nq);
      for (;;) {
        if (tid == 0) s_item = (int)atomicAdd(qhead + x, 1u);
        __syncthreads();
        const int q = s_item;
        __syncthreads();  // every wave has read s_item before it is rewritten
        if (q >= len) break;
        const int4 item = dense_g[lo + (q + rot < len ? q + rot : q + rot - len)];
        grid_item<NT, SEPARABLE, R>(sacc, item, acc, N, T, ntx, rec, vis, kern, scale);
      }
    }
  }
  // with one plane group the tile-major list is used as it is (nitems)
  const int nit = (int)(ngroups > 1 ? st->qcount[XCD_QUEUES] : *nitems);
  for (;;) {
    if (tid == 0) s_item = (int)atomicAdd(qhead + XCD_QUEUES, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= nit) break;
    grid_item<NT, SEPARABLE, R>(sacc, items[it], acc, N, T, ntx, rec, vis, kern, scale);
  }
}

// Slot in the list counted by *count for every lane with take set: one
// atomic per wavefront (the lowest taking lane adds the wave's count),
// slots in lane order.
__device__ __forceinline__ uint32_t wave_claim(unsigned int *count, bool take) {
  const unsigned long long m = __ballot(take);
  const int lane = threadIdx.x & 63;
  if (!m) return 0u;
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0u;
  if (lane == leader) base = atomicAdd(count, (unsigned int)__popcll(m));
  base = __shfl(base, leader, 64);
  return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// The gridding work lists.  A tile holding at least dense_min visibilities
// (with more than one plane group) is dense: its per-XCD range items are
// kept, compacted within each range's span of items_g into dense_g, with
// their count per queue; every other tile's tile-major items go to sparse.
// Order within a list follows the atomics; gridding sums are exact, so it
// does not change the result.  Every lane of a wavefront runs to the end
// (the claims are wave-wide).
__global__ void k_grid_lists(const int4 *__restrict__ items, const uint32_t *__restrict__ nitems,
                             const int4 *__restrict__ items_g,
                             const uint32_t *__restrict__ nitems_g,
                             const uint32_t *__restrict__ gstart, int ngroups,
                             const uint32_t *__restrict__ offset, int dense_min,
                             int4 *__restrict__ dense_g, int4 *__restrict__ sparse,
                             ScaleState *__restrict__ st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool split = ngroups > 1;
  const int nq = min(XCD_QUEUES, ngroups);
  // dense tiles' range items, per queue
  int4 item = make_int4(0, 0, 0, 0);
  int x = 0;  // the queue whose range holds item i
  bool want = false;
  if (split && i < (int)*nitems_g) {
    item = items_g[i];
    const int t = item.x;
    want = offset[(t + 1) * ngroups] - offset[t * ngroups] >= (uint32_t)dense_min;
    while (x + 1 < nq && (int)gstart[(int64_t)(x + 1) * ngroups / nq] <= i) x++;
  }
  for (;;) {  // one round per distinct queue among the wave's items
    const unsigned long long m = __ballot(want);
    if (!m) break;
    const int xl = __shfl(x, __ffsll((long long)m) - 1, 64);
    const bool mine = want && x == xl;
    const uint32_t pos = wave_claim(&st->qcount[xl], mine);
    if (mine) {
      dense_g[gstart[(int64_t)xl * ngroups / nq] + pos] = item;
      want = false;
    }
  }
  // the other tiles' tile-major items
  bool take = false;
  int4 it2 = make_int4(0, 0, 0, 0);
  if (i < (int)*nitems) {
    it2 = items[i];
    const int t = it2.x;
    take = !split || offset[(t + 1) * ngroups] - offset[t * ngroups] < (uint32_t)dense_min;
  }
  const uint32_t pos = wave_claim(&st->qcount[XCD_QUEUES], take);
  if (take) sparse[pos] = it2;
}

// grid[c] += acc[c] / 2^e, clearing nothing (the workspace is zeroed per call).
__global__ void k_grid_finish(float2 *__restrict__ grid,
                              const long long *__restrict__ acc, int64_t n,
                              const ScaleState *__restrict__ st) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long re = acc[2 * i], im = acc[2 * i + 1];
  if (re | im) {
    const double s = st->inv_scale;
    float2 g = grid[i];
    g.x += (float)((double)re * s);
    g.y += (float)((double)im * s);
    grid[i] = g;
  }
}

// Non-finite values, as synthesis.py:479-480 computes them in FP64: a NaN or
// infinite visibility makes every cell of its window non-finite, a NaN or
// infinite tap the cell it lands on.  The fixed-point pass cannot hold them
// (their contributions are meaningless there), so those cells are set to
// NaN afterwards.  Runs only when k_scale_maxima flagged a value.
__global__ void k_grid_nonfinite(float2 *__restrict__ grid, int N,
                                 const uint4 *__restrict__ rec, int64_t n,
                                 const float2 *__restrict__ vis,
                                 const float2 *__restrict__ kern, int separable,
                                 const ScaleState *__restrict__ st) {
  const unsigned flags = st->nonfinite;
  if (flags == 0) return;
  const float nan = __builtin_nanf("");
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint4 r = rec[i];
    const bool vbad = !finite2(vis[r.w]);
    if (!vbad && !(flags & 2u)) continue;
    const int S = (int)(r.y & 0xffu);
    const int xf = (int)((r.y >> 8) & 0xfffu), yf = (int)(r.y >> 20);
    const int x0 = (int)(r.x & 0xffffu), y0 = (int)(r.x >> 16);
    const float2 *kp = kern + r.z;
    for (int y = 0; y < S; ++y) {
      for (int x = 0; x < S; ++x) {
        const bool bad = vbad || (separable ? !(finite2(kp[yf * S + y]) && finite2(kp[xf * S + x]))
                                            : !finite2(kp[y * S + x]));
        if (bad) grid[(int64_t)(y0 + y) * N + x0 + x] = make_float2(nan, nan);
      }
    }
  }
}

}  // namespace

// Gridding of binned visibilities into grid (record w: index into vis).  No
// status read.
int grid_binned(croc_ctx *ctx, const BinPlan &plan, void *grid, int N,
                int64_t nv, const void *vis, const void *kern, int Q, int separable,
                int nplanes, const int64_t *plane_off, const int32_t *plane_s) {
  int rc;
  const int64_t cells = (int64_t)N * N;
  unsigned long long *acc =
      (unsigned long long *)ws_get(ctx, WS_ACC, sizeof(unsigned long long) * 2 * cells);
  ScaleState *st = (ScaleState *)ws_get(ctx, WS_SCALE, sizeof(ScaleState));
  if (!acc || !st) return CROC_ENOMEM;
  rc = zero_u32_async(ctx, (uint32_t *)acc, 4 * cells);
  if (rc) return rc;
  rc = zero_u32_async(ctx, (uint32_t *)st, sizeof(ScaleState) / 4);
  if (rc) return rc;
  k_scale_maxima<<<ctx->num_cu, 1024, 0, ctx->stream>>>(
      (const float2 *)vis, nv, (const float2 *)kern, plane_off, plane_s, nplanes, Q,
      separable, st);
  k_scale_pick<<<1, 1, 0, ctx->stream>>>(st, nv, separable);
  CROC_HIP(hipGetLastError());

  // Tiles of at least 1.5 chunks are split by plane-group range (measured on
  // wproj2048 with chunk 512: grid 1.18 / 1.15 / 1.17 ms for thresholds of
  // 512 / 768 / 1024 visibilities, 1.42 ms without the split; 256 is far
  // slower, the range items getting too short for their region flushes).
  const int dense_min = plan.chunk + plan.chunk / 2;
  int4 *lists = (int4 *)ws_get(ctx, WS_GRIDQ, 2 * sizeof(int4) * plan.max_items);
  if (!lists) return CROC_ENOMEM;
  int4 *dense_g = lists, *sparse = lists + plan.max_items;
  const int ng_lists = plan.items_g ? plan.ngroups : 1;
  if (ng_lists > 1) {
    k_grid_lists<<<(unsigned)((plan.max_items + 255) / 256), 256, 0, ctx->stream>>>(
        plan.items, plan.nitems, plan.items_g, plan.nitems_g, plan.gstart, ng_lists,
        plan.offset, dense_min, dense_g, sparse, st);
    CROC_HIP(hipGetLastError());
  } else {
    sparse = plan.items;  // nothing to split
  }
  CROC_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  tap_timer_reset(ctx, 0);
  if ((rc = tap_timer_begin(ctx, 0))) return rc;
  const float2 *kt = (const float2 *)kern;
  if (plan.R == TILE_R_SMALL) {
    const int G = persistent_blocks(ctx, 2);  // 63.5 KiB of LDS each
    if (separable)
      k_grid_tiles<GRID_NT, true, TILE_R_SMALL><<<G, GRID_NT, 0, ctx->stream>>>(
          acc, N, plan.T, plan.ntx, plan.rec, (const float2 *)vis, kt, sparse, plan.nitems, dense_g, plan.gstart, ng_lists, plan.qhead, st);
    else
      k_grid_tiles<GRID_NT, false, TILE_R_SMALL><<<G, GRID_NT, 0, ctx->stream>>>(
          acc, N, plan.T, plan.ntx, plan.rec, (const float2 *)vis, kt, sparse, plan.nitems, dense_g, plan.gstart, ng_lists, plan.qhead, st);
  } else {
    const int G = persistent_blocks(ctx, 1);
    if (separable)
      k_grid_tiles<GRID_NT, true, TILE_R><<<G, GRID_NT, 0, ctx->stream>>>(
          acc, N, plan.T, plan.ntx, plan.rec, (const float2 *)vis, kt, sparse, plan.nitems, dense_g, plan.gstart, ng_lists, plan.qhead, st);
    else
      k_grid_tiles<GRID_NT, false, TILE_R><<<G, GRID_NT, 0, ctx->stream>>>(
          acc, N, plan.T, plan.ntx, plan.rec, (const float2 *)vis, kt, sparse, plan.nitems, dense_g, plan.gstart, ng_lists, plan.qhead, st);
  }
  CROC_HIP(hipGetLastError());
  if ((rc = tap_timer_end(ctx, 0))) return rc;
  CROC_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  ctx->ev_grid = true;
  k_grid_finish<<<(unsigned)((cells + 255) / 256), 256, 0, ctx->stream>>>(
      (float2 *)grid, (const long long *)acc, cells, st);
  k_grid_nonfinite<<<ctx->num_cu, 256, 0, ctx->stream>>>(
      (float2 *)grid, N, plan.rec, nv, (const float2 *)vis, kt, separable, st);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

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

  BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_off, separable,
            plane_mode, vplane, wstep, wbin0, max_support};
  BinPlan plan;
  // both lists: the per-XCD one carries the dense tiles (see k_grid_tiles)
  rc = bin_visibilities(ctx, ctx->ws, a, max_support, ITEMS_BY_TILE | ITEMS_BY_GROUP, &plan);
  if (rc) return rc;
  rc = grid_binned(ctx, plan, grid, N, nv, vis, kern, Q, separable, nplanes,
                   plane_off, plane_s);
  if (rc) return rc;
  rc = read_status(ctx, plan);
  return rc;
}
