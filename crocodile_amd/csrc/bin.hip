// Tile binning of visibilities (shared by gridding and degridding).
//
// Each visibility is located in FP64 exactly as crocodile's frac_coords does
// (synthesis.py:419-449), its kernel plane is selected (make_gcf :451-461,
// w-bin of w_cache_imaging :793) and resolved to the offset of its kernel
// table, and it is assigned to the bin (tile, plane-group): the tile holds
// the first cell of its S x S window, the plane-group is the L2-sized slice
// of the kernel bank its table lies in.  A counting sort lays the records out
// bin by bin (tile-major), and the scan cuts them into work items of at most
// `chunk` visibilities for the persistent tile kernels, either per tile
// (gridding) or per bin in group-major order (degridding).  Block-local LDS
// histograms keep global atomics to one per (block, bin) pair, which matters
// for the strongly centre-peaked uv densities of real arrays.
#include "internal.h"

namespace {

constexpr int BIN_NT = 1024;
constexpr int BIN_PER = 4;
constexpr int BIN_VPB = BIN_NT * BIN_PER;
constexpr int HIST_MAX = 32768;  // 128 KiB of LDS
constexpr uint32_t INVALID = 0xffffffffu;

struct BinGeom {
  int T, ntx, ngroups;
  int64_t group_elems;
};

__device__ __forceinline__ uint4 locate(const BinArgs &a, const BinGeom &g,
                                        int64_t i, uint32_t *st) {
  const double *pu = a.uvw + i * a.stride;
  const double px = grid_fraction(pu[0], a.lam);
  const double py = grid_fraction(pu[1], a.lam);
  int ix, xf, iy, yf;
  const bool okx = frac_coord_dev(a.N, a.Q, px, &ix, &xf);
  const bool oky = frac_coord_dev(a.N, a.Q, py, &iy, &yf);
  if (!(okx && oky)) {
    *st |= ST_DOMAIN;
    return make_uint4(0, 0, 0, INVALID);
  }
  int64_t plane = 0;
  if (a.plane_mode == CROC_PLANE_INDEX) {
    plane = a.vplane[i];
  } else if (a.plane_mode == CROC_PLANE_WBIN) {
    plane = wbin_dev(pu[2], a.wstep) - a.wbin0;
  }
  if (plane < 0 || plane >= a.nplanes) {
    *st |= ST_PLANE;
    return make_uint4(0, 0, 0, INVALID);
  }
  const int S = a.plane_s[plane];
  if (S < 1 || S > a.max_support) {  // LDS regions are sized for max_support
    *st |= ST_SUPPORT;
    return make_uint4(0, 0, 0, INVALID);
  }
  const int x0 = ix - S / 2, y0 = iy - S / 2;
  if (x0 < 0 || y0 < 0 || x0 + S > a.N || y0 + S > a.N) {
    *st |= ST_BOUNDS;
    return make_uint4(0, 0, 0, INVALID);
  }
  const int64_t poff = a.plane_off[plane];
  const int64_t toff =
      a.separable ? poff : poff + ((int64_t)yf * a.Q + xf) * S * S;
  const int64_t tend = toff + (a.separable ? (int64_t)a.Q * S : (int64_t)S * S);
  if (poff < 0 || tend > (int64_t)UINT32_MAX) {
    *st |= ST_BANK;
    return make_uint4(0, 0, 0, INVALID);
  }
  const int grp = (int)min((int64_t)g.ngroups - 1, toff / g.group_elems);
  const uint32_t tile = (uint32_t)((y0 / g.T) * g.ntx + x0 / g.T);
  return make_uint4((uint32_t)x0 | ((uint32_t)y0 << 16),
                    (uint32_t)S | ((uint32_t)xf << 8) | ((uint32_t)yf << 20),
                    (uint32_t)toff, tile * (uint32_t)g.ngroups + (uint32_t)grp);
}

__global__ __launch_bounds__(BIN_NT) void k_bin_count(
    BinArgs a, BinGeom g, int nbins, uint4 *__restrict__ loc,
    uint32_t *__restrict__ count, uint32_t *__restrict__ status) {
  __shared__ uint32_t hist[HIST_MAX];
  const bool lds = nbins <= HIST_MAX;
  const int tid = threadIdx.x;
  if (lds)
    for (int t = tid; t < nbins; t += BIN_NT) hist[t] = 0;
  __syncthreads();
  uint32_t st = 0;
  const int64_t base = (int64_t)blockIdx.x * BIN_VPB;
#pragma unroll
  for (int k = 0; k < BIN_PER; k++) {
    const int64_t i = base + k * BIN_NT + tid;
    if (i < a.nv) {
      const uint4 L = locate(a, g, i, &st);
      loc[i] = L;
      if (L.w != INVALID) {
        if (lds)
          atomicAdd(&hist[L.w], 1u);
        else
          atomicAdd(&count[L.w], 1u);
      }
    }
  }
  if (st) atomicOr(status, st);
  barrier_after_lds_atomics();
  if (lds)
    for (int t = tid; t < nbins; t += BIN_NT)
      if (hist[t]) atomicAdd(&count[t], hist[t]);
}

// Block-wide exclusive scan of one value per thread (Hillis-Steele in LDS);
// returns the thread's exclusive prefix.
__device__ uint32_t block_exclusive_scan(uint32_t x, uint32_t *sc) {
  const int tid = threadIdx.x;
  sc[tid] = x;
  __syncthreads();
  for (int off = 1; off < BIN_NT; off <<= 1) {
    const uint32_t a = (tid >= off) ? sc[tid - off] : 0u;
    __syncthreads();
    sc[tid] += a;
    __syncthreads();
  }
  const uint32_t r = sc[tid] - x;
  __syncthreads();
  return r;
}

// One block: exclusive scan of the bin counts (record offsets, tile-major),
// then the requested work-item lists: tile-major `items` and / or the
// per-XCD queues `items_g`.  The groups are cut into nq = min(8, ngroups)
// contiguous ranges; queue x lists, tile by tile, the records of range x
// (group-sorted within each item), and gstart[first group of range x] is
// its first item.  An error status empties the item lists so that no
// output is modified.
__global__ __launch_bounds__(BIN_NT) void k_bin_scan(
    int ntiles, int ngroups, int chunk, int order,
    const uint32_t *__restrict__ count, uint32_t *__restrict__ offset,
    uint32_t *__restrict__ cursor, int4 *__restrict__ items, int4 *__restrict__ items_g,
    uint32_t *__restrict__ gstart, uint32_t *__restrict__ nitems,
    uint32_t *__restrict__ nitems_g, const uint32_t *__restrict__ status,
    uint32_t *__restrict__ qheads) {
  __shared__ uint32_t sc[BIN_NT];
  const int tid = threadIdx.x;
  const int nbins = ntiles * ngroups;
  const bool failed = *status != 0;

  // record offsets, bins in tile-major order
  const int per = (nbins + BIN_NT - 1) / BIN_NT;
  const int b0 = min(nbins, tid * per), b1 = min(nbins, b0 + per);
  uint32_t cs = 0;
  for (int b = b0; b < b1; b++) cs += count[b];
  uint32_t co = block_exclusive_scan(cs, sc);
  for (int b = b0; b < b1; b++) {
    offset[b] = co;
    cursor[b] = co;
    co += count[b];
  }
  if (tid == BIN_NT - 1) offset[nbins] = co;
  __syncthreads();  // offsets visible to the whole block

  if (order & ITEMS_BY_TILE) {
    const int tper = (ntiles + BIN_NT - 1) / BIN_NT;
    const int t0 = min(ntiles, tid * tper), t1 = min(ntiles, t0 + tper);
    uint32_t is = 0;
    for (int t = t0; t < t1; t++) {
      const uint32_t c = offset[(t + 1) * ngroups] - offset[t * ngroups];
      is += (c + chunk - 1) / chunk;
    }
    uint32_t io = block_exclusive_scan(is, sc);
    for (int t = t0; t < t1; t++) {
      const uint32_t s0 = offset[t * ngroups], s1 = offset[(t + 1) * ngroups];
      for (uint32_t s = s0; s < s1; s += chunk)
        items[io++] = make_int4(t, (int)s, (int)min(s1, s + chunk), 0);
    }
    if (tid == BIN_NT - 1) *nitems = failed ? 0u : io;
  }
  if (order & ITEMS_BY_GROUP) {
    // queue-major: index q = x * ntiles + t covers tile t's groups of queue x
    const int nq = min(XCD_QUEUES, ngroups);
    const int nqs = nq * ntiles;
    const int qper = (nqs + BIN_NT - 1) / BIN_NT;
    const int q0 = min(nqs, tid * qper), q1 = min(nqs, q0 + qper);
    uint32_t is = 0;
    for (int q = q0; q < q1; q++) {
      const int x = q / ntiles, t = q - x * ntiles;
      const int ga = (int)((int64_t)x * ngroups / nq), gb = (int)((int64_t)(x + 1) * ngroups / nq);
      is += (offset[t * ngroups + gb] - offset[t * ngroups + ga] + chunk - 1) / chunk;
    }
    uint32_t io = block_exclusive_scan(is, sc);
    for (int q = q0; q < q1; q++) {
      const int x = q / ntiles, t = q - x * ntiles;
      const int ga = (int)((int64_t)x * ngroups / nq), gb = (int)((int64_t)(x + 1) * ngroups / nq);
      if (t == 0) gstart[ga] = failed ? 0u : io;
      const uint32_t s0 = offset[t * ngroups + ga], s1 = offset[t * ngroups + gb];
      for (uint32_t s = s0; s < s1; s += chunk)
        items_g[io++] = make_int4(t, (int)s, (int)min(s1, s + chunk), 0);
    }
    if (tid == BIN_NT - 1) {
      *nitems_g = failed ? 0u : io;
      gstart[ngroups] = failed ? 0u : io;
    }
  }
  if (tid == 0)
    for (int q = 0; q < XCD_QUEUES; q++) qheads[q] = 0;
}

__global__ __launch_bounds__(BIN_NT) void k_bin_scatter(
    int64_t nv, int nbins, const uint4 *__restrict__ loc,
    uint32_t *__restrict__ cursor, uint4 *__restrict__ rec,
    const float2 *__restrict__ vis, float2 *__restrict__ svis) {
  __shared__ uint32_t hist[HIST_MAX];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * BIN_VPB;
  uint4 L[BIN_PER];
  uint32_t rank[BIN_PER];
  if (nbins <= HIST_MAX) {
    for (int t = tid; t < nbins; t += BIN_NT) hist[t] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BIN_PER; k++) {
      const int64_t i = base + k * BIN_NT + tid;
      L[k] = (i < nv) ? loc[i] : make_uint4(0, 0, 0, INVALID);
      rank[k] = (L[k].w != INVALID) ? atomicAdd(&hist[L[k].w], 1u) : 0u;
    }
    __syncthreads();
    for (int t = tid; t < nbins; t += BIN_NT)
      if (hist[t]) hist[t] = atomicAdd(&cursor[t], hist[t]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BIN_PER; k++) {
      if (L[k].w == INVALID) continue;
      const int64_t i = base + k * BIN_NT + tid;
      const uint32_t pos = hist[L[k].w] + rank[k];
      rec[pos] = make_uint4(L[k].x, L[k].y, L[k].z, (uint32_t)i);
      if (vis) svis[pos] = vis[i];
    }
  } else {
#pragma unroll
    for (int k = 0; k < BIN_PER; k++) {
      const int64_t i = base + k * BIN_NT + tid;
      if (i >= nv) continue;
      const uint4 l = loc[i];
      if (l.w == INVALID) continue;
      const uint32_t pos = atomicAdd(&cursor[l.w], 1u);
      rec[pos] = make_uint4(l.x, l.y, l.z, (uint32_t)i);
      if (vis) svis[pos] = vis[i];
    }
  }
}

}  // namespace

int choose_tile(croc_ctx *ctx, int max_support) {
  if (ctx->tile) return ctx->tile;
  // Largest multiple of 16 with T + S - 1 inside the small region (measured
  // fastest on w-projection banks), else inside the large one.
  const int s16 = (max_support + 15) / 16 * 16;
  if (TILE_R_SMALL + 1 - s16 >= 16) return TILE_R_SMALL + 1 - s16;
  return TILE_R + 1 - s16;
}

int choose_region(int T, int max_support) {
  return T + max_support - 1 <= TILE_R_SMALL ? TILE_R_SMALL : TILE_R;
}

int bin_visibilities(croc_ctx *ctx, WsSet &set, const BinArgs &a, int max_support,
                     int order, BinPlan *plan) {
  CROC_REQUIRE(max_support >= 1 && max_support <= 64, CROC_EINVAL,
               "kernel support %d outside [1, 64]", max_support);
  CROC_REQUIRE(a.N >= 1 && a.N <= 65535, CROC_EINVAL, "grid size %d outside "
               "[1, 65535]", a.N);
  CROC_REQUIRE(a.Q >= 1 && a.Q <= 4096, CROC_EINVAL, "oversampling %d outside"
               " [1, 4096]", a.Q);
  CROC_REQUIRE(a.nv >= 0 && a.nv < (int64_t)INT32_MAX, CROC_EINVAL,
               "visibility count %lld outside [0, 2^31)", (long long)a.nv);
  CROC_REQUIRE(a.nplanes >= 1, CROC_EINVAL, "empty kernel bank");
  CROC_REQUIRE(a.max_support == max_support, CROC_EINVAL, "support mismatch");
  const int T = choose_tile(ctx, max_support);
  CROC_REQUIRE(T + max_support - 1 <= TILE_R, CROC_EINVAL,
               "tile %d too large for support %d", T, max_support);
  BinPlan &p = *plan;
  p.T = T;
  p.R = choose_region(T, max_support);
  p.ntx = (a.N + T - 1) / T;
  p.nty = p.ntx;
  p.ntiles = p.ntx * p.nty;
  p.chunk = ctx->chunk;

  // Plane-groups: slices of group_kib KiB of an upper bound of the bank's
  // extent, at most 64 (bin counts beyond the LDS histogram use global
  // atomics).
  const int64_t S = max_support;
  const int64_t bound = (int64_t)a.nplanes * (a.separable ? a.Q * S : (int64_t)a.Q * a.Q * S * S);
  int64_t ng = 1;
  if (ctx->group_kib > 0) {
    const int64_t gel = (int64_t)ctx->group_kib * 1024 / 8;
    ng = (bound + gel - 1) / gel;
    ng = std::min<int64_t>(ng, 64);
    ng = std::max<int64_t>(ng, 1);
  }
  p.ngroups = (int)ng;
  p.group_elems = std::max<int64_t>(1, (bound + ng - 1) / ng);
  p.nbins = p.ntiles * p.ngroups;
  p.max_items = a.nv / p.chunk + p.nbins + 1;

  p.loc = (uint4 *)ws_get_in(ctx, set, WS_LOC, sizeof(uint4) * a.nv);
  p.rec = (uint4 *)ws_get_in(ctx, set, WS_REC, sizeof(uint4) * a.nv);
  if (a.vis) p.svis = (float2 *)ws_get_in(ctx, set, WS_SVIS, sizeof(float2) * a.nv);
  uint32_t *tiles = (uint32_t *)ws_get_in(
      ctx, set, WS_TILES, sizeof(uint32_t) * (3 * (size_t)p.nbins + p.ngroups + 2));
  p.items = (order & ITEMS_BY_TILE)
                ? (int4 *)ws_get_in(ctx, set, WS_ITEMS, sizeof(int4) * p.max_items)
                : nullptr;
  p.items_g = (order & ITEMS_BY_GROUP)
                  ? (int4 *)ws_get_in(ctx, set, WS_ITEMS_G, sizeof(int4) * p.max_items)
                  : nullptr;
  uint32_t *small = (uint32_t *)ws_get_in(ctx, set, WS_SMALL, 256);
  if (!p.loc || !p.rec || (a.vis && !p.svis) || !tiles || !small ||
      (!p.items && (order & ITEMS_BY_TILE)) || (!p.items_g && (order & ITEMS_BY_GROUP)))
    return CROC_ENOMEM;
  p.count = tiles;
  p.offset = tiles + p.nbins;
  p.cursor = tiles + 2 * p.nbins + 1;
  p.gstart = tiles + 3 * p.nbins + 1;
  p.status = small;
  p.nitems = small + 1;
  p.nitems_g = small + 2;
  p.qhead = small + 8;  // XCD_QUEUES queue heads

  hipStream_t s = ctx->stream;
  int rc = zero_u32_async(ctx, p.count, p.nbins);
  if (rc) return rc;
  rc = zero_u32_async(ctx, small, 16);
  if (rc) return rc;
  const BinGeom g{T, p.ntx, p.ngroups, p.group_elems};
  const int nblk = (int)((a.nv + BIN_VPB - 1) / BIN_VPB);
  if (nblk > 0)
    k_bin_count<<<nblk, BIN_NT, 0, s>>>(a, g, p.nbins, p.loc, p.count, p.status);
  k_bin_scan<<<1, BIN_NT, 0, s>>>(p.ntiles, p.ngroups, p.chunk, order, p.count,
                                  p.offset, p.cursor, p.items, p.items_g, p.gstart,
                                  p.nitems, p.nitems_g, p.status, p.qhead);
  if (nblk > 0)
    k_bin_scatter<<<nblk, BIN_NT, 0, s>>>(a.nv, p.nbins, p.loc, p.cursor,
                                          p.rec, a.vis, p.svis);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

namespace {
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
}  // namespace

// A fill kernel rather than hipMemsetAsync, so the call sequence stays valid
// under stream capture.
int zero_u32_async(croc_ctx *ctx, uint32_t *p, int64_t n) {
  if (n <= 0) return CROC_OK;
  k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, n, 0u);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

int read_status(croc_ctx *ctx, const BinPlan &plan) {
  CROC_HIP(hipMemcpyAsync(ctx->host_status, plan.status, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return croc_status_to_code(*ctx->host_status);
}
