// Tile binning of visibilities (shared by gridding and degridding).
//
// Each visibility is located in FP64 exactly as crocodile's frac_coords does
// (synthesis.py:419-449), its kernel plane is selected (make_gcf :451-461,
// w-bin of w_cache_imaging :793) and resolved to the offset of its kernel
// table, and it is assigned to the bin (tile, plane-group): the tile holds
// the first cell of its S x S window, the plane-group is the L2-sized slice
// of the kernel bank its table lies in.  A counting sort lays the records
// out bin by bin (tile-major), and two scans cut them into work items of at
// most `chunk` visibilities for the persistent tile kernels: per tile
// (gridding), and per (XCD queue, tile) where each queue covers a contiguous
// range of plane groups (degridding).  Block-local LDS histograms keep
// global atomics to one per (block, bin) pair, which matters
// for the strongly centre-peaked uv densities of real arrays.
#include "locate.h"

#include <hipcub/hipcub.hpp>

namespace {

constexpr int BIN_NT = 1024;
constexpr int BIN_PER = 8;
constexpr int BIN_VPB = BIN_NT * BIN_PER;
constexpr int HIST_MAX = 32768;  // 128 KiB of LDS
__global__ __launch_bounds__(BIN_NT) void k_bin_count(
    BinArgs a, BinGeom g, int nbins, uint4 *__restrict__ loc,
    uint32_t *__restrict__ count, uint32_t *__restrict__ status) {
  __shared__ uint32_t hist[HIST_MAX];
  __shared__ unsigned long long s_ext;
  const bool lds = nbins <= HIST_MAX;
  const int tid = threadIdx.x;
  // Groups are equal slices of the bank's actual extent.
  if (g.ngroups > 1) {
    const int64_t ext = bank_extent(a.plane_off, a.plane_s, a.nplanes, a.Q, a.separable, &s_ext);
    g.group_elems = max((int64_t)1, (ext + g.ngroups - 1) / g.ngroups);
  }
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
      if (L.w != LOC_INVALID) {
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

// Item slots: slot t < ntiles is tile t of the tile-major list; slot ntiles
// + r * ntiles + t is tile t's records of range r (groups [r * ngroups / nr,
// (r + 1) * ngroups / nr)) in the range-major per-XCD list.  scount[slot] =
// items the slot needs; the record offsets are copied to the scatter
// cursors.
__global__ void k_item_counts(int ntiles, int ngroups, int nr, int chunk, int order, int nbins,
                              const uint32_t *__restrict__ offset,
                              uint32_t *__restrict__ cursor, uint32_t *__restrict__ scount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbins) cursor[i] = offset[i];
  const int nslots = ntiles * (1 + nr);
  if (i >= nslots) return;
  uint32_t c = 0;
  if (i < ntiles) {
    if (order & ITEMS_BY_TILE) c = offset[(i + 1) * ngroups] - offset[i * ngroups];
  } else if (order & ITEMS_BY_GROUP) {
    const int q = i - ntiles, r = q / ntiles, t = q - r * ntiles;
    const int ga = (int)((int64_t)r * ngroups / nr), gb = (int)((int64_t)(r + 1) * ngroups / nr);
    c = offset[t * ngroups + gb] - offset[t * ngroups + ga];
  }
  scount[i] = (c + chunk - 1) / chunk;
}

// Slot s's first record and end record.
__device__ __forceinline__ void slot_range(int s, int ntiles, int ngroups, int nr,
                                           const uint32_t *__restrict__ offset, int *t,
                                           uint32_t *s0, uint32_t *s1) {
  if (s < ntiles) {
    *t = s;
    *s0 = offset[s * ngroups];
    *s1 = offset[(s + 1) * ngroups];
  } else {
    const int q = s - ntiles, r = q / ntiles;
    *t = q - r * ntiles;
    const int ga = (int)((int64_t)r * ngroups / nr), gb = (int)((int64_t)(r + 1) * ngroups / nr);
    *s0 = offset[*t * ngroups + ga];
    *s1 = offset[*t * ngroups + gb];
  }
}

// One thread per item (items of all slots numbered by soff, the tile-major
// list first): finds its slot by binary search over soff and writes {tile,
// first, end}, the per-XCD list renumbered from 0.  Threads below nr + 1
// also write gstart for the first group of each range and the item counts,
// and the first XCD_QUEUES reset the queue heads.  An error status leaves
// both lists empty, so that no output is modified.
__global__ void k_item_write(int ntiles, int ngroups, int nr, int chunk, int max_total,
                             const uint32_t *__restrict__ offset,
                             const uint32_t *__restrict__ scount,
                             const uint32_t *__restrict__ soff, int4 *__restrict__ items,
                             int4 *__restrict__ items_g, uint32_t *__restrict__ gstart,
                             uint32_t *__restrict__ nitems, uint32_t *__restrict__ nitems_g,
                             const uint32_t *__restrict__ status,
                             uint32_t *__restrict__ qheads) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nslots = ntiles * (1 + nr);
  const bool failed = *status != 0;
  const uint32_t base_g = soff[ntiles];  // items of the tile-major list
  const uint32_t total = soff[nslots - 1] + scount[nslots - 1];
  if (i < XCD_QUEUES) qheads[i] = 0;
  if (i <= nr) {
    if (i == nr) {
      *nitems = failed ? 0u : base_g;
      *nitems_g = failed ? 0u : total - base_g;
      gstart[ngroups] = failed ? 0u : total - base_g;
    } else {
      gstart[(int64_t)i * ngroups / nr] = failed ? 0u : soff[ntiles + i * ntiles] - base_g;
    }
  }
  if (failed || i >= max_total || (uint32_t)i >= total) return;
  // slot s: soff[s] <= i < soff[s] + scount[s] (the last slot whose
  // offset is <= i; empty slots share their successor's offset)
  int lo = 0, hi = nslots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (soff[mid] <= (uint32_t)i) lo = mid; else hi = mid - 1;
  }
  int t;
  uint32_t s0, s1;
  slot_range(lo, ntiles, ngroups, nr, offset, &t, &s0, &s1);
  const uint32_t k = (uint32_t)i - soff[lo];
  const uint32_t a = s0 + k * (uint32_t)chunk;
  const int4 item = make_int4(t, (int)a, (int)min(s1, a + (uint32_t)chunk), 0);
  if (lo < ntiles)
    items[i] = item;
  else
    items_g[i - base_g] = item;
}

__global__ __launch_bounds__(BIN_NT) void k_bin_scatter(
    int64_t nv, int nbins, const uint4 *__restrict__ loc,
    uint32_t *__restrict__ cursor, uint4 *__restrict__ rec) {
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
      L[k] = (i < nv) ? loc[i] : make_uint4(0, 0, 0, LOC_INVALID);
      rank[k] = (L[k].w != LOC_INVALID) ? atomicAdd(&hist[L[k].w], 1u) : 0u;
    }
    __syncthreads();
    for (int t = tid; t < nbins; t += BIN_NT)
      if (hist[t]) hist[t] = atomicAdd(&cursor[t], hist[t]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BIN_PER; k++) {
      if (L[k].w == LOC_INVALID) continue;
      const int64_t i = base + k * BIN_NT + tid;
      const uint32_t pos = hist[L[k].w] + rank[k];
      rec[pos] = make_uint4(L[k].x, L[k].y, L[k].z, (uint32_t)i);
    }
  } else {
#pragma unroll
    for (int k = 0; k < BIN_PER; k++) {
      const int64_t i = base + k * BIN_NT + tid;
      if (i >= nv) continue;
      const uint4 l = loc[i];
      if (l.w == LOC_INVALID) continue;
      const uint32_t pos = atomicAdd(&cursor[l.w], 1u);
      rec[pos] = make_uint4(l.x, l.y, l.z, (uint32_t)i);
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

  // Plane-groups: their number is set from an upper bound of the bank's
  // extent (every plane at max_support), group_kib KiB per group, or
  // automatically (group_kib 0) one per XCD queue when that bound exceeds
  // one XCD's L2 (4 MiB), else one; at most 64, and as many as keep the bin
  // histogram in LDS.  The count kernel then cuts the bank's actual extent
  // into that many equal groups.
  const int64_t S = max_support;
  const int64_t bound = (int64_t)a.nplanes * (a.separable ? a.Q * S : (int64_t)a.Q * a.Q * S * S);
  int64_t ng;
  if (ctx->group_kib > 0) {
    const int64_t gel = (int64_t)ctx->group_kib * 1024 / 8;
    ng = (bound + gel - 1) / gel;
  } else {
    ng = 8 * bound > ((int64_t)4 << 20) ? XCD_QUEUES : 1;
  }
  ng = std::min<int64_t>(ng, 64);
  ng = std::min<int64_t>(ng, std::max<int64_t>(1, HIST_MAX / p.ntiles));
  ng = std::max<int64_t>(ng, 1);
  p.ngroups = (int)ng;
  p.group_elems = std::max<int64_t>(1, (bound + ng - 1) / ng);
  p.nbins = p.ntiles * p.ngroups;
  p.max_items = a.nv / p.chunk + p.nbins + 1;

  p.loc = (uint4 *)ws_get_in(ctx, set, WS_LOC, sizeof(uint4) * a.nv);
  p.rec = (uint4 *)ws_get_in(ctx, set, WS_REC, sizeof(uint4) * a.nv);
  // Ranges of the per-XCD list: one per group, at least one per queue.
  const int nr = p.ngroups;
  const int nslots = p.ntiles * (1 + nr);
  size_t scan_bytes = 0, t_bytes = 0;
  CROC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t_bytes, (const uint32_t *)nullptr,
                                            (uint32_t *)nullptr, p.nbins + 1, ctx->stream));
  scan_bytes = t_bytes;
  CROC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t_bytes, (const uint32_t *)nullptr,
                                            (uint32_t *)nullptr, nslots, ctx->stream));
  scan_bytes = std::max(scan_bytes, t_bytes);
  // count, offset [nbins + 1]; cursor [nbins]; gstart [ngroups + 1];
  // scount, soff [nslots]
  uint32_t *tiles = (uint32_t *)ws_get_in(
      ctx, set, WS_TILES,
      sizeof(uint32_t) * (3 * (size_t)p.nbins + p.ngroups + 3 + 2 * (size_t)nslots));
  void *scan_tmp = ws_get_in(ctx, set, WS_SCAN, scan_bytes + 16);
  p.items = (order & ITEMS_BY_TILE)
                ? (int4 *)ws_get_in(ctx, set, WS_ITEMS, sizeof(int4) * p.max_items)
                : nullptr;
  p.items_g = (order & ITEMS_BY_GROUP)
                  ? (int4 *)ws_get_in(ctx, set, WS_ITEMS_G, sizeof(int4) * p.max_items)
                  : nullptr;
  uint32_t *small = (uint32_t *)ws_get_in(ctx, set, WS_SMALL, 256);
  if (!p.loc || !p.rec || !tiles || !scan_tmp || !small ||
      (!p.items && (order & ITEMS_BY_TILE)) || (!p.items_g && (order & ITEMS_BY_GROUP)))
    return CROC_ENOMEM;
  p.count = tiles;
  p.offset = tiles + p.nbins + 1;
  p.cursor = tiles + 2 * p.nbins + 2;
  p.gstart = tiles + 3 * p.nbins + 2;
  uint32_t *scount = p.gstart + p.ngroups + 1;
  uint32_t *soff = scount + nslots;
  p.status = small;
  p.nitems = small + 1;
  p.nitems_g = small + 2;
  p.qhead = small + 8;  // XCD_QUEUES per-XCD queue heads, then the shared one

  hipStream_t s = ctx->stream;
  int rc = zero_u32_async(ctx, p.count, p.nbins + 1);
  if (rc) return rc;
  rc = zero_u32_async(ctx, small, 24);
  if (rc) return rc;
  const BinGeom g{T, p.ntx, p.ngroups, p.group_elems};
  const int nblk = (int)((a.nv + BIN_VPB - 1) / BIN_VPB);
  if (nblk > 0)
    k_bin_count<<<nblk, BIN_NT, 0, s>>>(a, g, p.nbins, p.loc, p.count, p.status);
  // record offsets (bins tile-major, offset[nbins] = total), then the items
  t_bytes = scan_bytes;
  CROC_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, t_bytes, p.count, p.offset,
                                            p.nbins + 1, s));
  const int nth = std::max(p.nbins, nslots);
  k_item_counts<<<(nth + 255) / 256, 256, 0, s>>>(p.ntiles, p.ngroups, nr, p.chunk, order,
                                                  p.nbins, p.offset, p.cursor, scount);
  t_bytes = scan_bytes;
  CROC_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, t_bytes, scount, soff, nslots, s));
  // at most max_items per list
  const int max_total = (int)std::min<int64_t>(2 * p.max_items, INT32_MAX);
  const int nthr = std::max(max_total, std::max(nr + 1, XCD_QUEUES));
  k_item_write<<<(nthr + 255) / 256, 256, 0, s>>>(
      p.ntiles, p.ngroups, nr, p.chunk, max_total, p.offset, scount, soff, p.items, p.items_g,
      p.gstart, p.nitems, p.nitems_g, p.status, p.qhead);
  if (nblk > 0)
    k_bin_scatter<<<nblk, BIN_NT, 0, s>>>(a.nv, p.nbins, p.loc, p.cursor,
                                          p.rec);
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
