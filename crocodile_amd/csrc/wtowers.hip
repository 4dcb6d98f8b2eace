// W-towers imaging and predict (examples/notebooks/wtowers.ipynb,
// wtowers-predict.ipynb; examples/grid/grid.c:482-710) on CDNA4.
//
// Geometry (shared with the FP64 oracle): chunk = sg - margin; a
// visibility's chunk is (floor(u theta + 0.5) + N/2) / chunk per axis; the
// chunk's sub-grid starts at x0 = cx*chunk - margin/2 and is centred on
// u_mid = (x0 + sg/2 - N/2) / theta; the w-plane is rint(w / wstep).
//
// Imaging, per call:
//   1. bin visibilities by job = (chunk, w-plane): stable radix sort of
//      (job key, visibility index), run-length encoding of the sorted keys;
//      the job schedules are made on the host from the job table;
//   2. grid the jobs, in batches of sub-grid slots, each part of a job on
//      its own sg x sg slot: separable kernels of support <= 16 on the
//      matrix cores (k_tw_grid_jobs_mfma), others in two passes of sg/2
//      rows held in LDS as exact 64-bit fixed point (k_tw_grid_jobs);
//      multi-part jobs summed (k_tw_sum_parts);
//   3. the job images ifft(sub-grid) * exp(-2 pi i wp wstep (1 - n)) formed
//      in LDS and summed per chunk in registers, in job order
//      (k_tw_image_runs); with separable kernels of support <= 16 the jobs
//      of at most TW_SMALL_VIS visibilities skip step 2 and are gridded
//      inside this kernel;
//   4. forward FFT of every chunk image;
//   5. add the chunk sub-grids into the grid in four passes by chunk parity,
//      so overlapping margins are never written concurrently (deterministic).
// Predict reverses it: cut + inverse FFT per chunk; per job the forward FFT
// with exp(+2 pi i wp wstep (1 - n)) applied to the input, then degrid:
// with separable kernels of support <= 16 both in one kernel from LDS
// (k_tw_predict_sep), otherwise batched transforms into slots and
// k_tw_degrid_jobs.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "fft128.h"
#include "taps.h"

namespace {

constexpr int SG = 128;
constexpr int SGN = SG * SG;
constexpr int TW_NT = 1024;
#ifndef JOB_BATCH
#define JOB_BATCH 16384  // sub-grid slots per batch (2 GiB of 128^2 complex64)
#endif
constexpr int PART_VIS = 4096;     // visibilities per work part of a job
constexpr uint32_t INVALID = 0xffffffffu;

struct TowerArgs {
  int N, sg, margin, chunk, ncx;
  double theta, wstep;
  int wp_min, nwp;
  int Q, S, nearest;
  // w-kernel set of the residual w - wp wstep (grid.c's w_kernel_data):
  // plane floor((w - wp wstep - wk_min) / wk_step + 0.5) of nwk; one plane
  // (nwk = 1) serves every visibility
  int nwk;
  double wk_min, wk_step;
  const double *uvw;
  int64_t nv, stride;
};

// Range of w-planes rint(w / wstep): grid-stride, reduced per block before
// one global atomic pair per block.
__global__ void k_tw_wrange(const double *__restrict__ uvw, int64_t nv, int64_t stride,
                            double wstep, int *mm, uint32_t *status) {
  int lo = INT_MAX, hi = INT_MIN;
  bool bad = false;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += step) {
    const double q = rint(uvw[i * stride + 2] / wstep);
    if (!(fabs(q) < 1e9)) {  // checked before any integer conversion
      bad = true;
      continue;
    }
    lo = min(lo, (int)q);
    hi = max(hi, (int)q);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off, 64));
    hi = max(hi, __shfl_xor(hi, off, 64));
  }
  __shared__ int s_lo, s_hi;
  if (threadIdx.x == 0) {
    s_lo = INT_MAX;
    s_hi = INT_MIN;
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && lo <= hi) {
    atomicMin(&s_lo, lo);
    atomicMax(&s_hi, hi);
  }
  barrier_after_lds_atomics();
  if (bad) atomicOr(status, (uint32_t)ST_PLANE);
  if (threadIdx.x == 0 && s_lo <= s_hi) {
    atomicMin(&mm[0], s_lo);
    atomicMax(&mm[1], s_hi);
  }
}

// Zero the listed chunk images (blockIdx.y = list entry).
__global__ void k_tw_zero_chunks(const int32_t *__restrict__ chunks, float2 *__restrict__ cimg) {
  float4 *d = (float4 *)(cimg + (int64_t)chunks[blockIdx.y] * SGN);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SGN / 2; i += gridDim.x * blockDim.x)
    d[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Sub-grid window of visibility i; returns the job key or INVALID.
__device__ uint4 tw_locate(const TowerArgs &a, int64_t i, uint32_t *st) {
#pragma clang fp contract(off)
  const double *pu = a.uvw + i * a.stride;
  const double u = pu[0], v = pu[1];
  const int64_t cellx = (int64_t)floor(u * a.theta + 0.5) + a.N / 2;
  const int64_t celly = (int64_t)floor(v * a.theta + 0.5) + a.N / 2;
  const int64_t cx = floordiv64(cellx, a.chunk), cy = floordiv64(celly, a.chunk);
  if (cx < 0 || cy < 0 || cx >= a.ncx || cy >= a.ncx) {
    *st |= ST_BOUNDS;
    return make_uint4(0, 0, 0, INVALID);
  }
  const int wp = (int)wbin_dev(pu[2], a.wstep);
  const int64_t x0 = cx * a.chunk - a.margin / 2, y0 = cy * a.chunk - a.margin / 2;
  const double lam_sub = (double)a.sg / a.theta;
  const double px = (u - (double)(x0 + a.sg / 2 - a.N / 2) / a.theta) / lam_sub;
  const double py = (v - (double)(y0 + a.sg / 2 - a.N / 2) / a.theta) / lam_sub;
  int wx, wy, xf = 0, yf = 0;
  if (a.nearest) {
    // grid(): cell sg/2 + floor(0.5 + sg p), p < 0.5 (synthesis.py:396-399)
    if (!(px < 0.5 && py < 0.5)) {
      *st |= ST_DOMAIN;
      return make_uint4(0, 0, 0, INVALID);
    }
    wx = a.sg / 2 + (int)floor(0.5 + (double)a.sg * px);
    wy = a.sg / 2 + (int)floor(0.5 + (double)a.sg * py);
    if (wx < 0 || wy < 0 || wx >= a.sg || wy >= a.sg) {
      *st |= ST_BOUNDS;
      return make_uint4(0, 0, 0, INVALID);
    }
  } else {
    int ix, iy;
    if (!frac_coord_dev(a.sg, a.Q, px, &ix, &xf) || !frac_coord_dev(a.sg, a.Q, py, &iy, &yf)) {
      *st |= ST_DOMAIN;
      return make_uint4(0, 0, 0, INVALID);
    }
    wx = ix - a.S / 2;
    wy = iy - a.S / 2;
    if (wx < 0 || wy < 0 || wx + a.S > a.sg || wy + a.S > a.sg) {
      *st |= ST_BOUNDS;
      return make_uint4(0, 0, 0, INVALID);
    }
  }
  // w-kernel plane of the residual w relative to the plane's w (grid.c:286-289)
  int wk = 0;
  if (a.nwk > 1) {
    const double dw = pu[2] - (double)wp * a.wstep;
    const double q = floor((dw - a.wk_min) / a.wk_step + 0.5);
    if (!(q >= 0.0 && q < (double)a.nwk)) {
      *st |= ST_PLANE;
      return make_uint4(0, 0, 0, INVALID);
    }
    wk = (int)q;
  }
  const uint32_t key =
      (uint32_t)(((int64_t)cy * a.ncx + cx) * a.nwp + (wp - a.wp_min));
  return make_uint4((uint32_t)wx | ((uint32_t)wy << 16), (uint32_t)xf | ((uint32_t)yf << 16),
                    (uint32_t)wk, key);
}

// Located records, and (job key, visibility index) pairs for the sort; an
// invalid visibility gets key `bad`, above every job key (the call then
// fails on the status before any job is used).
__global__ void k_tw_locate(TowerArgs a, uint32_t bad, uint4 *__restrict__ loc,
                            uint32_t *__restrict__ keys, uint32_t *__restrict__ idx,
                            uint32_t *__restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nv) return;
  uint32_t st = 0;
  const uint4 L = tw_locate(a, i, &st);
  loc[i] = L;
  keys[i] = L.w == INVALID ? bad : L.w;
  idx[i] = (uint32_t)i;
  if (st) atomicOr(status, st);
}

// Records (and visibilities) in job order: rec[pos] = (window, offsets,
// w-kernel plane, visibility index) of the pos-th visibility of the sorted
// order.
__global__ void k_tw_gather(int64_t nv, const uint32_t *__restrict__ sidx,
                            const uint4 *__restrict__ loc, const float2 *__restrict__ vis,
                            uint4 *__restrict__ rec, float2 *__restrict__ svis) {
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= nv) return;
  const uint32_t i = sidx[pos];
  const uint4 L = loc[i];
  rec[pos] = make_uint4(L.x, L.y, L.z, i);
  if (vis) svis[pos] = vis[i];
}

// jobs[r] = (key, first record, end record) of run r of the sorted keys.
__global__ void k_tw_jobs(const uint32_t *__restrict__ ukeys, const uint32_t *__restrict__ counts,
                          const uint32_t *__restrict__ starts, int nruns,
                          int4 *__restrict__ jobs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  jobs[r] = make_int4((int)ukeys[r], (int)starts[r], (int)(starts[r] + counts[r]), 0);
}

struct TwScale {
  unsigned int max_vis_bits, max_k_bits;  // float bits of max |re|, |im|
  double scale, inv_scale;                // 2^e, 2^-e
  unsigned int nonfinite;  // bit 0: a visibility, bit 1: a kernel value is NaN or infinite
  unsigned int pad;
};

__global__ void k_tw_maxima(const float2 *__restrict__ vis, int64_t nv,
                            const float2 *__restrict__ kern, int64_t nk, TwScale *fs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool bv = false, bk = false;
  block_max_to(array_max(vis, nv, i0, stride, bv), &fs->max_vis_bits);
  block_max_to(array_max(kern, nk, i0, stride, bk), &fs->max_k_bits);
  if (bv || bk) atomicOr(&fs->nonfinite, (bv ? 1u : 0u) | (bk ? 2u : 0u));
}

// 2^e with (nv + 1) * c * 2^e < 2^62 (no sum overflows) and c * 2^e < 2^50
// (fixed() is exact), c = 2 * max|v| * max|k| bounding one contribution.
__global__ void k_tw_pick(TwScale *fs, int64_t nv, int nearest, int separable) {
  const double mv = (double)__uint_as_float(fs->max_vis_bits);
  double mk = nearest ? 1.0 : (double)__uint_as_float(fs->max_k_bits);
  if (separable && !nearest) mk = 2.0 * mk * mk;
  const double one = 2.0 * mv * mk;
  const double bound = one * (double)(nv + 1);
  int e = 40;
  if (bound > 0.0 && isfinite(bound)) {
    int ex, ex1;
    frexp(bound, &ex);
    frexp(one, &ex1);
    e = min(61 - ex, 50 - ex1);
  }
  e = max(-1000, min(e, 1000));
  fs->scale = ldexp(1.0, e);
  fs->inv_scale = ldexp(1.0, -e);
}

// Grid job parts into sub-grid slots: parts[b] = (slot, first record, end
// record); one workgroup per part, in passes of GRID_ROWS rows: each
// wavefront adds the taps of one visibility at a time that lie in the pass's
// rows (a contiguous tap range) to the int64 region, and the pass ends by
// converting the region to complex64 in sub[slot] and zeroing it for the
// next pass (it is zeroed once at the start of the launch).  Region rows are
// padded by 8 bytes (GRID_RU64 u64 per row): with the bare 2 KiB pitch the
// rows of a 64-tap block would all fall on the same LDS banks.  (Passes of
// 32 rows, two workgroups per compute unit, measured slower: every pass
// lists the part's visibilities again.)
constexpr int GRID_ROWS = SG / 2;
constexpr int GRID_RU64 = 2 * SG + 1;
template <bool SEPARABLE>
__global__ __launch_bounds__(TW_NT) void k_tw_grid_jobs(
    const int4 *__restrict__ parts, int nb, const uint4 *__restrict__ rec,
    const float2 *__restrict__ svis, const float2 *__restrict__ kern, int Q, int S,
    int64_t tsize, int nearest, const TwScale *__restrict__ fs, float2 *__restrict__ sub,
    uint32_t *__restrict__ qc) {
  constexpr int ROWS = GRID_ROWS, HC = ROWS * SG;
  __shared__ unsigned long long acc[ROWS * GRID_RU64];  // re, im interleaved (128 KiB + pads)
  __shared__ uint16_t s_list[PART_VIS];       // the pass's visibilities (offsets in the part)
  __shared__ int s_n;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int NW = TW_NT / 64;
  const double scale = fs->scale, inv = fs->inv_scale;
  const int SS = S * S;
  __shared__ int s_b;
  // The region starts zeroed; each pass's read-out zeroes it again.
  for (int c = tid; c < ROWS * GRID_RU64; c += TW_NT) acc[c] = 0ull;
  for (;;) {  // parts taken from the launch's queue counter qc
    if (tid == 0) s_b = (int)atomicAdd(qc, 1u);
    __syncthreads();
    const int b = s_b;
    if (b >= nb) break;
    const int4 job = parts[b];
    const int jbeg = __builtin_amdgcn_readfirstlane(job.y);
    const int jend = __builtin_amdgcn_readfirstlane(job.z);
    float2 *out = sub + (int64_t)__builtin_amdgcn_readfirstlane(job.x) * SGN;
    for (int r0 = 0; r0 < SG; r0 += ROWS) {
      if (tid == 0) s_n = 0;
      __syncthreads();
      if (nearest) {
        for (int j = jbeg + tid; j < jend; j += TW_NT) {
          const uint4 r = rec[j];
          const int wy = (int)(r.x >> 16) - r0;
          if (wy >= 0 && wy < ROWS) {
            const float2 v = svis[j];
            unsigned long long *cell = acc + wy * GRID_RU64 + 2 * (int)(r.x & 0xffffu);
            atomicAdd(cell, fixed(v.x, scale));
            atomicAdd(cell + 1, fixed(v.y, scale));
          }
        }
      } else {
        // Walk the part in windows of PART_VIS records.  Each window first
        // lists its visibilities whose kernel window meets the pass's rows
        // (order is immaterial: integer sums), so that the walk pays its
        // per-visibility set-up once per visibility and pass it touches.
        for (int jb = jbeg; jb < jend; jb += PART_VIS) {
          const int je = min(jend, jb + PART_VIS);
          for (int j0 = jb; j0 < je; j0 += TW_NT) {
            const int j = j0 + tid;
            bool hit = false;
            if (j < je) {
              const int wy = (int)(rec[j].x >> 16);
              hit = wy < r0 + ROWS && wy + S > r0;
            }
            const unsigned long long bal = __ballot(hit);
            int base = 0;
            if (lane == 0 && bal) base = atomicAdd(&s_n, __popcll(bal));
            base = __shfl(base, 0, 64);
            if (hit)
              s_list[base + __popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)(j - jb);
          }
          barrier_after_lds_atomics();
          const int nhit = __builtin_amdgcn_readfirstlane(s_n);
          // The wave's listed visibilities k = wave, + NW, ..., records and
          // values fetched 64 at a time (lane l: the l-th of the batch) and
          // handed out with readlane: no scalar load, and so no wait for the
          // wave's own LDS atomics, before each visibility.
          for (int kb = wave; kb < nhit; kb += 64 * NW) {
            const int kl = kb + lane * NW;
            uint4 rb = make_uint4(0, 0, 0, 0);
            float2 vb = make_float2(0.f, 0.f);
            if (kl < nhit) {
              const int j = jb + (int)s_list[kl];
              rb = rec[j];
              vb = svis[j];
            }
            const int nbat = min(64, (nhit - kb + NW - 1) / NW);
            for (int it = 0; it < nbat; it++) {
              const uint4 r = lane_u4(rb, it);
              const float2 v = lane_f2(vb, it);
              const int wy = (int)(r.x >> 16), wx = (int)(r.x & 0xffffu);
              // window rows wy + i in the pass: i in [ilo, ihi)
              const int ilo = max(0, r0 - wy), ihi = min(S, r0 + ROWS - wy);
              const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
              const float2 *kt = kern + (int64_t)r.z * tsize;  // its w-kernel plane
              const float2 *k2 = kt + (int64_t)(yf * Q + xf) * SS;
              const float2 *ky = kt + (int64_t)yf * S, *kx = kt + (int64_t)xf * S;
              // cell of tap t: (wy - r0) * SG + wx + t + i * (SG - S), in
              // [0, HC) for the pass's rows, at byte 16 * cell + 8 * its row
              spread_range<SEPARABLE>(acc, (wy - r0) * SG + wx, SG, S, ilo * S, ihi * S, lane,
                                      k2, ky, kx, (double)v.x * scale, (double)v.y * scale, 8u,
                                      8u * (uint32_t)(wy - r0));
            }
          }
          __syncthreads();  // the list is read; the next window rewrites it
          if (tid == 0) s_n = 0;
          __syncthreads();
        }
      }
      barrier_after_lds_atomics();
      for (int c = tid; c < HC; c += TW_NT) {
        const int ca = (c / SG) * GRID_RU64 + 2 * (c % SG);
        const long long re = (long long)acc[ca], im = (long long)acc[ca + 1];
        acc[ca] = 0ull;
        acc[ca + 1] = 0ull;
        // streaming store: the sub-grids must not evict the kernel table from L2
        store_stream(out + r0 * SG + c,
                     make_float2((float)((double)re * inv), (float)((double)im * inv)));
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Gridding of job parts with separable kernels on the matrix cores.
//
// With a separable kernel a visibility adds the outer product
// (v ky) (x) kx to its S x S window, so the part's sub-grid is
//   G = sum_k a_k b_k^T,  a_k = ky_k placed at rows [wy_k, wy_k + S),
//                         b_k = v_k kx_k placed at columns [wx_k, wx_k + S),
// a product of a 128 x K and a K x 128 matrix over the K visibilities.  Cut
// into 16 x 16 blocks, block (by, bx) only needs the visibilities whose
// window meets it (at most 2 x 2 blocks each for S <= 16), and for those
// G_blk = A_blk B_blk is a dense complex product with K = the block's
// visibilities: 16 x 16 x 32 steps of v_mfma_f32_16x16x32_f16, each FP32
// operand split as hi + lo in fp16 (x = hi + lo to 22 bits; the lo * lo
// term, 2^-22 relative, is dropped) and accumulated in FP32.  No LDS atomics:
// every block belongs to one wavefront, which writes it once.  Operands:
// A (ky) comes from a copy of the kernel set pre-split into fp16 (hi, lo)
// pairs scaled by sk; B = v kx is split on the fly, scaled by sv sk.
// Blocks are taken dynamically by the part's wavefronts; a wavefront lists
// the block's visibilities in index order into a ring (ballot compaction)
// and runs the matrix products 32 at a time, so the sums are deterministic.
typedef _Float16 tw_half8 __attribute__((ext_vector_type(8)));
typedef float tw_floatx4 __attribute__((ext_vector_type(4)));

constexpr int MG_WIN = 1024;   // visibilities staged in LDS per window of a part
constexpr int MG_RING = 128;   // per-wavefront ring of listed visibilities
constexpr int MG_NW = TW_NT / 64;  // wavefronts per workgroup
constexpr int MG_NB = 64;      // 8 x 8 blocks of 16 x 16 cells

// Largest power of two 2^-e with x * 2^-e <= 1 (x > 0 finite), else 1.
__device__ __forceinline__ float unit_scale(float x) {
  if (!(x > 0.f) || !isfinite(x)) return 1.f;
  int e;
  frexpf(x, &e);  // x < 2^e
  return ldexpf(1.f, -e);
}

// hi / lo fp16 split of the kernel set, scaled by sk = unit_scale(max |k|):
// ksplit[i] = (re hi, re lo, im hi, im lo).
__global__ void k_tw_split_tables(const float2 *__restrict__ kern, int64_t n,
                                  const TwScale *__restrict__ fs, uint2 *__restrict__ ksplit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float sk = unit_scale(__uint_as_float(fs->max_k_bits));
  const float2 k = kern[i];
  const float re = k.x * sk, im = k.y * sk;
  const _Float16 rh = (_Float16)re, ih = (_Float16)im;
  const _Float16 rl = (_Float16)(re - (float)rh), il = (_Float16)(im - (float)ih);
  union {
    _Float16 h[4];
    uint2 u;
  } o;
  o.h[0] = rh;
  o.h[1] = rl;
  o.h[2] = ih;
  o.h[3] = il;
  ksplit[i] = o.u;
}

__device__ __forceinline__ tw_floatx4 tw_mfma(tw_half8 a, tw_half8 b, tw_floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// One workgroup per part, in windows of MG_WIN visibilities (S <= 16).
// Per window:
//   1. stage the records in LDS, with each visibility's B row, v kx split
//      into fp16 (hi, lo) pairs (16 entries, zero beyond S), and count the
//      visibilities by bucket, the block holding their window's first cell
//      (per wavefront, each owning a contiguous range of the window);
//   2. bucket starts, and each wavefront's offset within each bucket;
//   3. stable scatter: bucket lists in visibility order (ballots per
//      bucket within each batch of 64, wavefront ranges in order);
//   4. blocks taken dynamically by the wavefronts: the visibilities that
//      can reach block (by, bx) start in buckets (by - 1 .. by, bx - 1 ..
//      bx); each is tested against the block and listed in a ring, 32 at a
//      time going through the matrix products, A from the split kernel set
//      in memory, B from LDS.
// The lists' order depends only on the data, so the sums are
// deterministic; a block is written (first window) or added to (later
// windows) by the one wavefront holding it.
__global__ __launch_bounds__(TW_NT) void k_tw_grid_jobs_mfma(
    const int4 *__restrict__ parts, int nb, const uint4 *__restrict__ rec,
    const float2 *__restrict__ svis, const float2 *__restrict__ kern,
    const uint2 *__restrict__ ksplit, int S, int64_t tsize, const TwScale *__restrict__ fs,
    float2 *__restrict__ sub, uint32_t *__restrict__ qc) {
  __shared__ uint2 s_rec[MG_WIN];       // {wx | wy << 16, ky offset}
  __shared__ uint2 s_bv[MG_WIN][16];    // B rows: v kx (hi, lo) pairs, scaled by sv sk
  __shared__ uint8_t s_key[MG_WIN];
  __shared__ uint16_t s_idx[MG_WIN];    // window positions, bucket by bucket
  // listed visibilities: {wx | wy << 7 | position << 14, ky offset}
  __shared__ uint2 s_ring[MG_NW][MG_RING];
  __shared__ int s_cnt[MG_NW][MG_NB];   // per-wavefront bucket counts, then cursors
  __shared__ int s_bstart[MG_NB + 1];
  __shared__ int s_b, s_blk;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m = lane & 15;          // A row / B column of this lane
  const int kk = 8 * (lane >> 4);   // first of the lane's 8 K slots
  const unsigned long long lt = (1ull << lane) - 1ull;
  const float sk = unit_scale(__uint_as_float(fs->max_k_bits));
  const float sv = unit_scale(__uint_as_float(fs->max_vis_bits));
  const float out_scale = 1.f / (sv * sk * sk);
  uint2 *ring = s_ring[wave];
  ring[lane] = make_uint2(0u, 0u);  // slots past the listed ones are read, then ignored
  ring[lane + 64] = make_uint2(0u, 0u);
  for (;;) {  // parts taken from the launch's queue counter qc
    if (tid == 0) s_b = (int)atomicAdd(qc, 1u);
    __syncthreads();
    const int b = s_b;
    if (b >= nb) break;
    const int4 job = parts[b];
    float2 *out = sub + (int64_t)job.x * SGN;
    for (int jb = job.y, win = 0; jb < job.z; jb += MG_WIN, win++) {
      const int n = min(job.z - jb, MG_WIN);
      const int per = (n + MG_NW - 1) / MG_NW;
      const int lo = min(n, wave * per), hi = min(n, lo + per);  // this wavefront's range
      s_cnt[tid >> 6][tid & 63] = 0;  // MG_NW x MG_NB = TW_NT counters
      __syncthreads();
      // 1. per 64 visibilities of the wavefront's range: every lane loads one
      // record and value (records, bucket counts), then the B rows are
      // built 16 lanes per visibility, the table rows' loads of eight
      // visibilities issued back to back (offsets and values taken from the
      // owner lane), so each batch waits for memory twice, not per row.
      {
        const int c = lane & 15, sub = lane >> 4;
        const int cs = min(c, S - 1);
        for (int j0 = lo; j0 < hi; j0 += 64) {
          const int j = j0 + lane;
          const int cnt = min(64, hi - j0);
          uint4 r = make_uint4(0u, 0u, 0u, 0u);
          float2 v = make_float2(0.f, 0.f);
          if (j < hi) {
            r = rec[jb + j];
            v = svis[jb + j];
            const uint32_t yf = r.y >> 16;
            s_rec[j] = make_uint2(r.x, r.z * (uint32_t)tsize + yf * (uint32_t)S);
            const int key = (int)((r.x >> 20) << 3) + (int)((r.x & 0xffffu) >> 4);  // (wy/16, wx/16)
            s_key[j] = (uint8_t)key;
            atomicAdd(&s_cnt[wave][key], 1);
          }
          const uint32_t kxo = r.z * (uint32_t)tsize + (r.y & 0xffffu) * (uint32_t)S;
          for (int st = 0; st < 16; st += 8) {
            float2 kv[8], vv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const int src = 4 * (st + q) + sub;  // owner lane of visibility j0 + src
              const uint32_t o = (uint32_t)__shfl((int)kxo, src, 64);
              vv[q] = make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
              kv[q] = kern[o + (uint32_t)cs];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const int src = 4 * (st + q) + sub;
              if (src >= cnt) continue;
              uint2 o = make_uint2(0u, 0u);
              if (c < S) {
                const float2 k = kv[q], w = vv[q];
                const float br = (w.x * k.x - w.y * k.y) * (sv * sk);
                const float bi = (w.x * k.y + w.y * k.x) * (sv * sk);
                union {
                  _Float16 h[4];
                  uint2 u;
                } h;
                h.h[0] = (_Float16)br;
                h.h[1] = (_Float16)(br - (float)h.h[0]);
                h.h[2] = (_Float16)bi;
                h.h[3] = (_Float16)(bi - (float)h.h[2]);
                o = h.u;
              }
              s_bv[j0 + src][c] = o;
            }
          }
        }
      }
      barrier_after_lds_atomics();
      // 2. bucket starts (exclusive scan over buckets) and wavefront offsets
      if (tid < MG_NB) {
        int tot = 0;
        for (int w = 0; w < MG_NW; w++) tot += s_cnt[w][tid];
        int inc = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int y = __shfl_up(inc, off, 64);
          if (lane >= off) inc += y;
        }
        int run = inc - tot;
        s_bstart[tid] = run;
        if (tid == MG_NB - 1) s_bstart[MG_NB] = inc;
        for (int w = 0; w < MG_NW; w++) {
          const int c = s_cnt[w][tid];
          s_cnt[w][tid] = run;
          run += c;
        }
      }
      __syncthreads();
      // 3. stable scatter of the wavefront's range
      for (int j0 = lo; j0 < hi; j0 += 64) {
        const int j = j0 + lane;
        bool todo = j < hi;
        const int key = todo ? (int)s_key[j] : 0;
        for (;;) {
          const unsigned long long pend = __ballot(todo);
          if (!pend) break;
          const int k =
              __builtin_amdgcn_readfirstlane(__shfl(key, __ffsll((long long)pend) - 1, 64));
          const unsigned long long mk = __ballot(todo && key == k);
          const int cur = __builtin_amdgcn_readfirstlane(s_cnt[wave][k]);
          if (todo && key == k) {
            s_idx[cur + __popcll(mk & lt)] = (uint16_t)j;
            todo = false;
          }
          if (lane == 0) s_cnt[wave][k] = cur + __popcll(mk);
        }
      }
      if (tid == 0) s_blk = 0;
      __syncthreads();
      // 4. blocks
      for (;;) {
        int blk = 0;
        if (lane == 0) blk = atomicAdd(&s_blk, 1);
        blk = __builtin_amdgcn_readfirstlane(__shfl(blk, 0, 64));
        if (blk >= MG_NB) break;
        const int by = blk >> 3, bx = blk & 7;
        const int r0 = 16 * by, c0 = 16 * bx;
        tw_floatx4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
        uint32_t head = 0, tail = 0;
        // K-group of 32 listed visibilities starting at ring position h
        // (nvalid of them real, the rest zero)
        auto group = [&](uint32_t h, int nvalid) {
          // branch-free: every slot loads (clamped addresses), then selects
          uint2 ka[8], kb[8];
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const uint2 e = ring[(h + kk + i) & (MG_RING - 1)];
            const int wx = (int)(e.x & 127u), wy = (int)((e.x >> 7) & 127u);
            const int j = (int)(e.x >> 14);
            const int ry = r0 + m - wy, cx = c0 + m - wx;
            const bool ok = kk + i < nvalid;
            const uint2 a = ksplit[e.y + (uint32_t)min(max(ry, 0), S - 1)];
            const uint2 c = s_bv[j][cx & 15];
            const bool va = ok && ry >= 0 && ry < S, vb = ok && cx >= 0 && cx < 16;
            ka[i] = make_uint2(va ? a.x : 0u, va ? a.y : 0u);
            kb[i] = make_uint2(vb ? c.x : 0u, vb ? c.y : 0u);
          }
          tw_half8 arh, arl, aih, ail, brh, brl, bih, bil;
#pragma unroll
          for (int i = 0; i < 8; i++) {
            union {
              uint2 u;
              _Float16 h[4];
            } a, c;
            a.u = ka[i];
            c.u = kb[i];
            arh[i] = a.h[0];
            arl[i] = a.h[1];
            aih[i] = a.h[2];
            ail[i] = a.h[3];
            brh[i] = c.h[0];
            brl[i] = c.h[1];
            bih[i] = c.h[2];
            bil[i] = c.h[3];
          }
          const tw_half8 naih = -aih, nail = -ail;
          cr = tw_mfma(arh, brh, cr);
          cr = tw_mfma(arh, brl, cr);
          cr = tw_mfma(arl, brh, cr);
          cr = tw_mfma(naih, bih, cr);
          cr = tw_mfma(naih, bil, cr);
          cr = tw_mfma(nail, bih, cr);
          ci = tw_mfma(arh, bih, ci);
          ci = tw_mfma(arh, bil, ci);
          ci = tw_mfma(arl, bih, ci);
          ci = tw_mfma(aih, brh, ci);
          ci = tw_mfma(aih, brl, ci);
          ci = tw_mfma(ail, brh, ci);
        };
        for (int sy = max(0, by - 1); sy <= by; sy++) {
          for (int sx = max(0, bx - 1); sx <= bx; sx++) {
            const int key = 8 * sy + sx;
            const int p0 = s_bstart[key], p1 = s_bstart[key + 1];
            for (int q0 = p0; q0 < p1; q0 += 64) {
              const int q = q0 + lane;
              bool mem = false;
              int j = 0;
              uint2 rj = make_uint2(0u, 0u);
              if (q < p1) {
                j = s_idx[q];
                rj = s_rec[j];
                const int wx = (int)(rj.x & 0xffffu), wy = (int)(rj.x >> 16);
                mem = wy < r0 + 16 && wy + S > r0 && wx < c0 + 16 && wx + S > c0;
                rj.x = (uint32_t)wx | ((uint32_t)wy << 7) | ((uint32_t)j << 14);
              }
              const unsigned long long bal = __ballot(mem);
              if (mem) ring[(tail + __popcll(bal & lt)) & (MG_RING - 1)] = rj;
              tail += (uint32_t)__popcll(bal);
              while (tail - head >= 32) {
                group(head, 32);
                head += 32;
              }
            }
          }
        }
        if (tail > head) group(head, (int)(tail - head));
        // D element (row 4 (lane / 16) + q, column lane % 16)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          float2 *o = out + (r0 + 4 * (lane >> 4) + q) * SG + c0 + m;
          float2 val = make_float2(cr[q] * out_scale, ci[q] * out_scale);
          if (win) {
            const float2 old = *o;
            val.x += old.x;
            val.y += old.y;
          }
          *o = val;
        }
      }
      __syncthreads();  // the window's records are read; the next one rewrites them
    }
  }
}

// sub[first] += sub[first + 1 .. first + n - 1], parts in order: multi[m] =
// (first slot, n) of a job gridded in n parts.  blockIdx.y = job,
// blockIdx.x = slice of the pixels (two per thread).
constexpr int ACC_SLICES = 16;
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__global__ void k_tw_sum_parts(const int2 *__restrict__ multi, float2 *__restrict__ sub) {
  const int2 m = multi[blockIdx.y];
  float4 *dst = (float4 *)(sub + (int64_t)m.x * SGN);
  constexpr int PER = SGN / 2 / ACC_SLICES;
  for (int i = blockIdx.x * PER + threadIdx.x; i < (blockIdx.x + 1) * PER; i += blockDim.x) {
    float4 acc = dst[i];
    int k = 1;
    // four parts in flight, added in part order
    for (; k + 4 <= m.y; k += 4) {
      float4 t[4];
#pragma unroll
      for (int q = 0; q < 4; q++) t[q] = ((const float4 *)(sub + (int64_t)(m.x + k + q) * SGN))[i];
#pragma unroll
      for (int q = 0; q < 4; q++) acc = add4(acc, t[q]);
    }
    for (; k < m.y; k++) acc = add4(acc, ((const float4 *)(sub + (int64_t)(m.x + k) * SGN))[i]);
    dst[i] = acc;
  }
}

// cimg[chunk] += sum of the batch's job images of that chunk, jobs in order
// (deterministic).  groups[g] = (chunk id, first job, end job) of the batch,
// jslot[job] = its sub-grid slot; blockIdx.y = group, blockIdx.x = slice of
// the pixels (two per thread), so every group spreads over ACC_SLICES
// workgroups.
__global__ void k_tw_chunk_accum(const int4 *__restrict__ groups, const int32_t *__restrict__ jslot,
                                 const float2 *__restrict__ sub, float2 *__restrict__ cimg) {
  const int4 g = groups[blockIdx.y];
  float4 *dst = (float4 *)(cimg + (int64_t)g.x * SGN);
  const float4 *src = (const float4 *)sub;
  constexpr int PER = SGN / 2 / ACC_SLICES;  // float4 (pixel pairs) per slice
  for (int i = blockIdx.x * PER + threadIdx.x; i < (blockIdx.x + 1) * PER; i += blockDim.x) {
    float4 acc = dst[i];
    int b = g.y;
    // eight job images in flight, added in job order
    for (; b + 8 <= g.z; b += 8) {
      float4 s[8];
#pragma unroll
      for (int k = 0; k < 8; k++) s[k] = src[(int64_t)jslot[b + k] * (SGN / 2) + i];
#pragma unroll
      for (int k = 0; k < 8; k++) acc = add4(acc, s[k]);
    }
    for (; b < g.z; b++) acc = add4(acc, src[(int64_t)jslot[b] * (SGN / 2) + i]);
    dst[i] = acc;
  }
}

// Jobs of at most TW_SMALL_VIS visibilities are gridded by the summing
// kernel itself (separable kernels of support S <= 16), never through a
// sub-grid slot; their records are staged TW_STAGE at a time.
#ifndef TW_SMALL_VIS
#define TW_SMALL_VIS 128
#endif
constexpr int TW_STAGE = 64;
// fewest lanes per visibility in the fused predict's degridding rounds
#ifndef TW_DEG_TMIN
#define TW_DEG_TMIN 8
#endif

// Job images summed per chunk without writing them out.  runs[r] = (chunk,
// first, end, out).  Each job's image
//   ifft(G) * exp(-i kw[job] phi)
// is formed with the arithmetic of k_fft2d_128 (rows then columns, the
// 1 / 128^2 scale, centring sign and phase applied as the last column pass
// leaves LDS) and added, in job order, to the thread's 16 output pixels in
// registers (the last pass gives every thread the same pixels for every
// job).  out < 0: cimg[chunk] += the run's sum; otherwise the sum goes to
// sub[out], which only this run touches (a later k_tw_chunk_accum adds
// those sums to cimg in run order).  Runs are taken from the queue counter
// qc.  The run's jobs are jl[first .. end - 1]; the job sub-grid G:
//   SMALL = false: G = sub[jslot[job]] as the gridding kernels left it; the
//     next job's rows are loaded during the current transform;
//   SMALL = true: (jobs[] the record ranges, at most TW_SMALL_VIS records
//     each) TW_STAGE records at a time: their v kx
//     and ky rows staged in LDS, then each thread sums the taps falling on
//     its own row elements (row y, columns x = j2 mod 8: at most two per
//     window) into the LDS sub-grid in record order,
//       G[wy + dy][wx + dx] += (v kx[dx]) ky[dy]
//     (the taps of synthesis.py:456), before the transform reads them.
template <bool SMALL>
__global__ __launch_bounds__(FFT_NT) void k_tw_image_runs(
    const int4 *__restrict__ runs, int nruns, const int32_t *__restrict__ jl,
    const int32_t *__restrict__ jslot, const int32_t *__restrict__ kw,
    const double *__restrict__ phi, float2 *sub, float2 *__restrict__ cimg,
    uint32_t *__restrict__ qc, const int4 *__restrict__ jobs, const uint4 *__restrict__ rec,
    const float2 *__restrict__ svis, const float2 *__restrict__ kern, int S, int64_t tsize) {
  constexpr int NS = SMALL ? TW_STAGE : 1;
  __shared__ float2 lds[FN * FP];
  __shared__ float2 tw[FN];
  __shared__ float2 s_vk[NS][16], s_ky[NS][16];  // v kx and ky rows (zero beyond S)
  __shared__ uint32_t s_pos[NS];                 // window corners wx | wy << 16
  __shared__ int s_need[SMALL ? FN : 1];         // small jobs: rows holding taps,
  __shared__ int s_rows[SMALL ? FN : 1];         // listed in order
  __shared__ int s_nrows, s_run;
  const int tid = threadIdx.x;
  constexpr float sgn = 1.0f;  // inverse transform
  constexpr float out_scale = 1.0f / (float)SGN;
  fft128_twiddles(tw);
  if (SMALL && tid < FN) s_need[tid] = 0;
  const int y = tid >> 3, j2 = tid & 7;
  for (;;) {
    if (tid == 0) s_run = (int)atomicAdd(qc, 1u);
    __syncthreads();
    const int run = s_run;
    if (run >= nruns) break;
    const int4 rn = runs[run];
    float2 acc[PASS_B][8];
#pragma unroll
    for (int r = 0; r < PASS_B; r++)
#pragma unroll
      for (int k2 = 0; k2 < 8; k2++) acc[r][k2] = make_float2(0.f, 0.f);
    float2 nxt[SMALL ? 1 : 16];  // the thread's row elements of the next job's sub-grid
    if constexpr (!SMALL) {
      const float2 *row = sub + (int64_t)jslot[jl[rn.y]] * SGN + y * FN + j2;
#pragma unroll
      for (int j1 = 0; j1 < 16; j1++) nxt[j1] = row[8 * j1];
    }
    for (int b = rn.y; b < rn.z; b++) {
      int zero = 0;
      asm volatile("" : "+s"(zero));
      const float2 *twz = tw + zero;
      const int job = jl[b];
      int nrows = FN;  // rows through the row passes (small jobs: those holding taps)
      if constexpr (SMALL) {
        const int4 jr = jobs[job];
        const int n = jr.z - jr.y;
        __syncthreads();  // tw written; the previous job's LDS reads done
        // the thread's own row elements (row y, columns 8 j1 + j2, centring
        // sign (-1)^(y + j2)) cleared, then summed in record order
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) lds[y * FP + 8 * j1 + j2] = make_float2(0.f, 0.f);
        const float neg = ((y + j2) & 1) ? -1.0f : 1.0f;
        for (int c0 = 0; c0 < n; c0 += TW_STAGE) {
          const int m = min(TW_STAGE, n - c0);
          if (c0 > 0) __syncthreads();  // the previous records' reads done
          if (tid < 16 * m) {
            const int i = tid >> 4, c = tid & 15;
            const uint4 r = rec[jr.y + c0 + i];
            const float2 w = svis[jr.y + c0 + i];
            const float2 *kt = kern + (int64_t)r.z * tsize;
            float2 kx = make_float2(0.f, 0.f), ky = make_float2(0.f, 0.f);
            if (c < S) {
              kx = kt[(int64_t)(r.y & 0xffffu) * S + c];
              ky = kt[(int64_t)(r.y >> 16) * S + c];
              s_need[(int)(r.x >> 16) + c] = 1;  // the window's rows
            }
            s_vk[i][c] = cmul(w, kx);
            s_ky[i][c] = ky;
            if (c == 0) s_pos[i] = r.x;
          }
          __syncthreads();
          for (int i = 0; i < m; i++) {
            const uint32_t pos = s_pos[i];
            const int wx = (int)(pos & 0xffffu), dy = y - (int)(pos >> 16);
            if ((unsigned)dy < (unsigned)S) {
              const float2 kyv = s_ky[i][dy];
              const int o = (j2 - wx) & 7;  // first window column of this thread
              float2 *cell = lds + y * FP + wx + o;
              if (o < S) {
                const float2 t = cmul(s_vk[i][o], kyv);
                cell[0] = make_float2(cell[0].x + neg * t.x, cell[0].y + neg * t.y);
              }
              if (o + 8 < S) {
                const float2 t = cmul(s_vk[i][o + 8], kyv);
                cell[8] = make_float2(cell[8].x + neg * t.x, cell[8].y + neg * t.y);
              }
            }
          }
        }
        if (tid < 64) {
          // rows holding taps listed in order (ballots over two rows per
          // lane); this wavefront alone reads the marks, and clears them
          const int f0 = s_need[2 * tid], f1 = s_need[2 * tid + 1];
          const unsigned long long m0 = __ballot(f0 != 0), m1 = __ballot(f1 != 0);
          const unsigned long long lt = (1ull << tid) - 1ull;
          const int pos = __popcll(m0 & lt) + __popcll(m1 & lt);
          if (f0) s_rows[pos] = 2 * tid;
          if (f1) s_rows[pos + (f0 ? 1 : 0)] = 2 * tid + 1;
          if (tid == 0) s_nrows = __popcll(m0) + __popcll(m1);
          s_need[2 * tid] = 0;
          s_need[2 * tid + 1] = 0;
        }
        __syncthreads();
        nrows = s_nrows;
        rows_pass_a(lds, s_rows, nrows, twz, sgn);  // rows without taps stay zero
      } else {
        float2 v[16];
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) v[j1] = nxt[j1];
        __syncthreads();  // tw written; the previous job's LDS reads done
        const float neg = ((y + j2) & 1) ? -1.0f : 1.0f;
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) v[j1] = make_float2(neg * v[j1].x, neg * v[j1].y);
        fft_reg<16>(v, sgn);
        float2 *p = lds + y * FP + j2;
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) {
          const float2 w = twz[(j2 * k1) & 127];
          p[8 * k1] = cmul(v[k1], make_float2(w.x, sgn * w.y));
        }
        if (b + 1 < rn.z) {
          const float2 *row = sub + (int64_t)jslot[jl[b + 1]] * SGN + y * FN + j2;
#pragma unroll
          for (int j1 = 0; j1 < 16; j1++) nxt[j1] = row[8 * j1];
        }
      }
      __syncthreads();
      {
        // row pass B, written back to the natural positions once all are read
        float2 v[PASS_B][8];
        rows_pass_b(lds, SMALL ? s_rows : nullptr, nrows, sgn, v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PASS_B; r++) {
          int k, k1;
          rows_pass_b_item(r, k, k1);
          if (k < nrows) {
            const int line = SMALL ? s_rows[k] : k;
#pragma unroll
            for (int k2 = 0; k2 < 8; k2++) lds[line * FP + k1 + 16 * k2] = v[r][k2];
          }
        }
      }
      __syncthreads();
      cols_pass_a(lds, twz, sgn);  // columns
      __syncthreads();
      {
        // column pass B: item (column x, k1) holds rows y = k1 + 16 k2,
        // which share the parity of x + k1
        const double kout = -(double)kw[job];
#pragma unroll
        for (int r = 0; r < PASS_B; r++) {
          int x, k1;
          cols_pass_b_item(r, x, k1);
          double ph[8];
#pragma unroll
          for (int k2 = 0; k2 < 8; k2++) ph[k2] = phi[zero + (k1 + 16 * k2) * FN + x];
          float2 v[8];
          const float2 *p = lds + x + 8 * k1 * FP;
#pragma unroll
          for (int jj = 0; jj < 8; jj++) v[jj] = p[jj * FP];
          fft_reg<8>(v, sgn);
          const float sc = ((x + k1) & 1) ? -out_scale : out_scale;
#pragma unroll
          for (int k2 = 0; k2 < 8; k2++) {
            const float2 t = rotate(make_float2(v[k2].x * sc, v[k2].y * sc), kout, ph[k2]);
            acc[r][k2] = make_float2(acc[r][k2].x + t.x, acc[r][k2].y + t.y);
          }
        }
      }
    }
    float2 *dst = rn.w < 0 ? cimg + (int64_t)rn.x * SGN : sub + (int64_t)rn.w * SGN;
#pragma unroll
    for (int r = 0; r < PASS_B; r++) {
      int x, k1;
      cols_pass_b_item(r, x, k1);
#pragma unroll
      for (int k2 = 0; k2 < 8; k2++) {
        const int pix = (k1 + 16 * k2) * FN + x;
        float2 o = acc[r][k2];
        if (rn.w < 0) {
          const float2 c = dst[pix];
          o = make_float2(c.x + o.x, c.y + o.y);
        }
        dst[pix] = o;
      }
    }
  }
}

// grid[window] += chunk_sub[chunk] (clipped); chunk ids of one parity class.
__global__ void k_tw_add(const int32_t *__restrict__ chunks, const float2 *__restrict__ chunk_sub,
                         float2 *__restrict__ grid, int N, int chunk, int margin, int ncx) {
  const int c = chunks[blockIdx.y];
  const int cx = c % ncx, cy = c / ncx;
  const int x0 = cx * chunk - margin / 2, y0 = cy * chunk - margin / 2;
  const float2 *s = chunk_sub + (int64_t)c * SGN;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SGN; i += gridDim.x * blockDim.x) {
    const int y = y0 + i / SG, x = x0 + i % SG;
    if (x < 0 || y < 0 || x >= N || y >= N) continue;
    float2 g = grid[(int64_t)y * N + x];
    g.x += s[i].x;
    g.y += s[i].y;
    grid[(int64_t)y * N + x] = g;
  }
}

// Non-finite values: a NaN or infinite visibility makes its job's sub-grid
// image non-finite after the inverse transform, and with it the chunk's
// stacked image and the chunk's whole sub-grid region of the grid
// (wtowers.ipynb imaging cell, grid.c:604-678, in FP64).  The passes below
// mark such chunks and set their regions to NaN after the adds; a
// non-finite kernel value marks every chunk.  They run only when
// k_tw_maxima flagged a value.
__global__ void k_tw_nonfinite_mark(const int4 *__restrict__ jobs, int njobs, int nwp,
                                    const float2 *__restrict__ svis,
                                    const TwScale *__restrict__ fs, uint32_t *__restrict__ cflag) {
  const unsigned flags = fs->nonfinite;
  if (flags == 0) return;
  for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
    const int4 job = jobs[j];
    const int c = job.x / nwp;
    if (flags & 2u) {
      if (threadIdx.x == 0) cflag[c] = 1u;
      continue;
    }
    for (int i = job.y + (int)threadIdx.x; i < job.z; i += blockDim.x) {
      if (!finite2(svis[i])) {
        cflag[c] = 1u;
        break;
      }
    }
  }
}

__global__ void k_tw_nonfinite_fill(const int32_t *__restrict__ chunks,
                                    const uint32_t *__restrict__ cflag, float2 *__restrict__ grid,
                                    int N, int chunk, int margin, int ncx,
                                    const TwScale *__restrict__ fs) {
  if (fs->nonfinite == 0) return;
  const int c = chunks[blockIdx.y];
  if (!cflag[c]) return;
  const int cx = c % ncx, cy = c / ncx;
  const int x0 = cx * chunk - margin / 2, y0 = cy * chunk - margin / 2;
  const float nan = __builtin_nanf("");
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SGN; i += gridDim.x * blockDim.x) {
    const int y = y0 + i / SG, x = x0 + i % SG;
    if (x < 0 || y < 0 || x >= N || y >= N) continue;
    grid[(int64_t)y * N + x] = make_float2(nan, nan);
  }
}

// cimg[chunk] = grid window (zero outside the grid).
__global__ void k_tw_cut(const int32_t *__restrict__ chunks, const float2 *__restrict__ grid,
                         float2 *__restrict__ cimg, int N, int chunk, int margin, int ncx) {
  const int c = chunks[blockIdx.y];
  const int cx = c % ncx, cy = c / ncx;
  const int x0 = cx * chunk - margin / 2, y0 = cy * chunk - margin / 2;
  float2 *d = cimg + (int64_t)c * SGN;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SGN; i += gridDim.x * blockDim.x) {
    const int y = y0 + i / SG, x = x0 + i % SG;
    d[i] = (x < 0 || y < 0 || x >= N || y >= N) ? make_float2(0.f, 0.f)
                                                 : grid[(int64_t)y * N + x];
  }
}

// Separable degridding from a sub-grid window in LDS (row pitch `pitch`),
// the taps k = ky[y] kx[x] of synthesis.py:456, in FP32:
//   v = sum_y ky[y] r[y],  r[y] = sum_x kx[x] G[y][x].
// Support 16: r[y] summed over x in order; the 16 products ky[y] r[y]
// summed as a binary tree ((p0 + p1) + (p2 + p3)) + ...  degrid_rows16<R>
// gives the subtree of R consecutive rows (R = 16 / lanes per visibility),
// the lanes' subtrees then being added by shuffles, so every split gives
// the same value bit for bit.  Rows kx and ky start at multiples of 16
// elements: 16-byte loads.
__device__ __forceinline__ float2 tap_row16(const float2 *g, const float2 (&kxv)[16]) {
  float2 gv[16];  // the row's cells read before the products: reads in flight together
#pragma unroll
  for (int x = 0; x < 16; x++) gv[x] = g[x];
  float rr = 0.f, ri = 0.f;
#pragma unroll
  for (int x = 0; x < 16; x++) {
    rr = fmaf(gv[x].x, kxv[x].x, fmaf(-gv[x].y, kxv[x].y, rr));
    ri = fmaf(gv[x].x, kxv[x].y, fmaf(gv[x].y, kxv[x].x, ri));
  }
  return make_float2(rr, ri);
}

template <int R>
__device__ __forceinline__ float2 degrid_rows16(const float2 *g, int pitch, const float2 *kx,
                                                const float2 *ky) {
  float2 kxv[16], kyv[R];
  const float4 *kx4 = (const float4 *)kx;
#pragma unroll
  for (int x = 0; x < 8; x++) {
    const float4 a = kx4[x];
    kxv[2 * x] = make_float2(a.x, a.y);
    kxv[2 * x + 1] = make_float2(a.z, a.w);
  }
  if constexpr (R == 1) {
    kyv[0] = ky[0];
  } else {
    const float4 *ky4 = (const float4 *)ky;
#pragma unroll
    for (int y = 0; y < R / 2; y++) {
      const float4 b = ky4[y];
      kyv[2 * y] = make_float2(b.x, b.y);
      kyv[2 * y + 1] = make_float2(b.z, b.w);
    }
  }
  // the tree built as a binary counter: acc[l] holds a finished subtree of
  // 2^l rows waiting for its right neighbour
  float2 acc[5];
#pragma unroll
  for (int y = 0; y < R; y++) {
    const float2 t = tap_row16(g + y * pitch, kxv);
    float2 p = make_float2(fmaf(t.x, kyv[y].x, -(t.y * kyv[y].y)),
                           fmaf(t.x, kyv[y].y, t.y * kyv[y].x));
#pragma unroll
    for (int l = 0; l < 5; l++) {
      if ((y >> l) & 1) {
        p = make_float2(acc[l].x + p.x, acc[l].y + p.y);
      } else {
        acc[l] = p;
        break;
      }
    }
  }
  constexpr int top = R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : R == 2 ? 1 : 0;
  return acc[top];
}

// Any support S <= 16: rows and cells in order, v accumulated row by row.
__device__ __forceinline__ float2 degrid_sep_lds(const float2 *g, int pitch, const float2 *kx,
                                                 const float2 *ky, int S) {
  float2 kxv[16], kyv[16];
#pragma unroll
  for (int x = 0; x < 16; x++) {
    kxv[x] = x < S ? kx[x] : make_float2(0.f, 0.f);
    kyv[x] = x < S ? ky[x] : make_float2(0.f, 0.f);
  }
  float ar = 0.f, ai = 0.f;
#pragma unroll
  for (int y = 0; y < 16; y++) {
    if (y < S) {
      float rr = 0.f, ri = 0.f;
#pragma unroll
      for (int x = 0; x < 16; x++) {
        if (x < S) {
          const float2 gv = g[y * pitch + x];
          rr = fmaf(gv.x, kxv[x].x, fmaf(-gv.y, kxv[x].y, rr));
          ri = fmaf(gv.x, kxv[x].y, fmaf(gv.y, kxv[x].x, ri));
        }
      }
      ar = fmaf(rr, kyv[y].x, fmaf(-ri, kyv[y].y, ar));
      ai = fmaf(rr, kyv[y].y, fmaf(ri, kyv[y].x, ai));
    }
  }
  return make_float2(ar, ai);
}

// One round of support-16 degridding with T lanes per visibility: the
// visibilities j0 + tid / T (below j1), lane tid % T taking window rows
// [R (tid % T), R (tid % T + 1)), R = 16 / T.  Every lane takes part in
// the shuffles.
template <int T>
__device__ __forceinline__ void degrid_round16(const float2 *lds, int j0, int j1,
                                               const uint4 *__restrict__ rec,
                                               const float2 *__restrict__ kern, int64_t tsize,
                                               float2 *__restrict__ vis_out) {
  constexpr int R = 16 / T;
  const int tid = threadIdx.x, j = j0 + tid / T, sub = tid % T;
  float2 v = make_float2(0.f, 0.f);
  uint4 r = make_uint4(0u, 0u, 0u, 0u);
  if (j < j1) {
    r = rec[j];
    const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
    const int wx = (int)(r.x & 0xffffu), wy = (int)(r.x >> 16);
    const float2 *kt = kern + (int64_t)r.z * tsize;
    v = degrid_rows16<R>(lds + (wy + R * sub) * FP + wx, FP, kt + (int64_t)xf * 16,
                         kt + (int64_t)yf * 16 + R * sub);
  }
#pragma unroll
  for (int off = 1; off < T; off <<= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  if (j < j1 && sub == 0) store_stream(vis_out + r.w, v);
}

// Predict with a separable kernel of support S <= 16 (SF = 16: S fixed at
// 16, degrid_round16; SF = 0: any S, degrid_sep_lds), the plane sub-grids
// never leaving LDS.  items[i] = (chunk, w-plane wp, first record, end
// record); a run is the items roff[r] .. roff[r + 1] - 1, runs taken from
// the queue counter qc by persistent workgroups.  Per item:
//   1. the sub-grid fft(chunk image * exp(+i wp phi)) formed in LDS with the
//      arithmetic of k_fft2d_128 (centring sign and phase applied to the
//      input as it is read, the output sign as the last pass writes back),
//      columns first, then the rows the item's windows cover (all rows
//      when it holds more than FFT_NT visibilities); the chunk image's
//      columns are read at each item (consecutive items mostly share it);
//   2. the item's visibilities degridded from LDS, several lanes per
//      visibility; the window lies inside the sub-grid (S <= margin,
//      checked at binning).
template <int SF>
__global__ __launch_bounds__(FFT_NT) void k_tw_predict_sep(
    const int4 *__restrict__ items, const int32_t *__restrict__ roff, int nruns,
    const float2 *__restrict__ cimg, const double *__restrict__ phi,
    const uint4 *__restrict__ rec, const float2 *__restrict__ kern, int S, int64_t tsize,
    float2 *__restrict__ vis_out, uint32_t *__restrict__ qc) {
  __shared__ float2 lds[FN * FP];
  __shared__ float2 tw[FN];
  __shared__ int s_need[FN];  // rows some window of the item covers
  __shared__ int s_rows[FN];  // those rows, ascending
  __shared__ int s_nrows, s_run;
  const int tid = threadIdx.x;
  constexpr float sgn = -1.0f;  // forward transform
  fft128_twiddles(tw);
  if (tid < FN) s_need[tid] = 0;
  // Column pass A: the thread owns column x and its elements y = 8 j1 + c2,
  // which share the parity of x + c2 (the (-1)^(x + y) centring sign).
  const int x = tid & (FN - 1), c2 = tid >> 7;
  for (;;) {
    if (tid == 0) s_run = (int)atomicAdd(qc, 1u);
    __syncthreads();
    const int run = s_run;
    if (run >= nruns) break;
    const int i1 = roff[run + 1];
    for (int i = roff[run]; i < i1; i++) {
      const int4 it = items[i];
      // an offset the compiler cannot see through keeps the twiddle and
      // phase-table loads inside the loop (as in k_fft2d_128)
      int zero = 0;
      asm volatile("" : "+s"(zero));
      const float2 *twz = tw + zero;
      {
        float2 v[16];
        double ph[16];
        const float2 *col = cimg + (int64_t)it.x * SGN + c2 * FN + x;
        const double *pcol = phi + zero + c2 * FN + x;
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) {
          v[j1] = col[8 * j1 * FN];
          ph[j1] = pcol[8 * j1 * FN];
        }
        __syncthreads();  // tw written; the previous item's LDS reads done
        // the rows the item's windows cover (s_need was cleared after the
        // previous item's row passes)
        const int n = it.w - it.z;
        if (n > FFT_NT) {
          if (tid < FN) s_need[tid] = 1;
        } else if (tid < n) {
          const int wy = (int)(rec[it.z + tid].x >> 16);
          for (int k = 0; k < S; k++) s_need[wy + k] = 1;
        }
        const float neg = ((x + c2) & 1) ? -1.0f : 1.0f;
        const double kin = (double)it.y;
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) {
          const float2 t = rotate(v[j1], kin, ph[j1]);
          v[j1] = make_float2(neg * t.x, neg * t.y);
        }
        fft_reg<16>(v, sgn);
        float2 *p = lds + x + c2 * FP;
#pragma unroll
        for (int k1 = 0; k1 < 16; k1++) {
          const float2 w = twz[(c2 * k1) & 127];
          p[8 * k1 * FP] = cmul(v[k1], make_float2(w.x, sgn * w.y));
        }
      }
      __syncthreads();
      if (tid < 64) {
        // the covered rows listed in order (ballots over two rows per lane)
        const int f0 = s_need[2 * tid], f1 = s_need[2 * tid + 1];
        const unsigned long long m0 = __ballot(f0 != 0), m1 = __ballot(f1 != 0);
        const unsigned long long lt = (1ull << tid) - 1ull;
        const int pos = __popcll(m0 & lt) + __popcll(m1 & lt);
        if (f0) s_rows[pos] = 2 * tid;
        if (f1) s_rows[pos + (f0 ? 1 : 0)] = 2 * tid + 1;
        if (tid == 0) s_nrows = __popcll(m0) + __popcll(m1);
      }
      {
        // column pass B, written back to the natural positions once all are read
        float2 v[PASS_B][8];
        cols_pass_b(lds, sgn, v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PASS_B; r++) {
          int line, k1;
          cols_pass_b_item(r, line, k1);
#pragma unroll
          for (int k2 = 0; k2 < 8; k2++) lds[(k1 + 16 * k2) * FP + line] = v[r][k2];
        }
      }
      __syncthreads();
      {
        // row passes on the listed rows, 32 rows per half-wavefront
        rows_pass_a(lds, s_rows, s_nrows, twz, sgn);
      }
      __syncthreads();
      {
        // row pass B back to LDS with the output sign: item (row y, k1)
        // holds columns k1 + 16 k2, which share the parity of y + k1
        const int nrows = s_nrows;
        float2 v[PASS_B][8];
        rows_pass_b(lds, s_rows, nrows, sgn, v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PASS_B; r++) {
          int k, k1;
          rows_pass_b_item(r, k, k1);
          if (k < nrows) {
            const int line = s_rows[k];
            const float sc = ((line + k1) & 1) ? -1.0f : 1.0f;
#pragma unroll
            for (int k2 = 0; k2 < 8; k2++)
              lds[line * FP + k1 + 16 * k2] = make_float2(v[r][k2].x * sc, v[r][k2].y * sc);
          }
        }
      }
      __syncthreads();
      if (tid < FN) s_need[tid] = 0;  // read again only after the next item's first barrier
      if constexpr (SF == 16) {
        // rounds of FFT_NT / T visibilities with T lanes each: at least
        // TW_DEG_TMIN (the lanes of one visibility read different window
        // rows, on different LDS banks), more when few visibilities are left
        for (int j0 = it.z; j0 < it.w;) {
          const int rem = it.w - j0;
          if (TW_DEG_TMIN <= 1 && rem > FFT_NT / 2) {
            degrid_round16<1>(lds, j0, it.w, rec, kern, tsize, vis_out);
            j0 += FFT_NT;
          } else if (TW_DEG_TMIN <= 2 && rem > FFT_NT / 4) {
            degrid_round16<2>(lds, j0, it.w, rec, kern, tsize, vis_out);
            j0 += FFT_NT / 2;
          } else if (TW_DEG_TMIN <= 4 && rem > FFT_NT / 8) {
            degrid_round16<4>(lds, j0, it.w, rec, kern, tsize, vis_out);
            j0 += FFT_NT / 4;
          } else if (TW_DEG_TMIN <= 8 && rem > FFT_NT / 16) {
            degrid_round16<8>(lds, j0, it.w, rec, kern, tsize, vis_out);
            j0 += FFT_NT / 8;
          } else {
            degrid_round16<16>(lds, j0, it.w, rec, kern, tsize, vis_out);
            j0 += FFT_NT / 16;
          }
        }
      } else {
        for (int j = it.z + tid; j < it.w; j += FFT_NT) {
          const uint4 r = rec[j];
          const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
          const int wx = (int)(r.x & 0xffffu), wy = (int)(r.x >> 16);
          const float2 *kt = kern + (int64_t)r.z * tsize;
          store_stream(vis_out + r.w, degrid_sep_lds(lds + wy * FP + wx, FP,
                                                     kt + (int64_t)xf * S,
                                                     kt + (int64_t)yf * S, S));
        }
      }
    }
  }
}

// Degrid job parts from their sub-grids: parts[b] = (slot, first record,
// end record), one workgroup per part staging sub[slot] in LDS.
template <bool SEPARABLE>
__global__ __launch_bounds__(TW_NT) void k_tw_degrid_jobs(
    const int4 *__restrict__ parts, int nb, const uint4 *__restrict__ rec,
    const float2 *__restrict__ sub, const float2 *__restrict__ kern, int Q, int S,
    int64_t tsize, int nearest, float2 *__restrict__ vis_out, uint32_t *__restrict__ qc) {
  // Rows padded by 128 bytes (DEG_PITCH cells): the bare 1 KiB pitch would
  // put the rows of a 64-tap block on the same LDS banks.
  constexpr int DEG_PITCH = SG + 16;
  __shared__ float2 sg[SG * DEG_PITCH];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int NW = TW_NT / 64;
  const int SS = S * S;
  __shared__ int s_b;
  for (;;) {  // parts taken from the launch's queue counter qc
    if (tid == 0) s_b = (int)atomicAdd(qc, 1u);
    __syncthreads();
    const int b = s_b;
    if (b >= nb) break;
    const int4 job = parts[b];
    const float2 *src = sub + (int64_t)job.x * SGN;
    // stage the sub-grid, eight streaming loads in flight per thread
    static_assert(SGN % (8 * TW_NT) == 0, "staging covers the sub-grid exactly");
    for (int i0 = tid; i0 < SGN; i0 += 8 * TW_NT) {
      float2 gv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) gv[k] = load_stream(src + i0 + k * TW_NT);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int i = i0 + k * TW_NT;
        sg[(i / SG) * DEG_PITCH + i % SG] = gv[k];
      }
    }
    __syncthreads();
    if (nearest) {
      for (int j = job.y + tid; j < job.z; j += TW_NT) {
        const uint4 r = rec[j];
        vis_out[r.w] = sg[(int)(r.x >> 16) * DEG_PITCH + (int)(r.x & 0xffffu)];
      }
    } else {
      const int jbeg = __builtin_amdgcn_readfirstlane(job.y);
      const int jend = __builtin_amdgcn_readfirstlane(job.z);
      // records of the wave's visibilities j = first + wave, + NW, ...
      // fetched 64 at a time and handed out with readlane
      for (int jb = jbeg + wave; jb < jend; jb += 64 * NW) {
        const int jl = jb + lane * NW;
        const uint4 rb = jl < jend ? rec[jl] : make_uint4(0, 0, 0, 0);
        const int nbat = min(64, (jend - jb + NW - 1) / NW);
        for (int it = 0; it < nbat; it++) {
          const uint4 r = lane_u4(rb, it);
          const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
          const int wy = (int)(r.x >> 16);
          const int base = wy * SG + (int)(r.x & 0xffffu);
          const float2 *kt = kern + (int64_t)r.z * tsize;  // its w-kernel plane
          const float2 *k2 = kt + (int64_t)(yf * Q + xf) * SS;
          const float2 *ky = kt + (int64_t)yf * S, *kx = kt + (int64_t)xf * S;
          float sr = 0.f, si = 0.f;
          gather_range<SEPARABLE>(sg, base, SG, S, 0, SS, lane, k2, ky, kx, sr, si,
                                  8u * (DEG_PITCH - SG), 8u * (DEG_PITCH - SG) * (uint32_t)wy);
          sr = wave_sum_to_63(sr);
          si = wave_sum_to_63(si);
          if (lane == 63) store_stream(vis_out + r.w, make_float2(sr, si));
        }
      }
    }
    __syncthreads();
  }
}

// phi[pixel] = 2 pi wstep (1 - sqrt(1 - l^2 - m^2)) on the sub-grid's image
// coordinates l = theta (x - sg/2) / sg (kernel_coordinates, synthesis.py:250).
__global__ void k_tw_phi(double *phi, double theta, double wstep) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= SGN) return;
  const int y = i / SG, x = i % SG;
  const double l = ((double)(x - SG / 2) / (double)SG) * theta;
  const double m = ((double)(y - SG / 2) / (double)SG) * theta;
  const double r2 = l * l + m * m;
  phi[i] = r2 < 1.0 ? 2.0 * M_PI * wstep * (1.0 - sqrt(1.0 - r2)) : 0.0;
}

// Gridding schedule of a binning, made on the host once and kept on the
// device: jobs cut into parts of at most PART_VIS visibilities, batches of
// whole jobs holding at most JOB_BATCH part slots.  Per batch: its parts
// (slot, first, end) and multi-part jobs (first slot, parts); per job its
// w-plane and slot; the chunks in four parity classes for the final adds.
//
// A batch's job images are summed per chunk by k_tw_image_runs over runs
// (chunk, first, end, out) of at most `per` of its jobs of one chunk
// (entries first .. end - 1 of the job list jl), longest runs first; a
// chunk's later runs leave their sums in slots listed in pslot, added to
// the chunk image in run order by k_tw_chunk_accum over rgroups (chunk,
// first, end).  With `split`, jobs of at most TW_SMALL_VIS visibilities are
// left out of the batches and summed by the small-job form of
// k_tw_image_runs before them: runs small.r0 .. + small.nr, partial sums in
// slots 0, 1, ... (the batches' slots are not yet in use).
struct GridSched {
  struct Batch {
    size_t p0, np, m0, nm, r0, nr, q0, nq;
  };
  std::vector<Batch> batches;
  Batch small{};  // split: the small jobs' runs and partial-sum groups
  bool split = false;
  int max_slots = 0;
  int32_t *kw = nullptr, *chunks = nullptr, *jslot = nullptr, *par = nullptr, *pslot = nullptr;
  int32_t *jl = nullptr;
  int4 *parts = nullptr, *runs = nullptr, *rgroups = nullptr;
  int2 *multi = nullptr;
  size_t npar[4] = {0, 0, 0, 0};
};

// Degridding schedule: batches of at most JOB_BATCH jobs (job b of a batch in
// slot b), every job cut into parts of at most PART_VIS visibilities sharing
// its slot; per job its w-plane and chunk.  For k_tw_predict_sep: the jobs
// as items (chunk, w-plane, first, end) in job order, cut into nruns runs
// of about equal cost (items roff[r] .. roff[r + 1] - 1); a job that straddles
// a run boundary becomes one item in each run, each transforming it.
struct DegridSched {
  int jb = 0;
  std::vector<size_t> poff;  // first part of each batch, and the end
  int32_t *kw = nullptr, *jchunk = nullptr, *chunks = nullptr;
  int4 *parts = nullptr;
  int4 *items = nullptr;
  int32_t *roff = nullptr;
  int nruns = 0;
  bool fused = false;
};

// Cost of a predict item in visibility units: TW_FFT_COST for its sub-grid
// transform plus one per visibility.  Runs of about 1 / TW_RUNS_PER_CU of a
// compute unit's share: the queue evens out the model's error, and a job
// larger than a run is cut over several runs (each transforming it).
#ifndef TW_FFT_COST
#define TW_FFT_COST 2048
#endif
#ifndef TW_RUNS_PER_CU
#define TW_RUNS_PER_CU 16
#endif
constexpr int TW_MIN_SPLIT = 256;  // fewest visibilities of an item cut at a run's end

void tw_predict_runs(const std::vector<int4> &hjobs, const std::vector<int32_t> &jchunk,
                     const std::vector<int32_t> &jwp, int num_cu, std::vector<int4> &items,
                     std::vector<int32_t> &roff) {
  const int nruns = std::max(1, TW_RUNS_PER_CU * num_cu);
  items.clear();
  items.reserve(hjobs.size() + 2 * (size_t)nruns + 16);
  roff.assign(1, 0);
  roff.reserve(2 * (size_t)nruns + 16);
  // integer costs: the loop is a chain through acc
  int64_t total = 0;
  for (const int4 &j : hjobs) total += TW_FFT_COST + (int64_t)(j.z - j.y);
  const int64_t target = total / nruns;
  int64_t acc = 0;  // cost of the open run
  for (size_t k = 0; k < hjobs.size(); k++) {
    const int4 j = hjobs[k];
    const int chunk = jchunk[k], wp = jwp[k];
    for (int v = j.y; v < j.z;) {
      if (acc > 0 && acc + TW_FFT_COST + TW_MIN_SPLIT > target) {
        roff.push_back((int32_t)items.size());
        acc = 0;
      }
      const int64_t room = std::max<int64_t>(TW_MIN_SPLIT, target - acc - TW_FFT_COST);
      const int e = (int64_t)(j.z - v) > room ? v + (int)room : j.z;
      items.push_back(make_int4(chunk, wp, v, e));
      acc += TW_FFT_COST + (int64_t)(e - v);
      v = e;
    }
  }
  if ((int)items.size() > roff.back()) roff.push_back((int32_t)items.size());
}

struct Plan {
  TowerArgs a;
  WsSet *keep = nullptr;  // where rec, svis and jobs live (null: the context's set)
  uint4 *loc = nullptr, *rec = nullptr;
  float2 *svis = nullptr;
  uint32_t *small = nullptr;
  int4 *jobs = nullptr;
  int njobs = 0;
  std::vector<int4> hjobs;
  std::vector<int32_t> jchunk, jwp;  // per job its chunk and w-plane
  std::vector<int32_t> chunks;  // non-empty chunk ids, ascending
  GridSched gs, gs_split;  // without / with the small jobs split off
  DegridSched ds;
  bool gs_ready = false, gss_ready = false, ds_ready = false;
  int kind = 0;  // plans: 0 no kernel (nearest cell), 1 separable, 2 two-dimensional
};

int tw_sync_status(croc_ctx *ctx, uint32_t *st) {
  CROC_HIP(hipMemcpyAsync(ctx->host_status, st, 4, hipMemcpyDeviceToHost, ctx->stream));
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return croc_status_to_code(*ctx->host_status);
}

// Bin all visibilities into (chunk, plane) jobs; fills p (host job list too).
int tw_bin(croc_ctx *ctx, Plan &p, const float2 *vis) {
  TowerArgs &a = p.a;
  hipStream_t s = ctx->stream;
  p.small = (uint32_t *)ws_get(ctx, WS_SMALL, 256);
  if (!p.small) return CROC_ENOMEM;
  int rc = zero_u32_async(ctx, p.small, 16);
  if (rc) return rc;
  int *mm = (int *)(p.small + 8);
  const int init[2] = {INT_MAX, INT_MIN};
  CROC_HIP(hipMemcpyAsync(mm, init, sizeof init, hipMemcpyHostToDevice, s));
  const unsigned nb = (unsigned)((a.nv + 255) / 256);
  k_tw_wrange<<<(unsigned)std::min<int64_t>(nb, 2 * ctx->num_cu), 256, 0, s>>>(
      a.uvw, a.nv, a.stride, a.wstep, mm, p.small);
  CROC_HIP(hipGetLastError());
  int hmm[2];
  CROC_HIP(hipMemcpyAsync(hmm, mm, sizeof hmm, hipMemcpyDeviceToHost, s));
  if ((rc = tw_sync_status(ctx, p.small))) return rc;
  a.wp_min = hmm[0];
  a.nwp = hmm[1] - hmm[0] + 1;
  const int64_t nkeys = (int64_t)a.ncx * a.ncx * a.nwp;
  CROC_REQUIRE(nkeys < ((int64_t)1 << 31), CROC_EINVAL,
               "too many (chunk, w-plane) pairs (%lld); increase wstep", (long long)nkeys);

  const uint32_t bad = (uint32_t)nkeys;
  int end_bit = 1;
  while (end_bit < 32 && ((uint64_t)1 << end_bit) <= bad) end_bit++;
  const size_t n = (size_t)a.nv;
  uint32_t *keys = nullptr, *skeys = nullptr, *idx = nullptr, *sidx = nullptr;
  uint32_t *ukeys = nullptr, *counts = nullptr, *starts = nullptr, *nruns_d = nullptr;
  size_t t_sort = 0, t_rle = 0, t_scan = 0;
  CROC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t_sort, keys, skeys, idx, sidx, (int)n,
                                              0, end_bit, s));
  CROC_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, t_rle, skeys, ukeys, counts, nruns_d,
                                                 (int)n, s));
  CROC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, counts, starts, (int)n, s));
  const size_t t_bytes = std::max(t_sort, std::max(t_rle, t_scan));
  char *w = (char *)ws_get(ctx, WS_SORT, sizeof(uint32_t) * (7 * n + 64) + t_bytes + 256);
  WsSet &keep = p.keep ? *p.keep : ctx->ws;
  p.loc = (uint4 *)ws_get(ctx, WS_LOC, sizeof(uint4) * a.nv);
  p.rec = (uint4 *)ws_get_in(ctx, keep, WS_REC, sizeof(uint4) * a.nv);
  if (vis) p.svis = (float2 *)ws_get_in(ctx, keep, WS_SVIS, sizeof(float2) * a.nv);
  p.jobs = (int4 *)ws_get_in(ctx, keep, WS_ITEMS, sizeof(int4) * (a.nv + 1));
  if (!w || !p.loc || !p.rec || (vis && !p.svis) || !p.jobs) return CROC_ENOMEM;
  keys = (uint32_t *)w;
  skeys = keys + n;
  idx = skeys + n;
  sidx = idx + n;
  ukeys = sidx + n;
  counts = ukeys + n;
  starts = counts + n;
  nruns_d = starts + n;
  void *temp = (void *)(((uintptr_t)(nruns_d + 16) + 255) & ~(uintptr_t)255);
  k_tw_locate<<<nb, 256, 0, s>>>(a, bad, p.loc, keys, idx, p.small);
  CROC_HIP(hipGetLastError());
  if ((rc = tw_sync_status(ctx, p.small))) return rc;
  size_t tb = t_bytes;
  CROC_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, keys, skeys, idx, sidx, (int)n, 0,
                                              end_bit, s));
  tb = t_bytes;
  CROC_HIP(hipcub::DeviceRunLengthEncode::Encode(temp, tb, skeys, ukeys, counts, nruns_d, (int)n,
                                                 s));
  k_tw_gather<<<nb, 256, 0, s>>>(a.nv, sidx, p.loc, vis, p.rec, p.svis);
  CROC_HIP(hipGetLastError());
  int nruns = 0;
  CROC_HIP(hipMemcpyAsync(&nruns, nruns_d, sizeof(int), hipMemcpyDeviceToHost, s));
  CROC_HIP(hipStreamSynchronize(s));
  tb = t_bytes;
  CROC_HIP(hipcub::DeviceScan::ExclusiveSum(temp, tb, counts, starts, nruns, s));
  k_tw_jobs<<<(unsigned)((nruns + 255) / 256), 256, 0, s>>>(ukeys, counts, starts, nruns, p.jobs);
  CROC_HIP(hipGetLastError());
  p.njobs = nruns;
  p.hjobs.resize(p.njobs);
  if (p.njobs) {
    CROC_HIP(hipMemcpyAsync(p.hjobs.data(), p.jobs, sizeof(int4) * p.njobs,
                            hipMemcpyDeviceToHost, s));
    CROC_HIP(hipStreamSynchronize(s));
  }
  p.chunks.clear();
  p.jchunk.resize(p.njobs);
  p.jwp.resize(p.njobs);
  for (int j = 0; j < p.njobs; j++) {
    const int c = p.hjobs[j].x / a.nwp;
    p.jchunk[j] = c;
    p.jwp[j] = p.hjobs[j].x - c * a.nwp + a.wp_min;
    if (p.chunks.empty() || p.chunks.back() != c) p.chunks.push_back(c);
  }
  return CROC_OK;
}

int tw_setup(croc_ctx *ctx, Plan &p, int N, const double *uvw, int64_t nv, int64_t stride,
             double theta, double wstep, int subgrid, int margin, bool has_kernel, int Q,
             int S, int nwk, double wk_min, double wk_step) {
  CROC_REQUIRE(subgrid == SG, CROC_EINVAL, "subgrid size must be %d", SG);
  CROC_REQUIRE(margin >= 0 && margin % 2 == 0 && margin <= subgrid / 2, CROC_EINVAL,
               "margin must be even and in [0, subgrid/2]");
  CROC_REQUIRE(N >= 2 && theta > 0 && wstep > 0 && stride >= 3, CROC_EINVAL, "bad arguments");
  CROC_REQUIRE(nv < (int64_t)INT32_MAX, CROC_EINVAL, "too many visibilities");
  CROC_REQUIRE(!has_kernel || (Q >= 1 && Q <= 65535 && S >= 1 && S <= margin),
               CROC_EINVAL, "kernel support %d must not exceed the margin %d", S, margin);
  CROC_REQUIRE(nwk >= 1 && (has_kernel || nwk == 1), CROC_EINVAL,
               "w-kernel plane count must be >= 1 (1 without a kernel)");
  // the set must cover the whole w-increment (grid.c:493-496)
  CROC_REQUIRE(nwk == 1 || (wk_step > 0 && wk_min <= -wstep / 2 &&
                            wk_min + (nwk - 1) * wk_step >= wstep / 2),
               CROC_EINVAL, "the w-kernel set [%g, %g] does not cover a whole w-increment "
               "[%g, %g]", wk_min, wk_min + (nwk - 1) * wk_step, -wstep / 2, wstep / 2);
  TowerArgs &a = p.a;
  a.N = N;
  a.sg = subgrid;
  a.margin = margin;
  a.chunk = subgrid - margin;
  a.ncx = N / a.chunk + 1;
  a.theta = theta;
  a.wstep = wstep;
  a.Q = has_kernel ? Q : 1;
  a.S = has_kernel ? S : 1;
  a.nearest = !has_kernel;
  a.nwk = nwk;
  a.wk_min = wk_min;
  a.wk_step = wk_step > 0 ? wk_step : 1.0;
  a.uvw = uvw;
  a.nv = nv;
  a.stride = stride;
  return CROC_OK;
}

// A schedule's device block filled from its host image: the arrays are
// copied into the context's pinned staging buffer at their device offsets
// and the block goes up in one copy, waited for before send() returns (the
// host vectors and the staging buffer are then free).
struct Upload {
  croc_ctx *ctx;
  char *d, *h;
  size_t n;
  Upload(croc_ctx *c, char *dev, size_t bytes)
      : ctx(c), d(dev), h((char *)host_stage(c, bytes)), n(bytes) {}
  template <class T>
  void put(T *dst, const std::vector<T> &v) {
    if (!v.empty()) memcpy(h + ((char *)dst - d), v.data(), sizeof(T) * v.size());
  }
  int send() {
    CROC_HIP(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, ctx->stream));
    CROC_HIP(hipStreamSynchronize(ctx->stream));
    return CROC_OK;
  }
};

// Runs over the job-list entries e0 .. e1 - 1 (chunk[e] their chunks, in
// order by chunk): at most `per` jobs each, the first of each chunk adding
// to the chunk image, later ones to the next partial slot (*nslot).
void tw_sum_runs(const std::vector<int32_t> &jl, const std::vector<int32_t> &jslot,
                 const std::vector<int32_t> &jchunk, size_t e0, size_t e1, int per, bool slot_of_job,
                 int *nslot, std::vector<int4> &runs, std::vector<int4> &rgroups,
                 std::vector<int32_t> &pslot, GridSched::Batch &bt) {
  bt.r0 = runs.size();
  bt.q0 = rgroups.size();
  for (size_t e = e0; e < e1;) {
    const int c = jchunk[jl[e]];
    size_t f = e;
    while (f < e1 && jchunk[jl[f]] == c) f++;
    const size_t q = pslot.size();
    for (size_t b = e; b < f; b += per) {
      int out = -1;
      if (b != e) {
        out = slot_of_job ? jslot[jl[b]] : (*nslot)++;
        pslot.push_back(out);
      }
      runs.push_back(make_int4(c, (int)b, (int)std::min(f, b + per), out));
    }
    if (pslot.size() > q) rgroups.push_back(make_int4(c, (int)q, (int)pslot.size(), 0));
    e = f;
  }
  bt.nr = runs.size() - bt.r0;
  bt.nq = rgroups.size() - bt.q0;
  // longest first, stable: a counting sort by length (at most `per`)
  std::vector<size_t> cnt(per + 2, 0);
  for (size_t r = bt.r0; r < runs.size(); r++) cnt[per - (runs[r].z - runs[r].y)]++;
  size_t at = bt.r0;
  for (int l = 0; l <= per; l++) {
    const size_t c = cnt[l];
    cnt[l] = at;
    at += c;
  }
  std::vector<int4> sorted(bt.nr);
  for (size_t r = bt.r0; r < runs.size(); r++)
    sorted[cnt[per - (runs[r].z - runs[r].y)]++ - bt.r0] = runs[r];
  std::copy(sorted.begin(), sorted.end(), runs.begin() + bt.r0);
}

// Build the gridding schedule of p (split: the small jobs apart, p.gs_split
// in slot WS_TOWER; else p.gs in slot WS_TILES) and upload it into `set`.
int tw_grid_sched(croc_ctx *ctx, Plan &p, WsSet &set, bool split) {
  GridSched &g = split ? p.gs_split : p.gs;
  g = GridSched();
  g.split = split;
  const TowerArgs &a = p.a;
  std::vector<int4> parts, runs, rgroups;
  std::vector<int2> multi;
  std::vector<int32_t> jslot(p.njobs, -1), jl, pslot;
  const std::vector<int32_t> &jchunk = p.jchunk;
  parts.reserve(p.njobs + 64);
  runs.reserve(p.njobs + 64);
  rgroups.reserve(p.chunks.size() + 64);
  jl.reserve(p.njobs);
  pslot.reserve(p.njobs);
  // small jobs first, in job order
  if (split) {
    for (int j = 0; j < p.njobs; j++)
      if (p.hjobs[j].z - p.hjobs[j].y <= TW_SMALL_VIS) jl.push_back(j);
    int nslot = 0;
    const int per = std::max(1, (int)((jl.size() + 2 * ctx->num_cu - 1) / (2 * ctx->num_cu)));
    tw_sum_runs(jl, jslot, jchunk, 0, jl.size(), per, false, &nslot, runs, rgroups, pslot,
                g.small);
    g.max_slots = nslot;
  }
  // batches of the other jobs
  int slots = 0;
  size_t e0 = jl.size();
  auto close_batch = [&]() {
    GridSched::Batch &bt = g.batches.back();
    bt.np = parts.size() - bt.p0;
    bt.nm = multi.size() - bt.m0;
    const int nj = (int)(jl.size() - e0);
    const int per = std::min(8, std::max(1, (nj + 2 * ctx->num_cu - 1) / (2 * ctx->num_cu)));
    tw_sum_runs(jl, jslot, jchunk, e0, jl.size(), per, true, nullptr, runs, rgroups, pslot, bt);
    e0 = jl.size();
  };
  for (int j = 0; j < p.njobs; j++) {
    const int4 job = p.hjobs[j];
    const int n = job.z - job.y;
    if (split && n <= TW_SMALL_VIS) continue;
    const int np = std::min(std::max(1, (n + PART_VIS - 1) / PART_VIS), JOB_BATCH / 4);
    if (g.batches.empty() || slots + np > JOB_BATCH) {
      if (!g.batches.empty()) close_batch();
      g.batches.push_back(GridSched::Batch{parts.size(), 0, multi.size(), 0, 0, 0, 0, 0});
      slots = 0;
    }
    jslot[j] = slots;
    jl.push_back(j);
    const int len = (n + np - 1) / np;
    for (int k = 0; k < np; k++)
      parts.push_back(make_int4(slots + k, job.y + k * len,
                                std::min(job.z, job.y + (k + 1) * len), 0));
    if (np > 1) multi.push_back(make_int2(slots, np));
    slots += np;
    g.max_slots = std::max(g.max_slots, slots);
  }
  if (!g.batches.empty()) close_batch();
  // parity classes of the chunks: one add launch each, stream order keeps
  // overlapping margins of different classes apart
  std::vector<int32_t> par[4], allpar;
  for (int c : p.chunks) par[(c % a.ncx) % 2 + 2 * ((c / a.ncx) % 2)].push_back(c);
  for (int q = 0; q < 4; q++) {
    allpar.insert(allpar.end(), par[q].begin(), par[q].end());
    g.npar[q] = par[q].size();
  }
  const size_t n32 = 2 * (size_t)p.njobs + 2 * p.chunks.size() + pslot.size() + jl.size();
  const size_t bytes = sizeof(int4) * (parts.size() + runs.size() + rgroups.size()) +
                       sizeof(int2) * multi.size() + sizeof(int32_t) * n32 + 256;
  char *d = (char *)ws_get_in(ctx, set, split ? WS_TOWER : WS_TILES, bytes);
  if (!d) return CROC_ENOMEM;
  g.parts = (int4 *)d;
  g.runs = g.parts + parts.size();
  g.rgroups = g.runs + runs.size();
  g.multi = (int2 *)(g.rgroups + rgroups.size());
  g.kw = (int32_t *)(g.multi + multi.size());
  g.jslot = g.kw + p.njobs;
  g.chunks = g.jslot + p.njobs;
  g.par = g.chunks + p.chunks.size();
  g.pslot = g.par + p.chunks.size();
  g.jl = g.pslot + pslot.size();
  Upload up(ctx, d, bytes);
  if (!up.h) return CROC_ENOMEM;
  up.put(g.parts, parts);
  up.put(g.runs, runs);
  up.put(g.rgroups, rgroups);
  up.put(g.multi, multi);
  up.put(g.kw, p.jwp);
  up.put(g.jslot, jslot);
  up.put(g.chunks, p.chunks);
  up.put(g.par, allpar);
  up.put(g.pslot, pslot);
  up.put(g.jl, jl);
  int rc = up.send();
  if (rc) return rc;
  (split ? p.gss_ready : p.gs_ready) = true;
  return CROC_OK;
}

// Build the degridding schedule of p and upload it into `set` (slot WS_GRIDQ).
// fused: the schedule of k_tw_predict_sep (items and runs), else that of
// the batched transforms and k_tw_degrid_jobs (batches and parts).
int tw_degrid_sched(croc_ctx *ctx, Plan &p, WsSet &set, bool fused) {
  DegridSched &g = p.ds;
  g = DegridSched();
  g.fused = fused;
  g.jb = std::min(p.njobs, JOB_BATCH);
  std::vector<int4> parts;
  for (int j0 = 0; !fused && j0 < p.njobs; j0 += g.jb) {
    g.poff.push_back(parts.size());
    for (int b = 0; b < std::min(g.jb, p.njobs - j0); b++) {
      const int4 job = p.hjobs[j0 + b];
      for (int v = job.y; v < job.z; v += PART_VIS)
        parts.push_back(make_int4(b, v, std::min(job.z, v + PART_VIS), 0));
    }
  }
  g.poff.push_back(parts.size());
  std::vector<int4> items;
  std::vector<int32_t> roff(1, 0);
  if (fused) tw_predict_runs(p.hjobs, p.jchunk, p.jwp, ctx->num_cu, items, roff);
  g.nruns = (int)roff.size() - 1;
  const size_t bytes = sizeof(int4) * (parts.size() + items.size() + 32) +
                       sizeof(int32_t) * (2 * (size_t)p.njobs + p.chunks.size() + roff.size()) +
                       256;
  char *d = (char *)ws_get_in(ctx, set, WS_GRIDQ, bytes);
  if (!d) return CROC_ENOMEM;
  g.parts = (int4 *)d;
  g.items = g.parts + parts.size() + 16;
  g.kw = (int32_t *)(g.items + items.size() + 16);
  g.jchunk = g.kw + p.njobs;
  g.chunks = g.jchunk + p.njobs;
  g.roff = g.chunks + p.chunks.size();
  Upload up(ctx, d, bytes);
  if (!up.h) return CROC_ENOMEM;
  up.put(g.parts, parts);
  up.put(g.items, items);
  up.put(g.kw, p.jwp);
  up.put(g.jchunk, p.jchunk);
  up.put(g.chunks, p.chunks);
  up.put(g.roff, roff);
  int rc = up.send();
  if (rc) return rc;
  p.ds_ready = true;
  return CROC_OK;
}

int tw_grid_run(croc_ctx *ctx, const Plan &p, void *grid, const void *kern, int separable);
int tw_degrid_run(croc_ctx *ctx, const Plan &p, const void *grid, void *vis_out,
                  const void *kern, int separable);

}  // namespace

extern "C" int croc_wtowers_grid(croc_ctx *ctx, void *grid, int N, const double *uvw,
                                 int64_t nv, int64_t stride, double theta, const void *vis,
                                 double wstep, int subgrid, int margin, const void *kern,
                                 int Q, int S, int separable, int nwk, double wk_min,
                                 double wk_step) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid, CROC_EINVAL, "null grid");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && vis, CROC_EINVAL, "null uvw / vis");
  Plan p;
  if ((rc = tw_setup(ctx, p, N, uvw, nv, stride, theta, wstep, subgrid, margin,
                     kern != nullptr, Q, S, nwk, wk_min, wk_step)))
    return rc;
  if ((rc = tw_bin(ctx, p, (const float2 *)vis))) return rc;
  if ((rc = tw_grid_sched(ctx, p, ctx->ws, !p.a.nearest && separable && S <= 16))) return rc;
  return tw_grid_run(ctx, p, grid, kern, separable);
}

namespace {

// Imaging from binned jobs (p.svis holds the visibilities in record order,
// p.gs the schedule).  Enqueues kernels and event records only: no host
// synchronisation, no allocation once the context's buffers have this
// size, no memset.
int tw_grid_run(croc_ctx *ctx, const Plan &p, void *grid, const void *kern, int separable) {
  int rc;
  if (p.njobs == 0) return CROC_OK;
  const TowerArgs &a = p.a;
  hipStream_t s = ctx->stream;

  // fixed-point scale, Fresnel phase table, a part queue per batch
  const int64_t Q = a.Q, S = a.S;
  const int64_t tsize = separable ? Q * S : Q * Q * S * S;  // elements per w-kernel plane
  const int64_t nk = kern ? tsize * a.nwk : 0;
  // Separable kernels of support <= 16 grid on the matrix cores
  // (k_tw_grid_jobs_mfma), the small jobs in the summing kernel.
  const bool mfma = !a.nearest && separable && a.S <= 16;
  CROC_REQUIRE(mfma ? p.gss_ready : p.gs_ready, CROC_EINVAL, "gridding schedule missing");
  const GridSched &gs = mfma ? p.gs_split : p.gs;
  char *aux = (char *)ws_get(ctx, WS_SCALE, 256 + sizeof(double) * SGN +
                                                sizeof(uint32_t) * (2 * gs.batches.size() + 64) +
                                                sizeof(uint32_t) * (size_t)a.ncx * a.ncx);
  const size_t cimg_bytes = sizeof(float2) * SGN * (size_t)a.ncx * a.ncx;
  float2 *stack =
      (float2 *)ws_get(ctx, WS_SUB, cimg_bytes + sizeof(float2) * SGN * (size_t)gs.max_slots);
  if (!aux || !stack) return CROC_ENOMEM;
  TwScale *fs = (TwScale *)aux;
  double *phi = (double *)(aux + 256);
  uint32_t *qc = (uint32_t *)(phi + SGN);  // part queues, then summing-run queues
  uint32_t *cflag = qc + 2 * gs.batches.size() + 64;  // chunks to set non-finite
  float2 *cimg = stack, *sub = stack + (size_t)SGN * a.ncx * a.ncx;
  if ((rc = zero_u32_async(ctx, (uint32_t *)fs, sizeof(TwScale) / 4))) return rc;
  k_tw_maxima<<<ctx->num_cu, 1024, 0, s>>>(p.svis, a.nv, (const float2 *)kern, nk, fs);
  k_tw_pick<<<1, 1, 0, s>>>(fs, a.nv, a.nearest, separable);
  // the matrix-core gridding reads an fp16 (hi, lo) copy of the kernel set
  uint2 *ksplit = nullptr;
  if (mfma) {
    CROC_REQUIRE(nk < ((int64_t)1 << 32), CROC_EINVAL, "kernel set beyond 2^32 elements");
    ksplit = (uint2 *)ws_get(ctx, WS_KSPLIT, sizeof(uint2) * (size_t)nk);
    if (!ksplit) return CROC_ENOMEM;
    k_tw_split_tables<<<(unsigned)((nk + 255) / 256), 256, 0, s>>>((const float2 *)kern, nk, fs,
                                                                  ksplit);
  }
  k_tw_phi<<<SGN / 256, 256, 0, s>>>(phi, a.theta, a.wstep);
  CROC_HIP(hipGetLastError());
  // zero the chunk images that will receive data
  k_tw_zero_chunks<<<dim3(16, (unsigned)p.chunks.size()), 256, 0, s>>>(gs.chunks, cimg);
  CROC_HIP(hipGetLastError());

  if ((rc = zero_u32_async(ctx, qc, 2 * (int64_t)gs.batches.size() + 1))) return rc;
  CROC_HIP(hipEventRecord(ctx->ev[0], s));
  tap_timer_reset(ctx, 0);
  if (gs.small.nr) {
    // small jobs gridded, transformed and summed per chunk; partial sums in
    // slots 0, 1, ... before the batches take the slots
    k_tw_image_runs<true><<<(unsigned)std::min<size_t>(gs.small.nr, (size_t)ctx->num_cu), FFT_NT,
                            0, s>>>(gs.runs + gs.small.r0, (int)gs.small.nr, gs.jl, gs.jslot,
                                    gs.kw, phi, sub, cimg, qc + 2 * gs.batches.size(), p.jobs,
                                    p.rec, p.svis, (const float2 *)kern, a.S, tsize);
    CROC_HIP(hipGetLastError());
    if (gs.small.nq) {
      k_tw_chunk_accum<<<dim3(ACC_SLICES, (unsigned)gs.small.nq), 256, 0, s>>>(
          gs.rgroups + gs.small.q0, gs.pslot, sub, cimg);
      CROC_HIP(hipGetLastError());
    }
  }
  for (size_t bi = 0; bi < gs.batches.size(); bi++) {
    const GridSched::Batch &bt = gs.batches[bi];
    const int G = (int)std::min<size_t>(bt.np, (size_t)ctx->num_cu);  // one per CU (LDS)
    if ((rc = tap_timer_begin(ctx, 0))) return rc;
    if (ksplit)
      k_tw_grid_jobs_mfma<<<G, TW_NT, 0, s>>>(gs.parts + bt.p0, (int)bt.np, p.rec, p.svis,
                                              (const float2 *)kern, ksplit, a.S, tsize, fs, sub,
                                              qc + bi);
    else if (a.nearest || !separable)
      k_tw_grid_jobs<false><<<G, TW_NT, 0, s>>>(gs.parts + bt.p0, (int)bt.np, p.rec, p.svis,
                                                (const float2 *)kern, a.Q, a.S, tsize,
                                                a.nearest, fs, sub, qc + bi);
    else
      k_tw_grid_jobs<true><<<G, TW_NT, 0, s>>>(gs.parts + bt.p0, (int)bt.np, p.rec, p.svis,
                                               (const float2 *)kern, a.Q, a.S, tsize,
                                               a.nearest, fs, sub, qc + bi);
    CROC_HIP(hipGetLastError());
    if ((rc = tap_timer_end(ctx, 0))) return rc;
    if (bt.nm) {
      k_tw_sum_parts<<<dim3(ACC_SLICES, (unsigned)bt.nm), 256, 0, s>>>(gs.multi + bt.m0, sub);
      CROC_HIP(hipGetLastError());
    }
    // images ifft(sub[slot]) * exp(-i wp phi) summed into the chunk images
    k_tw_image_runs<false><<<(unsigned)std::min<size_t>(bt.nr, (size_t)ctx->num_cu), FFT_NT, 0,
                             s>>>(gs.runs + bt.r0, (int)bt.nr, gs.jl, gs.jslot, gs.kw, phi, sub,
                                  cimg, qc + gs.batches.size() + bi, nullptr, nullptr, nullptr,
                                  nullptr, 0, 0);
    CROC_HIP(hipGetLastError());
    if (bt.nq) {
      k_tw_chunk_accum<<<dim3(ACC_SLICES, (unsigned)bt.nq), 256, 0, s>>>(gs.rgroups + bt.q0,
                                                                         gs.pslot, sub, cimg);
      CROC_HIP(hipGetLastError());
    }
  }
  // chunk sub-grids = fft(chunk image), then parity-ordered adds
  rc = fft2d_launch(ctx, cimg, cimg, SG, (int64_t)p.chunks.size(), SGN, -1, 1.0f, nullptr,
                    nullptr, 0, nullptr, nullptr, 0, gs.chunks, gs.chunks);
  if (rc) return rc;
  size_t po = 0;
  for (int q = 0; q < 4; q++) {
    if (gs.npar[q] == 0) continue;
    dim3 g(16, (unsigned)gs.npar[q]);
    k_tw_add<<<g, 256, 0, s>>>(gs.par + po, cimg, (float2 *)grid, a.N, a.chunk, a.margin,
                               a.ncx);
    CROC_HIP(hipGetLastError());
    po += gs.npar[q];
  }
  // chunks of non-finite values (only when flagged)
  if ((rc = zero_u32_async(ctx, cflag, (int64_t)a.ncx * a.ncx))) return rc;
  k_tw_nonfinite_mark<<<(unsigned)std::min(p.njobs, 4 * ctx->num_cu), 256, 0, s>>>(
      p.jobs, p.njobs, a.nwp, p.svis, fs, cflag);
  k_tw_nonfinite_fill<<<dim3(16, (unsigned)p.chunks.size()), 256, 0, s>>>(
      gs.chunks, cflag, (float2 *)grid, a.N, a.chunk, a.margin, a.ncx, fs);
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipEventRecord(ctx->ev[1], s));
  ctx->ev_grid = true;
  return CROC_OK;
}

}  // namespace

extern "C" int croc_wtowers_degrid(croc_ctx *ctx, const void *grid, int N, const double *uvw,
                                   int64_t nv, int64_t stride, double theta, void *vis_out,
                                   double wstep, int subgrid, int margin, const void *kern,
                                   int Q, int S, int separable, int nwk, double wk_min,
                                   double wk_step) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid, CROC_EINVAL, "null grid");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && vis_out, CROC_EINVAL, "null uvw / vis_out");
  Plan p;
  if ((rc = tw_setup(ctx, p, N, uvw, nv, stride, theta, wstep, subgrid, margin,
                     kern != nullptr, Q, S, nwk, wk_min, wk_step)))
    return rc;
  if ((rc = tw_bin(ctx, p, nullptr))) return rc;
  if ((rc = tw_degrid_sched(ctx, p, ctx->ws, kern && separable && S <= 16))) return rc;
  return tw_degrid_run(ctx, p, grid, vis_out, kern, separable);
}

namespace {

// Predict from binned jobs (p.ds the schedule).  Enqueues kernels and event
// records only, like tw_grid_run.
int tw_degrid_run(croc_ctx *ctx, const Plan &p, const void *grid, void *vis_out,
                  const void *kern, int separable) {
  int rc;
  if (p.njobs == 0) return CROC_OK;
  CROC_REQUIRE(p.ds_ready, CROC_EINVAL, "degridding schedule missing");
  const TowerArgs &a = p.a;
  const DegridSched &ds = p.ds;
  hipStream_t s = ctx->stream;
  char *aux = (char *)ws_get(ctx, WS_SCALE, 256 + sizeof(double) * SGN +
                                                sizeof(uint32_t) * (ds.poff.size() + 64));
  const int64_t tsize = separable ? (int64_t)a.Q * a.S : (int64_t)a.Q * a.Q * a.S * a.S;
  // Separable kernels of support <= 16 transform and degrid each plane
  // sub-grid in LDS (k_tw_predict_sep): no sub-grid slots.
  const bool sep16 = !a.nearest && separable && a.S <= 16;
  CROC_REQUIRE(sep16 == ds.fused, CROC_EINVAL, "degridding schedule made for another kernel kind");
  const size_t cimg_bytes = sizeof(float2) * SGN * (size_t)a.ncx * a.ncx;
  float2 *stack = (float2 *)ws_get(
      ctx, WS_SUB, cimg_bytes + (sep16 ? 0 : sizeof(float2) * SGN * (size_t)ds.jb));
  if (!aux || !stack) return CROC_ENOMEM;
  double *phi = (double *)(aux + 256);
  uint32_t *qc = (uint32_t *)(phi + SGN);  // a part queue per batch
  float2 *cimg = stack, *sub = stack + (size_t)SGN * a.ncx * a.ncx;
  k_tw_phi<<<SGN / 256, 256, 0, s>>>(phi, a.theta, a.wstep);
  if ((rc = zero_u32_async(ctx, qc, (int64_t)ds.poff.size()))) return rc;
  dim3 g(16, (unsigned)p.chunks.size());
  CROC_HIP(hipEventRecord(ctx->ev[2], s));
  tap_timer_reset(ctx, 1);
  k_tw_cut<<<g, 256, 0, s>>>(ds.chunks, (const float2 *)grid, cimg, a.N, a.chunk, a.margin,
                             a.ncx);
  CROC_HIP(hipGetLastError());
  rc = fft2d_launch(ctx, cimg, cimg, SG, (int64_t)p.chunks.size(), SGN, +1, 1.0f / (float)SGN,
                    nullptr, nullptr, 0, nullptr, nullptr, 0, ds.chunks, ds.chunks);
  if (rc) return rc;
  if (sep16) {
    const int G = std::min(ds.nruns, ctx->num_cu);  // one per CU (LDS)
    if ((rc = tap_timer_begin(ctx, 1))) return rc;
    if (a.S == 16)
      k_tw_predict_sep<16><<<G, FFT_NT, 0, s>>>(ds.items, ds.roff, ds.nruns, cimg, phi, p.rec,
                                                (const float2 *)kern, a.S, tsize,
                                                (float2 *)vis_out, qc);
    else
      k_tw_predict_sep<0><<<G, FFT_NT, 0, s>>>(ds.items, ds.roff, ds.nruns, cimg, phi, p.rec,
                                               (const float2 *)kern, a.S, tsize,
                                               (float2 *)vis_out, qc);
    CROC_HIP(hipGetLastError());
    if ((rc = tap_timer_end(ctx, 1))) return rc;
  }
  for (int j0 = 0, bi = 0; !sep16 && j0 < p.njobs; j0 += ds.jb, bi++) {
    const int nb = std::min(ds.jb, p.njobs - j0);
    // plane sub-grid = fft(chunk image * exp(+i wp phi))
    rc = fft2d_launch(ctx, cimg, sub, SG, nb, SGN, -1, 1.0f, phi, ds.kw + j0, +1, nullptr, nullptr,
                      0, ds.jchunk + j0, nullptr);
    if (rc) return rc;
    const int npart = (int)(ds.poff[bi + 1] - ds.poff[bi]);
    const int G = std::min(npart, ctx->num_cu);  // one per CU (LDS)
    if ((rc = tap_timer_begin(ctx, 1))) return rc;
    if (a.nearest || !separable)
      k_tw_degrid_jobs<false><<<G, TW_NT, 0, s>>>(ds.parts + ds.poff[bi], npart, p.rec, sub,
                                                  (const float2 *)kern, a.Q, a.S, tsize,
                                                  a.nearest, (float2 *)vis_out, qc + bi);
    else
      k_tw_degrid_jobs<true><<<G, TW_NT, 0, s>>>(ds.parts + ds.poff[bi], npart, p.rec, sub,
                                                 (const float2 *)kern, a.Q, a.S, tsize,
                                                 a.nearest, (float2 *)vis_out, qc + bi);
    CROC_HIP(hipGetLastError());
    if ((rc = tap_timer_end(ctx, 1))) return rc;
  }
  CROC_HIP(hipEventRecord(ctx->ev[3], s));
  ctx->ev_degrid = true;
  return CROC_OK;
}

}  // namespace

// W-towers plans: the (chunk, w-plane) binning of one set of coordinates,
// reused by any number of imaging and predict calls on them (the one-shot
// calls bin on every call).  croc_wtowers_plan_bin validates every
// coordinate, as the one-shot calls do.
struct croc_tw_plan {
  int device = 0;
  WsSet ws;  // records, visibilities in record order, jobs
  Plan p;
  bool ready = false;
};

namespace {

// svis[pos] = vis[rec[pos].w]: the visibilities in the plan's record order.
__global__ void k_tw_vis_order(int64_t nv, const uint4 *__restrict__ rec,
                               const float2 *__restrict__ vis, float2 *__restrict__ svis) {
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos < nv) svis[pos] = vis[rec[pos].w];
}

int tw_plan_check(croc_ctx *ctx, const croc_tw_plan *plan, const void *kern, int separable) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(plan != nullptr, CROC_EINVAL, "null plan");
  CROC_REQUIRE(plan->ready, CROC_EINVAL, "plan holds no successful binning");
  CROC_REQUIRE(plan->device == ctx->device, CROC_EINVAL, "plan belongs to another device");
  CROC_REQUIRE((kern != nullptr) == !plan->p.a.nearest, CROC_EINVAL,
               plan->p.a.nearest ? "plan was binned without a kernel"
                                 : "plan was binned for a kernel, none given");
  CROC_REQUIRE(!kern || (separable != 0) == (plan->p.kind == 1), CROC_EINVAL,
               plan->p.kind == 1 ? "plan was binned for a separable kernel"
                                 : "plan was binned for a two-dimensional kernel");
  return CROC_OK;
}

}  // namespace

extern "C" int croc_wtowers_plan_create(croc_ctx *ctx, croc_tw_plan **out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(out != nullptr, CROC_EINVAL, "null output pointer");
  croc_tw_plan *p = new croc_tw_plan();
  p->device = ctx->device;
  *out = p;
  return CROC_OK;
}

extern "C" int croc_wtowers_plan_destroy(croc_ctx *ctx, croc_tw_plan *plan) {
  if (!plan) return CROC_OK;
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  (void)hipStreamSynchronize(ctx->stream);
  ws_free(plan->ws);
  delete plan;
  return CROC_OK;
}

extern "C" int croc_wtowers_plan_bin(croc_ctx *ctx, croc_tw_plan *plan, int N,
                                     const double *uvw, int64_t nv, int64_t stride,
                                     double theta, double wstep, int subgrid, int margin,
                                     int has_kernel, int Q, int S, int nwk, double wk_min,
                                     double wk_step) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(plan != nullptr, CROC_EINVAL, "null plan");
  CROC_REQUIRE(plan->device == ctx->device, CROC_EINVAL, "plan belongs to another device");
  CROC_REQUIRE(nv >= 0, CROC_EINVAL, "negative visibility count");
  CROC_REQUIRE(nv == 0 || uvw, CROC_EINVAL, "null uvw");
  plan->ready = false;
  Plan &p = plan->p;
  p = Plan();
  p.keep = &plan->ws;
  CROC_REQUIRE(has_kernel >= 0 && has_kernel <= 2, CROC_EINVAL,
               "has_kernel must be 0 (none), 1 (separable) or 2 (two-dimensional)");
  if ((rc = tw_setup(ctx, p, N, uvw, nv, stride, theta, wstep, subgrid, margin,
                     has_kernel != 0, Q, S, nwk, wk_min, wk_step)))
    return rc;
  if (nv > 0 && (rc = tw_bin(ctx, p, nullptr))) return rc;
  // the schedules of the kernel kind: separable kernels of support <= 16
  // split the small jobs off and fuse the predict
  p.kind = has_kernel;
  const bool sep16 = has_kernel == 1 && S <= 16;
  if (nv > 0 && (rc = tw_grid_sched(ctx, p, plan->ws, sep16))) return rc;
  if (nv > 0 && (rc = tw_degrid_sched(ctx, p, plan->ws, sep16))) return rc;
  plan->ready = true;
  return CROC_OK;
}

extern "C" int croc_wtowers_grid_plan(croc_ctx *ctx, croc_tw_plan *plan, void *grid,
                                      const void *vis, const void *kern, int separable) {
  int rc = tw_plan_check(ctx, plan, kern, separable);
  if (rc) return rc;
  CROC_REQUIRE(grid, CROC_EINVAL, "null grid");
  Plan &p = plan->p;
  if (p.njobs == 0) return CROC_OK;
  CROC_REQUIRE(vis, CROC_EINVAL, "null vis");
  p.svis = (float2 *)ws_get_in(ctx, plan->ws, WS_SVIS, sizeof(float2) * p.a.nv);
  if (!p.svis) return CROC_ENOMEM;
  k_tw_vis_order<<<(unsigned)((p.a.nv + 255) / 256), 256, 0, ctx->stream>>>(
      p.a.nv, p.rec, (const float2 *)vis, p.svis);
  CROC_HIP(hipGetLastError());
  return tw_grid_run(ctx, p, grid, kern, separable);
}

extern "C" int croc_wtowers_degrid_plan(croc_ctx *ctx, croc_tw_plan *plan, const void *grid,
                                        void *vis_out, const void *kern, int separable) {
  int rc = tw_plan_check(ctx, plan, kern, separable);
  if (rc) return rc;
  CROC_REQUIRE(grid, CROC_EINVAL, "null grid");
  if (plan->p.njobs == 0) return CROC_OK;
  CROC_REQUIRE(vis_out, CROC_EINVAL, "null vis_out");
  return tw_degrid_run(ctx, plan->p, grid, vis_out, kern, separable);
}
