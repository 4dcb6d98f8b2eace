// Internal declarations shared by the libcrocgpu translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/crocgpu.h"

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
void croc_set_error(const char *fmt, ...);

#define CROC_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      croc_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                     __FILE__, __LINE__);                                     \
      return CROC_EHIP;                                                       \
    }                                                                         \
  } while (0)

#define CROC_REQUIRE(cond, code, ...)                                         \
  do {                                                                        \
    if (!(cond)) {                                                            \
      croc_set_error(__VA_ARGS__);                                            \
      return (code);                                                          \
    }                                                                         \
  } while (0)

// Status bits raised by device kernels (OR-ed into a device word).
enum : uint32_t {
  ST_DOMAIN = 1u,  // coordinate outside [-0.5, 0.5)
  ST_BOUNDS = 2u,  // kernel window outside the grid
  ST_PLANE = 4u,   // plane index outside the bank
  ST_SUPPORT = 8u, // a plane's support exceeds the declared maximum
  ST_BANK = 16u,   // a kernel table lies beyond element 2^32 of the bank
  ST_RANGE = 32u,  // a baseline's time / frequency window or array range is invalid
};

int croc_status_to_code(uint32_t st);

// ---------------------------------------------------------------------------
// Context: stream, reusable device workspace, pinned status word, events.
// ---------------------------------------------------------------------------
enum WsSlot {
  WS_LOC = 0,   // per-visibility located record (uint4)
  WS_REC,       // tile-sorted record (uint4)
  WS_SVIS,      // tile-sorted visibilities (float2)
  WS_TILES,     // count / offset / cursor per tile
  WS_ITEMS,     // work items (int4), tile-major
  WS_ITEMS_G,   // work items (int4), plane-group-major
  WS_SMALL,     // status, item count, queue heads
  WS_SCAN,      // binning scan temporaries
  WS_FF,        // w-kernel far fields (double2)
  WS_TMP,       // oversampling intermediate (double2)
  WS_ACC,       // int64 fixed-point gridding accumulator [N][N][2]
  WS_SCALE,     // fixed-point scale state
  WS_DFT_FWD,   // fp16-split centred DFT matrix, forward sign
  WS_DFT_INV,   // fp16-split centred DFT matrix, inverse sign
  WS_SUB,       // w-towers sub-grid stack
  WS_TOWER,     // w-towers job / chunk tables
  WS_SORT,      // w-towers radix sort keys, indices, run tables, temporaries
  WS_GRIDQ,     // gridding work lists: dense tiles' range items, sparse tiles' items
  WS_KSPLIT,    // w-towers: fp16 (hi, lo) split of a separable kernel set
  WS_BL,        // baseline layout: flattened coordinates, planes, visibilities
  WS_BL_BANK,   // baseline layout: kernel-bank plane tables, unit kernel, FP32 table
  WS_BL_WIN,    // baseline layout: per-subgrid windows of a batched streamer call
  WS_COUNT
};

// A set of grow-only device buffers, one per slot.  With an arena the slots
// are carved from that fixed device buffer instead (no allocation ever; a
// request beyond its end fails with CROC_ENOMEM and a message).
struct WsSet {
  void *ws[WS_COUNT] = {};
  size_t ws_size[WS_COUNT] = {};
  char *arena = nullptr;
  size_t arena_size = 0, arena_used = 0;
};

struct croc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  WsSet ws;                         // the context's own workspace
  uint32_t *host_status = nullptr;  // pinned
  hipEvent_t ev[4] = {};
  // ev[0], ev[1] bracket the last gridding kernel(s), ev[2], ev[3] the last
  // degridding; the durations are read on demand (croc_last_kernel_ms)
  bool ev_grid = false, ev_degrid = false;
  // Event pairs around every launch of the tap kernels (the gridding /
  // degridding kernels proper: k_grid_tiles, k_tw_grid_jobs, ...) of the
  // last gridding (dir 0) and degridding (dir 1) call: tev[dir][2k],
  // tev[dir][2k + 1] bracket launch k of tev_n[dir].
  std::vector<hipEvent_t> tev[2];
  int tev_n[2] = {0, 0};
  // pinned host staging of uploads (grow-only; reused once the copy that
  // read it has been waited for)
  void *stage = nullptr;
  size_t stage_size = 0;
  int num_cu = 256;
  int tile = 0;     // 0: automatic
  int chunk = 512;  // visibilities per work item
  int group_kib = 0;  // kernel-bank KiB per plane-group (0: automatic)
};

// Grow-only workspace; returns nullptr on failure (error already set).
// ws_get uses the context's set; ws_get_in a given one (a plan's).
void *ws_get(croc_ctx *ctx, WsSlot slot, size_t bytes);
void *ws_get_in(croc_ctx *ctx, WsSet &set, WsSlot slot, size_t bytes);
void ws_free(WsSet &set);

// Make the context's device current.
int ctx_enter(croc_ctx *ctx);

// The context's pinned host staging buffer of at least `bytes` (nullptr on
// failure, error set); the caller waits for the copy that reads it before
// the next use.
void *host_stage(croc_ctx *ctx, size_t bytes);

// Tap-kernel launch timers: tap_timer_reset at the start of a gridding
// (dir 0) or degridding (dir 1) call, tap_timer_begin / tap_timer_end
// around each launch of its tap kernel.
void tap_timer_reset(croc_ctx *ctx, int dir);
int tap_timer_begin(croc_ctx *ctx, int dir);
int tap_timer_end(croc_ctx *ctx, int dir);

// The context's one-word status (WS_SMALL), cleared on its stream, and its
// read-back: waits for the stream and maps the status bits to a CROC_E* code.
int small_status(croc_ctx *ctx, uint32_t **st);
int finish_status(croc_ctx *ctx, uint32_t *st);

// Device address of the allocation-free entry points' static scratch array
// on the current device, looked up once per device (noalloc.hip).
int noalloc_init(int device);

// ---------------------------------------------------------------------------
// FP64 coordinate arithmetic.  Every product and sum is rounded separately
// (FMA contraction is switched off in these functions): the results must
// equal NumPy's bit for bit, ties included.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t floordiv64(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
  return q;
}

// p = u / lam (IEEE division, or u itself when lam == 1).
__device__ __forceinline__ double grid_fraction(double u, double lam) {
#pragma clang fp contract(off)
  return lam == 1.0 ? u : u / lam;
}

// synthesis.py:419-435: x = N//2 + p*N; fx = around(-x*Q) (half-even);
// returns false when p is outside [-0.5, 0.5).
__device__ __forceinline__ bool frac_coord_dev(int N, int Q, double p, int *ix,
                                               int *xf) {
#pragma clang fp contract(off)
  if (!(p >= -0.5 && p < 0.5)) return false;
  double x = (double)(N / 2) + p * (double)N;
  double fx = rint(-x * (double)Q);
  int64_t f = (int64_t)fx;
  int64_t q = floordiv64(f, Q);
  *ix = (int)(-q);
  *xf = (int)(f - q * Q);
  return true;
}

// numpy.round(w / wstep) (half-even).
__device__ __forceinline__ int64_t wbin_dev(double w, double wstep) {
#pragma clang fp contract(off)
  return (int64_t)rint(w / wstep);
}

// synthesis.py:399 / :414: N//2 + floor(0.5 + N p); a negative index wraps
// once as NumPy indexing does, anything else outside the grid is an error.
__device__ __forceinline__ bool nearest_cell(int N, double p, int *c) {
#pragma clang fp contract(off)
  const double f = floor(0.5 + (double)N * p);
  int64_t k = (int64_t)(N / 2) + (int64_t)f;
  if (k < 0) k += N;
  if (k < 0 || k >= N) return false;
  *c = (int)k;
  return true;
}

// Workgroup barrier after LDS atomics whose results are not used
// (ds_add_f32 / ds_add_u32 without return).  The compiler does not wait for
// such operations before s_barrier, so another wavefront could read a cell
// before the atomic has landed; drain the LDS counter explicitly first.
__device__ __forceinline__ void barrier_after_lds_atomics() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __syncthreads();
}

// Extent of the kernel bank in elements: the largest plane_off[p] + table
// size of plane p (Q*Q*S*S, or Q*S separable), 0 for none.  Called by
// every thread of a workgroup; s_ext is the workgroup's LDS cell for the
// maximum.
__device__ __forceinline__ int64_t bank_extent(const int64_t *plane_off, const int32_t *plane_s,
                                               int nplanes, int Q, int separable,
                                               unsigned long long *s_ext) {
  if (threadIdx.x == 0) *s_ext = 0ull;
  __syncthreads();
  int64_t e = 0;
  for (int p = threadIdx.x; p < nplanes; p += blockDim.x) {
    const int64_t S = plane_s[p];
    e = max(e, plane_off[p] + (separable ? Q * S : (int64_t)Q * Q * S * S));
  }
  if (e > 0) atomicMax(s_ext, (unsigned long long)e);
  barrier_after_lds_atomics();
  return (int64_t)*s_ext;
}

// ---------------------------------------------------------------------------
// Binning of visibilities by the grid tile holding their kernel window's
// first cell (shared by gridding and degridding).
// ---------------------------------------------------------------------------
struct BinArgs {
  int N;
  const double *uvw;
  int64_t nv, stride;
  double lam;
  int Q;
  int nplanes;
  const int32_t *plane_s;
  const int64_t *plane_off;
  int separable;
  int plane_mode;
  const int32_t *vplane;
  double wstep;
  int wbin0;
  int max_support;    // every plane_s must lie in [1, max_support]
};

// Tile-kernel record of one visibility (16 bytes):
//   x  window origin x0 | y0 << 16
//   y  support S | xf << 8 | yf << 20      (S <= 64, Q <= 4096)
//   z  kernel-bank element offset of its table: the 2D table [S][S] of
//      (yf, xf) in plane p, or plane p's separable table [Q][S]
//   w  visibility index (before the scatter: its bin)
// Bins are (tile, plane-group) pairs, bin = tile * ngroups + group, where a
// plane-group is a slice of group_elems consecutive kernel-bank elements, so
// that the tables one group uses fit in an XCD's L2.
struct BinPlan {
  int T = 64, R = 95, ntx = 0, nty = 0, ntiles = 0, chunk = 4096;
  int ngroups = 1, nbins = 0;
  int64_t group_elems = 0;
  int64_t max_items = 0;
  uint4 *loc = nullptr, *rec = nullptr;
  uint32_t *count = nullptr, *offset = nullptr, *cursor = nullptr;
  int4 *items = nullptr;    // tile-major (gridding)
  int4 *items_g = nullptr;  // per-XCD queues (degridding)
  uint32_t *gstart = nullptr;  // [ngroups + 1]: first item of each queue's first group
  uint32_t *status = nullptr, *nitems = nullptr, *nitems_g = nullptr, *qhead = nullptr;
};

// Work queues of the degridding kernel: the plane groups are split into
// XCD_QUEUES contiguous ranges, one per XCD (each XCD has its own L2), so
// that every XCD reads kernel taps from its own slice of the bank.
constexpr int XCD_QUEUES = 8;

// XCD this workgroup runs on (hardware register XCC_ID, gfx940 and later).
__device__ __forceinline__ int xcd_id() {
  return (int)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & (XCD_QUEUES - 1));
}

// Region edges the tile kernels hold in LDS: a tile of T cells with
// windows of support up to S needs T + S - 1 <= R.  The smaller region lets
// more workgroups share a compute unit (grid: 63.5 vs 141 KiB).
constexpr int TILE_R = 95;
constexpr int TILE_R_SMALL = 63;

// round-half-even(x * 2^e) as a 64-bit two's complement integer, for
// |x * 2^e| < 2^51.  x * 2^e is exact, and adding 1.5 * 2^52 in the same fma
// rounds it once to an integer held in the low mantissa bits of a double in
// [2^52, 2^53): the bit pattern is 0x4338000000000000 + round(x * 2^e).
// Three instructions instead of the emulated double -> int64 conversion.
__device__ __forceinline__ unsigned long long fixed(float x, double scale) {
  const double d = __builtin_fma((double)x, scale, 6755399441055744.0);  // 1.5 * 2^52
  return (unsigned long long)__double_as_longlong(d) - 0x4338000000000000ull;
}

// round(a * b + c * d) as a 64-bit integer, for |a * b| + |c * d| < 2^50
// (a and c carry the fixed-point scale 2^e): the inner fma rounds c * d +
// 1.5 * 2^52 once to an integer, the outer adds a * b and rounds once more,
// so the result is within 1 of the exact value.  Products of FP32 inputs
// are exact in FP64; four fma and two integer adds per complex tap.
__device__ __forceinline__ unsigned long long fixed2(double a, double b, double c, double d) {
  const double s = __builtin_fma(a, b, __builtin_fma(c, d, 6755399441055744.0));
  return (unsigned long long)__double_as_longlong(s) - 0x4338000000000000ull;
}

// Tap t of an S x S window (row-major, S <= 64, t < S * S) lies in row
// t / S.  tap_div(S) returns m with 2^20 / S + 0.875 <= m <= 2^20 / S +
// 2.125 (v_rcp_f32 is within 1 ulp); then (t * m) >> 20 == t / S, because
// the excess t * (m - 2^20 / S) / 2^20 < 4096 * 2.125 / 2^20 stays below
// 1 / S.  t * m < 2^32 and both factors fit the 24-bit multiplier.
__device__ __forceinline__ uint32_t tap_div(int S) {
  return (uint32_t)(1048576.0f * __builtin_amdgcn_rcpf((float)S)) + 2u;
}
__device__ __forceinline__ int tap_row(int t, uint32_t m) {
  return (int)(__umul24((uint32_t)t, m) >> 20);
}

// p[off / 8] for a wave-uniform table pointer and a 32-bit byte offset
// (scalar base + vector offset addressing, no 64-bit address arithmetic).
__device__ __forceinline__ float2 ldk(const float2 *p, uint32_t off) {
  return *(const float2 *)((const char *)p + off);
}

// Load / store of a complex64 with the non-temporal hint: data read or
// written once stays out of the way of the kernel tables in L2.
typedef float croc_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 load_stream(const float2 *p) {
  const croc_v2f v = __builtin_nontemporal_load((const croc_v2f *)p);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void store_stream(float2 *p, float2 v) {
  croc_v2f w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, (croc_v2f *)p);
}

// max(|re|, |im|) of a finite value; 0 and `bad` set for a NaN or infinite
// component (NaN fails every comparison).
__device__ __forceinline__ float cabs_max(float2 a, bool &bad) {
  const float x = fabsf(a.x), y = fabsf(a.y);
  const bool fin = x <= 3.402823466e38f && y <= 3.402823466e38f;
  bad |= !fin;
  return fin ? fmaxf(x, y) : 0.f;
}

__device__ __forceinline__ bool finite2(float2 a) {
  return fabsf(a.x) <= 3.402823466e38f && fabsf(a.y) <= 3.402823466e38f;
}

// max |re|, |im| of the finite values among a[i0], a[i0 + stride], ...
// (`bad` set when one is not finite): eight independent loads in flight per
// thread per round (the scan is latency-bound otherwise).
__device__ __forceinline__ float array_max(const float2 *__restrict__ a, int64_t n, int64_t i0,
                                           int64_t stride, bool &bad) {
  float m[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int64_t i = i0;
  for (; i + 7 * stride < n; i += 8 * stride) {
    float2 x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = a[i + k * stride];
#pragma unroll
    for (int k = 0; k < 8; k++) m[k] = fmaxf(m[k], cabs_max(x[k], bad));
  }
  for (; i < n; i += stride) m[0] = fmaxf(m[0], cabs_max(a[i], bad));
#pragma unroll
  for (int k = 1; k < 8; k++) m[0] = fmaxf(m[0], m[k]);
  return m[0];
}

// Block-wide maximum of non-negative floats, combined into *dst (float
// bits; non-negative floats order like their bit patterns) with one global
// atomic per block: thousands of wavefronts hitting one address serialise.
// blockDim.x must be a multiple of 64, at most 1024.
__device__ __forceinline__ void block_max_to(float v, unsigned int *dst) {
  __shared__ unsigned int s_bm;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  if (threadIdx.x == 0) s_bm = 0u;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicMax(&s_bm, __float_as_uint(v));
  barrier_after_lds_atomics();
  if (threadIdx.x == 0 && s_bm) atomicMax(dst, s_bm);
  __syncthreads();
}

// Values every lane of the wavefront loaded from the same address (the
// record of the visibility the wave works on): kept in scalar registers, so
// the arithmetic derived from them runs on the scalar unit.
__device__ __forceinline__ uint4 uniform_u4(uint4 r) {
  return make_uint4(__builtin_amdgcn_readfirstlane(r.x), __builtin_amdgcn_readfirstlane(r.y),
                    __builtin_amdgcn_readfirstlane(r.z), __builtin_amdgcn_readfirstlane(r.w));
}

__device__ __forceinline__ float2 uniform_f2(float2 v) {
  return make_float2(__int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.x))),
                     __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.y))));
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

int choose_tile(croc_ctx *ctx, int max_support);
int choose_region(int T, int max_support);
// Tile-kernel launches on binned visibilities (grid.hip, degrid.hip); they
// read no status.  Record w holds the visibility's index into vis.
int grid_binned(croc_ctx *ctx, const BinPlan &plan, void *grid, int N,
                int64_t nv, const void *vis, const void *kern, int Q, int separable,
                int nplanes, const int64_t *plane_off, const int32_t *plane_s);
int degrid_binned(croc_ctx *ctx, const BinPlan &plan, const void *grid, int N, void *vis_out,
                  const void *kern, int separable);
// Work-item lists (bit mask): ITEMS_BY_TILE cuts each tile's records (all
// groups) into items (gridding: one LDS flush per item); ITEMS_BY_GROUP
// splits the plane groups into one contiguous range per XCD queue and cuts
// each tile's records of a range into items (degridding: each XCD reads
// kernel taps from its own slice of the bank).
enum { ITEMS_BY_TILE = 1, ITEMS_BY_GROUP = 2 };
int bin_visibilities(croc_ctx *ctx, WsSet &set, const BinArgs &a, int max_support,
                     int order, BinPlan *plan);
int read_status(croc_ctx *ctx, const BinPlan &plan);
int persistent_blocks(croc_ctx *ctx, int per_cu);
int zero_u32_async(croc_ctx *ctx, uint32_t *p, int64_t n);

// Batched centred 2D DFT (dft.hip): out[oi(b)] = out_scale * D in[ii(b)] D
// with optional per-pixel modulation exp(i sign k[b] phi[pixel]) of the
// input and / or output; ii / oi are in_idx / out_idx or the identity.
// Same contract on the vector ALUs (fft.hip): radix 16 x 8 FFT in LDS, FP32,
// n = 128 only.
int fft2d_launch(croc_ctx *ctx, const void *in, void *out, int n, int64_t batch,
                 int64_t batch_stride, int sign, float out_scale, const double *phi_in,
                 const int32_t *k_in, int sign_in, const double *phi_out,
                 const int32_t *k_out, int sign_out, const int32_t *in_idx,
                 const int32_t *out_idx);
int dft2d_launch(croc_ctx *ctx, const void *in, void *out, int n, int64_t batch,
                 int64_t batch_stride, int sign, float out_scale, const double *phi_in,
                 const int32_t *k_in, int sign_in, const double *phi_out,
                 const int32_t *k_out, int sign_out, const int32_t *in_idx,
                 const int32_t *out_idx);
