// The reference's native visibility layout -- baselines with their own dump
// and channel axes, uvw in metres (examples/grid/grid.h:14-62, struct
// bl_data / vis_data) -- and the functions of examples/grid over it:
//   weight / degrid_simple / grid_simple     grid.c:119-177
//   degrid_conv_uv / degrid_conv_bl / degrid_conv   grid.c:179-275
//   grid_wprojection (w_project)             grid.c:277-322
//   streamer_degrid_worker                   streamer.c:366-486
//
// The separable degridder is a kernel of its own (below).  The gridding
// functions reuse the binned tile gridder (exact int64 fixed point): a
// first pass turns every visibility's grid.c coordinate into the grid
// coordinate p whose synthesis.py frac_coord gives the same cell and
// sub-offset, so the tile path lands each visibility exactly where grid.c
// does.
#include "internal.h"

namespace {

constexpr double C_LIGHT = 299792458.0;  // grid.h:51
constexpr int BL_NT = 256;

struct BlView {
  int64_t nbl, uvw_rows, freq_len, nvis;
  const int32_t *nt, *nf;
  const int64_t *uvw0, *f0, *v0;
  const double *uvw_m, *freq;
  float2 *vis;
};

BlView view(const croc_bl_set *s) {
  return BlView{s->nbl,        s->uvw_rows,   s->freq_len,   s->nvis,
                s->time_count, s->freq_count, s->uvw_first,  s->freq_first,
                s->vis_first,  s->uvw_m,      s->freq,       (float2 *)s->vis};
}

// grid.h:53-60: uvw_m * f / c, the product rounded before the division.
__device__ __forceinline__ double bl_lambda(const BlView &s, int64_t row, int64_t fi, int k) {
#pragma clang fp contract(off)
  const double prod = s.uvw_m[3 * row + k] * s.freq[fi];
  return prod / C_LIGHT;
}

// grid.c:35-53 for x = theta * u: ox = rint((N/2 - x) Q), cell N - ox/Q,
// sub-offset ox % Q (C truncating division).  False when x lies beyond the
// grid's upper edge (ox < 0: a negative sub-offset) or is not finite.
__device__ __forceinline__ bool frac_coord_c(int N, int Q, double x, int64_t *ix, int *fx) {
#pragma clang fp contract(off)
  const double s = ((double)(N / 2) - x) * (double)Q;
  if (!(fabs(s) < 4.0e18)) return false;
  const int64_t ox = (int64_t)rint(s);
  if (ox < 0) return false;
  *ix = (int64_t)N - ox / Q;
  *fx = (int)(ox % Q);
  return true;
}

// Window of a visibility at (x, y) = theta * (u, v): origin (x0, y0) and
// sub-offsets; false when any of its S x S cells leaves the grid.
__device__ __forceinline__ bool bl_window(int N, int Q, int S, double theta, double u, double v,
                                          int *x0, int *y0, int *fx, int *fy) {
#pragma clang fp contract(off)
  int64_t ix, iy;
  if (!frac_coord_c(N, Q, theta * u, &ix, fx) || !frac_coord_c(N, Q, theta * v, &iy, fy))
    return false;
  const int64_t ax = ix - S / 2, ay = iy - S / 2;
  if (ax < 0 || ay < 0 || ax + S > N || ay + S > N) return false;
  *x0 = (int)ax;
  *y0 = (int)ay;
  return true;
}

// grid.c:22-32: nearest cell floor(theta * u + .5) + N/2; false outside.
__device__ __forceinline__ bool bl_cell(int N, double theta, double u, int *c) {
#pragma clang fp contract(off)
  const double t = floor(theta * u + .5);
  if (!(fabs(t) < 4.0e18)) return false;
  const int64_t k = (int64_t)t + N / 2;
  if (k < 0 || k >= N) return false;
  *c = (int)k;
  return true;
}

// ---------------------------------------------------------------------------
// Work decomposition: chunks of visibilities (a baseline's rectangle of
// dumps x channels), one workgroup per chunk at a time.
// ---------------------------------------------------------------------------
enum ChunkMode {
  CH_FULL = 0,   // chunk b = baseline b, all dumps and channels; out = vis layout
  CH_WINDOW,     // chunk b = baseline b, [t0, t1) x [f0, f1); chunk-local out index
  CH_LIST,       // chunks from a device list (b, t0, t1, f0, f1); out = queue slots
  CH_ROWS,       // a plain coordinate list uv[n][2] cut into chunks of `rows` entries
};

enum Select {
  SEL_ALL = 0,   // degrid every visibility (degrid_conv_uv)
  SEL_WINDOW,    // degrid_conv_bl: inside [min, max) or untouched
  SEL_MIRROR,    // streamer: inside, else the mirrored coordinate, else 0
};

struct Chunks {
  int mode;
  int64_t n;
  const int32_t *list;
  int t0, t1, f0, f1;       // CH_WINDOW
  int64_t slot;             // CH_LIST: queue elements per chunk
  int slot_row;             // CH_LIST: channels per slot row (freq_chunk)
  int slot_t, slot_f;       // CH_LIST: time_chunk, freq_chunk (extent limits)
  const double *uv;         // CH_ROWS
  int64_t nrows, rows;      // CH_ROWS
};

struct Chunk {
  int64_t b;
  int t0, t1, f0, f1;
  int64_t out0;  // output index of (t0, f0)
  int64_t orow;  // output elements per dump
  bool ok;
};

__device__ __forceinline__ Chunk chunk_at(const BlView &s, const Chunks &c, int64_t k) {
  Chunk r{0, 0, 0, 0, 0, 0, 0, true};
  if (c.mode == CH_ROWS) {
    r.t0 = 0;
    r.t1 = (int)min(c.rows, c.nrows - k * c.rows);
    r.f0 = 0;
    r.f1 = 1;
    r.out0 = k * c.rows;
    r.orow = 1;
    r.b = k;
    return r;
  }
  if (c.mode == CH_LIST) {
    const int32_t *e = c.list + 5 * k;
    r.b = e[0];
    r.t0 = e[1];
    r.t1 = e[2];
    r.f0 = e[3];
    r.f1 = e[4];
    r.out0 = k * c.slot;
    r.orow = c.slot_row;
    r.ok = r.t1 - r.t0 <= c.slot_t && r.f1 - r.f0 <= c.slot_f;
  } else {
    r.b = k;
    if (c.mode == CH_FULL) {
      r.t1 = s.nt[k];
      r.f1 = s.nf[k];
    } else {
      r.t0 = c.t0;
      r.t1 = c.t1;
      r.f0 = c.f0;
      r.f1 = c.f1;
    }
  }
  if (r.b < 0 || r.b >= s.nbl) {
    r.ok = false;
    return r;
  }
  const int64_t nt = s.nt[r.b], nf = s.nf[r.b];
  r.ok = r.ok && r.t0 >= 0 && r.f0 >= 0 && r.t0 <= r.t1 && r.f0 <= r.f1 && r.t1 <= nt &&
         r.f1 <= nf && nt >= 0 && nf >= 0 && s.uvw0[r.b] >= 0 && s.f0[r.b] >= 0 &&
         s.v0[r.b] >= 0 && s.uvw0[r.b] + nt <= s.uvw_rows && s.f0[r.b] + nf <= s.freq_len &&
         s.v0[r.b] + nt * nf <= s.nvis;
  if (c.mode == CH_FULL) {
    r.out0 = s.v0[r.b];
    r.orow = nf;
  } else if (c.mode == CH_WINDOW) {
    r.out0 = s.v0[r.b];
    r.orow = r.f1 - r.f0;
  }
  return r;
}

// (u, v) of visibility (t, f) of a chunk, in wavelengths.
__device__ __forceinline__ void chunk_uv(const BlView &s, const Chunks &c, const Chunk &r, int t,
                                         int f, double *u, double *v) {
  if (c.mode == CH_ROWS) {
    const int64_t i = r.out0 + t;
    *u = c.uv[2 * i];
    *v = c.uv[2 * i + 1];
    return;
  }
  const int64_t row = s.uvw0[r.b] + t, fi = s.f0[r.b] + f;
  *u = bl_lambda(s, row, fi, 0);
  *v = bl_lambda(s, row, fi, 1);
}

struct Box {
  int sel;
  double du, dv, min_u, max_u, min_v, max_v;
};

// What happens to a visibility: 0 untouched, 1 degrid at (*pu, *pv), 2
// degrid there and conjugate, 3 write zero (streamer.c:418-438,
// grid.c:247-255).
__device__ __forceinline__ int classify(const Box &b, double u, double v, double *pu,
                                        double *pv) {
#pragma clang fp contract(off)
  if (b.sel == SEL_ALL) {
    *pu = u;
    *pv = v;
    return 1;
  }
  if (b.sel == SEL_WINDOW) {
    if (u < b.min_u || u >= b.max_u || v < b.min_v || v >= b.max_v) return 0;
    *pu = u - b.du;
    *pv = v - b.dv;
    return 1;
  }
  if (u >= fmax(b.min_u, 0.0) && u < b.max_u && v >= b.min_v && v < b.max_v) {
    *pu = u - b.du;
    *pv = v - b.dv;
    return 1;
  }
  if (-u >= fmax(b.min_u, 2.2250738585072014e-308) && -u < b.max_u && -v >= b.min_v &&
      -v < b.max_v) {
    *pu = -u - b.du;
    *pv = -v - b.dv;
    return 2;
  }
  return 3;
}

struct DegridArgs {
  BlView s;
  Chunks c;
  Box box;
  const float2 *grid;
  int N, S, Q;
  double theta;
  const double *kern;  // real [Q][S]
  float2 *out;
  unsigned long long *count;  // [n chunks] degridded visibilities (check pass)
  uint32_t *status;
};

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long x,
                                                            unsigned long long *red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
  __syncthreads();
  return t;
}

// Validation pass: ranges, windows, per-chunk counts.  Nothing written but
// counts and status.
__global__ __launch_bounds__(BL_NT) void k_bl_degrid_check(DegridArgs a) {
  __shared__ unsigned long long red[BL_NT / 64];
  uint32_t st = 0;
  for (int64_t k = blockIdx.x; k < a.c.n; k += gridDim.x) {
    const Chunk r = chunk_at(a.s, a.c, k);
    unsigned long long n = 0;
    if (!r.ok) {
      st |= ST_RANGE;
    } else {
      const int nfw = r.f1 - r.f0;
      const int64_t m = (int64_t)(r.t1 - r.t0) * nfw;
      for (int64_t j = threadIdx.x; j < m; j += BL_NT) {
        const int t = r.t0 + (int)(j / nfw), f = r.f0 + (int)(j % nfw);
        double u, v, pu, pv;
        chunk_uv(a.s, a.c, r, t, f, &u, &v);
        const int kind = classify(a.box, u, v, &pu, &pv);
        if (kind == 1 || kind == 2) {
          int x0, y0, fx, fy;
          if (!bl_window(a.N, a.Q, a.S, a.theta, pu, pv, &x0, &y0, &fx, &fy)) st |= ST_BOUNDS;
          n++;
        }
      }
    }
    n = block_sum_u64(n, red);
    if (threadIdx.x == 0 && a.count) a.count[k] = n;
  }
  if (st) atomicOr(a.status, st);
}

// Degridding: G lanes per visibility (G a power of two >= S); lane x sums
// column x of the window over the rows weighted by k[fy][y], the group then
// sums the columns weighted by k[fx][x].  Rows are read as G consecutive
// complex64 cells.  FP32 arithmetic on the FP64 table's values.
template <int G>
__global__ __launch_bounds__(BL_NT) void k_bl_degrid(DegridArgs a) {
  constexpr int V = BL_NT / G;
  const int lane = threadIdx.x & (G - 1);
  const int slot = threadIdx.x / G;
  const int S = a.S, N = a.N;
  for (int64_t k = blockIdx.x; k < a.c.n; k += gridDim.x) {
    const Chunk r = chunk_at(a.s, a.c, k);
    const int nfw = r.f1 - r.f0;
    const int64_t m = (int64_t)(r.t1 - r.t0) * nfw;
    for (int64_t j = slot; j < m; j += V) {
      const int t = r.t0 + (int)(j / nfw), f = r.f0 + (int)(j % nfw);
      const int64_t o = r.out0 + (int64_t)(t - r.t0) * r.orow + (f - r.f0);
      double u, v, pu, pv;
      chunk_uv(a.s, a.c, r, t, f, &u, &v);
      const int kind = classify(a.box, u, v, &pu, &pv);
      if (kind == 0) continue;
      if (kind == 3) {
        if (lane == 0) a.out[o] = make_float2(0.f, 0.f);
        continue;
      }
      int x0 = 0, y0 = 0, fx = 0, fy = 0;
      bl_window(N, a.Q, S, a.theta, pu, pv, &x0, &y0, &fx, &fy);
      float ar = 0.f, ai = 0.f;
      if (lane < S) {
        const float2 *col = a.grid + (int64_t)y0 * N + x0 + lane;
        const double *ky = a.kern + (int64_t)S * fy;
        for (int y = 0; y < S; y++) {
          const float ky_ = (float)ky[y];
          const float2 g = col[(int64_t)y * N];
          ar = fmaf(ky_, g.x, ar);
          ai = fmaf(ky_, g.y, ai);
        }
        const float kx = (float)a.kern[(int64_t)S * fx + lane];
        ar *= kx;
        ai *= kx;
      }
#pragma unroll
      for (int w = G / 2; w > 0; w >>= 1) {
        ar += __shfl_xor(ar, w);
        ai += __shfl_xor(ai, w);
      }
      if (lane == 0) a.out[o] = make_float2(ar, kind == 2 ? -ai : ai);
    }
  }
}

// ---------------------------------------------------------------------------
// Nearest-cell functions and the coordinate pass of the gridding functions.
// ---------------------------------------------------------------------------
enum FlatMode { FLAT_NEAREST = 0, FLAT_WPROJ = 1 };

struct FlatArgs {
  BlView s;
  int mode;
  int N, Q, S;
  double theta, w_min, w_step;
  int plane_count;
  double *p;          // [nvis][2]: grid coordinate of synthesis.py's frac_coord
  int32_t *plane;     // [nvis] (FLAT_WPROJ)
  float2 *vis_c;      // [nvis]: vis, conjugated for FLAT_WPROJ
  uint32_t *status;
};

__global__ void k_bl_flat_fill(FlatArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.s.nvis) return;
  a.p[2 * i] = 0.0;
  a.p[2 * i + 1] = 0.0;
  if (a.plane) a.plane[i] = 0;
  a.vis_c[i] = make_float2(0.f, 0.f);
}

// p with x = N//2 + p N equal to ix - fx/Q: synthesis.py's frac_coord
// returns (ix, fx) for it (the value is exact for power-of-two N and Q,
// and within a few ulps of it otherwise, far from any rounding boundary).
__device__ __forceinline__ double p_of(int N, int Q, int64_t ix, int fx) {
#pragma clang fp contract(off)
  const double x = (double)(ix - N / 2) - (double)fx / (double)Q;
  return x / (double)N;
}

// w_project's plane, grid.c:286: floor((w - w_min) / w_step + .5).
__device__ __forceinline__ double w_plane(double w, double w_min, double w_step) {
#pragma clang fp contract(off)
  return floor((w - w_min) / w_step + .5);
}

__global__ __launch_bounds__(BL_NT) void k_bl_flatten(FlatArgs a) {
  uint32_t st = 0;
  Chunks c{};
  c.mode = CH_FULL;
  c.n = a.s.nbl;
  for (int64_t k = blockIdx.x; k < a.s.nbl; k += gridDim.x) {
    const Chunk r = chunk_at(a.s, c, k);
    if (!r.ok) {
      st |= ST_RANGE;
      continue;
    }
    const int nfw = r.f1;
    const int64_t m = (int64_t)r.t1 * nfw;
    for (int64_t j = threadIdx.x; j < m; j += BL_NT) {
      const int t = (int)(j / nfw), f = (int)(j % nfw);
      const int64_t i = r.out0 + j;
      double u, v;
      chunk_uv(a.s, c, r, t, f, &u, &v);
      double px, py;
      const float2 vis = a.s.vis[i];
      if (a.mode == FLAT_NEAREST) {
        int cx, cy;
        if (!bl_cell(a.N, a.theta, u, &cx) || !bl_cell(a.N, a.theta, v, &cy)) {
          st |= ST_BOUNDS;
          continue;
        }
        px = p_of(a.N, 1, cx, 0);
        py = p_of(a.N, 1, cy, 0);
        a.vis_c[i] = vis;
      } else {
        int x0, y0, fx, fy;
        if (!bl_window(a.N, a.Q, a.S, a.theta, u, v, &x0, &y0, &fx, &fy)) {
          st |= ST_BOUNDS;
          continue;
        }
        px = p_of(a.N, a.Q, x0 + a.S / 2, fx);
        py = p_of(a.N, a.Q, y0 + a.S / 2, fy);
        const double q = w_plane(bl_lambda(a.s, a.s.uvw0[r.b] + t, a.s.f0[r.b] + f, 2),
                                 a.w_min, a.w_step);
        if (!(q >= 0.0 && q < (double)a.plane_count)) {
          st |= ST_PLANE;
          continue;
        }
        a.plane[i] = (int32_t)q;
        a.vis_c[i] = make_float2(vis.x, -vis.y);
      }
      if (!(px >= -0.5 && px < 0.5 && py >= -0.5 && py < 0.5)) {
        st |= ST_BOUNDS;
        continue;
      }
      a.p[2 * i] = px;
      a.p[2 * i + 1] = py;
    }
  }
  if (st) atomicOr(a.status, st);
}

__global__ void k_conj_inplace(float2 *g, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i].y = -g[i].y;
}

// Kernel-bank plane tables of a w-kernel set (plane k at k Q^2 S^2) and the
// unit 1 x 1 kernel of grid_simple.
__global__ void k_bl_bank(int64_t *plane_off, int32_t *plane_s, int nplanes, int Q, int S,
                          float2 *unit) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nplanes) {
    plane_off[k] = (int64_t)k * Q * Q * S * S;
    plane_s[k] = S;
  }
  if (k == 0 && unit) *unit = make_float2(1.f, 0.f);
}

enum CellOp { CELL_COUNT = 0, CELL_DIVIDE, CELL_GATHER, CELL_CHECK };

struct CellArgs {
  BlView s;
  int op, N;
  double theta;
  uint32_t *wgrid;
  const float2 *grid;
  uint32_t *status;
};

__global__ __launch_bounds__(BL_NT) void k_bl_cells(CellArgs a) {
  uint32_t st = 0;
  Chunks c{};
  c.mode = CH_FULL;
  c.n = a.s.nbl;
  for (int64_t k = blockIdx.x; k < a.s.nbl; k += gridDim.x) {
    const Chunk r = chunk_at(a.s, c, k);
    if (!r.ok) {
      st |= ST_RANGE;
      continue;
    }
    const int nfw = r.f1;
    const int64_t m = (int64_t)r.t1 * nfw;
    for (int64_t j = threadIdx.x; j < m; j += BL_NT) {
      const int t = (int)(j / nfw), f = (int)(j % nfw);
      double u, v;
      chunk_uv(a.s, c, r, t, f, &u, &v);
      int cx, cy;
      if (!bl_cell(a.N, a.theta, u, &cx) || !bl_cell(a.N, a.theta, v, &cy)) {
        st |= ST_BOUNDS;
        continue;
      }
      const int64_t cell = (int64_t)cy * a.N + cx, i = r.out0 + j;
      if (a.op == CELL_COUNT) {
        atomicAdd(&a.wgrid[cell], 1u);
      } else if (a.op == CELL_DIVIDE) {
        const float n = (float)a.wgrid[cell];
        const float2 x = a.s.vis[i];
        a.s.vis[i] = make_float2(x.x / n, x.y / n);
      } else if (a.op == CELL_GATHER) {
        a.s.vis[i] = a.grid[cell];
      }
    }
  }
  if (st) atomicOr(a.status, st);
}

int status_word(croc_ctx *ctx, uint32_t **st) {
  *st = (uint32_t *)ws_get(ctx, WS_SMALL, 256);
  if (!*st) return CROC_ENOMEM;
  return zero_u32_async(ctx, *st, 1);
}

int read_word(croc_ctx *ctx, uint32_t *st) {
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipMemcpyAsync(ctx->host_status, st, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return croc_status_to_code(*ctx->host_status);
}

int blocks_for(croc_ctx *ctx, int64_t n) {
  const int64_t cap = (int64_t)ctx->num_cu * 8;
  return (int)(n < 1 ? 1 : (n < cap ? n : cap));
}

int check_set(const croc_bl_set *s) {
  CROC_REQUIRE(s, CROC_EINVAL, "null baseline set");
  CROC_REQUIRE(s->nbl >= 0 && s->nvis >= 0 && s->uvw_rows >= 0 && s->freq_len >= 0,
               CROC_EINVAL, "negative baseline-set size");
  CROC_REQUIRE(s->nbl == 0 || (s->time_count && s->freq_count && s->uvw_first &&
                               s->freq_first && s->vis_first && s->uvw_m && s->freq &&
                               s->vis),
               CROC_EINVAL, "null array in the baseline set");
  return CROC_OK;
}

int check_kernel(const double *kern, int S, int Q) {
  CROC_REQUIRE(kern, CROC_EINVAL, "null kernel");
  CROC_REQUIRE(S >= 1 && S <= 64, CROC_EINVAL, "kernel support must lie in [1, 64]");
  CROC_REQUIRE(Q >= 1 && Q <= (1 << 20), CROC_EINVAL, "oversampling must lie in [1, 2^20]");
  return CROC_OK;
}

// Check pass, then the degridding pass.  count: device [nchunks] or null.
int degrid_run(croc_ctx *ctx, DegridArgs a, uint64_t *ndeg_total) {
  if (a.c.n == 0) {
    if (ndeg_total) *ndeg_total = 0;
    return CROC_OK;
  }
  uint32_t *st;
  int rc = status_word(ctx, &st);
  if (rc) return rc;
  a.status = st;
  const bool own = a.count == nullptr;
  if (own) {
    a.count = (unsigned long long *)ws_get(ctx, WS_BL, sizeof(unsigned long long) * a.c.n);
    if (!a.count) return CROC_ENOMEM;
  }
  const int nb = blocks_for(ctx, a.c.n);
  k_bl_degrid_check<<<nb, BL_NT, 0, ctx->stream>>>(a);
  if ((rc = read_word(ctx, st))) return rc;
  if (ndeg_total) {
    // the counts are small: sum them on the host
    std::vector<unsigned long long> h(a.c.n);
    CROC_HIP(hipMemcpyAsync(h.data(), a.count, sizeof(unsigned long long) * a.c.n,
                            hipMemcpyDeviceToHost, ctx->stream));
    CROC_HIP(hipStreamSynchronize(ctx->stream));
    unsigned long long t = 0;
    for (auto x : h) t += x;
    *ndeg_total = t;
  }
  CROC_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
  tap_timer_reset(ctx, 1);
  if ((rc = tap_timer_begin(ctx, 1))) return rc;
  const int S = a.S;
  if (S <= 4)
    k_bl_degrid<4><<<nb, BL_NT, 0, ctx->stream>>>(a);
  else if (S <= 8)
    k_bl_degrid<8><<<nb, BL_NT, 0, ctx->stream>>>(a);
  else if (S <= 16)
    k_bl_degrid<16><<<nb, BL_NT, 0, ctx->stream>>>(a);
  else if (S <= 32)
    k_bl_degrid<32><<<nb, BL_NT, 0, ctx->stream>>>(a);
  else
    k_bl_degrid<64><<<nb, BL_NT, 0, ctx->stream>>>(a);
  CROC_HIP(hipGetLastError());
  if ((rc = tap_timer_end(ctx, 1))) return rc;
  CROC_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
  ctx->ev_degrid = true;
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}

uint64_t degrid_flops(uint64_t n, int S) { return n * 4ull * (uint64_t)(1 + S) * (uint64_t)S; }

// Coordinate pass + the binned tile gridder (+ conjugation for w_project).
int grid_run(croc_ctx *ctx, FlatArgs a, void *uvgrid, const void *kern, int nplanes,
             int S, int Q) {
  const int64_t nv = a.s.nvis;
  uint32_t *st;
  int rc = status_word(ctx, &st);
  if (rc) return rc;
  a.status = st;
  const size_t p_bytes = ((size_t)16 * nv + 255) & ~(size_t)255;
  const size_t pl_bytes = ((size_t)4 * nv + 255) & ~(size_t)255;
  char *buf = (char *)ws_get(ctx, WS_BL, p_bytes + pl_bytes + (size_t)8 * nv);
  if (!buf) return CROC_ENOMEM;
  a.p = (double *)buf;
  a.plane = a.mode == FLAT_WPROJ ? (int32_t *)(buf + p_bytes) : nullptr;
  a.vis_c = (float2 *)(buf + p_bytes + pl_bytes);
  const int np = a.mode == FLAT_WPROJ ? nplanes : 1;
  char *bank = (char *)ws_get(ctx, WS_BL_BANK, (size_t)12 * np + 1024);
  if (!bank) return CROC_ENOMEM;
  int64_t *plane_off = (int64_t *)bank;
  int32_t *plane_s = (int32_t *)(bank + (((size_t)8 * np + 255) & ~(size_t)255));
  float2 *unit = (float2 *)(bank + (((size_t)12 * np + 255 + 256) & ~(size_t)255));
  k_bl_bank<<<(np + 255) / 256, 256, 0, ctx->stream>>>(plane_off, plane_s, np, Q, S,
                                                         a.mode == FLAT_NEAREST ? unit : nullptr);
  k_bl_flat_fill<<<(unsigned)((nv + 255) / 256), 256, 0, ctx->stream>>>(a);
  k_bl_flatten<<<blocks_for(ctx, a.s.nbl), BL_NT, 0, ctx->stream>>>(a);
  if ((rc = read_word(ctx, st))) return rc;
  const int64_t cells = (int64_t)a.N * a.N;
  if (a.mode == FLAT_WPROJ) {
    k_conj_inplace<<<(unsigned)((cells + 255) / 256), 256, 0, ctx->stream>>>(
        (float2 *)uvgrid, cells);
    CROC_HIP(hipGetLastError());
    rc = croc_convgrid(ctx, uvgrid, a.N, a.p, nv, 2, 1.0, a.vis_c, kern, Q, 0, nplanes,
                       plane_off, plane_s, S, CROC_PLANE_INDEX, a.plane, 0.0, 0);
    // conjugate back whatever the gridder returned (the grid is unchanged
    // on its errors)
    k_conj_inplace<<<(unsigned)((cells + 255) / 256), 256, 0, ctx->stream>>>(
        (float2 *)uvgrid, cells);
    CROC_HIP(hipGetLastError());
    if (rc) return rc;
  } else {
    rc = croc_convgrid(ctx, uvgrid, a.N, a.p, nv, 2, 1.0, a.vis_c, unit, 1, 0, 1, plane_off,
                       plane_s, 1, CROC_PLANE_SINGLE, nullptr, 0.0, 0);
    if (rc) return rc;
  }
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}

uint64_t set_visibilities(croc_ctx *ctx, const croc_bl_set *s, int *rc) {
  // total number of visibilities of the set: sum over baselines of
  // time_count * freq_count (for the flop counts)
  *rc = CROC_OK;
  if (s->nbl == 0) return 0;
  std::vector<int32_t> nt(s->nbl), nf(s->nbl);
  if (hipMemcpyAsync(nt.data(), s->time_count, 4 * s->nbl, hipMemcpyDeviceToHost,
                     ctx->stream) != hipSuccess ||
      hipMemcpyAsync(nf.data(), s->freq_count, 4 * s->nbl, hipMemcpyDeviceToHost,
                     ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    croc_set_error("reading the baseline set's counts failed");
    *rc = CROC_EHIP;
    return 0;
  }
  uint64_t n = 0;
  for (int64_t b = 0; b < s->nbl; b++) n += (uint64_t)max(nt[b], 0) * (uint64_t)max(nf[b], 0);
  return n;
}

}  // namespace

extern "C" int croc_bl_weight(croc_ctx *ctx, uint32_t *wgrid, int N, double theta,
                              const croc_bl_set *vis) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(vis))) return rc;
  CROC_REQUIRE(wgrid && N >= 1, CROC_EINVAL, "weight grid and N >= 1 required");
  uint32_t *st;
  if ((rc = status_word(ctx, &st))) return rc;
  CellArgs a{view(vis), CELL_CHECK, N, theta, wgrid, nullptr, st};
  const int nb = blocks_for(ctx, vis->nbl);
  if (vis->nbl) k_bl_cells<<<nb, BL_NT, 0, ctx->stream>>>(a);
  if ((rc = read_word(ctx, st))) return rc;
  if ((rc = zero_u32_async(ctx, wgrid, (int64_t)N * N))) return rc;
  if (vis->nbl == 0) {
    CROC_HIP(hipStreamSynchronize(ctx->stream));
    return CROC_OK;
  }
  a.op = CELL_COUNT;
  k_bl_cells<<<nb, BL_NT, 0, ctx->stream>>>(a);
  a.op = CELL_DIVIDE;
  k_bl_cells<<<nb, BL_NT, 0, ctx->stream>>>(a);
  return read_word(ctx, st);
}

extern "C" int croc_bl_degrid_simple(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                                     const croc_bl_set *vis, uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(vis))) return rc;
  CROC_REQUIRE(uvgrid && N >= 1, CROC_EINVAL, "grid and N >= 1 required");
  uint32_t *st;
  if ((rc = status_word(ctx, &st))) return rc;
  CellArgs a{view(vis), CELL_CHECK, N, theta, nullptr, (const float2 *)uvgrid, st};
  const int nb = blocks_for(ctx, vis->nbl);
  if (vis->nbl) k_bl_cells<<<nb, BL_NT, 0, ctx->stream>>>(a);
  if ((rc = read_word(ctx, st))) return rc;
  a.op = CELL_GATHER;
  if (vis->nbl) k_bl_cells<<<nb, BL_NT, 0, ctx->stream>>>(a);
  if ((rc = read_word(ctx, st))) return rc;
  if (flops) {
    const uint64_t n = set_visibilities(ctx, vis, &rc);
    if (rc) return rc;
    *flops = 2 * n;
  }
  return CROC_OK;
}

extern "C" int croc_bl_grid_simple(croc_ctx *ctx, void *uvgrid, int N, double theta,
                                   const croc_bl_set *vis, uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(vis))) return rc;
  CROC_REQUIRE(uvgrid && N >= 2 && N <= 65536, CROC_EINVAL, "grid and 2 <= N <= 65536 required");
  if (vis->nvis > 0) {
    FlatArgs a{};
    a.s = view(vis);
    a.mode = FLAT_NEAREST;
    a.N = N;
    a.Q = 1;
    a.S = 1;
    a.theta = theta;
    if ((rc = grid_run(ctx, a, uvgrid, nullptr, 1, 1, 1))) return rc;
  }
  if (flops) {
    const uint64_t n = set_visibilities(ctx, vis, &rc);
    if (rc) return rc;
    *flops = 2 * n;
  }
  return CROC_OK;
}

extern "C" int croc_grid_wprojection(croc_ctx *ctx, void *uvgrid, int N, double theta,
                                     const croc_bl_set *vis, const void *wkern,
                                     int plane_count, double w_min, double w_step, int S, int Q,
                                     uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(vis))) return rc;
  CROC_REQUIRE(uvgrid && wkern && N >= 2 && N <= 65536, CROC_EINVAL,
               "grid, kernels and 2 <= N <= 65536 required");
  CROC_REQUIRE(S >= 1 && S <= 64 && Q >= 1 && Q <= 4095 && plane_count >= 1, CROC_EINVAL,
               "need 1 <= S <= 64, 1 <= Q <= 4095, plane_count >= 1");
  CROC_REQUIRE((int64_t)plane_count * Q * Q * S * S <= (int64_t)UINT32_MAX, CROC_EINVAL,
               "w-kernel set larger than 2^32 complex elements");
  CROC_REQUIRE(w_step > 0, CROC_EINVAL, "w_step must be positive");
  if (vis->nvis > 0) {
    FlatArgs a{};
    a.s = view(vis);
    a.mode = FLAT_WPROJ;
    a.N = N;
    a.Q = Q;
    a.S = S;
    a.theta = theta;
    a.w_min = w_min;
    a.w_step = w_step;
    a.plane_count = plane_count;
    if ((rc = grid_run(ctx, a, uvgrid, wkern, plane_count, S, Q))) return rc;
  }
  if (flops) {
    const uint64_t n = set_visibilities(ctx, vis, &rc);
    if (rc) return rc;
    *flops = n * 8ull * (uint64_t)S * (uint64_t)S;
  }
  return CROC_OK;
}

extern "C" int croc_degrid_conv_uv(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                                   const double *uv, int64_t n, const double *kern, int S,
                                   int Q, void *vis_out, uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_kernel(kern, S, Q))) return rc;
  CROC_REQUIRE(uvgrid && N >= 1 && n >= 0 && (n == 0 || (uv && vis_out)), CROC_EINVAL,
               "null array argument");
  DegridArgs a{};
  a.c.mode = CH_ROWS;
  a.c.uv = uv;
  a.c.nrows = n;
  a.c.rows = 1024;
  a.c.n = (n + a.c.rows - 1) / a.c.rows;
  a.box.sel = SEL_ALL;
  a.grid = (const float2 *)uvgrid;
  a.N = N;
  a.S = S;
  a.Q = Q;
  a.theta = theta;
  a.kern = kern;
  a.out = (float2 *)vis_out;
  uint64_t nd = 0;
  if ((rc = degrid_run(ctx, a, &nd))) return rc;
  if (flops) *flops = degrid_flops(nd, S);
  return CROC_OK;
}

extern "C" int croc_degrid_conv_bl(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                                   double d_u, double d_v, double min_u, double max_u,
                                   double min_v, double max_v, const croc_bl_set *bl,
                                   int time0, int time1, int freq0, int freq1,
                                   const double *kern, int S, int Q, uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(bl))) return rc;
  if ((rc = check_kernel(kern, S, Q))) return rc;
  CROC_REQUIRE(uvgrid && N >= 1, CROC_EINVAL, "grid and N >= 1 required");
  CROC_REQUIRE(time0 >= 0 && freq0 >= 0 && time0 <= time1 && freq0 <= freq1, CROC_EINVAL,
               "need 0 <= time0 <= time1 and 0 <= freq0 <= freq1");
  DegridArgs a{};
  a.s = view(bl);
  a.c.mode = CH_WINDOW;
  a.c.n = bl->nbl;
  a.c.t0 = time0;
  a.c.t1 = time1;
  a.c.f0 = freq0;
  a.c.f1 = freq1;
  a.box = Box{SEL_WINDOW, d_u, d_v, min_u, max_u, min_v, max_v};
  a.grid = (const float2 *)uvgrid;
  a.N = N;
  a.S = S;
  a.Q = Q;
  a.theta = theta;
  a.kern = kern;
  a.out = (float2 *)bl->vis;
  uint64_t nd = 0;
  if ((rc = degrid_run(ctx, a, &nd))) return rc;
  if (flops) *flops = degrid_flops(nd, S);
  return CROC_OK;
}

extern "C" int croc_degrid_conv(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                                double d_u, double d_v, const croc_bl_set *vis,
                                const double *kern, int S, int Q, uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(vis))) return rc;
  if ((rc = check_kernel(kern, S, Q))) return rc;
  CROC_REQUIRE(uvgrid && N >= 1 && theta > 0, CROC_EINVAL, "grid, N >= 1, theta > 0 required");
  DegridArgs a{};
  a.s = view(vis);
  a.c.mode = CH_FULL;
  a.c.n = vis->nbl;
  {
#pragma clang fp contract(off)
    const double h = (double)N / theta / 2;  // grid.c:268-271
    a.box = Box{SEL_WINDOW, d_u, d_v, d_u - h, d_u + h, d_v - h, d_v + h};
  }
  a.grid = (const float2 *)uvgrid;
  a.N = N;
  a.S = S;
  a.Q = Q;
  a.theta = theta;
  a.kern = kern;
  a.out = (float2 *)vis->vis;
  uint64_t nd = 0;
  if ((rc = degrid_run(ctx, a, &nd))) return rc;
  if (flops) *flops = degrid_flops(nd, S);
  return CROC_OK;
}

extern "C" int croc_streamer_degrid(croc_ctx *ctx, const void *subgrid, int subgrid_size,
                                    double theta, double mid_u, double mid_v, double min_u,
                                    double max_u, double min_v, double max_v,
                                    const croc_bl_set *bl, const int32_t *chunks,
                                    int64_t nchunks, int time_chunk, int freq_chunk,
                                    const double *kern, int S, int Q, void *vis_queue,
                                    int64_t *ndegrid, uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(bl))) return rc;
  if ((rc = check_kernel(kern, S, Q))) return rc;
  CROC_REQUIRE(subgrid && subgrid_size >= 1, CROC_EINVAL, "subgrid and size >= 1 required");
  CROC_REQUIRE(nchunks >= 0 && (nchunks == 0 || (chunks && vis_queue)), CROC_EINVAL,
               "null chunk list / visibility queue");
  CROC_REQUIRE(time_chunk >= 1 && freq_chunk >= 1, CROC_EINVAL,
               "time_chunk and freq_chunk must be positive");
  DegridArgs a{};
  a.s = view(bl);
  a.c.mode = CH_LIST;
  a.c.n = nchunks;
  a.c.list = chunks;
  a.c.slot = (int64_t)time_chunk * freq_chunk;
  a.c.slot_row = freq_chunk;
  a.c.slot_t = time_chunk;
  a.c.slot_f = freq_chunk;
  a.box = Box{SEL_MIRROR, mid_u, mid_v, min_u, max_u, min_v, max_v};
  a.grid = (const float2 *)subgrid;
  a.N = subgrid_size;
  a.S = S;
  a.Q = Q;
  a.theta = theta;
  a.kern = kern;
  a.out = (float2 *)vis_queue;
  a.count = (unsigned long long *)ndegrid;
  uint64_t nd = 0;
  if ((rc = degrid_run(ctx, a, flops ? &nd : nullptr))) return rc;
  if (flops) *flops = degrid_flops(nd, S);
  return CROC_OK;
}
