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
// grid's upper edge (ox < 0: a negative sub-offset) or so far outside that
// no grid of at most 65536 cells with Q <= 2^20 could hold it.  The
// quotient comes from the FP64 reciprocal of Q (within 2^-12 of the true
// quotient for ox < 2^40), corrected by the remainder: exact.
__device__ __forceinline__ bool frac_coord_c(int N, int Q, double x, int64_t *ix, int *fx) {
#pragma clang fp contract(off)
  const double s = ((double)(N / 2) - x) * (double)Q;
  if (!(fabs(s) < 0x1p40)) return false;
  const double so = rint(s);
  if (so < 0.0) return false;
  const int64_t ox = (int64_t)so;
  int64_t q = (int64_t)(so * (1.0 / (double)Q));
  int64_t rem = ox - q * Q;
  if (rem < 0) {
    q -= 1;
    rem += Q;
  } else if (rem >= Q) {
    q += 1;
    rem -= Q;
  }
  *ix = (int64_t)N - q;
  *fx = (int)rem;
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
  CH_LIST,       // chunks from a device list ([sub,] b, t0, t1, f0, f1); out = queue slots
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
  int cols;                 // CH_LIST: 5 (b, t0, t1, f0, f1) or 6 (subgrid first)
  int nsub;                 // CH_LIST with 6 columns: subgrids
  int t0, t1, f0, f1;       // CH_WINDOW
  int64_t slot;             // CH_LIST: queue elements per chunk
  int slot_row;             // CH_LIST: channels per slot row (freq_chunk)
  int slot_t, slot_f;       // CH_LIST: time_chunk, freq_chunk (extent limits)
  const double *uv;         // CH_ROWS
  int64_t nrows, rows;      // CH_ROWS
};

struct Chunk {
  int64_t b;
  int sub;       // subgrid of the chunk (6-column lists), else 0
  int t0, t1, f0, f1;
  int64_t out0;  // output index of (t0, f0)
  int64_t orow;  // output elements per dump
  bool ok;
};

__device__ __forceinline__ Chunk chunk_at(const BlView &s, const Chunks &c, int64_t k) {
  Chunk r{0, 0, 0, 0, 0, 0, 0, 0, true};
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
    const int32_t *e = c.list + c.cols * k;
    if (c.cols == 6) r.sub = *e++;
    r.b = e[0];
    r.t0 = e[1];
    r.t1 = e[2];
    r.f0 = e[3];
    r.f1 = e[4];
    r.out0 = k * c.slot;
    r.orow = c.slot_row;
    r.ok = r.t1 - r.t0 <= c.slot_t && r.f1 - r.f0 <= c.slot_f && r.sub >= 0 &&
           (c.cols == 5 || r.sub < c.nsub);
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
  const double *win;   // [nsub][6] per-subgrid mirror windows (6-column lists) or null
  const float2 *grid;  // subgrid 0; subgrid s at grid + s * N * N
  int N, S, Q;
  double theta;
  const double *kern;  // real [Q][S]
  const float *kf;     // the same rounded to FP32 (degrid pass)
  float2 *out;
  unsigned long long *count;  // [n chunks] degridded visibilities
  unsigned long long *total;  // their sum
  int count_in_degrid;        // counts made by the degrid pass (else the check pass)
  uint32_t *status;
};

// Window and subgrid of a chunk.
__device__ __forceinline__ Box chunk_box(const DegridArgs &a, const Chunk &r) {
  if (!a.win) return a.box;
  const double *w = a.win + 6 * r.sub;
  return Box{SEL_MIRROR, w[0], w[1], w[2], w[3], w[4], w[5]};
}

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
  unsigned long long tot = 0;
  for (int64_t k = blockIdx.x; k < a.c.n; k += gridDim.x) {
    const Chunk r = chunk_at(a.s, a.c, k);
    const Box box = chunk_box(a, r);
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
        const int kind = classify(box, u, v, &pu, &pv);
        if (kind == 1 || kind == 2) {
          int x0, y0, fx, fy;
          if (!bl_window(a.N, a.Q, a.S, a.theta, pu, pv, &x0, &y0, &fx, &fy)) st |= ST_BOUNDS;
          n++;
        }
      }
    }
    n = block_sum_u64(n, red);
    if (threadIdx.x == 0) {
      a.count[k] = n;
      tot += n;
    }
  }
  if (threadIdx.x == 0 && tot) atomicAdd(a.total, tot);
  if (st) atomicOr(a.status, st);
}

// Validation of the chunk ranges only, one thread per chunk (when every
// window the chunks can select provably lies inside its subgrid).
__global__ void k_bl_chunk_check(DegridArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.c.n) return;
  const Chunk r = chunk_at(a.s, a.c, k);
  a.count[k] = 0;
  if (!r.ok) atomicOr(a.status, (uint32_t)ST_RANGE);
}

// Degridding.  Each wavefront takes 64 consecutive visibilities of a chunk:
// first every lane works out the coordinates of one of them (FP64 uvw
// scaling, window rule, frac_coord), then groups of G lanes (G a power of
// two >= S) degrid them, 64 / G at a time, fetching each visibility's
// window, sub-offsets and output index from the lane that computed them.
// In a group, lane x sums column x of the window over the rows weighted by
// k[fy][y]; the group then sums the columns weighted by k[fx][x].  Rows are
// read as S consecutive complex64 cells.  FP32 arithmetic on the table's
// values rounded to FP32 (kf).
template <int G>
__global__ __launch_bounds__(BL_NT) void k_bl_degrid(DegridArgs a) {
  constexpr int V = 64 / G;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = lane & (G - 1), grp = lane / G;
  const int S = a.S, N = a.N;
  unsigned long long tot = 0;
  for (int64_t k = blockIdx.x; k < a.c.n; k += gridDim.x) {
    const Chunk r = chunk_at(a.s, a.c, k);
    const Box box = chunk_box(a, r);
    const float2 *grid = a.grid + (int64_t)r.sub * N * N;
    const int nfw = r.f1 - r.f0;
    const int64_t m = (int64_t)(r.t1 - r.t0) * nfw;
    unsigned long long nk = 0;
    for (int64_t jb = (int64_t)wave * 64; jb < m; jb += BL_NT) {
      // this lane's visibility: window origin, sub-offsets, kind, output
      uint32_t xy = 0u, fxv = 0u, kfy = 0u;
      int64_t o = 0;
      const int64_t j = jb + lane;
      if (j < m) {
        const int t = r.t0 + (int)(j / nfw), f = r.f0 + (int)(j % nfw);
        o = r.out0 + (int64_t)(t - r.t0) * r.orow + (f - r.f0);
        double u, v, pu, pv;
        chunk_uv(a.s, a.c, r, t, f, &u, &v);
        const int kind = classify(box, u, v, &pu, &pv);
        int x0 = 0, y0 = 0, fx = 0, fy = 0;
        if (kind == 1 || kind == 2) bl_window(N, a.Q, S, a.theta, pu, pv, &x0, &y0, &fx, &fy);
        xy = (uint32_t)x0 | ((uint32_t)y0 << 16);
        fxv = (uint32_t)fx;
        kfy = (uint32_t)fy | ((uint32_t)kind << 28);
      }
      nk += (unsigned long long)__popcll(__ballot((kfy >> 28) == 1u || (kfy >> 28) == 2u));
      const int nr = (int)((min<int64_t>(64, m - jb) + V - 1) / V);  // wave-uniform
      for (int rr = 0; rr < nr; rr++) {
        const int q = rr * V + grp;
        const uint32_t xy_q = __shfl(xy, q), fx_q = __shfl(fxv, q), kfy_q = __shfl(kfy, q);
        const uint32_t olo = __shfl((uint32_t)o, q), ohi = __shfl((uint32_t)(o >> 32), q);
        const int kind = (int)(kfy_q >> 28);
        float ar = 0.f, ai = 0.f;
        if ((kind == 1 || kind == 2) && x < S) {
          const int x0 = (int)(xy_q & 0xffffu), y0 = (int)(xy_q >> 16);
          const float2 *col = grid + (int64_t)y0 * N + x0 + x;
          const float *ky = a.kf + (int64_t)S * (kfy_q & 0xfffffffu);
          // all G row loads issue before the products: rows past S re-read
          // row S - 1 (in the window) with weight 0
          float2 g[G];
          float w[G];
#pragma unroll
          for (int y = 0; y < G; y++) {
            const int yy = y < S ? y : S - 1;
            g[y] = col[(int64_t)yy * N];
            w[y] = y < S ? ky[yy] : 0.f;
          }
#pragma unroll
          for (int y = 0; y < G; y++) {
            ar = fmaf(w[y], g[y].x, ar);
            ai = fmaf(w[y], g[y].y, ai);
          }
          const float kx = a.kf[(int64_t)S * fx_q + x];
          ar *= kx;
          ai *= kx;
        }
#pragma unroll
        for (int w = G / 2; w > 0; w >>= 1) {
          ar += __shfl_xor(ar, w);
          ai += __shfl_xor(ai, w);
        }
        if (x == 0 && kind != 0) {
          const int64_t oq = (int64_t)(((uint64_t)ohi << 32) | olo);
          a.out[oq] = kind == 3 ? make_float2(0.f, 0.f)
                                : make_float2(ar, kind == 2 ? -ai : ai);
        }
      }
    }
    if (a.count_in_degrid && lane == 0 && nk) atomicAdd(&a.count[k], nk);
    tot += nk;
  }
  if (a.count_in_degrid && lane == 0 && tot) atomicAdd(a.total, tot);
}

__global__ void k_to_float(const double *in, float *out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
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

// Status word (offset 0) and a 64-bit counter (offset 8), cleared.
int status_word(croc_ctx *ctx, uint32_t **st) {
  *st = (uint32_t *)ws_get(ctx, WS_SMALL, 256);
  if (!*st) return CROC_ENOMEM;
  return zero_u32_async(ctx, *st, 4);
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
// True when every window a visibility selected by [min, max) x [min, max)
// (shifted by (du, dv)) can take lies inside an N x N grid: then no window
// needs a per-visibility bounds check.  The window of x = theta (u - du)
// cells starts at N/2 + x - S/2 (within a cell) and ends S - 1 cells
// later; the margins below cover that cell and the rounding.
bool box_inside(int N, int S, double theta, double du, double dv, double min_u, double max_u,
                double min_v, double max_v) {
  const double lo = -(double)(N / 2) + S + 2, hi = (double)(N / 2) - S - 2;
  const double b[4] = {theta * (min_u - du), theta * (max_u - du), theta * (min_v - dv),
                       theta * (max_v - dv)};
  for (double x : b)
    if (!(x >= lo && x <= hi)) return false;
  return theta > 0;
}

// Validation, then the degridding pass (queued; the call waits only for
// the validation, and for the total when ndeg_total is wanted).  safe:
// every selectable window lies inside its grid (box_inside), so only the
// chunk ranges are validated and the degrid pass counts; otherwise a
// per-visibility pass checks every window and counts.
int degrid_run(croc_ctx *ctx, DegridArgs a, bool safe, uint64_t *ndeg_total) {
  if (a.c.n == 0) {
    if (ndeg_total) *ndeg_total = 0;
    return CROC_OK;
  }
  uint32_t *st;
  int rc = status_word(ctx, &st);
  if (rc) return rc;
  a.status = st;
  a.total = (unsigned long long *)((char *)st + 8);
  if (!a.count) {
    a.count = (unsigned long long *)ws_get(ctx, WS_BL, sizeof(unsigned long long) * a.c.n);
    if (!a.count) return CROC_ENOMEM;
  }
  a.count_in_degrid = safe ? 1 : 0;
  const int nb = blocks_for(ctx, a.c.n);
  if (safe)
    k_bl_chunk_check<<<(unsigned)((a.c.n + 255) / 256), 256, 0, ctx->stream>>>(a);
  else
    k_bl_degrid_check<<<nb, BL_NT, 0, ctx->stream>>>(a);
  if ((rc = read_word(ctx, st))) return rc;
  {
    const int64_t nk = (int64_t)a.Q * a.S;
    float *kf = (float *)ws_get(ctx, WS_BL_BANK, sizeof(float) * nk);
    if (!kf) return CROC_ENOMEM;
    k_to_float<<<(unsigned)((nk + 255) / 256), 256, 0, ctx->stream>>>(a.kern, kf, nk);
    a.kf = kf;
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
  if (ndeg_total) {
    unsigned long long t = 0;
    CROC_HIP(hipMemcpyAsync(&t, a.total, sizeof(t), hipMemcpyDeviceToHost, ctx->stream));
    CROC_HIP(hipStreamSynchronize(ctx->stream));
    *ndeg_total = t;
  }
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
  if ((rc = degrid_run(ctx, a, false, flops ? &nd : nullptr))) return rc;
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
  const bool safe = box_inside(N, S, theta, d_u, d_v, min_u, max_u, min_v, max_v);
  if ((rc = degrid_run(ctx, a, safe, flops ? &nd : nullptr))) return rc;
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
  const bool safe = box_inside(N, S, theta, d_u, d_v, a.box.min_u, a.box.max_u, a.box.min_v,
                               a.box.max_v);
  if ((rc = degrid_run(ctx, a, safe, flops ? &nd : nullptr))) return rc;
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
  a.c.cols = 5;
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
  const bool safe =
      box_inside(subgrid_size, S, theta, mid_u, mid_v, min_u, max_u, min_v, max_v);
  if ((rc = degrid_run(ctx, a, safe, flops ? &nd : nullptr))) return rc;
  if (flops) *flops = degrid_flops(nd, S);
  return CROC_OK;
}

extern "C" int croc_streamer_degrid_batch(croc_ctx *ctx, const void *subgrids, int nsub,
                                          int subgrid_size, double theta,
                                          const double *windows, const croc_bl_set *bl,
                                          const int32_t *chunks, int64_t nchunks,
                                          int time_chunk, int freq_chunk, const double *kern,
                                          int S, int Q, void *vis_queue, int64_t *ndegrid,
                                          uint64_t *flops) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_set(bl))) return rc;
  if ((rc = check_kernel(kern, S, Q))) return rc;
  CROC_REQUIRE(subgrids && windows && nsub >= 1 && subgrid_size >= 1, CROC_EINVAL,
               "subgrids, windows, nsub >= 1 and size >= 1 required");
  CROC_REQUIRE(nchunks >= 0 && (nchunks == 0 || (chunks && vis_queue)), CROC_EINVAL,
               "null chunk list / visibility queue");
  CROC_REQUIRE(time_chunk >= 1 && freq_chunk >= 1, CROC_EINVAL,
               "time_chunk and freq_chunk must be positive");
  bool safe = true;
  for (int k = 0; k < nsub; k++) {
    const double *w = windows + 6 * k;
    safe = safe && box_inside(subgrid_size, S, theta, w[0], w[1], w[2], w[3], w[4], w[5]);
  }
  double *win = (double *)ws_get(ctx, WS_BL_WIN, sizeof(double) * 6 * nsub);
  if (!win) return CROC_ENOMEM;
  CROC_HIP(hipMemcpyAsync(win, windows, sizeof(double) * 6 * nsub, hipMemcpyHostToDevice,
                          ctx->stream));
  DegridArgs a{};
  a.s = view(bl);
  a.c.mode = CH_LIST;
  a.c.n = nchunks;
  a.c.list = chunks;
  a.c.cols = 6;
  a.c.nsub = nsub;
  a.c.slot = (int64_t)time_chunk * freq_chunk;
  a.c.slot_row = freq_chunk;
  a.c.slot_t = time_chunk;
  a.c.slot_f = freq_chunk;
  a.win = win;
  a.grid = (const float2 *)subgrids;
  a.N = subgrid_size;
  a.S = S;
  a.Q = Q;
  a.theta = theta;
  a.kern = kern;
  a.out = (float2 *)vis_queue;
  a.count = (unsigned long long *)ndegrid;
  uint64_t nd = 0;
  if ((rc = degrid_run(ctx, a, safe, flops ? &nd : nullptr))) return rc;
  if (flops) *flops = degrid_flops(nd, S);
  return CROC_OK;
}
