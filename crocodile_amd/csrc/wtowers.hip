// W-towers imaging and predict (examples/notebooks/wtowers.ipynb,
// wtowers-predict.ipynb; examples/grid/grid.c:482-710) on CDNA4.
//
// Geometry (shared with the FP64 oracle): chunk = sg - margin; a
// visibility's chunk is (floor(u theta + 0.5) + N/2) / chunk per axis; the
// chunk's sub-grid starts at x0 = cx*chunk - margin/2 and is centred on
// u_mid = (x0 + sg/2 - N/2) / theta; the w-plane is rint(w / wstep).
//
// Imaging, per call:
//   1. bin visibilities by job = (chunk, w-plane) with a counting sort;
//   2. grid each job on its own sg x sg sub-grid: the workgroup holds the
//      sub-grid in LDS as int32 fixed point for batches of at most 32
//      visibilities (integer LDS atomics), and folds every batch into
//      Kahan-compensated FP32 registers (each thread owns sg*sg/1024 cells);
//   3. inverse DFT of every job sub-grid on the matrix cores, with the
//      Fresnel factor exp(-2 pi i wp wstep (1 - n)) applied to the output;
//   4. sum the job images of each chunk (one workgroup per chunk and batch,
//      in job order: deterministic);
//   5. forward DFT of every chunk image;
//   6. add the chunk sub-grids into the grid in four passes by chunk parity,
//      so overlapping margins are never written concurrently (deterministic).
// Predict reverses it: cut + inverse DFT per chunk, forward DFT with
// exp(+2 pi i wp wstep (1 - n)) applied to the input per job, degrid.
#include <algorithm>
#include <climits>
#include <vector>

#include "internal.h"

namespace {

constexpr int SG = 128;
constexpr int SGN = SG * SG;
constexpr int TW_NT = 1024;
constexpr int TW_BATCH_VIS = 32;   // visibilities per int32 accumulation batch
constexpr int JOB_BATCH = 2048;    // sub-grids transformed per launch
constexpr uint32_t INVALID = 0xffffffffu;

struct TowerArgs {
  int N, sg, margin, chunk, ncx;
  double theta, wstep;
  int wp_min, nwp;
  int Q, S, nearest;
  const double *uvw;
  int64_t nv, stride;
};

__global__ void k_tw_wrange(const double *__restrict__ uvw, int64_t nv, int64_t stride,
                            double wstep, int *mm, uint32_t *status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double q = rint(uvw[i * stride + 2] / wstep);
  if (!(fabs(q) < 1e9)) {  // checked before any integer conversion
    atomicOr(status, (uint32_t)ST_PLANE);
    return;
  }
  atomicMin(&mm[0], (int)q);
  atomicMax(&mm[1], (int)q);
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
  const uint32_t key =
      (uint32_t)(((int64_t)cy * a.ncx + cx) * a.nwp + (wp - a.wp_min));
  return make_uint4((uint32_t)wx | ((uint32_t)wy << 16), (uint32_t)xf | ((uint32_t)yf << 16),
                    0u, key);
}

__global__ void k_tw_locate(TowerArgs a, uint4 *__restrict__ loc, uint32_t *__restrict__ count,
                            uint32_t *__restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nv) return;
  uint32_t st = 0;
  const uint4 L = tw_locate(a, i, &st);
  loc[i] = L;
  if (L.w != INVALID) atomicAdd(&count[L.w], 1u);
  if (st) atomicOr(status, st);
}

// --- three-kernel exclusive scan of key counts, emitting non-empty jobs ---
constexpr int SCAN_NT = 1024, SCAN_PER = 16, SCAN_SPAN = SCAN_NT * SCAN_PER;

__device__ __forceinline__ void block_excl_scan2(uint32_t &a, uint32_t &b, uint32_t *sa,
                                                 uint32_t *sb, uint32_t *ta, uint32_t *tb) {
  const int tid = threadIdx.x;
  sa[tid] = a;
  sb[tid] = b;
  __syncthreads();
  for (int off = 1; off < SCAN_NT; off <<= 1) {
    uint32_t x = 0, y = 0;
    if (tid >= off) {
      x = sa[tid - off];
      y = sb[tid - off];
    }
    __syncthreads();
    sa[tid] += x;
    sb[tid] += y;
    __syncthreads();
  }
  *ta = sa[SCAN_NT - 1];
  *tb = sb[SCAN_NT - 1];
  a = sa[tid] - a;
  b = sb[tid] - b;
  __syncthreads();
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_blocks(const uint32_t *__restrict__ count,
                                                         int64_t n, uint32_t *__restrict__ bsum) {
  __shared__ uint32_t sa[SCAN_NT], sb[SCAN_NT];
  const int64_t base = (int64_t)blockIdx.x * SCAN_SPAN + (int64_t)threadIdx.x * SCAN_PER;
  uint32_t c = 0, nz = 0;
  for (int k = 0; k < SCAN_PER; k++) {
    const int64_t i = base + k;
    if (i < n) {
      c += count[i];
      nz += count[i] ? 1u : 0u;
    }
  }
  uint32_t tc, tn;
  block_excl_scan2(c, nz, sa, sb, &tc, &tn);
  if (threadIdx.x == 0) {
    bsum[2 * blockIdx.x] = tc;
    bsum[2 * blockIdx.x + 1] = tn;
  }
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_top(uint32_t *__restrict__ bsum, int nb,
                                                      uint32_t *__restrict__ totals) {
  __shared__ uint32_t sa[SCAN_NT], sb[SCAN_NT];
  uint32_t carry_c = 0, carry_n = 0;
  for (int b0 = 0; b0 < nb; b0 += SCAN_NT) {
    const int b = b0 + threadIdx.x;
    uint32_t c = b < nb ? bsum[2 * b] : 0u, z = b < nb ? bsum[2 * b + 1] : 0u;
    uint32_t tc, tn;
    block_excl_scan2(c, z, sa, sb, &tc, &tn);
    if (b < nb) {
      bsum[2 * b] = c + carry_c;
      bsum[2 * b + 1] = z + carry_n;
    }
    carry_c += tc;
    carry_n += tn;
  }
  if (threadIdx.x == 0) {
    totals[0] = carry_c;
    totals[1] = carry_n;
  }
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(
    const uint32_t *__restrict__ count, int64_t n, const uint32_t *__restrict__ bsum,
    uint32_t *__restrict__ cursor, int4 *__restrict__ jobs) {
  __shared__ uint32_t sa[SCAN_NT], sb[SCAN_NT];
  const int64_t base = (int64_t)blockIdx.x * SCAN_SPAN + (int64_t)threadIdx.x * SCAN_PER;
  uint32_t c = 0, nz = 0;
  for (int k = 0; k < SCAN_PER; k++) {
    const int64_t i = base + k;
    if (i < n) {
      c += count[i];
      nz += count[i] ? 1u : 0u;
    }
  }
  uint32_t tc, tn;
  block_excl_scan2(c, nz, sa, sb, &tc, &tn);
  uint32_t off = bsum[2 * blockIdx.x] + c, job = bsum[2 * blockIdx.x + 1] + nz;
  for (int k = 0; k < SCAN_PER; k++) {
    const int64_t i = base + k;
    if (i < n) {
      const uint32_t ci = count[i];
      cursor[i] = off;
      if (ci) jobs[job++] = make_int4((int)i, (int)off, (int)(off + ci), 0);
      off += ci;
    }
  }
}

__global__ void k_tw_scatter(int64_t nv, const uint4 *__restrict__ loc,
                             uint32_t *__restrict__ cursor, uint4 *__restrict__ rec,
                             const float2 *__restrict__ vis, float2 *__restrict__ svis) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const uint4 L = loc[i];
  if (L.w == INVALID) return;
  const uint32_t pos = atomicAdd(&cursor[L.w], 1u);
  rec[pos] = make_uint4(L.x, L.y, 0u, (uint32_t)i);
  if (vis) svis[pos] = vis[i];
}

struct FixScale {
  unsigned int max_vis_bits, max_k_bits;
  float scale, inv_scale;
};

__global__ void k_tw_maxima(const float2 *__restrict__ vis, int64_t nv,
                            const float2 *__restrict__ kern, int64_t nk, FixScale *fs) {
  float mv = 0.f, mk = 0.f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride)
    mv = fmaxf(mv, fmaxf(fabsf(vis[i].x), fabsf(vis[i].y)));
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += stride)
    mk = fmaxf(mk, fmaxf(fabsf(kern[i].x), fabsf(kern[i].y)));
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    mv = fmaxf(mv, __shfl_xor(mv, off, 64));
    mk = fmaxf(mk, __shfl_xor(mk, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&fs->max_vis_bits, __float_as_uint(mv));
    atomicMax(&fs->max_k_bits, __float_as_uint(mk));
  }
}

// 2^e with 2 * max|v| * max|k| * TW_BATCH_VIS * 2^e < 2^31.
__global__ void k_tw_pick(FixScale *fs, int nearest, int separable) {
  const double mv = (double)__uint_as_float(fs->max_vis_bits);
  double mk = nearest ? 1.0 : (double)__uint_as_float(fs->max_k_bits);
  if (separable && !nearest) mk = 2.0 * mk * mk;
  const double bound = 2.0 * mv * mk * (double)TW_BATCH_VIS;
  int e = 20;
  if (bound > 0.0 && isfinite(bound)) {
    int ex;
    frexp(bound, &ex);
    e = 30 - ex;
  }
  e = max(-120, min(e, 120));
  fs->scale = ldexpf(1.f, e);
  fs->inv_scale = ldexpf(1.f, -e);
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Grid jobs [j0, j0 + nb) into sub-grids sub[b] (sg x sg complex64).
template <bool SEPARABLE>
__global__ __launch_bounds__(TW_NT) void k_tw_grid_jobs(
    const int4 *__restrict__ jobs, int j0, int nb, const uint4 *__restrict__ rec,
    const float2 *__restrict__ svis, const float2 *__restrict__ kern, int Q, int S,
    int nearest, const FixScale *__restrict__ fs, float2 *__restrict__ sub) {
  constexpr int PER = SGN / TW_NT;
  __shared__ int sre[SGN], sim[SGN];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int NW = TW_NT / 64;
  const float scale = fs->scale, inv = fs->inv_scale;
  const int SS = S * S;
  const float invS = 1.0f / (float)S;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int4 job = jobs[j0 + b];
    float ar[PER], ai[PER], cr[PER], ci[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      ar[k] = ai[k] = cr[k] = ci[k] = 0.f;
      sre[tid + k * TW_NT] = 0;
      sim[tid + k * TW_NT] = 0;
    }
    __syncthreads();
    for (int v0 = job.y; v0 < job.z; v0 += TW_BATCH_VIS) {
      const int v1 = min(job.z, v0 + TW_BATCH_VIS);
      if (nearest) {
        for (int j = v0 + tid; j < v1; j += TW_NT) {
          const uint4 r = rec[j];
          const float2 v = svis[j];
          const int c = (int)(r.x >> 16) * SG + (int)(r.x & 0xffffu);
          atomicAdd(&sre[c], __float2int_rn(v.x * scale));
          atomicAdd(&sim[c], __float2int_rn(v.y * scale));
        }
      } else {
        for (int j = v0 + wave; j < v1; j += NW) {
          const uint4 r = rec[j];
          const float2 v = svis[j];
          const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
          const int base = (int)(r.x >> 16) * SG + (int)(r.x & 0xffffu);
          const float2 *k2 = kern + (int64_t)(yf * Q + xf) * SS;
          const float2 *ky = kern + (int64_t)yf * S, *kx = kern + (int64_t)xf * S;
          for (int t = lane; t < SS; t += 64) {
            const int i = (int)(((float)t + 0.5f) * invS);
            float2 kv;
            if (!SEPARABLE) {
              kv = k2[t];
            } else {
              const float2 p = ky[i], q = kx[t - i * S];
              kv = make_float2(p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x);
            }
            const int c = base + i * (SG - S) + t;
            atomicAdd(&sre[c], __float2int_rn((v.x * kv.x - v.y * kv.y) * scale));
            atomicAdd(&sim[c], __float2int_rn((v.x * kv.y + v.y * kv.x) * scale));
          }
        }
      }
      barrier_after_lds_atomics();
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int c = tid + k * TW_NT;
        kahan_add(ar[k], cr[k], (float)sre[c] * inv);
        kahan_add(ai[k], ci[k], (float)sim[c] * inv);
        sre[c] = 0;
        sim[c] = 0;
      }
      __syncthreads();
    }
    float2 *out = sub + (int64_t)b * SGN;
#pragma unroll
    for (int k = 0; k < PER; k++)
      out[tid + k * TW_NT] = make_float2(ar[k] - cr[k], ai[k] - ci[k]);
    __syncthreads();
  }
}

// cimg[chunk] += sum of sub[b] over the batch jobs of that chunk.
// groups[g] = (chunk id, first local job, end local job).
__global__ void k_tw_chunk_accum(const int4 *__restrict__ groups, const float2 *__restrict__ sub,
                                 float2 *__restrict__ cimg) {
  const int4 g = groups[blockIdx.x];
  float2 *dst = cimg + (int64_t)g.x * SGN;
  for (int i = threadIdx.x; i < SGN; i += blockDim.x) {
    float2 acc = dst[i];
    for (int b = g.y; b < g.z; b++) {
      const float2 s = sub[(int64_t)b * SGN + i];
      acc.x += s.x;
      acc.y += s.y;
    }
    dst[i] = acc;
  }
}

// grid[window] += csub[chunk] (clipped); chunk ids of one parity class.
__global__ void k_tw_add(const int32_t *__restrict__ chunks, const float2 *__restrict__ csub,
                         float2 *__restrict__ grid, int N, int chunk, int margin, int ncx) {
  const int c = chunks[blockIdx.y];
  const int cx = c % ncx, cy = c / ncx;
  const int x0 = cx * chunk - margin / 2, y0 = cy * chunk - margin / 2;
  const float2 *s = csub + (int64_t)c * SGN;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SGN; i += gridDim.x * blockDim.x) {
    const int y = y0 + i / SG, x = x0 + i % SG;
    if (x < 0 || y < 0 || x >= N || y >= N) continue;
    float2 g = grid[(int64_t)y * N + x];
    g.x += s[i].x;
    g.y += s[i].y;
    grid[(int64_t)y * N + x] = g;
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

// Degrid the visibilities of jobs [j0, j0 + nb) from their sub-grids.
template <bool SEPARABLE>
__global__ __launch_bounds__(TW_NT) void k_tw_degrid_jobs(
    const int4 *__restrict__ jobs, int j0, int nb, const uint4 *__restrict__ rec,
    const float2 *__restrict__ sub, const float2 *__restrict__ kern, int Q, int S, int nearest,
    float2 *__restrict__ vis_out) {
  __shared__ float2 sg[SGN];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int NW = TW_NT / 64;
  const int SS = S * S;
  const float invS = 1.0f / (float)S;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int4 job = jobs[j0 + b];
    const float2 *src = sub + (int64_t)b * SGN;
    for (int i = tid; i < SGN; i += TW_NT) sg[i] = src[i];
    __syncthreads();
    if (nearest) {
      for (int j = job.y + tid; j < job.z; j += TW_NT) {
        const uint4 r = rec[j];
        vis_out[r.w] = sg[(int)(r.x >> 16) * SG + (int)(r.x & 0xffffu)];
      }
    } else {
      for (int j = job.y + wave; j < job.z; j += NW) {
        const uint4 r = rec[j];
        const int xf = (int)(r.y & 0xffffu), yf = (int)(r.y >> 16);
        const int base = (int)(r.x >> 16) * SG + (int)(r.x & 0xffffu);
        const float2 *k2 = kern + (int64_t)(yf * Q + xf) * SS;
        const float2 *ky = kern + (int64_t)yf * S, *kx = kern + (int64_t)xf * S;
        float sr = 0.f, si = 0.f;
        for (int t = lane; t < SS; t += 64) {
          const int i = (int)(((float)t + 0.5f) * invS);
          float2 kv;
          if (!SEPARABLE) {
            kv = k2[t];
          } else {
            const float2 p = ky[i], q = kx[t - i * S];
            kv = make_float2(p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x);
          }
          const float2 g = sg[base + i * (SG - S) + t];
          sr += g.x * kv.x - g.y * kv.y;
          si += g.x * kv.y + g.y * kv.x;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          sr += __shfl_xor(sr, off, 64);
          si += __shfl_xor(si, off, 64);
        }
        if (lane == 0) vis_out[r.w] = make_float2(sr, si);
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

struct Plan {
  TowerArgs a;
  uint4 *loc = nullptr, *rec = nullptr;
  float2 *svis = nullptr;
  uint32_t *count = nullptr, *cursor = nullptr, *bsum = nullptr, *small = nullptr;
  int4 *jobs = nullptr;
  int njobs = 0;
  std::vector<int4> hjobs;
  std::vector<int32_t> chunks;  // non-empty chunk ids, ascending
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
  k_tw_wrange<<<nb, 256, 0, s>>>(a.uvw, a.nv, a.stride, a.wstep, mm, p.small);
  CROC_HIP(hipGetLastError());
  int hmm[2];
  CROC_HIP(hipMemcpyAsync(hmm, mm, sizeof hmm, hipMemcpyDeviceToHost, s));
  if ((rc = tw_sync_status(ctx, p.small))) return rc;
  a.wp_min = hmm[0];
  a.nwp = hmm[1] - hmm[0] + 1;
  const int64_t nkeys = (int64_t)a.ncx * a.ncx * a.nwp;
  CROC_REQUIRE(nkeys < ((int64_t)1 << 31), CROC_EINVAL,
               "too many (chunk, w-plane) pairs (%lld); increase wstep", (long long)nkeys);

  const int64_t nblk = (nkeys + SCAN_SPAN - 1) / SCAN_SPAN;
  char *t = (char *)ws_get(ctx, WS_TOWER,
                           sizeof(uint32_t) * (2 * nkeys + 2 * nblk + 64) + 256);
  p.loc = (uint4 *)ws_get(ctx, WS_LOC, sizeof(uint4) * a.nv);
  p.rec = (uint4 *)ws_get(ctx, WS_REC, sizeof(uint4) * a.nv);
  if (vis) p.svis = (float2 *)ws_get(ctx, WS_SVIS, sizeof(float2) * a.nv);
  p.jobs = (int4 *)ws_get(ctx, WS_ITEMS, sizeof(int4) * (a.nv + 1));
  if (!t || !p.loc || !p.rec || (vis && !p.svis) || !p.jobs) return CROC_ENOMEM;
  p.count = (uint32_t *)t;
  p.cursor = p.count + nkeys;
  p.bsum = p.cursor + nkeys;
  uint32_t *totals = p.bsum + 2 * nblk;
  if ((rc = zero_u32_async(ctx, p.count, nkeys))) return rc;
  k_tw_locate<<<nb, 256, 0, s>>>(a, p.loc, p.count, p.small);
  k_scan_blocks<<<(unsigned)nblk, SCAN_NT, 0, s>>>(p.count, nkeys, p.bsum);
  k_scan_top<<<1, SCAN_NT, 0, s>>>(p.bsum, (int)nblk, totals);
  k_scan_apply<<<(unsigned)nblk, SCAN_NT, 0, s>>>(p.count, nkeys, p.bsum, p.cursor, p.jobs);
  k_tw_scatter<<<nb, 256, 0, s>>>(a.nv, p.loc, p.cursor, p.rec, vis, p.svis);
  CROC_HIP(hipGetLastError());
  uint32_t htot[2];
  CROC_HIP(hipMemcpyAsync(htot, totals, sizeof htot, hipMemcpyDeviceToHost, s));
  if ((rc = tw_sync_status(ctx, p.small))) return rc;
  p.njobs = (int)htot[1];
  p.hjobs.resize(p.njobs);
  if (p.njobs) {
    CROC_HIP(hipMemcpyAsync(p.hjobs.data(), p.jobs, sizeof(int4) * p.njobs,
                            hipMemcpyDeviceToHost, s));
    CROC_HIP(hipStreamSynchronize(s));
  }
  p.chunks.clear();
  for (const int4 &j : p.hjobs) {
    const int c = j.x / a.nwp;
    if (p.chunks.empty() || p.chunks.back() != c) p.chunks.push_back(c);
  }
  return CROC_OK;
}

int tw_setup(croc_ctx *ctx, Plan &p, int N, const double *uvw, int64_t nv, int64_t stride,
             double theta, double wstep, int subgrid, int margin, const void *kern, int Q,
             int S) {
  CROC_REQUIRE(subgrid == SG, CROC_EINVAL, "subgrid size must be %d", SG);
  CROC_REQUIRE(margin >= 0 && margin % 2 == 0 && margin <= subgrid / 2, CROC_EINVAL,
               "margin must be even and in [0, subgrid/2]");
  CROC_REQUIRE(N >= 2 && theta > 0 && wstep > 0 && stride >= 3, CROC_EINVAL, "bad arguments");
  CROC_REQUIRE(nv < (int64_t)INT32_MAX, CROC_EINVAL, "too many visibilities");
  CROC_REQUIRE(kern == nullptr || (Q >= 1 && Q <= 65535 && S >= 1 && S <= margin),
               CROC_EINVAL, "kernel support %d must not exceed the margin %d", S, margin);
  TowerArgs &a = p.a;
  a.N = N;
  a.sg = subgrid;
  a.margin = margin;
  a.chunk = subgrid - margin;
  a.ncx = N / a.chunk + 1;
  a.theta = theta;
  a.wstep = wstep;
  a.Q = kern ? Q : 1;
  a.S = kern ? S : 1;
  a.nearest = kern == nullptr;
  a.uvw = uvw;
  a.nv = nv;
  a.stride = stride;
  return CROC_OK;
}

int upload_i32(croc_ctx *ctx, int32_t *dst, const std::vector<int32_t> &v) {
  if (v.empty()) return CROC_OK;
  CROC_HIP(hipMemcpyAsync(dst, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  return CROC_OK;
}

}  // namespace

extern "C" int croc_wtowers_grid(croc_ctx *ctx, void *grid, int N, const double *uvw,
                                 int64_t nv, int64_t stride, double theta, const void *vis,
                                 double wstep, int subgrid, int margin, const void *kern,
                                 int Q, int S, int separable) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid, CROC_EINVAL, "null grid");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && vis, CROC_EINVAL, "null uvw / vis");
  Plan p;
  if ((rc = tw_setup(ctx, p, N, uvw, nv, stride, theta, wstep, subgrid, margin, kern, Q, S)))
    return rc;
  if ((rc = tw_bin(ctx, p, (const float2 *)vis))) return rc;
  if (p.njobs == 0) return CROC_OK;
  const TowerArgs &a = p.a;
  hipStream_t s = ctx->stream;

  // fixed-point scale, Fresnel phase table, per-job plane multipliers
  const int64_t nk = kern ? (separable ? (int64_t)Q * S : (int64_t)Q * Q * S * S) : 0;
  char *aux = (char *)ws_get(ctx, WS_SCALE, 256 + sizeof(double) * SGN +
                                                sizeof(int32_t) * (3 * (size_t)p.njobs + 64));
  const size_t cimg_bytes = sizeof(float2) * SGN * (size_t)a.ncx * a.ncx;
  const int jb = std::min(p.njobs, JOB_BATCH);
  float2 *stack = (float2 *)ws_get(ctx, WS_SUB, cimg_bytes + sizeof(float2) * SGN * (size_t)jb);
  if (!aux || !stack) return CROC_ENOMEM;
  FixScale *fs = (FixScale *)aux;
  double *phi = (double *)(aux + 256);
  int32_t *kw = (int32_t *)(phi + SGN);
  int32_t *dchunks = kw + p.njobs;
  float2 *cimg = stack, *sub = stack + (size_t)SGN * a.ncx * a.ncx;
  std::vector<int32_t> hk(p.njobs);
  for (int j = 0; j < p.njobs; j++) hk[j] = p.hjobs[j].x % a.nwp + a.wp_min;
  if ((rc = upload_i32(ctx, kw, hk))) return rc;
  if ((rc = upload_i32(ctx, dchunks, p.chunks))) return rc;
  if ((rc = zero_u32_async(ctx, (uint32_t *)fs, 4))) return rc;
  k_tw_maxima<<<4 * ctx->num_cu, 256, 0, s>>>((const float2 *)vis, nv, (const float2 *)kern,
                                              nk, fs);
  k_tw_pick<<<1, 1, 0, s>>>(fs, a.nearest, separable);
  k_tw_phi<<<SGN / 256, 256, 0, s>>>(phi, a.theta, a.wstep);
  CROC_HIP(hipGetLastError());
  // zero the chunk images that will receive data
  for (int c : p.chunks)
    if ((rc = zero_u32_async(ctx, (uint32_t *)(cimg + (size_t)c * SGN), 2 * SGN))) return rc;

  std::vector<int4> groups;
  int4 *dgroups = (int4 *)ws_get(ctx, WS_TILES, sizeof(int4) * (size_t)(p.chunks.size() + 64));
  if (!dgroups) return CROC_ENOMEM;
  CROC_HIP(hipEventRecord(ctx->ev[0], s));
  for (int j0 = 0; j0 < p.njobs; j0 += jb) {
    const int nb = std::min(jb, p.njobs - j0);
    const int G = std::min(nb, 2 * ctx->num_cu);
    if (a.nearest || !separable)
      k_tw_grid_jobs<false><<<G, TW_NT, 0, s>>>(p.jobs, j0, nb, p.rec, p.svis,
                                                (const float2 *)kern, a.Q, a.S, a.nearest, fs,
                                                sub);
    else
      k_tw_grid_jobs<true><<<G, TW_NT, 0, s>>>(p.jobs, j0, nb, p.rec, p.svis,
                                               (const float2 *)kern, a.Q, a.S, a.nearest, fs,
                                               sub);
    CROC_HIP(hipGetLastError());
    // image = ifft(sub) * exp(-i wp phi)
    rc = dft2d_launch(ctx, sub, sub, SG, nb, SGN, +1, 1.0f / (float)SGN, nullptr, nullptr, 0,
                      phi, kw + j0, -1, nullptr, nullptr);
    if (rc) return rc;
    groups.clear();
    for (int b = 0; b < nb; b++) {
      const int c = p.hjobs[j0 + b].x / a.nwp;
      if (groups.empty() || groups.back().x != c)
        groups.push_back(make_int4(c, b, b + 1, 0));
      else
        groups.back().z = b + 1;
    }
    CROC_HIP(hipMemcpyAsync(dgroups, groups.data(), sizeof(int4) * groups.size(),
                            hipMemcpyHostToDevice, s));
    k_tw_chunk_accum<<<(unsigned)groups.size(), 256, 0, s>>>(dgroups, sub, cimg);
    CROC_HIP(hipGetLastError());
    CROC_HIP(hipStreamSynchronize(s));  // groups buffer is reused
  }
  // chunk sub-grids = fft(chunk image), then parity-ordered adds
  rc = dft2d_launch(ctx, cimg, cimg, SG, (int64_t)p.chunks.size(), SGN, -1, 1.0f, nullptr,
                    nullptr, 0, nullptr, nullptr, 0, dchunks, dchunks);
  if (rc) return rc;
  std::vector<int32_t> par[4];
  for (int c : p.chunks) par[(c % a.ncx) % 2 + 2 * ((c / a.ncx) % 2)].push_back(c);
  int32_t *dpar = dchunks + p.chunks.size();  // room: 3 * njobs >= njobs + chunks
  for (int q = 0; q < 4; q++) {
    if (par[q].empty()) continue;
    if ((rc = upload_i32(ctx, dpar, par[q]))) return rc;
    dim3 g(16, (unsigned)par[q].size());
    k_tw_add<<<g, 256, 0, s>>>(dpar, cimg, (float2 *)grid, a.N, a.chunk, a.margin, a.ncx);
    CROC_HIP(hipGetLastError());
    CROC_HIP(hipStreamSynchronize(s));  // dpar is reused
  }
  CROC_HIP(hipEventRecord(ctx->ev[1], s));
  CROC_HIP(hipStreamSynchronize(s));
  (void)hipEventElapsedTime(&ctx->grid_ms, ctx->ev[0], ctx->ev[1]);
  return CROC_OK;
}

extern "C" int croc_wtowers_degrid(croc_ctx *ctx, const void *grid, int N, const double *uvw,
                                   int64_t nv, int64_t stride, double theta, void *vis_out,
                                   double wstep, int subgrid, int margin, const void *kern,
                                   int Q, int S, int separable) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid, CROC_EINVAL, "null grid");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && vis_out, CROC_EINVAL, "null uvw / vis_out");
  Plan p;
  if ((rc = tw_setup(ctx, p, N, uvw, nv, stride, theta, wstep, subgrid, margin, kern, Q, S)))
    return rc;
  if ((rc = tw_bin(ctx, p, nullptr))) return rc;
  if (p.njobs == 0) return CROC_OK;
  const TowerArgs &a = p.a;
  hipStream_t s = ctx->stream;
  char *aux = (char *)ws_get(ctx, WS_SCALE, 256 + sizeof(double) * SGN +
                                                sizeof(int32_t) * (3 * (size_t)p.njobs + 64));
  const size_t cimg_bytes = sizeof(float2) * SGN * (size_t)a.ncx * a.ncx;
  const int jb = std::min(p.njobs, JOB_BATCH);
  float2 *stack = (float2 *)ws_get(ctx, WS_SUB, cimg_bytes + sizeof(float2) * SGN * (size_t)jb);
  if (!aux || !stack) return CROC_ENOMEM;
  double *phi = (double *)(aux + 256);
  int32_t *kw = (int32_t *)(phi + SGN);
  int32_t *jchunk = kw + p.njobs;
  int32_t *dchunks = jchunk + p.njobs;
  float2 *cimg = stack, *sub = stack + (size_t)SGN * a.ncx * a.ncx;
  std::vector<int32_t> hk(p.njobs), hc(p.njobs);
  for (int j = 0; j < p.njobs; j++) {
    hk[j] = p.hjobs[j].x % a.nwp + a.wp_min;
    hc[j] = p.hjobs[j].x / a.nwp;
  }
  if ((rc = upload_i32(ctx, kw, hk))) return rc;
  if ((rc = upload_i32(ctx, jchunk, hc))) return rc;
  if ((rc = upload_i32(ctx, dchunks, p.chunks))) return rc;
  k_tw_phi<<<SGN / 256, 256, 0, s>>>(phi, a.theta, a.wstep);
  dim3 g(16, (unsigned)p.chunks.size());
  CROC_HIP(hipEventRecord(ctx->ev[2], s));
  k_tw_cut<<<g, 256, 0, s>>>(dchunks, (const float2 *)grid, cimg, a.N, a.chunk, a.margin, a.ncx);
  CROC_HIP(hipGetLastError());
  rc = dft2d_launch(ctx, cimg, cimg, SG, (int64_t)p.chunks.size(), SGN, +1, 1.0f / (float)SGN,
                    nullptr, nullptr, 0, nullptr, nullptr, 0, dchunks, dchunks);
  if (rc) return rc;
  for (int j0 = 0; j0 < p.njobs; j0 += jb) {
    const int nb = std::min(jb, p.njobs - j0);
    // plane sub-grid = fft(chunk image * exp(+i wp phi))
    rc = dft2d_launch(ctx, cimg, sub, SG, nb, SGN, -1, 1.0f, phi, kw + j0, +1, nullptr, nullptr,
                      0, jchunk + j0, nullptr);
    if (rc) return rc;
    const int G = std::min(nb, 2 * ctx->num_cu);
    if (a.nearest || !separable)
      k_tw_degrid_jobs<false><<<G, TW_NT, 0, s>>>(p.jobs, j0, nb, p.rec, sub,
                                                  (const float2 *)kern, a.Q, a.S, a.nearest,
                                                  (float2 *)vis_out);
    else
      k_tw_degrid_jobs<true><<<G, TW_NT, 0, s>>>(p.jobs, j0, nb, p.rec, sub,
                                                 (const float2 *)kern, a.Q, a.S, a.nearest,
                                                 (float2 *)vis_out);
    CROC_HIP(hipGetLastError());
  }
  CROC_HIP(hipEventRecord(ctx->ev[3], s));
  CROC_HIP(hipStreamSynchronize(s));
  (void)hipEventElapsedTime(&ctx->degrid_ms, ctx->ev[2], ctx->ev[3]);
  return CROC_OK;
}
