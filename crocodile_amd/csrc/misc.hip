// frac_coords, nearest-cell grid / degrid and make_grid_hermitian.
#include "internal.h"

namespace {

__global__ void k_frac_coords(const double *__restrict__ uvw, int64_t nv,
                              int64_t stride, double lam, int Nh, int Nw, int Q,
                              int32_t *__restrict__ x, int32_t *__restrict__ xf,
                              int32_t *__restrict__ y, int32_t *__restrict__ yf,
                              uint32_t *status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double px = grid_fraction(uvw[i * stride], lam);
  const double py = grid_fraction(uvw[i * stride + 1], lam);
  int ix = 0, ixf = 0, iy = 0, iyf = 0;
  const bool ok = frac_coord_dev(Nw, Q, px, &ix, &ixf) &&
                  frac_coord_dev(Nh, Q, py, &iy, &iyf);
  if (!ok) atomicOr(status, (uint32_t)ST_DOMAIN);
  x[i] = ix;
  xf[i] = ixf;
  y[i] = iy;
  yf[i] = iyf;
}

__global__ void k_grid_nearest(float2 *__restrict__ grid, int N,
                               const double *__restrict__ uvw, int64_t nv,
                               int64_t stride, double lam,
                               const float2 *__restrict__ vis, int apply,
                               uint32_t *status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double px = grid_fraction(uvw[i * stride], lam);
  const double py = grid_fraction(uvw[i * stride + 1], lam);
  int x, y;
  // numpy.max(p) < 0.5 covers every column of p, w included.
  const bool wok = stride < 3 || grid_fraction(uvw[i * stride + 2], lam) < 0.5;
  if (!(px < 0.5 && py < 0.5 && wok)) {
    atomicOr(status, (uint32_t)ST_DOMAIN);
    return;
  }
  if (!nearest_cell(N, px, &x) || !nearest_cell(N, py, &y)) {
    atomicOr(status, (uint32_t)ST_BOUNDS);
    return;
  }
  if (!apply) return;
  float *g = (float *)(grid + (int64_t)y * N + x);
  unsafeAtomicAdd(g, vis[i].x);
  unsafeAtomicAdd(g + 1, vis[i].y);
}

__global__ void k_degrid_nearest(const float2 *__restrict__ grid, int N,
                                 const double *__restrict__ uvw, int64_t nv,
                                 int64_t stride, double lam,
                                 float2 *__restrict__ vis, int apply,
                                 uint32_t *status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double px = grid_fraction(uvw[i * stride], lam);
  const double py = grid_fraction(uvw[i * stride + 1], lam);
  int x, y;
  // numpy.max(p) < 0.5 covers every column of p, w included.
  const bool wok = stride < 3 || grid_fraction(uvw[i * stride + 2], lam) < 0.5;
  if (!(px < 0.5 && py < 0.5 && wok)) {
    atomicOr(status, (uint32_t)ST_DOMAIN);
    return;
  }
  if (!nearest_cell(N, px, &x) || !nearest_cell(N, py, &y)) {
    atomicOr(status, (uint32_t)ST_BOUNDS);
    return;
  }
  if (apply) vis[i] = grid[(int64_t)y * N + x];
}

// g[c] += conj(g[mirror(c)]) with the mirror taken from the original values:
// each (cell, mirror) pair is handled by the thread of its lower index.
// Even N mirrors rows / columns 1..N-1 about N/2 (row / column 0 untouched).
template <typename C>
__global__ void k_hermitian(C *__restrict__ g, int N) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)N * N) return;
  const int i = (int)(c / N), j = (int)(c - (int64_t)i * N);
  int mi, mj;
  if (N % 2 == 0) {
    if (i == 0 || j == 0) return;
    mi = N - i;
    mj = N - j;
  } else {
    mi = N - 1 - i;
    mj = N - 1 - j;
  }
  const int64_t mc = (int64_t)mi * N + mj;
  if (mc < c) return;
  const C a = g[c];
  if (mc == c) {
    g[c].x = a.x + a.x;
    g[c].y = a.y - a.y;
    return;
  }
  const C b = g[mc];
  g[c].x = a.x + b.x;
  g[c].y = a.y - b.y;
  g[mc].x = b.x + a.x;
  g[mc].y = b.y - a.y;
}

// Non-square grids: each visibility's cell and sub-offset by synthesis.py's
// rule for an [Nh][Nw] grid -- frac_coords (:438-449, per-axis sizes) for
// the convolution, or the nearest cell of grid / degrid (:389-416, which
// take N = Nh for both axes and index [y, x] with NumPy's negative wrap) --
// checked against that grid (convolution window rows y - gh//2 ..
// y + (gh+1)//2 - 1, columns likewise), then re-expressed as the coordinate
// of a square [Nb][Nb] grid holding the [Nh][Nw] one at rows oy.., columns
// ox..: the square-grid kernels land every tap on the same cell there.
__global__ void k_remap(const double *__restrict__ uvw, int64_t nv, int64_t stride,
                        double lam, int Nh, int Nw, int Q, int gh, int gw, int nearest,
                        int Nb, int oy, int ox, double *__restrict__ p,
                        uint32_t *status) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double px = grid_fraction(uvw[i * stride], lam);
  const double py = grid_fraction(uvw[i * stride + 1], lam);
  int64_t cx, cy;
  int fx = 0, fy = 0;
  if (nearest) {
    // numpy.max(p) < 0.5 covers every column of p, w included
    const bool wok = stride < 3 || grid_fraction(uvw[i * stride + 2], lam) < 0.5;
    if (!(px < 0.5 && py < 0.5 && wok)) {
      atomicOr(status, (uint32_t)ST_DOMAIN);
      return;
    }
    const double tx = floor(0.5 + (double)Nh * px), ty = floor(0.5 + (double)Nh * py);
    if (!(fabs(tx) < 1e15 && fabs(ty) < 1e15)) {
      atomicOr(status, (uint32_t)ST_BOUNDS);
      return;
    }
    cx = (int64_t)(Nh / 2) + (int64_t)tx;
    cy = (int64_t)(Nh / 2) + (int64_t)ty;
    if (cx < 0) cx += Nw;
    if (cy < 0) cy += Nh;
    if (cx < 0 || cx >= Nw || cy < 0 || cy >= Nh) {
      atomicOr(status, (uint32_t)ST_BOUNDS);
      return;
    }
  } else {
    int ix, iy;
    if (!frac_coord_dev(Nw, Q, px, &ix, &fx) || !frac_coord_dev(Nh, Q, py, &iy, &fy)) {
      atomicOr(status, (uint32_t)ST_DOMAIN);
      return;
    }
    if (ix - gw / 2 < 0 || ix + (gw + 1) / 2 > Nw || iy - gh / 2 < 0 || iy + (gh + 1) / 2 > Nh) {
      atomicOr(status, (uint32_t)ST_BOUNDS);
      return;
    }
    cx = ix;
    cy = iy;
  }
  p[2 * i] = ((double)(cx + ox - Nb / 2) - (double)fx / (double)Q) / (double)Nb;
  p[2 * i + 1] = ((double)(cy + oy - Nb / 2) - (double)fy / (double)Q) / (double)Nb;
}

}  // namespace

int small_status(croc_ctx *ctx, uint32_t **st) {
  *st = (uint32_t *)ws_get(ctx, WS_SMALL, 256);
  if (!*st) return CROC_ENOMEM;
  return zero_u32_async(ctx, *st, 1);
}

int finish_status(croc_ctx *ctx, uint32_t *st) {
  CROC_HIP(hipMemcpyAsync(ctx->host_status, st, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return croc_status_to_code(*ctx->host_status);
}

extern "C" int croc_frac_coords(croc_ctx *ctx, const double *uvw, int64_t nv,
                                int64_t stride, double lam, int Nh, int Nw,
                                int Q, int32_t *x, int32_t *xf, int32_t *y,
                                int32_t *yf) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(Q >= 1 && Nh >= 1 && Nw >= 1 && stride >= 2 && lam > 0,
               CROC_EINVAL, "bad arguments");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && x && xf && y && yf, CROC_EINVAL, "null array");
  uint32_t *st;
  if ((rc = small_status(ctx, &st))) return rc;
  k_frac_coords<<<(unsigned)((nv + 255) / 256), 256, 0, ctx->stream>>>(
      uvw, nv, stride, lam, Nh, Nw, Q, x, xf, y, yf, st);
  CROC_HIP(hipGetLastError());
  return finish_status(ctx, st);
}

extern "C" int croc_remap_coords(croc_ctx *ctx, const double *uvw, int64_t nv, int64_t stride,
                                 double lam, int Nh, int Nw, int Q, int gh, int gw,
                                 int nearest, int Nb, int oy, int ox, double *p_out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(Q >= 1 && Nh >= 1 && Nw >= 1 && stride >= 2 && lam > 0 && gh >= 1 && gw >= 1,
               CROC_EINVAL, "bad arguments");
  CROC_REQUIRE(Nb >= 2 && Nb % 2 == 0 && oy >= 0 && ox >= 0 && oy + Nh <= Nb && ox + Nw <= Nb,
               CROC_EINVAL, "the square grid must be even and hold the [Nh][Nw] grid");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && p_out, CROC_EINVAL, "null array");
  uint32_t *st;
  if ((rc = small_status(ctx, &st))) return rc;
  k_remap<<<(unsigned)((nv + 255) / 256), 256, 0, ctx->stream>>>(
      uvw, nv, stride, lam, Nh, Nw, Q, gh, gw, nearest, Nb, oy, ox, p_out, st);
  CROC_HIP(hipGetLastError());
  return finish_status(ctx, st);
}

extern "C" int croc_grid_nearest(croc_ctx *ctx, void *grid, int N,
                                 const double *uvw, int64_t nv, int64_t stride,
                                 double lam, const void *vis) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(N >= 1 && stride >= 2 && lam > 0 && grid, CROC_EINVAL,
               "bad arguments");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && vis, CROC_EINVAL, "null array");
  uint32_t *st;
  if ((rc = small_status(ctx, &st))) return rc;
  const unsigned nb = (unsigned)((nv + 255) / 256);
  // Validate everything first so that an error leaves the grid untouched.
  k_grid_nearest<<<nb, 256, 0, ctx->stream>>>((float2 *)grid, N, uvw, nv,
                                              stride, lam,
                                              (const float2 *)vis, 0, st);
  CROC_HIP(hipGetLastError());
  rc = finish_status(ctx, st);
  if (rc) return rc;
  k_grid_nearest<<<nb, 256, 0, ctx->stream>>>((float2 *)grid, N, uvw, nv,
                                              stride, lam,
                                              (const float2 *)vis, 1, st);
  CROC_HIP(hipGetLastError());
  return finish_status(ctx, st);
}

extern "C" int croc_degrid_nearest(croc_ctx *ctx, const void *grid, int N,
                                   const double *uvw, int64_t nv,
                                   int64_t stride, double lam, void *vis_out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(N >= 1 && stride >= 2 && lam > 0 && grid, CROC_EINVAL,
               "bad arguments");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(uvw && vis_out, CROC_EINVAL, "null array");
  uint32_t *st;
  if ((rc = small_status(ctx, &st))) return rc;
  const unsigned nb = (unsigned)((nv + 255) / 256);
  k_degrid_nearest<<<nb, 256, 0, ctx->stream>>>(
      (const float2 *)grid, N, uvw, nv, stride, lam, (float2 *)vis_out, 0, st);
  CROC_HIP(hipGetLastError());
  rc = finish_status(ctx, st);
  if (rc) return rc;
  k_degrid_nearest<<<nb, 256, 0, ctx->stream>>>(
      (const float2 *)grid, N, uvw, nv, stride, lam, (float2 *)vis_out, 1, st);
  CROC_HIP(hipGetLastError());
  return finish_status(ctx, st);
}

extern "C" int croc_make_hermitian(croc_ctx *ctx, void *grid, int N,
                                   int is_double) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid && N >= 1, CROC_EINVAL, "bad arguments");
  const int64_t n = (int64_t)N * N;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (is_double)
    k_hermitian<double2><<<nb, 256, 0, ctx->stream>>>((double2 *)grid, N);
  else
    k_hermitian<float2><<<nb, 256, 0, ctx->stream>>>((float2 *)grid, N);
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}
