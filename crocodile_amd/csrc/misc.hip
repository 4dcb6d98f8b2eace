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

}  // namespace

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
