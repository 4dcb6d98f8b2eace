// AW-towers imaging (examples/notebooks/awtowers.ipynb cells 17, 19, 22):
// w-towers whose (chunk, w-plane) images are built per antenna pair, each
// pre-gridded with the anti-aliasing kernel on a twice-larger sub-grid at
// twice the field of view, transformed, cut to the middle, corrected by the
// anti-aliasing windows and multiplied by the pair's A-screen ratio
// aimage[a2] / aimage[a1] in image space before the pairs are summed.
//
// Work unit: a group = (chunk, w-plane, antenna pair), groups sorted by that
// key.  Three launches per batch of groups plus two per call:
//   k_awt_pregrid  visibilities of the batch onto [G][M][M] (M = 2S) grids
//                  with the separable anti-aliasing kernel (one thread per
//                  visibility and kernel row);
//   (batched inverse FFT of the stack, rocFFT)
//   k_awt_apply    per chunk and sub-grid pixel: the sum over the chunk's
//                  groups of pixel * ratio * wkern^-wp, times the window;
//   (batched forward FFT of the chunk images, rocFFT)
//   k_awt_place    chunk grids added into the grid, clipped at its edges.
// With even S and M = 2S the centring signs of the M-grid's inverse
// transform and of the S-grid's forward transform cancel on the cut-out
// pixels, so k_awt_apply applies none; the pre-grid carries the input sign
// of the inverse transform and k_awt_place the output sign of the forward
// one.
#include "internal.h"

namespace {

// One (visibility, kernel row): p = rel / lam_sub, cell and sub-offset by
// synthesis.py's frac_coord on the M grid, row y - K//2 + i of the window
// += gcf[yf][i] * gcf[xf][j] * v (synthesis.py:451-481, separable table).
// With apply == 0 only the checks run (every visibility of the call).
__global__ void k_awt_pregrid(float2 *__restrict__ stack, int M,
                              const double *__restrict__ rel, int64_t v0, int64_t v1,
                              double lam_sub, int Q, int K,
                              const float2 *__restrict__ gcf,
                              const float2 *__restrict__ vis,
                              const int32_t *__restrict__ gid, int32_t g0, int apply,
                              uint32_t *status) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t v = v0 + t / K;
  const int i = (int)(t % K);
  if (v >= v1) return;
  const double px = grid_fraction(rel[v * 3], lam_sub);
  const double py = grid_fraction(rel[v * 3 + 1], lam_sub);
  int x, xf, y, yf;
  if (!frac_coord_dev(M, Q, px, &x, &xf) || !frac_coord_dev(M, Q, py, &y, &yf)) {
    if (i == 0) atomicOr(status, (uint32_t)ST_DOMAIN);
    return;
  }
  const int xs = x - K / 2, ys = y - K / 2;
  if (xs < 0 || ys < 0 || xs + K > M || ys + K > M) {
    if (i == 0) atomicOr(status, (uint32_t)ST_BOUNDS);
    return;
  }
  if (!apply) return;
  const float2 a = vis[v];
  const float2 ky = gcf[(int64_t)yf * K + i];
  // ky * v, then times each column's tap
  float2 r;
  r.x = ky.x * a.x - ky.y * a.y;
  r.y = ky.x * a.y + ky.y * a.x;
  const int row = ys + i;
  float *g = (float *)(stack + ((int64_t)(gid[v] - g0) * M + row) * M + xs);
  for (int j = 0; j < K; ++j) {
    const float2 kx = gcf[(int64_t)xf * K + j];
    float ox = kx.x * r.x - kx.y * r.y;
    float oy = kx.x * r.y + kx.y * r.x;
    if ((row + xs + j) & 1) {
      ox = -ox;
      oy = -oy;
    }
    unsafeAtomicAdd(g + 2 * j, ox);
    unsafeAtomicAdd(g + 2 * j + 1, oy);
  }
}

// bins[c0 + k][y][x] += win[y][x] * sum over groups g in [start[k],
// start[k + 1]) of stack[g][y + S/2][x + S/2] * ascreen[a2[g]] / ascreen[a1[g]]
// * wkern^-wp[g], wkern = 2^lmag exp(2 pi i turns) per pixel; the power's
// phase is reduced to turns in FP64 before the FP32 sincospi.
__global__ void k_awt_apply(float2 *__restrict__ bins, int c0,
                            const float2 *__restrict__ stack, int S,
                            const int32_t *__restrict__ start,
                            const int32_t *__restrict__ gwp,
                            const int32_t *__restrict__ ga1,
                            const int32_t *__restrict__ ga2,
                            const float2 *__restrict__ ascreen,
                            const float *__restrict__ win,
                            const double *__restrict__ turns,
                            const float *__restrict__ lmag) {
  const int SS = S * S;
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= SS) return;
  const int k = blockIdx.y;
  const int y = pix / S, x = pix - y * S;
  const int M = 2 * S;
  const int64_t off = (int64_t)(y + S / 2) * M + (x + S / 2);
  const double tw = turns[pix];
  const float lm = lmag[pix];
  double ax = 0.0, ay = 0.0;
  const int gs = start[k], ge = start[k + 1];
  for (int g = gs; g < ge; ++g) {
    const float2 v = stack[(int64_t)g * M * M + off];
    const float2 p = ascreen[(int64_t)ga1[g] * SS + pix];
    const float2 q = ascreen[(int64_t)ga2[g] * SS + pix];
    // q / p = q conj(p) / |p|^2
    const float inv = 1.0f / (p.x * p.x + p.y * p.y);
    const float rx = (q.x * p.x + q.y * p.y) * inv;
    const float ry = (q.y * p.x - q.x * p.y) * inv;
    const int wp = gwp[g];
    const double t = -(double)wp * tw;
    float s, c;
    sincospif((float)(2.0 * (t - rint(t))), &s, &c);
    const float mag = lm == 0.0f ? 1.0f : exp2f(-(float)wp * lm);
    const float fx = c * mag, fy = s * mag;
    // v * r * f
    const float vx = v.x * rx - v.y * ry;
    const float vy = v.x * ry + v.y * rx;
    ax += (double)(vx * fx - vy * fy);
    ay += (double)(vx * fy + vy * fx);
  }
  const float w = win[pix];
  float2 *o = bins + (int64_t)(c0 + k) * SS + pix;
  float2 b = *o;
  b.x += (float)ax * w;
  b.y += (float)ay * w;
  *o = b;
}

// grid[y0 + y][x0 + x] += (-1)^(x+y) * bins[c][y][x] inside the grid.
__global__ void k_awt_place(float2 *__restrict__ grid, int N,
                            const float2 *__restrict__ bins, int S,
                            const int32_t *__restrict__ corner) {
  const int SS = S * S;
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= SS) return;
  const int c = blockIdx.y;
  const int y = pix / S, x = pix - y * S;
  const int gx = corner[2 * c] + x, gy = corner[2 * c + 1] + y;
  if (gx < 0 || gy < 0 || gx >= N || gy >= N) return;
  float2 b = bins[(int64_t)c * SS + pix];
  if ((x + y) & 1) {
    b.x = -b.x;
    b.y = -b.y;
  }
  float *g = (float *)(grid + (int64_t)gy * N + gx);
  unsafeAtomicAdd(g, b.x);
  unsafeAtomicAdd(g + 1, b.y);
}

}  // namespace

extern "C" int croc_awt_pregrid(croc_ctx *ctx, void *stack, int M, const double *rel,
                                int64_t v0, int64_t v1, double lam_sub, int Q, int K,
                                const void *gcf, const void *vis, const int32_t *gid,
                                int32_t g0, int check_only) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(M >= 2 && M % 2 == 0 && Q >= 1 && K >= 1 && K <= M && lam_sub > 0 &&
                   v0 >= 0 && v1 >= v0,
               CROC_EINVAL, "bad arguments");
  if (v1 == v0) return CROC_OK;
  CROC_REQUIRE(rel && gcf && vis && gid && (check_only || stack), CROC_EINVAL, "null array");
  uint32_t *st;
  if ((rc = small_status(ctx, &st))) return rc;
  const int64_t nt = (v1 - v0) * K;
  k_awt_pregrid<<<(unsigned)((nt + 255) / 256), 256, 0, ctx->stream>>>(
      (float2 *)stack, M, rel, v0, v1, lam_sub, Q, K, (const float2 *)gcf,
      (const float2 *)vis, gid, g0, check_only ? 0 : 1, st);
  CROC_HIP(hipGetLastError());
  if (check_only) return finish_status(ctx, st);
  return CROC_OK;
}

extern "C" int croc_awt_apply(croc_ctx *ctx, void *bins, int c0, int nc, const void *stack,
                              int S, const int32_t *start, const int32_t *gwp,
                              const int32_t *ga1, const int32_t *ga2, const void *ascreen,
                              const float *win, const double *turns, const float *lmag) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(S >= 2 && S % 2 == 0 && c0 >= 0 && nc >= 0 && nc < 65536, CROC_EINVAL,
               "bad arguments");
  if (nc == 0) return CROC_OK;
  CROC_REQUIRE(bins && stack && start && gwp && ga1 && ga2 && ascreen && win && turns && lmag,
               CROC_EINVAL, "null array");
  dim3 grid((unsigned)((S * S + 255) / 256), (unsigned)nc);
  k_awt_apply<<<grid, 256, 0, ctx->stream>>>((float2 *)bins, c0, (const float2 *)stack, S,
                                             start, gwp, ga1, ga2,
                                             (const float2 *)ascreen, win, turns, lmag);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

extern "C" int croc_awt_place(croc_ctx *ctx, void *grid, int N, const void *bins, int nc,
                              int S, const int32_t *corner) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(N >= 1 && S >= 2 && S % 2 == 0 && nc >= 0 && nc < 65536, CROC_EINVAL,
               "bad arguments");
  if (nc == 0) return CROC_OK;
  CROC_REQUIRE(grid && bins && corner, CROC_EINVAL, "null array");
  dim3 g((unsigned)((S * S + 255) / 256), (unsigned)nc);
  k_awt_place<<<g, 256, 0, ctx->stream>>>((float2 *)grid, N, (const float2 *)bins, S, corner);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}
