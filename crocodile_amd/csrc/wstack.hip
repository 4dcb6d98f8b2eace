// W-stacking (examples/notebooks/wstacking.ipynb, cell 10, JSON lines
// 191-240): visibilities binned by w-plane round-half-even(w / wstep), each
// plane gridded nearest-cell onto its own full grid, transformed, divided by
// the plane's Fresnel pattern w_kernel_function(l, m, mean w) and summed in
// image space.
//
// A batch of planes is one [P][N][N] stack: k_wstack_grid fills it in one
// pass over the visibilities (a plane -> slot table selects the batch), a
// batched inverse FFT runs over the stack (rocFFT), and k_wstack_add reads
// the P transformed planes of each pixel once, applies each plane's Fresnel
// phase and adds the sum into the FP64 image.  For even N the centring
// shifts of the centred FFT are folded into both kernels as the
// checkerboard sign (-1)^(x+y) (the product of the two axes' global factors
// exp(i pi N / 2) is 1), so no shift pass touches the stack.
#include "internal.h"

namespace {

// One visibility onto its plane's slot of the stack, nearest cell
// (synthesis.py:389-401: cell N//2 + floor(0.5 + N p), p = uvw / lam, every
// column of p, w included, below 0.5).  With apply == 0 only the checks run.
__global__ void k_wstack_grid(float2 *__restrict__ stack, int N,
                              const double *__restrict__ uvw, int64_t nv,
                              int64_t stride, double lam, double wstep,
                              int64_t plane0, int ntab,
                              const int32_t *__restrict__ slot_of,
                              const float2 *__restrict__ vis, int centred,
                              int apply, uint32_t *status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double w = uvw[i * stride + 2];
  const int64_t t = wbin_dev(w, wstep) - plane0;
  if (t < 0 || t >= ntab) {
    atomicOr(status, (uint32_t)ST_PLANE);
    return;
  }
  const int slot = slot_of[t];
  if (slot < 0) return;
  const double px = grid_fraction(uvw[i * stride], lam);
  const double py = grid_fraction(uvw[i * stride + 1], lam);
  if (!(px < 0.5 && py < 0.5 && grid_fraction(w, lam) < 0.5)) {
    atomicOr(status, (uint32_t)ST_DOMAIN);
    return;
  }
  int x, y;
  if (!nearest_cell(N, px, &x) || !nearest_cell(N, py, &y)) {
    atomicOr(status, (uint32_t)ST_BOUNDS);
    return;
  }
  if (!apply) return;
  float2 v = vis[i];
  if (centred && ((x + y) & 1)) {
    v.x = -v.x;
    v.y = -v.y;
  }
  float *g = (float *)(stack + ((int64_t)slot * N + y) * N + x);
  unsafeAtomicAdd(g, v.x);
  unsafeAtomicAdd(g + 1, v.y);
}

// image[y][x] += sign * sum_p plane_p[y][x] * exp(-2 pi i midw[p] n1) with
// n1 = 1 - sqrt(1 - l^2 - m^2), l = theta (x - N//2) / N, m likewise in y
// (synthesis.py:66-80 coordinates2, :273-293 w_kernel_function; dividing by
// exp(2 pi i w n1) is multiplying by its conjugate).  The phase is reduced
// to turns in FP64 (midw n1 - rint(midw n1)) before the FP32 sincospi, so
// large |w| n1 keep their accuracy.
__global__ void k_wstack_add(double2 *__restrict__ image,
                             const float2 *__restrict__ stack, int P, int N,
                             double theta, const double *__restrict__ midw,
                             int centred) {
#pragma clang fp contract(off)
  const int64_t NN = (int64_t)N * N;
  const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= NN) return;
  const int y = (int)(pix / N), x = (int)(pix - (int64_t)y * N);
  const double l = theta * ((double)(x - N / 2) / (double)N);
  const double m = theta * ((double)(y - N / 2) / (double)N);
  const double n1 = 1.0 - sqrt(1.0 - (l * l + m * m));
  double ax = 0.0, ay = 0.0;
  const float2 *src = stack + pix;
  int p = 0;
  for (; p + 4 <= P; p += 4) {
    float2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = src[(int64_t)(p + k) * NN];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double t = midw[p + k] * n1;
      float s, c;
      sincospif((float)(2.0 * (t - rint(t))), &s, &c);
      ax += (double)v[k].x * c + (double)v[k].y * s;
      ay += (double)v[k].y * c - (double)v[k].x * s;
    }
  }
  for (; p < P; ++p) {
    const float2 v = src[(int64_t)p * NN];
    const double t = midw[p] * n1;
    float s, c;
    sincospif((float)(2.0 * (t - rint(t))), &s, &c);
    ax += (double)v.x * c + (double)v.y * s;
    ay += (double)v.y * c - (double)v.x * s;
  }
  if (centred && ((x + y) & 1)) {
    ax = -ax;
    ay = -ay;
  }
  double2 o = image[pix];
  o.x += ax;
  o.y += ay;
  image[pix] = o;
}

}  // namespace

extern "C" int croc_wstack_grid(croc_ctx *ctx, void *stack, int P, int N,
                                const double *uvw, int64_t nv, int64_t stride,
                                double lam, double wstep, int64_t plane0,
                                int ntab, const int32_t *slot_of,
                                const void *vis, int centred) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(N >= 1 && P >= 1 && stride >= 3 && lam > 0 && wstep > 0 && ntab >= 1,
               CROC_EINVAL, "bad arguments");
  CROC_REQUIRE(!centred || N % 2 == 0, CROC_EINVAL,
               "the folded centring sign needs an even grid");
  CROC_REQUIRE((int64_t)P * N * N < ((int64_t)1 << 46), CROC_EINVAL, "stack too large");
  if (nv == 0) return CROC_OK;
  CROC_REQUIRE(stack && uvw && vis && slot_of, CROC_EINVAL, "null array");
  uint32_t *st;
  if ((rc = small_status(ctx, &st))) return rc;
  const unsigned nb = (unsigned)((nv + 255) / 256);
  // Validate everything first so that an error leaves the stack untouched.
  k_wstack_grid<<<nb, 256, 0, ctx->stream>>>((float2 *)stack, N, uvw, nv, stride, lam,
                                             wstep, plane0, ntab, slot_of,
                                             (const float2 *)vis, centred, 0, st);
  CROC_HIP(hipGetLastError());
  if ((rc = finish_status(ctx, st))) return rc;
  k_wstack_grid<<<nb, 256, 0, ctx->stream>>>((float2 *)stack, N, uvw, nv, stride, lam,
                                             wstep, plane0, ntab, slot_of,
                                             (const float2 *)vis, centred, 1, st);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

extern "C" int croc_wstack_add(croc_ctx *ctx, void *image, const void *stack, int P,
                               int N, double theta, const double *midw, int centred) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(image && stack && midw && N >= 1 && P >= 1 && theta > 0, CROC_EINVAL,
               "bad arguments");
  CROC_REQUIRE(!centred || N % 2 == 0, CROC_EINVAL,
               "the folded centring sign needs an even grid");
  // synthesis.py:290: the corner of the image (the most negative l and m)
  // must lie inside the unit circle.
  const double c = theta * ((double)(0 - N / 2) / (double)N);
  CROC_REQUIRE(c * c + c * c < 1.0, CROC_EDOMAIN, "Error in image coordinate system");
  const int64_t NN = (int64_t)N * N;
  k_wstack_add<<<(unsigned)((NN + 255) / 256), 256, 0, ctx->stream>>>(
      (double2 *)image, (const float2 *)stack, P, N, theta, midw, centred);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}
