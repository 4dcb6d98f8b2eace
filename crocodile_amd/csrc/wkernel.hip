// W-kernel generation and kernel oversampling in FP64.
//
// crocodile builds an oversampled kernel by zero-padding the F x F far field
// to F*Q, taking a centred FFT and picking every Q-th sample
// (synthesis.py:330-354, :160-197).  Only the Q*S x Q*S frequencies around
// the centre survive the extraction, so the same numbers are the zoom DFT
//
//   K[yf][xf][y][x] = sum_{m,l} ff[m][l] E(ky, m) E(kx, l),
//   E(k, n) = exp(-2 pi i k (n - F//2) / (F Q)),
//   ky = Q (y - S//2) + yf,  kx = Q (x - S//2) + xf,
//
// evaluated here as two dense passes (rows, then columns) with the twiddle
// phase reduced exactly in integers before the FP64 sincospi.  The far field
// of a w-kernel, exp(2 pi i w (1 - sqrt(1 - l^2 - m^2))) on the coordinates
// of kernel_coordinates (synthesis.py:250-293), is evaluated in FP64 on the
// device, for all w-planes of a bank in one launch.
#include <string.h>

#include <vector>

#include "internal.h"

namespace {

// exp(-2 pi i j / M) with j already reduced to [0, M).
__device__ __forceinline__ double2 twiddle(int64_t j, int64_t M) {
  double s, c;
  sincospi(-2.0 * (double)j / (double)M, &s, &c);
  return make_double2(c, s);
}

__device__ __forceinline__ int64_t mod64(int64_t a, int64_t m) {
  int64_t r = a % m;
  return r < 0 ? r + m : r;
}

struct FFArgs {
  double theta, dl, dm;
  double T00, T01, T10, T11;
  int has_T;
  int invert;
  int F;
};

// Far field of plane p: ff[p][m][l].  Coordinates follow coordinates2 and
// kernel_coordinates: l = theta * ((l_idx - F//2) / F), likewise m; then the
// optional transform (l, m) <- (T00 l + T10 m, T01 l + T11 m) and shift.
__global__ void k_wfarfield(FFArgs a, const double *__restrict__ w,
                            const double *__restrict__ aterm, double2 *__restrict__ ff,
                            uint32_t *status) {
#pragma clang fp contract(off)
  const int F = a.F;
  const int p = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= F * F) return;
  const int mi = idx / F, li = idx - mi * F;
  double l = ((double)(li - F / 2) / (double)F) * a.theta;
  double m = ((double)(mi - F / 2) / (double)F) * a.theta;
  if (a.has_T) {
    const double l2 = a.T00 * l + a.T10 * m;
    const double m2 = a.T01 * l + a.T11 * m;
    l = l2;
    m = m2;
  }
  l = (l + a.dl);
  m = (m + a.dm);
  const double r2 = l * l + m * m;
  if (!(r2 < 1.0)) {
    atomicOr(status, 1u);
    ff[(int64_t)p * F * F + idx] = make_double2(0.0, 0.0);
    return;
  }
  const double ph = 1.0 - sqrt(1.0 - r2);
  const double phase = (2.0 * M_PI * w[p]) * ph;
  double s, c;
  sincos(phase, &s, &c);
  double2 v = make_double2(c, s);
  if (a.invert) {
    // 1 / (c + i s) by Smith's method, as complex division does.
    if (fabs(c) >= fabs(s)) {
      const double rt = s / c, den = c + s * rt;
      v = make_double2(1.0 / den, -rt / den);
    } else {
      const double rt = c / s, den = s + c * rt;
      v = make_double2(rt / den, -1.0 / den);
    }
  }
  if (aterm) {  // image-plane A-term and taper: the AW kernel's far field
    const double t = aterm[idx];
    v = make_double2(v.x * t, v.y * t);
  }
  ff[(int64_t)p * F * F + idx] = v;
}

// Row pass: tmp[p][m][kx] = sum_l ff[p][m][l] E(kx - Q(S//2), l), kx in [0, QS).
__global__ void k_zoom_rows(const double2 *__restrict__ ff, int64_t ff_stride,
                            int F, int Q, const int32_t *__restrict__ Sp,
                            double2 *__restrict__ tmp, int64_t tmp_stride) {
  const int p = blockIdx.y;
  const int S = Sp[p];
  const int QS = Q * S;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= F * QS) return;
  const int m = idx / QS, kidx = idx - m * QS;
  const int64_t k = (int64_t)kidx - (int64_t)Q * (S / 2);
  const int64_t M = (int64_t)F * Q;
  const double2 *row = ff + p * ff_stride + (int64_t)m * F;
  double sr = 0.0, si = 0.0;
  for (int l = 0; l < F; l++) {
    const double2 t = twiddle(mod64(k * (l - F / 2), M), M);
    const double2 f = row[l];
    sr += f.x * t.x - f.y * t.y;
    si += f.x * t.y + f.y * t.x;
  }
  tmp[p * tmp_stride + idx] = make_double2(sr, si);
}

// Column pass and re-layout to [yf][xf][y][x].
template <typename OutT>
__global__ void k_zoom_cols(const double2 *__restrict__ tmp, int64_t tmp_stride,
                            int F, int Q, const int32_t *__restrict__ Sp,
                            const int64_t *__restrict__ out_off,
                            OutT *__restrict__ out) {
  const int p = blockIdx.y;
  const int S = Sp[p];
  const int QS = Q * S;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)QS * QS) return;
  const int kyi = (int)(idx / QS), kxi = (int)(idx - (int64_t)kyi * QS);
  const int64_t ky = (int64_t)kyi - (int64_t)Q * (S / 2);
  const int64_t M = (int64_t)F * Q;
  const double2 *col = tmp + p * tmp_stride + kxi;
  double sr = 0.0, si = 0.0;
  for (int m = 0; m < F; m++) {
    const double2 t = twiddle(mod64(ky * (m - F / 2), M), M);
    const double2 f = col[(int64_t)m * QS];
    sr += f.x * t.x - f.y * t.y;
    si += f.x * t.y + f.y * t.x;
  }
  const int y = kyi / Q, yf = kyi - y * Q;
  const int x = kxi / Q, xf = kxi - x * Q;
  const int64_t o = out_off[p] + (((int64_t)yf * Q + xf) * S + y) * S + x;
  out[o].x = sr;
  out[o].y = si;
}

// 1D: out[xf][x] = sum_l ff[l] E(Q x + xf - Q(S//2), l); one workgroup per
// output sample, its threads summing strided terms (FP64), then a
// workgroup reduction: the far field may be a whole grid row (F in the
// thousands) while there are only Q * S outputs.
constexpr int ZOOM1D_NT = 256;
template <typename OutT>
__global__ __launch_bounds__(ZOOM1D_NT) void k_zoom_1d(const double2 *__restrict__ ff, int F,
                                                      int Q, int S, OutT *__restrict__ out) {
  __shared__ double s_re[ZOOM1D_NT], s_im[ZOOM1D_NT];
  const int64_t kidx = blockIdx.x;
  const int64_t k = kidx - (int64_t)Q * (S / 2);
  const int64_t M = (int64_t)F * Q;
  double sr = 0.0, si = 0.0;
  for (int l = threadIdx.x; l < F; l += ZOOM1D_NT) {
    const double2 t = twiddle(mod64(k * (l - F / 2), M), M);
    const double2 f = ff[l];
    sr += f.x * t.x - f.y * t.y;
    si += f.x * t.y + f.y * t.x;
  }
  s_re[threadIdx.x] = sr;
  s_im[threadIdx.x] = si;
  __syncthreads();
  for (int h = ZOOM1D_NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_re[threadIdx.x] += s_re[threadIdx.x + h];
      s_im[threadIdx.x] += s_im[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int x = (int)(kidx / Q), xf = (int)(kidx - (int64_t)x * Q);
    out[(int64_t)xf * S + x].x = s_re[0];
    out[(int64_t)xf * S + x].y = s_im[0];
  }
}

int oversample_planes(croc_ctx *ctx, const double2 *ff, int64_t ff_stride,
                      int F, int Q, int nplanes, const int32_t *dS, int maxS,
                      const int64_t *dOff, void *out, int out_double) {
  const int64_t maxQS = (int64_t)Q * maxS;
  const int64_t tmp_stride = (int64_t)F * maxQS;
  double2 *tmp = (double2 *)ws_get(ctx, WS_TMP,
                                   sizeof(double2) * tmp_stride * nplanes);
  if (!tmp) return CROC_ENOMEM;
  hipStream_t s = ctx->stream;
  dim3 g1((unsigned)((tmp_stride + 255) / 256), nplanes);
  k_zoom_rows<<<g1, 256, 0, s>>>(ff, ff_stride, F, Q, dS, tmp, tmp_stride);
  dim3 g2((unsigned)((maxQS * maxQS + 255) / 256), nplanes);
  if (out_double)
    k_zoom_cols<double2><<<g2, 256, 0, s>>>(tmp, tmp_stride, F, Q, dS, dOff,
                                            (double2 *)out);
  else
    k_zoom_cols<float2><<<g2, 256, 0, s>>>(tmp, tmp_stride, F, Q, dS, dOff,
                                           (float2 *)out);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

}  // namespace

extern "C" int croc_w_kernels(croc_ctx *ctx, double theta, const double *w,
                              int nw, int npix_ff, const int32_t *S, int Q,
                              int invert, double dl, double dm,
                              const double *T, const int64_t *plane_off,
                              void *out, int out_double, const double *aterm) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(w && S && plane_off && out, CROC_EINVAL, "null argument");
  CROC_REQUIRE(nw >= 1 && nw <= 65535 && npix_ff >= 1 && Q >= 1, CROC_EINVAL,
               "bad sizes nw=%d F=%d Q=%d", nw, npix_ff, Q);
  int maxS = 0;
  for (int p = 0; p < nw; p++) {
    CROC_REQUIRE(S[p] >= 1, CROC_EINVAL, "support must be positive");
    // synthesis.py:371
    CROC_REQUIRE(npix_ff > S[p] || (npix_ff == S[p] && Q == 1), CROC_EINVAL,
                 "NpixFF must exceed NpixKern (or equal it with Qpx == 1)");
    maxS = S[p] > maxS ? S[p] : maxS;
  }
  const int F = npix_ff;
  const int64_t ffsz = (int64_t)F * F;
  // Parameter block: w[nw] | off[nw] | S[nw] | status
  size_t pbytes = sizeof(double) * nw + sizeof(int64_t) * nw + sizeof(int32_t) * nw + 64;
  char *par = (char *)ws_get(ctx, WS_SMALL, 256 + pbytes);
  double2 *ff = (double2 *)ws_get(ctx, WS_FF, sizeof(double2) * ffsz * nw);
  if (!par || !ff) return CROC_ENOMEM;
  double *dw = (double *)(par + 256);
  int64_t *doff = (int64_t *)(dw + nw);
  int32_t *dS = (int32_t *)(doff + nw);
  uint32_t *dst = (uint32_t *)(((uintptr_t)(dS + nw) + 63) & ~(uintptr_t)63);
  std::vector<char> host(pbytes, 0);
  memcpy(host.data(), w, sizeof(double) * nw);
  memcpy(host.data() + sizeof(double) * nw, plane_off, sizeof(int64_t) * nw);
  memcpy(host.data() + (sizeof(double) + sizeof(int64_t)) * nw, S,
         sizeof(int32_t) * nw);
  CROC_HIP(hipMemcpyAsync(dw, host.data(), (char *)dst - (char *)dw,
                          hipMemcpyHostToDevice, ctx->stream));
  rc = zero_u32_async(ctx, dst, 1);
  if (rc) return rc;

  FFArgs fa;
  fa.theta = theta;
  fa.dl = dl;
  fa.dm = dm;
  fa.has_T = T != nullptr;
  fa.T00 = T ? T[0] : 1.0;
  fa.T01 = T ? T[1] : 0.0;
  fa.T10 = T ? T[2] : 0.0;
  fa.T11 = T ? T[3] : 1.0;
  fa.invert = invert;
  fa.F = F;
  dim3 g((unsigned)((ffsz + 255) / 256), nw);
  k_wfarfield<<<g, 256, 0, ctx->stream>>>(fa, dw, aterm, ff, dst);
  CROC_HIP(hipGetLastError());
  rc = oversample_planes(ctx, ff, ffsz, F, Q, nw, dS, maxS, doff, out,
                         out_double);
  if (rc) return rc;
  uint32_t st = 0;
  CROC_HIP(hipMemcpyAsync(ctx->host_status, dst, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  st = *ctx->host_status;
  CROC_REQUIRE(st == 0, CROC_EDOMAIN,
               "Error in image coordinate system: l^2 + m^2 >= 1");
  return CROC_OK;
}

extern "C" int croc_kernel_oversample(croc_ctx *ctx, const void *ff, int n,
                                      int ndim, int Q, int s, void *out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(ff && out, CROC_EINVAL, "null argument");
  CROC_REQUIRE(n >= 1 && Q >= 1 && s >= 1, CROC_EINVAL, "bad sizes");
  CROC_REQUIRE(ndim == 1 || ndim == 2, CROC_EINVAL, "ndim must be 1 or 2");
  // extract_oversampled needs the N*Q array to hold every extracted sample
  // (synthesis.py:184, :194 would otherwise give a short slice).
  const int64_t M = (int64_t)n * Q;
  const int64_t start = M / 2 - (int64_t)Q * (s / 2);
  CROC_REQUIRE(start >= 0 && start + (int64_t)Q * (s - 1) + (Q - 1) < M,
               CROC_EINVAL, "support %d too large for far field %d", s, n);
  if (ndim == 1) {
    const int64_t QS = (int64_t)Q * s;
    k_zoom_1d<double2><<<(unsigned)QS, ZOOM1D_NT, 0, ctx->stream>>>(
        (const double2 *)ff, n, Q, s, (double2 *)out);
    CROC_HIP(hipGetLastError());
    CROC_HIP(hipStreamSynchronize(ctx->stream));
    return CROC_OK;
  }
  char *par = (char *)ws_get(ctx, WS_SMALL, 512);
  if (!par) return CROC_ENOMEM;
  int64_t *doff = (int64_t *)(par + 256);
  int32_t *dS = (int32_t *)(par + 320);
  int64_t off0 = 0;
  CROC_HIP(hipMemcpyAsync(doff, &off0, sizeof off0, hipMemcpyHostToDevice,
                          ctx->stream));
  CROC_HIP(hipMemcpyAsync(dS, &s, sizeof s, hipMemcpyHostToDevice,
                          ctx->stream));
  rc = oversample_planes(ctx, (const double2 *)ff, (int64_t)n * n, n, Q, 1,
                         dS, s, doff, out, 1);
  if (rc) return rc;
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}
