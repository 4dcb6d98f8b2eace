// AW-kernel generation: crocodile/synthesis.py:701-734 (aw_kernel_fn)
// convolves the two antennas' A-kernels and then the result with every
// oversampled w sub-kernel, each with scipy.signal.convolve2d(mode='same').
// 'same' keeps the first operand's shape and takes the centred part of the
// full convolution:  out[i][j] = sum_{k,l} a[k][l] b[i + sy - k][j + sx - l],
// (sy, sx) = ((bh - 1) / 2, (bw - 1) / 2), b outside its extent = 0.
// One thread per output element, FP64, for nb kernels b sharing one a.
#include "internal.h"

namespace {

__global__ void k_convolve_same(const double2 *__restrict__ a, int ah, int aw,
                                const double2 *__restrict__ b, int bh, int bw, int64_t nb,
                                double2 *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)ah * aw;
  if (idx >= per * nb) return;
  const int64_t n = idx / per;
  const int r = (int)(idx - n * per);
  const int i = r / aw, j = r - i * aw;
  const int sy = (bh - 1) / 2, sx = (bw - 1) / 2;
  const double2 *bn = b + n * (int64_t)bh * bw;
  double sr = 0.0, si = 0.0;
  for (int k = 0; k < ah; k++) {
    const int y = i + sy - k;
    if (y < 0 || y >= bh) continue;
    for (int l = 0; l < aw; l++) {
      const int x = j + sx - l;
      if (x < 0 || x >= bw) continue;
      const double2 p = a[k * aw + l], q = bn[y * bw + x];
      sr += p.x * q.x - p.y * q.y;
      si += p.x * q.y + p.y * q.x;
    }
  }
  out[idx] = make_double2(sr, si);
}

}  // namespace

extern "C" int croc_convolve_same(croc_ctx *ctx, const void *a, int ah, int aw, const void *b,
                                  int bh, int bw, int64_t nb, void *out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(a && b && out, CROC_EINVAL, "null array");
  CROC_REQUIRE(ah >= 1 && aw >= 1 && bh >= 1 && bw >= 1 && nb >= 0, CROC_EINVAL, "bad shape");
  const int64_t total = (int64_t)ah * aw * nb;
  if (total == 0) return CROC_OK;
  k_convolve_same<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
      (const double2 *)a, ah, aw, (const double2 *)b, bh, bw, nb, (double2 *)out);
  CROC_HIP(hipGetLastError());
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}
