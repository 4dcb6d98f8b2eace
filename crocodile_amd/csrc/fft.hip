// Batched centred 2D FFT of 128 x 128 complex64 sub-grids on the vector
// ALUs, the sub-grid transform of the w-towers drivers.
//
// One workgroup per sub-grid holds it in LDS; each 1D transform of 128
// points is two passes, 16 x 8 (fft128.h).  Rows are transformed, then
// columns.  The centred transform of
// crocodile's fft / ifft (synthesis.py:83-106) is the plain DFT with input
// and output multiplied by (-1)^(x + y) (n = 128: the constant factor
// (-1)^(n/2) is 1); that sign, the per-pixel phase modulations of the
// w-towers (exp(i s k phi[pixel]) on input and / or output) and the output
// scale are applied while loading and storing.  Exact arithmetic is FP32,
// twiddles are rounded from FP64 values: ~1e-7 relative error.
#include "fft128.h"

namespace {

// Rows are transformed first, columns second.  The rows' pass A takes its
// input straight from memory and the columns' pass B writes straight to
// memory, so the sub-grid crosses LDS six times (three passes in place).
// The LDS sub-grid allows one workgroup per compute unit, so the workgroups
// are persistent (each takes a contiguous run of sub-grids) and load the
// next sub-grid's rows into registers while they transform the current one.
// Every sub-grid is read before it is written, by the one workgroup that
// transforms it, so in-place batches (in == out, injective indices) are
// safe.
template <bool MI, bool MO>
__global__ __launch_bounds__(FFT_NT) void k_fft2d_128(
    const float2 *in, float2 *out, int64_t batch, int64_t batch_stride,
    const int32_t *__restrict__ in_idx, const int32_t *__restrict__ out_idx, int sign,
    float out_scale, FModArgs mod_in, FModArgs mod_out) {
  __shared__ float2 lds[FN * FP];
  __shared__ float2 tw[FN];
  const int tid = threadIdx.x;
  const float sgn = sign < 0 ? -1.0f : 1.0f;
  // The phase tables are read in batches ahead of use so that their loads
  // overlap one another.
  constexpr bool mi = MI, mo = MO;  // input / output modulation
  fft128_twiddles(tw);
  // Row pass A: the thread owns row y and its elements x = 8 j1 + j2, which
  // share the parity of y + j2 (the (-1)^(x + y) centring sign).
  const int y = tid >> 3, j2 = tid & 7;
  float2 nxt[16];  // the thread's row elements of the workgroup's next sub-grid
  // Each workgroup takes a contiguous run of the batch.  When consecutive
  // sub-grids read the same input (in_idx equal, input and output distinct
  // buffers: the per-plane transforms of one chunk image), the rows already
  // in registers are reused instead of read again.
  const int64_t per = (batch + gridDim.x - 1) / gridDim.x;
  const int64_t bbeg = (int64_t)blockIdx.x * per;
  const int64_t bend = bbeg + per < batch ? bbeg + per : batch;
  const bool reuse = in_idx != nullptr && (const void *)in != (const void *)out;
  int64_t b = bbeg;
  if (b < bend) {
    const float2 *row = in + (int64_t)(in_idx ? in_idx[b] : b) * batch_stride + y * FN + j2;
#pragma unroll
    for (int j1 = 0; j1 < 16; j1++) nxt[j1] = row[8 * j1];
  }
  for (; b < bend; ++b) {
    float2 *Y = out + (int64_t)(out_idx ? out_idx[b] : b) * batch_stride;
    // An offset the compiler cannot see through (an empty asm, no code)
    // keeps the twiddle and phase-table loads inside the loop: hoisted out,
    // they would hold about a hundred registers for the workgroup's lifetime.
    int zero = 0;
    asm volatile("" : "+s"(zero));
    const float2 *twz = tw + zero;
    {
      float2 v[16];
      double ph[16];
#pragma unroll
      for (int j1 = 0; j1 < 16; j1++) v[j1] = nxt[j1];
      if (mi) {
        const double *prow = mod_in.phi + zero + y * FN + j2;
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) ph[j1] = prow[8 * j1];
      }
      __syncthreads();  // tw written; the previous sub-grid's LDS reads done
      const float neg = ((y + j2) & 1) ? -1.0f : 1.0f;
      const double kin = mi ? (double)mod_in.kplane[b] * (double)mod_in.sign : 0.0;
#pragma unroll
      for (int j1 = 0; j1 < 16; j1++) {
        const float2 t = mi ? rotate(v[j1], kin, ph[j1]) : v[j1];
        v[j1] = make_float2(neg * t.x, neg * t.y);
      }
      fft_reg<16>(v, sgn);
      float2 *p = lds + y * FP + j2;
#pragma unroll
      for (int k1 = 0; k1 < 16; k1++) {
        const float2 w = twz[(j2 * k1) & 127];
        p[8 * k1] = cmul(v[k1], make_float2(w.x, sgn * w.y));
      }
      // the next sub-grid's rows, in flight during the remaining passes
      const int64_t bn = b + 1;
      if (bn < bend && !(reuse && in_idx[bn] == in_idx[b])) {
        const float2 *row =
            in + (int64_t)(in_idx ? in_idx[bn] : bn) * batch_stride + y * FN + j2;
#pragma unroll
        for (int j1 = 0; j1 < 16; j1++) nxt[j1] = row[8 * j1];
      }
    }
    __syncthreads();
    {
      // Row pass B (32 rows per half-wavefront), written back to the
      // natural positions once all are read.
      float2 v[PASS_B][8];
      rows_pass_b(lds, nullptr, FN, sgn, v);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < PASS_B; r++) {
        int line, k1;
        rows_pass_b_item(r, line, k1);
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) lds[line * FP + k1 + 16 * k2] = v[r][k2];
      }
    }
    __syncthreads();
    cols_pass_a(lds, twz, sgn);  // columns
    __syncthreads();
    {
      // Column pass B to memory, one item at a time (the next sub-grid's
      // rows are live meanwhile): item (column x, k1) holds rows
      // y = k1 + 16 k2, which share the parity of x + k1.
      const double kout = mo ? (double)mod_out.kplane[b] * (double)mod_out.sign : 0.0;
#pragma unroll
      for (int r = 0; r < PASS_B; r++) {
        int x, k1;
        cols_pass_b_item(r, x, k1);
        double ph[8];
        if (mo) {
#pragma unroll
          for (int k2 = 0; k2 < 8; k2++) ph[k2] = mod_out.phi[zero + (k1 + 16 * k2) * FN + x];
        }
        float2 v[8];
        const float2 *p = lds + x + 8 * k1 * FP;
#pragma unroll
        for (int j2 = 0; j2 < 8; j2++) v[j2] = p[j2 * FP];
        fft_reg<8>(v, sgn);
        const float sc = ((x + k1) & 1) ? -out_scale : out_scale;
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) {
          const int pix = (k1 + 16 * k2) * FN + x;
          const float2 t = make_float2(v[k2].x * sc, v[k2].y * sc);
          Y[pix] = mo ? rotate(t, kout, ph[k2]) : t;
        }
      }
    }
  }
}

}  // namespace

int fft2d_launch(croc_ctx *ctx, const void *in, void *out, int n, int64_t batch,
                 int64_t batch_stride, int sign, float out_scale, const double *phi_in,
                 const int32_t *k_in, int sign_in, const double *phi_out,
                 const int32_t *k_out, int sign_out, const int32_t *in_idx,
                 const int32_t *out_idx) {
  CROC_REQUIRE(n == FN, CROC_EINVAL, "the sub-grid FFT is built for n = %d", FN);
  if (batch <= 0) return CROC_OK;
  FModArgs mi{phi_in, k_in, sign_in}, mo{phi_out, k_out, sign_out};
  const int G = (int)std::min<int64_t>(batch, ctx->num_cu);  // one per CU (LDS)
  const float2 *x = (const float2 *)in;
  float2 *y = (float2 *)out;
  if (phi_in && phi_out)
    k_fft2d_128<true, true><<<G, FFT_NT, 0, ctx->stream>>>(x, y, batch, batch_stride, in_idx,
                                                           out_idx, sign, out_scale, mi, mo);
  else if (phi_in)
    k_fft2d_128<true, false><<<G, FFT_NT, 0, ctx->stream>>>(x, y, batch, batch_stride, in_idx,
                                                            out_idx, sign, out_scale, mi, mo);
  else if (phi_out)
    k_fft2d_128<false, true><<<G, FFT_NT, 0, ctx->stream>>>(x, y, batch, batch_stride, in_idx,
                                                            out_idx, sign, out_scale, mi, mo);
  else
    k_fft2d_128<false, false><<<G, FFT_NT, 0, ctx->stream>>>(x, y, batch, batch_stride, in_idx,
                                                             out_idx, sign, out_scale, mi, mo);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

extern "C" int croc_fft2d_batched(croc_ctx *ctx, const void *in, void *out, int n,
                                  int64_t batch, int inverse) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(in && out, CROC_EINVAL, "null array");
  CROC_REQUIRE(batch >= 0 && batch <= 0x7fffffff, CROC_EINVAL, "bad batch");
  const float scale = inverse ? 1.0f / ((float)n * (float)n) : 1.0f;
  rc = fft2d_launch(ctx, in, out, n, batch, (int64_t)n * n, inverse ? +1 : -1, scale, nullptr,
                    nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr);
  if (rc) return rc;
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}
