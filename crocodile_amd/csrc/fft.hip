// Batched centred 2D FFT of 128 x 128 complex64 sub-grids on the vector
// ALUs, the sub-grid transform of the w-towers drivers.
//
// One workgroup per sub-grid holds it in LDS (rows padded to 129 elements
// so that row- and column-wise accesses spread over the banks).  Each 1D
// transform of 128 points is split as 128 = 16 x 8 (j = 8 j1 + j2,
// k = k1 + 16 k2):
//   pass A: for each j2, a 16-point DFT over j1 in registers, times the
//           twiddle W128^(j2 k1), written back in place;
//   pass B: for each k1, an 8-point DFT over j2 in registers, written to
//           the natural position k1 + 16 k2.
// Rows are transformed, then columns.  The centred transform of
// crocodile's fft / ifft (synthesis.py:83-106) is the plain DFT with input
// and output multiplied by (-1)^(x + y) (n = 128: the constant factor
// (-1)^(n/2) is 1); that sign, the per-pixel phase modulations of the
// w-towers (exp(i s k phi[pixel]) on input and / or output) and the output
// scale are applied while loading and storing.  Exact arithmetic is FP32,
// twiddles are rounded from FP64 values: ~1e-7 relative error.
#include "internal.h"

namespace {

constexpr int FN = 128;
constexpr int FP = FN + 1;  // padded LDS row
constexpr int FFT_NT = 1024;
constexpr int PASS_A = FN * 8 / FFT_NT;   // pass-A items (line, j2) per thread
constexpr int PASS_B = FN * 16 / FFT_NT;  // pass-B items (line, k1) per thread
static_assert(PASS_A * FFT_NT == FN * 8 && PASS_B * FFT_NT == FN * 16, "item split");

struct FModArgs {
  const double *phi;
  const int32_t *kplane;
  int sign;
};

// v * exp(i kf phi), kf = s k (the plane multiplier times the sign): the
// angle in turns, reduced to [-1/2, 1/2] in FP64, then the hardware sine
// and cosine of turns (v_sin_f32 / v_cos_f32: 1.24e-7 largest absolute
// error on [-1/2, 1/2], measured against FP64 over 2^20 points; a library
// sincosf of the FP32 angle in radians is 2.1e-7 there and several times
// the instructions).
__device__ __forceinline__ float2 rotate(float2 v, double kf, double phi) {
  const double t = kf * phi * 0.15915494309189535;
  const float f = (float)(t - rint(t));
  const float s = __builtin_amdgcn_sinf(f), c = __builtin_amdgcn_cosf(f);
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// cos / sin of 2 pi k / 16, k = 0..7 (rounded from FP64).
__device__ constexpr float C16[8] = {1.0f, 0.92387953251128674f, 0.70710678118654757f,
                                      0.38268343236508984f, 0.0f, -0.38268343236508973f,
                                      -0.70710678118654746f, -0.92387953251128674f};
__device__ constexpr float S16[8] = {0.0f, 0.38268343236508978f, 0.70710678118654746f,
                                      0.92387953251128674f, 1.0f, 0.92387953251128674f,
                                      0.70710678118654757f, 0.38268343236508989f};

// In-register radix-2 DFT of N <= 16 points: v[k] = sum_j v[j] W_N^(jk),
// W_N = exp(sgn 2 pi i / N).
template <int N>
__device__ __forceinline__ void fft_reg(float2 *v, float sgn) {
  if constexpr (N == 1) {
    return;
  } else {
    float2 e[N / 2], o[N / 2];
#pragma unroll
    for (int k = 0; k < N / 2; k++) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    fft_reg<N / 2>(e, sgn);
    fft_reg<N / 2>(o, sgn);
#pragma unroll
    for (int k = 0; k < N / 2; k++) {
      const int m = k * (16 / N);  // W_N^k = W_16^(k * 16 / N)
      const float2 w = make_float2(C16[m], sgn * S16[m]);
      const float2 t = cmul(o[k], w);
      v[k] = cadd(e[k], t);
      v[k + N / 2] = csub(e[k], t);
    }
  }
}

// Pass A of the 128-point transforms of all 128 LDS lines (line l starts at
// l * line_stride, elements at stride es) on items (line, j2), in place;
// tw[m] = W128^m.  LINE_FAST puts consecutive lanes on consecutive lines
// (columns: consecutive addresses), otherwise on consecutive items of one
// line (rows).
template <bool LINE_FAST>
__device__ __forceinline__ void pass_a(float2 *lds, int line_stride, int es, const float2 *tw,
                                       float sgn) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < PASS_A; r++) {
    const int it = tid + r * FFT_NT;
    const int line = LINE_FAST ? (it & 127) : (it >> 3);
    const int j2 = LINE_FAST ? (it >> 7) : (it & 7);
    float2 *p = lds + line * line_stride + j2 * es;
    float2 v[16];
#pragma unroll
    for (int j1 = 0; j1 < 16; j1++) v[j1] = p[8 * j1 * es];
    fft_reg<16>(v, sgn);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
      const float2 w = tw[(j2 * k1) & 127];
      p[8 * k1 * es] = cmul(v[k1], make_float2(w.x, sgn * w.y));
    }
  }
}

// Pass-B item r of the thread: (line, k1).
template <bool LINE_FAST>
__device__ __forceinline__ void pass_b_item(int r, int &line, int &k1) {
  const int it = threadIdx.x + r * FFT_NT;
  line = LINE_FAST ? (it & 127) : (it >> 4);
  k1 = LINE_FAST ? (it >> 7) : (it & 15);
}

// Pass B on items (line, k1): v[r][k2] = output k1 + 16 k2 of the thread's
// r-th line, read from LDS; the caller stores it.
template <bool LINE_FAST>
__device__ __forceinline__ void pass_b(const float2 *lds, int line_stride, int es, float sgn,
                                       float2 (&v)[PASS_B][8]) {
#pragma unroll
  for (int r = 0; r < PASS_B; r++) {
    int line, k1;
    pass_b_item<LINE_FAST>(r, line, k1);
    const float2 *p = lds + line * line_stride + 8 * k1 * es;
#pragma unroll
    for (int j2 = 0; j2 < 8; j2++) v[r][j2] = p[j2 * es];
    fft_reg<8>(v[r], sgn);
  }
}

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
  if (tid < FN) {
    double s, c;
    sincospi((double)tid / 64.0, &s, &c);  // W128^m = exp(sgn 2 pi i m / 128)
    tw[tid] = make_float2((float)c, (float)s);
  }
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
      // Row pass B, written back to the natural positions once all are read.
      float2 v[PASS_B][8];
      pass_b<false>(lds, FP, 1, sgn, v);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < PASS_B; r++) {
        int line, k1;
        pass_b_item<false>(r, line, k1);
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) lds[line * FP + k1 + 16 * k2] = v[r][k2];
      }
    }
    __syncthreads();
    pass_a<true>(lds, 1, FP, twz, sgn);  // columns
    __syncthreads();
    {
      // Column pass B to memory, one item at a time (the next sub-grid's
      // rows are live meanwhile): item (column x, k1) holds rows
      // y = k1 + 16 k2, which share the parity of x + k1.
      const double kout = mo ? (double)mod_out.kplane[b] * (double)mod_out.sign : 0.0;
#pragma unroll
      for (int r = 0; r < PASS_B; r++) {
        int x, k1;
        pass_b_item<true>(r, x, k1);
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
