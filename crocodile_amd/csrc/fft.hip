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

__device__ __forceinline__ float2 fmod_phase(float2 v, const FModArgs &md, int b, int pix) {
  if (!md.phi) return v;
  const double ang = (double)md.kplane[b] * md.phi[pix] * (double)md.sign;
  const double red = ang - 6.283185307179586 * rint(ang * 0.15915494309189535);
  float s, c;
  sincosf((float)red, &s, &c);
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

// One 128-point transform of all 128 LDS lines: line l starts at
// l * line_stride, elements at stride es.  Pass A on items (line, j2), pass
// B on items (line, k1); tw[m] = W128^m.  LINE_FAST puts consecutive lanes
// on consecutive lines (columns: consecutive addresses), otherwise on
// consecutive items of one line (rows).
template <bool LINE_FAST>
__device__ __forceinline__ void fft128_lines(float2 *lds, int line_stride, int es,
                                             const float2 *tw, float sgn) {
  const int tid = threadIdx.x;
  // pass A: 128 lines x 8 items
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
  __syncthreads();
  // pass B: 128 lines x 16 items; reads then writes
  float2 v[PASS_B][8];
  int base[PASS_B], k1s[PASS_B];
#pragma unroll
  for (int r = 0; r < PASS_B; r++) {
    const int it = tid + r * FFT_NT;
    const int line = LINE_FAST ? (it & 127) : (it >> 4);
    const int k1 = LINE_FAST ? (it >> 7) : (it & 15);
    base[r] = line * line_stride;
    k1s[r] = k1;
#pragma unroll
    for (int j2 = 0; j2 < 8; j2++) v[r][j2] = lds[base[r] + (8 * k1 + j2) * es];
    fft_reg<8>(v[r], sgn);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PASS_B; r++)
#pragma unroll
    for (int k2 = 0; k2 < 8; k2++) lds[base[r] + (k1s[r] + 16 * k2) * es] = v[r][k2];
  __syncthreads();
}

__global__ __launch_bounds__(FFT_NT) void k_fft2d_128(
    const float2 *in, float2 *out, int64_t batch_stride, const int32_t *__restrict__ in_idx,
    const int32_t *__restrict__ out_idx, int sign, float out_scale, FModArgs mod_in,
    FModArgs mod_out) {
  __shared__ float2 lds[FN * FP];
  __shared__ float2 tw[FN];
  const int b = blockIdx.x, tid = threadIdx.x;
  const float2 *X = in + (int64_t)(in_idx ? in_idx[b] : b) * batch_stride;
  float2 *Y = out + (int64_t)(out_idx ? out_idx[b] : b) * batch_stride;
  const float sgn = sign < 0 ? -1.0f : 1.0f;
  if (tid < FN) {
    double s, c;
    sincospi((double)tid / 64.0, &s, &c);  // W128^m = exp(sgn 2 pi i m / 128)
    tw[tid] = make_float2((float)c, (float)s);
  }
  // load: eight pixels in flight per thread per round
  for (int i0 = tid; i0 < FN * FN; i0 += 8 * FFT_NT) {
    float2 xv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) xv[k] = X[i0 + k * FFT_NT];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = i0 + k * FFT_NT, y = i >> 7, x = i & 127;
      float2 v = fmod_phase(xv[k], mod_in, b, i);
      if ((x + y) & 1) v = make_float2(-v.x, -v.y);
      lds[y * FP + x] = v;
    }
  }
  __syncthreads();
  fft128_lines<false>(lds, FP, 1, tw, sgn);  // rows
  fft128_lines<true>(lds, 1, FP, tw, sgn);   // columns
  for (int i = tid; i < FN * FN; i += FFT_NT) {
    const int y = i >> 7, x = i & 127;
    float2 v = lds[y * FP + x];
    const float s = ((x + y) & 1) ? -out_scale : out_scale;
    v = make_float2(v.x * s, v.y * s);
    Y[i] = fmod_phase(v, mod_out, b, i);
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
  k_fft2d_128<<<(unsigned)batch, FFT_NT, 0, ctx->stream>>>(
      (const float2 *)in, (float2 *)out, batch_stride, in_idx, out_idx, sign, out_scale, mi, mo);
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
