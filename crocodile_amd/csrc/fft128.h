// The 128-point building blocks of the batched 128 x 128 sub-grid FFT
// (fft.hip) shared with the kernels that transform a sub-grid in LDS and
// use it there (the w-towers predict).  Each 1D transform of 128 points is
// split as 128 = 16 x 8 (j = 8 j1 + j2, k = k1 + 16 k2):
//   pass A: for each j2, a 16-point DFT over j1 in registers, times the
//           twiddle W128^(j2 k1), written back in place;
//   pass B: for each k1, an 8-point DFT over j2 in registers, written to
//           the natural position k1 + 16 k2.
// LDS rows are padded to FP = 129 elements so that row- and column-wise
// accesses spread over the banks.
#pragma once

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

// Column passes of all 128 columns of a sub-grid in LDS (element (y, x) at
// y FP + x), consecutive lanes on consecutive columns (consecutive
// addresses: distinct banks).  Pass A on items (column x, j2), in place;
// tw[m] = W128^m.
__device__ __forceinline__ void cols_pass_a(float2 *lds, const float2 *tw, float sgn) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < PASS_A; r++) {
    const int it = tid + r * FFT_NT;
    const int x = it & 127, j2 = it >> 7;
    float2 *p = lds + x + j2 * FP;
    float2 v[16];
#pragma unroll
    for (int j1 = 0; j1 < 16; j1++) v[j1] = p[8 * j1 * FP];
    fft_reg<16>(v, sgn);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
      const float2 w = tw[(j2 * k1) & 127];
      p[8 * k1 * FP] = cmul(v[k1], make_float2(w.x, sgn * w.y));
    }
  }
}

// Column pass-B item r of the thread: (column x, k1).
__device__ __forceinline__ void cols_pass_b_item(int r, int &x, int &k1) {
  const int it = threadIdx.x + r * FFT_NT;
  x = it & 127;
  k1 = it >> 7;
}

// Column pass B on items (x, k1): v[r][k2] = output row k1 + 16 k2 of the
// thread's r-th column, read from LDS; the caller stores it.
__device__ __forceinline__ void cols_pass_b(const float2 *lds, float sgn,
                                            float2 (&v)[PASS_B][8]) {
#pragma unroll
  for (int r = 0; r < PASS_B; r++) {
    int x, k1;
    cols_pass_b_item(r, x, k1);
    const float2 *p = lds + x + 8 * k1 * FP;
#pragma unroll
    for (int j2 = 0; j2 < 8; j2++) v[r][j2] = p[j2 * FP];
    fft_reg<8>(v[r], sgn);
  }
}

// Row passes over the rows rows[0 .. nrows - 1] (rows == nullptr: row k
// is k) of a sub-grid in LDS, with 32 consecutive lanes on 32 rows at the
// same element offset: with the odd pitch FP they fall on distinct banks
// (a lane per row segment, 8 lanes to a row, puts 16 lanes of a pass-B
// read on 4 banks).  Items of pass A: (k, j2), k = 32 (t >> 8) +
// (t & 31), j2 = (t >> 5) & 7; of pass B: (k, k1) for it = t + r FFT_NT,
// k = 32 (it >> 9) + (it & 31), k1 = (it >> 5) & 15.  Wavefronts whose 32
// rows lie past nrows skip as a whole.
__device__ __forceinline__ int rowlist_line(const int *rows, int k) { return rows ? rows[k] : k; }

__device__ __forceinline__ void rows_pass_a(float2 *lds, const int *rows, int nrows,
                                            const float2 *tw, float sgn) {
  const int t = threadIdx.x;
  const int k = 32 * (t >> 8) + (t & 31), j2 = (t >> 5) & 7;
  if (__builtin_amdgcn_readfirstlane(32 * (t >> 8)) >= nrows || k >= nrows) return;
  float2 *p = lds + rowlist_line(rows, k) * FP + j2;
  float2 v[16];
#pragma unroll
  for (int j1 = 0; j1 < 16; j1++) v[j1] = p[8 * j1];
  fft_reg<16>(v, sgn);
#pragma unroll
  for (int k1 = 0; k1 < 16; k1++) {
    const float2 w = tw[(j2 * k1) & 127];
    p[8 * k1] = cmul(v[k1], make_float2(w.x, sgn * w.y));
  }
}

// Pass B item r of the thread: list entry k and k1 (k past nrows: none).
__device__ __forceinline__ void rows_pass_b_item(int r, int &k, int &k1) {
  const int it = threadIdx.x + r * FFT_NT;
  k = 32 * (it >> 9) + (it & 31);
  k1 = (it >> 5) & 15;
}

// Pass B reads: v[r][k2] = output k1 + 16 k2 of the thread's r-th item
// (zeros for wavefronts past the list); the caller writes them back after a
// barrier.
__device__ __forceinline__ void rows_pass_b(const float2 *lds, const int *rows, int nrows,
                                            float sgn, float2 (&v)[PASS_B][8]) {
#pragma unroll
  for (int r = 0; r < PASS_B; r++) {
    int k, k1;
    rows_pass_b_item(r, k, k1);
    const int kw = __builtin_amdgcn_readfirstlane(32 * ((threadIdx.x + r * FFT_NT) >> 9));
    if (kw < nrows) {
      const float2 *p = lds + rowlist_line(rows, min(k, nrows - 1)) * FP + 8 * k1;
#pragma unroll
      for (int j2 = 0; j2 < 8; j2++) v[r][j2] = p[j2];
      fft_reg<8>(v[r], sgn);
    } else {
#pragma unroll
      for (int j2 = 0; j2 < 8; j2++) v[r][j2] = make_float2(0.f, 0.f);
    }
  }
}

// W128^m = exp(2 pi i m / 128), m = 0..127, rounded from FP64 (the caller
// applies the sign of the transform to the imaginary part).
__device__ __forceinline__ void fft128_twiddles(float2 *tw) {
  if (threadIdx.x < FN) {
    double s, c;
    sincospi((double)threadIdx.x / 64.0, &s, &c);
    tw[threadIdx.x] = make_float2((float)c, (float)s);
  }
}

}  // namespace
