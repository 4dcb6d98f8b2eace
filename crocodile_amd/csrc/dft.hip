// Batched centred 2D DFT of n x n complex sub-grids on the matrix cores.
//
// Y = s * D X D with D[k][j] = exp(sign * 2 pi i (k - n/2)(j - n/2) / n) is
// numpy's fftshift(fft2(ifftshift(X))) (sign -1) or n^2 * the ifft variant
// (sign +1), i.e. crocodile's fft / ifft (synthesis.py:83-106) for one
// sub-grid.  The w-towers drivers transform every (chunk, w-plane) sub-grid
// this way.  The two passes are dense complex GEMMs:
//   pass 1  T = X D   (A = X rows from LDS, B = D from L2)
//   pass 2  Y = D T   (A = D rows from L2,  B = T^T from LDS)
// on v_mfma_f32_16x16x32_f16.  Each FP32 operand is split as hi + lo with
// hi = fp16(x), lo = fp16(x - hi) after scaling the whole sub-grid by a power
// of two so that max|x| ~ 2^14; products keep hi*hi + hi*lo + lo*hi
// ("FP16x3"), which carries ~22 significant bits -- FP32-level accuracy
// at the fp16 matrix rate.  (A bf16 split keeps only 8 + 8 bits, ~1e-5
// relative per product, too coarse for the 1e-5 error budget.)
//
// Work split (n = 128, 8 wavefronts): in pass 1 wavefront w owns the 16
// columns of tile-column w and sweeps the 8 tile-rows, so its D fragments
// are loaded once per k-step; in pass 2 it owns tile-row w.  LDS holds the
// split X (4 x 32 KB fp16) and then, overwritten, the split T^T.
//
// Optional per-pixel modulation exp(i * k * phi[pixel]) is applied to the
// input (mod_in) or to the output (mod_out), with k a per-sub-grid integer
// (the w-plane) and phi an FP64 table; the angle is reduced in FP64.
#include "internal.h"

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int DFT_N = 128;
constexpr int DFT_NW = DFT_N / 16;         // wavefronts per sub-grid
constexpr int DFT_NT = DFT_NW * 64;
constexpr int LDS_STRIDE = DFT_N + 8;      // fp16 elements per LDS row (16 B pad)
constexpr int PLANE = DFT_N * LDS_STRIDE;  // one split array

struct ModArgs {
  const double *phi;      // [n*n] phase per pixel (radians per unit k), or null
  const int32_t *kplane;  // [batch] integer multiplier per sub-grid
  int sign;               // +1 / -1
};

__device__ __forceinline__ half8 neg8(half8 a) { return -a; }

// Split four consecutive values (already scaled) and store them as one
// 8-byte word per plane at element o of hi / lo (o a multiple of 4).
__device__ __forceinline__ void split4_store(float a0, float a1, float a2, float a3,
                                             _Float16 *hi, _Float16 *lo, int o) {
  half4 h, l;
  h[0] = (_Float16)a0;
  l[0] = (_Float16)(a0 - (float)h[0]);
  h[1] = (_Float16)a1;
  l[1] = (_Float16)(a1 - (float)h[1]);
  h[2] = (_Float16)a2;
  l[2] = (_Float16)(a2 - (float)h[2]);
  h[3] = (_Float16)a3;
  l[3] = (_Float16)(a3 - (float)h[3]);
  *reinterpret_cast<half4 *>(hi + o) = h;
  *reinterpret_cast<half4 *>(lo + o) = l;
}

__device__ __forceinline__ floatx4 mfma(half8 a, half8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Complex multiply-accumulate of split fragments:
//   cr += ar*br - ai*bi,  ci += ar*bi + ai*br  (three products per real one)
__device__ __forceinline__ void cmac(floatx4 &cr, floatx4 &ci, half8 arh, half8 arl,
                                     half8 aih, half8 ail, half8 brh, half8 brl,
                                     half8 bih, half8 bil) {
  const half8 naih = neg8(aih), nail = neg8(ail);
  cr = mfma(arh, brh, cr);
  cr = mfma(arh, brl, cr);
  cr = mfma(arl, brh, cr);
  cr = mfma(naih, bih, cr);
  cr = mfma(naih, bil, cr);
  cr = mfma(nail, bih, cr);
  ci = mfma(arh, bih, ci);
  ci = mfma(arh, bil, ci);
  ci = mfma(arl, bih, ci);
  ci = mfma(aih, brh, ci);
  ci = mfma(aih, brl, ci);
  ci = mfma(ail, brh, ci);
}

__device__ __forceinline__ half8 ld8(const _Float16 *p) {
  return *reinterpret_cast<const half8 *>(p);
}

__device__ __forceinline__ float block_max(float v, float *red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float m = red[0];
#pragma unroll
  for (int i = 1; i < DFT_NW; i++) m = fmaxf(m, red[i]);
  __syncthreads();
  return m;
}

// 2^e putting m near 2^14 (fp16 max 65504); 1 for m == 0.
__device__ __forceinline__ float pow2_scale(float m) {
  if (!(m > 0.f) || !isfinite(m)) return 1.f;
  int ex;
  frexpf(m, &ex);  // m < 2^ex
  return ldexpf(1.f, 14 - ex);
}

__device__ __forceinline__ float2 modulate(float2 v, const ModArgs &md, int b, int pix) {
  if (!md.phi) return v;
  const double ang = (double)md.kplane[b] * md.phi[pix] * (double)md.sign;
  const double red = ang - 6.283185307179586 * rint(ang * 0.15915494309189535);
  float s, c;
  sincosf((float)red, &s, &c);
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

// D split tables: dsplit[0..3] = Re hi, Re lo, Im hi, Im lo, each [n][n].
__global__ void k_dft_matrix(_Float16 *dsplit, int n, int sign) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int k = idx / n, j = idx - k * n;
  const int64_t m = (int64_t)(k - n / 2) * (int64_t)(j - n / 2);
  int64_t r = m % n;
  if (r < 0) r += n;
  double s, c;
  sincospi(2.0 * (double)sign * (double)r / (double)n, &s, &c);
  _Float16 h, l;
  const float cf = (float)c, sf = (float)s;
  h = (_Float16)c;
  l = (_Float16)(c - (double)h);
  dsplit[idx] = h;
  dsplit[n * n + idx] = l;
  h = (_Float16)s;
  l = (_Float16)(s - (double)h);
  dsplit[2 * n * n + idx] = h;
  dsplit[3 * n * n + idx] = l;
  (void)cf;
  (void)sf;
}

__global__ __launch_bounds__(DFT_NT) void k_dft2d(
    const float2 *in, float2 *out, int64_t batch_stride,  // may alias (in place)
    const int32_t *__restrict__ in_idx, const int32_t *__restrict__ out_idx,
    const _Float16 *__restrict__ dsplit, float out_scale, ModArgs mod_in,
    ModArgs mod_out) {
  constexpr int n = DFT_N, NN = n * n;
  __shared__ _Float16 lds[4 * PLANE];
  __shared__ float red[DFT_NW];
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const float2 *X = in + (int64_t)(in_idx ? in_idx[b] : b) * batch_stride;
  float2 *Y = out + (int64_t)(out_idx ? out_idx[b] : b) * batch_stride;
  const _Float16 *Drh = dsplit, *Drl = dsplit + NN, *Dih = dsplit + 2 * NN,
                 *Dil = dsplit + 3 * NN;
  _Float16 *Lrh = lds, *Lrl = lds + PLANE, *Lih = lds + 2 * PLANE, *Lil = lds + 3 * PLANE;

  // ---- load X (optionally modulated) once into registers, scale to the
  // fp16 range, split ----
  // (each thread holds four consecutive elements of a row per step, loaded
  // as two 16-byte words and stored as one 8-byte word per split plane)
  constexpr int PER4 = NN / (4 * DFT_NT);
  static_assert(PER4 * 4 * DFT_NT == NN, "threads cover the sub-grid exactly");
  float2 xv[4 * PER4];
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < PER4; k++) {
    const int i = 4 * (tid + k * DFT_NT);
    const float4 p0 = reinterpret_cast<const float4 *>(X + i)[0];
    const float4 p1 = reinterpret_cast<const float4 *>(X + i)[1];
    xv[4 * k] = modulate(make_float2(p0.x, p0.y), mod_in, b, i);
    xv[4 * k + 1] = modulate(make_float2(p0.z, p0.w), mod_in, b, i + 1);
    xv[4 * k + 2] = modulate(make_float2(p1.x, p1.y), mod_in, b, i + 2);
    xv[4 * k + 3] = modulate(make_float2(p1.z, p1.w), mod_in, b, i + 3);
#pragma unroll
    for (int j = 0; j < 4; j++)
      m = fmaxf(m, fmaxf(fabsf(xv[4 * k + j].x), fabsf(xv[4 * k + j].y)));
  }
  const float s1 = pow2_scale(block_max(m, red));
#pragma unroll
  for (int k = 0; k < PER4; k++) {
    const int i = 4 * (tid + k * DFT_NT);
    const int y = i / n, x = i - y * n;
    const int o = y * LDS_STRIDE + x;
    const float2 *v = xv + 4 * k;
    split4_store(v[0].x * s1, v[1].x * s1, v[2].x * s1, v[3].x * s1, Lrh, Lrl, o);
    split4_store(v[0].y * s1, v[1].y * s1, v[2].y * s1, v[3].y * s1, Lih, Lil, o);
  }
  __syncthreads();

  // ---- pass 1: T = X D; wavefront w owns columns [16w, 16w + 16) ----
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  floatx4 tr[DFT_NW], ti[DFT_NW];
#pragma unroll
  for (int t = 0; t < DFT_NW; t++) tr[t] = ti[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  // B = D[k0 + lk .. +8][16w + lr]: D symmetric, read row 16w + lr; the
  // next k-step's fragments are loaded before this one's products.
  int dro = (16 * w + lr) * n + lk;
  half8 nrh = ld8(Drh + dro), nrl = ld8(Drl + dro), nih = ld8(Dih + dro), nil = ld8(Dil + dro);
  for (int k0 = 0; k0 < n; k0 += 32) {
    const half8 brh = nrh, brl = nrl, bih = nih, bil = nil;
    if (k0 + 32 < n) {
      dro += 32;
      nrh = ld8(Drh + dro);
      nrl = ld8(Drl + dro);
      nih = ld8(Dih + dro);
      nil = ld8(Dil + dro);
    }
#pragma unroll
    for (int t = 0; t < DFT_NW; t++) {
      const int xo = (16 * t + lr) * LDS_STRIDE + k0 + lk;  // A = X rows of tile-row t
      cmac(tr[t], ti[t], ld8(Lrh + xo), ld8(Lrl + xo), ld8(Lih + xo), ld8(Lil + xo), brh, brl,
           bih, bil);
    }
  }
  // rescale T to the fp16 range
  m = 0.f;
#pragma unroll
  for (int t = 0; t < DFT_NW; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) m = fmaxf(m, fmaxf(fabsf(tr[t][r]), fabsf(ti[t][r])));
  const float s2 = pow2_scale(block_max(m, red));  // (barrier: pass 1 reads done)
  // store T^T: element (row y = 16t + 4(lane>>4) + r, col x = 16w + lr) at
  // [x][y]; r = 0..3 are consecutive in y, one 8-byte word per plane
#pragma unroll
  for (int t = 0; t < DFT_NW; t++) {
    const int o = (16 * w + lr) * LDS_STRIDE + 16 * t + 4 * (lane >> 4);
    split4_store(tr[t][0] * s2, tr[t][1] * s2, tr[t][2] * s2, tr[t][3] * s2, Lrh, Lrl, o);
    split4_store(ti[t][0] * s2, ti[t][1] * s2, ti[t][2] * s2, ti[t][3] * s2, Lih, Lil, o);
  }
  __syncthreads();

  // ---- pass 2: Y = D T; wavefront w owns rows [16w, 16w + 16) ----
  floatx4 yr[DFT_NW], yi[DFT_NW];
#pragma unroll
  for (int t = 0; t < DFT_NW; t++) yr[t] = yi[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  int dao = (16 * w + lr) * n + lk;  // A = D rows of tile-row w, one k-step ahead
  half8 nah = ld8(Drh + dao), nal = ld8(Drl + dao), nbh = ld8(Dih + dao), nbl = ld8(Dil + dao);
  for (int k0 = 0; k0 < n; k0 += 32) {
    const half8 arh = nah, arl = nal, aih = nbh, ail = nbl;
    if (k0 + 32 < n) {
      dao += 32;
      nah = ld8(Drh + dao);
      nal = ld8(Drl + dao);
      nbh = ld8(Dih + dao);
      nbl = ld8(Dil + dao);
    }
#pragma unroll
    for (int t = 0; t < DFT_NW; t++) {
      const int to = (16 * t + lr) * LDS_STRIDE + k0 + lk;  // B = T[k][16t + lr] = T^T row
      cmac(yr[t], yi[t], arh, arl, aih, ail, ld8(Lrh + to), ld8(Lrl + to), ld8(Lih + to),
           ld8(Lil + to));
    }
  }
  const float inv = out_scale / (s1 * s2);
#pragma unroll
  for (int t = 0; t < DFT_NW; t++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 16 * w + 4 * (lane >> 4) + r, col = 16 * t + lr;
      const int i = row * n + col;
      Y[i] = modulate(make_float2(yr[t][r] * inv, yi[t][r] * inv), mod_out, b, i);
    }
  }
}

}  // namespace

// Device tables of the centred DFT matrix split, cached in the context.
int dft_tables(croc_ctx *ctx, int n, int sign, const _Float16 **out) {
  CROC_REQUIRE(n == DFT_N, CROC_EINVAL, "sub-grid size %d unsupported (only %d)", n, DFT_N);
  const WsSlot slot = sign < 0 ? WS_DFT_FWD : WS_DFT_INV;
  const size_t bytes = sizeof(_Float16) * 4 * n * n;
  const bool fresh = ctx->ws.ws_size[slot] < bytes;
  _Float16 *d = (_Float16 *)ws_get(ctx, slot, bytes);
  if (!d) return CROC_ENOMEM;
  if (fresh) {
    k_dft_matrix<<<(n * n + 255) / 256, 256, 0, ctx->stream>>>(d, n, sign);
    CROC_HIP(hipGetLastError());
  }
  *out = d;
  return CROC_OK;
}

int dft2d_launch(croc_ctx *ctx, const void *in, void *out, int n, int64_t batch,
                 int64_t batch_stride, int sign, float out_scale, const double *phi_in,
                 const int32_t *k_in, int sign_in, const double *phi_out,
                 const int32_t *k_out, int sign_out, const int32_t *in_idx,
                 const int32_t *out_idx) {
  const _Float16 *d = nullptr;
  int rc = dft_tables(ctx, n, sign, &d);
  if (rc) return rc;
  if (batch <= 0) return CROC_OK;
  ModArgs mi{phi_in, k_in, sign_in}, mo{phi_out, k_out, sign_out};
  k_dft2d<<<(unsigned)batch, DFT_NT, 0, ctx->stream>>>((const float2 *)in, (float2 *)out,
                                                        batch_stride, in_idx, out_idx, d,
                                                        out_scale, mi, mo);
  CROC_HIP(hipGetLastError());
  return CROC_OK;
}

extern "C" int croc_dft2d_batched(croc_ctx *ctx, const void *in, void *out, int n,
                                  int64_t batch, int inverse) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(in && out, CROC_EINVAL, "null array");
  CROC_REQUIRE(batch >= 0 && batch <= 0x7fffffff, CROC_EINVAL, "bad batch");
  const float scale = inverse ? 1.0f / ((float)n * (float)n) : 1.0f;
  rc = dft2d_launch(ctx, in, out, n, batch, (int64_t)n * n, inverse ? +1 : -1, scale, nullptr,
                    nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr);
  if (rc) return rc;
  CROC_HIP(hipStreamSynchronize(ctx->stream));
  return CROC_OK;
}
