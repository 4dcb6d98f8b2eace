/*
 * FP64 C restatement of the reference's native visibility-layout functions
 * -- TEST INFRASTRUCTURE (oracle).  Used by tests and by bench.py's
 * cpu_baseline leg only; the product library never links it.
 *
 * Follows examples/grid/grid.c and streamer.c:
 *   uvw_lambda          grid.h:53-60     (u_m * f / c, rounded as written)
 *   frac_coord          grid.c:35-53     (ox = rint((N/2 - u) * Q), x = N - ox/Q,
 *                                         fx = ox % Q, C truncating division)
 *   coord               grid.c:22-32     (floor(theta * u + .5) + N/2)
 *   frac_coord_sep_uv   grid.c:84-100
 *   weight              grid.c:119-141
 *   degrid_simple       grid.c:143-159
 *   grid_simple         grid.c:161-177
 *   degrid_conv_uv      grid.c:179-205   (the portable, non-AVX2 loop)
 *   degrid_conv_bl      grid.c:234-259
 *   degrid_conv         grid.c:261-275
 *   w_project           grid.c:277-305   (plane floor((w - w_min)/w_step + .5),
 *                                         += v * conj(k))
 *   grid_wprojection    grid.c:307-322
 *   streamer_degrid_worker  streamer.c:366-486 (mirror rule :418-438; the
 *                                         DFT spot checks :444-464 are left out)
 *
 * grid.c's frac_coord calls fesetround(3) first.  3 is not a rounding mode
 * of x86-64 glibc (FE_TONEAREST 0, FE_DOWNWARD 0x400, FE_UPWARD 0x800,
 * FE_TOWARDZERO 0xc00), so the call fails and the default round-half-even
 * mode stays: the reference's own T05_frac_coord vectors
 * (test_recombine.c:378-412) show exactly that.
 *
 * The visibility layout is a set of baselines in flat arrays (the device
 * library's croc_bl_set): baseline b has nt[b] dumps and nf[b] channels, uvw
 * rows uvw_m[3 * (uvw0[b] + t) + k] in metres, frequencies freq[f0[b] + f]
 * in Hz and visibilities vis[v0[b] + t * nf[b] + f] (grid.h:14-34 holds
 * the same fields per baseline).
 *
 * Out-of-range accesses that the C code does not check (a kernel window
 * leaving the grid, which there wraps into the neighbouring row or leaves the
 * array) are reported here as BL_EBOUNDS instead.
 *
 * Build with -ffp-contract=off so products and sums round as written.
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#define BL_OK 0
#define BL_EBOUNDS -2   /* window outside the grid                     */
#define BL_EARG -3      /* inconsistent arguments                      */

static const double SPEED_OF_LIGHT = 299792458.0;   /* grid.h:51 */

typedef struct {
    int64_t nbl;
    const int32_t *nt, *nf;
    const int64_t *uvw0, *f0, *v0;
    const double *uvw_m, *freq;
    double *vis;                 /* complex128 interleaved */
} bl_set;

static double uvw_lambda(const bl_set *s, int64_t b, int t, int f, int k) {
    volatile double prod = s->uvw_m[3 * (s->uvw0[b] + t) + k] * s->freq[s->f0[b] + f];
    return prod / SPEED_OF_LIGHT;
}

void bl_frac_coord(int grid_size, int oversample, double u, int64_t *x, int64_t *fx) {
    volatile double d = (double)(grid_size / 2) - u;
    volatile double s = d * (double)oversample;
    int64_t ox = (int64_t)nearbyint(s);
    *x = grid_size - ox / oversample;
    *fx = ox % oversample;
}

int bl_frac_coords(int grid_size, int oversample, const double *u, int64_t n,
                   int64_t *x, int64_t *fx) {
    for (int64_t i = 0; i < n; i++) bl_frac_coord(grid_size, oversample, u[i], &x[i], &fx[i]);
    return BL_OK;
}

static int coord(int grid_size, double theta, double u, double v, int64_t *idx) {
    volatile double tu = theta * u, tv = theta * v;
    int64_t x = (int64_t)floor(tu + .5), y = (int64_t)floor(tv + .5);
    x += grid_size / 2;
    y += grid_size / 2;
    if (x < 0 || y < 0 || x >= grid_size || y >= grid_size) return BL_EBOUNDS;
    *idx = y * grid_size + x;
    return BL_OK;
}

/* degrid_conv_uv: real separable kernel [Q][S], no conjugation. */
static int degrid_uv(const double *g, int N, double theta, double u, double v,
                     const double *kern, int S, int Q, double *out) {
    volatile double xu = theta * u, yv = theta * v;
    int64_t ix, ixf, iy, iyf;
    bl_frac_coord(N, Q, xu, &ix, &ixf);
    bl_frac_coord(N, Q, yv, &iy, &iyf);
    if (ixf < 0 || iyf < 0) return BL_EBOUNDS;
    int64_t x0 = ix - S / 2, y0 = iy - S / 2;
    if (x0 < 0 || y0 < 0 || x0 + S > N || y0 + S > N) return BL_EBOUNDS;
    const double *kx = kern + S * ixf, *ky = kern + S * iyf;
    double vr = 0, vi = 0;
    for (int y = 0; y < S; y++) {
        double sr = 0, si = 0;
        const double *row = g + 2 * ((y0 + y) * N + x0);
        for (int x = 0; x < S; x++) {
            sr += kx[x] * row[2 * x];
            si += kx[x] * row[2 * x + 1];
        }
        vr += ky[y] * sr;
        vi += ky[y] * si;
    }
    out[0] = vr;
    out[1] = vi;
    return BL_OK;
}

int bl_degrid_conv_uv(const double *g, int N, double theta, const double *uv, int64_t n,
                      const double *kern, int S, int Q, double *out) {
    for (int64_t i = 0; i < n; i++) {
        int rc = degrid_uv(g, N, theta, uv[2 * i], uv[2 * i + 1], kern, S, Q, out + 2 * i);
        if (rc) return rc;
    }
    return BL_OK;
}

/* degrid_conv_bl over every baseline of the set: window [time0, time1) x
 * [freq0, freq1), visibilities outside [min_u, max_u) x [min_v, max_v) are
 * left untouched, the others written to the chunk-local index
 * v0[b] + (time - time0) * (freq1 - freq0) + freq - freq0.  *ndeg counts the
 * degridded visibilities (flops = 4 (1 + S) S each). */
int bl_degrid_conv_bl(const double *g, int N, double theta, double d_u, double d_v,
                      double min_u, double max_u, double min_v, double max_v,
                      const bl_set *s, int time0, int time1, int freq0, int freq1,
                      const double *kern, int S, int Q, int64_t *ndeg) {
    int64_t n = 0;
    for (int64_t b = 0; b < s->nbl; b++) {
        if (time0 < 0 || freq0 < 0 || time1 > s->nt[b] || freq1 > s->nf[b]) return BL_EARG;
        for (int t = time0; t < time1; t++) {
            for (int f = freq0; f < freq1; f++) {
                double u = uvw_lambda(s, b, t, f, 0);
                if (u < min_u || u >= max_u) continue;
                double v = uvw_lambda(s, b, t, f, 1);
                if (v < min_v || v >= max_v) continue;
                volatile double uu = u - d_u, vv = v - d_v;
                int64_t o = s->v0[b] + (int64_t)(t - time0) * (freq1 - freq0) + (f - freq0);
                int rc = degrid_uv(g, N, theta, uu, vv, kern, S, Q, s->vis + 2 * o);
                if (rc) return rc;
                n++;
            }
        }
    }
    if (ndeg) *ndeg = n;
    return BL_OK;
}

/* streamer_degrid_worker for a list of chunks (b, it0, it1, if0, if1): chunk
 * c's visibility (time, freq) goes to out[c * time_chunk * freq_chunk +
 * (time - it0) * freq_chunk + freq - if0]; zero when it lies in neither the
 * window nor its mirror image.  ndeg[c] counts chunk c's degridded
 * visibilities. */
int bl_streamer_degrid(const double *g, int N, double theta, double mid_u, double mid_v,
                       double min_u, double max_u, double min_v, double max_v,
                       const bl_set *s, const int32_t *chunks, int64_t nchunks,
                       int time_chunk, int freq_chunk, const double *kern, int S, int Q,
                       double *out, int64_t *ndeg) {
    const double umin_pos = fmax(min_u, 0), umin_neg = fmax(min_u, DBL_MIN);
    for (int64_t c = 0; c < nchunks; c++) {
        const int32_t *ch = chunks + 5 * c;
        int64_t b = ch[0];
        int it0 = ch[1], it1 = ch[2], if0 = ch[3], if1 = ch[4];
        if (b < 0 || b >= s->nbl || it0 < 0 || if0 < 0 || it1 > s->nt[b] || if1 > s->nf[b] ||
            it1 - it0 > time_chunk || if1 - if0 > freq_chunk)
            return BL_EARG;
        int64_t n = 0;
        double *slot = out + 2 * c * (int64_t)time_chunk * freq_chunk;
        for (int t = it0; t < it1; t++) {
            for (int f = if0; f < if1; f++) {
                double u = uvw_lambda(s, b, t, f, 0);
                double v = uvw_lambda(s, b, t, f, 1);
                double *o = slot + 2 * ((int64_t)(t - it0) * freq_chunk + (f - if0));
                int rc;
                if (u >= umin_pos && u < max_u && v >= min_v && v < max_v) {
                    volatile double uu = u - mid_u, vv = v - mid_v;
                    rc = degrid_uv(g, N, theta, uu, vv, kern, S, Q, o);
                } else if (-u >= umin_neg && -u < max_u && -v >= min_v && -v < max_v) {
                    volatile double uu = -u - mid_u, vv = -v - mid_v;
                    rc = degrid_uv(g, N, theta, uu, vv, kern, S, Q, o);
                    o[1] = -o[1];
                } else {
                    o[0] = o[1] = 0;
                    continue;
                }
                if (rc) return rc;
                n++;
            }
        }
        if (ndeg) ndeg[c] = n;
    }
    return BL_OK;
}

/* w_project over every visibility: 2D w-kernels [plane][Q][Q][S][S],
 * grid += v * conj(k). */
int bl_grid_wprojection(double *g, int N, double theta, const bl_set *s, const double *wkern,
                        int plane_count, double w_min, double w_step, int S, int Q) {
    for (int64_t b = 0; b < s->nbl; b++) {
        for (int t = 0; t < s->nt[b]; t++) {
            for (int f = 0; f < s->nf[b]; f++) {
                volatile double xu = theta * uvw_lambda(s, b, t, f, 0);
                volatile double yv = theta * uvw_lambda(s, b, t, f, 1);
                int64_t ix, ixf, iy, iyf;
                bl_frac_coord(N, Q, xu, &ix, &ixf);
                bl_frac_coord(N, Q, yv, &iy, &iyf);
                if (ixf < 0 || iyf < 0) return BL_EBOUNDS;
                int64_t x0 = ix - S / 2, y0 = iy - S / 2;
                if (x0 < 0 || y0 < 0 || x0 + S > N || y0 + S > N) return BL_EBOUNDS;
                double w = uvw_lambda(s, b, t, f, 2);
                volatile double dw = (w - w_min) / w_step;
                int64_t p = (int64_t)floor(dw + .5);
                if (p < 0 || p >= plane_count) return BL_EBOUNDS;
                const double *k = wkern + 2 * (((p * Q + iyf) * Q + ixf) * (int64_t)S * S);
                const double *v = s->vis + 2 * (s->v0[b] + (int64_t)t * s->nf[b] + f);
                for (int y = 0; y < S; y++) {
                    for (int x = 0; x < S; x++) {
                        const double *kk = k + 2 * (y * S + x);
                        double *c = g + 2 * ((y0 + y) * N + x0 + x);
                        /* v * conj(k) */
                        volatile double re = v[0] * kk[0] + v[1] * kk[1];
                        volatile double im = v[1] * kk[0] - v[0] * kk[1];
                        c[0] += re;
                        c[1] += im;
                    }
                }
            }
        }
    }
    return BL_OK;
}

/* weight: count visibilities per nearest cell, then divide each by its
 * cell's count. */
int bl_weight(uint32_t *wgrid, int N, double theta, const bl_set *s) {
    memset(wgrid, 0, sizeof(uint32_t) * (size_t)N * N);
    for (int pass = 0; pass < 2; pass++) {
        for (int64_t b = 0; b < s->nbl; b++) {
            for (int t = 0; t < s->nt[b]; t++) {
                for (int f = 0; f < s->nf[b]; f++) {
                    int64_t idx;
                    int rc = coord(N, theta, uvw_lambda(s, b, t, f, 0),
                                   uvw_lambda(s, b, t, f, 1), &idx);
                    if (rc) return rc;
                    if (pass == 0) {
                        wgrid[idx]++;
                    } else {
                        double *v = s->vis + 2 * (s->v0[b] + (int64_t)t * s->nf[b] + f);
                        v[0] /= (double)wgrid[idx];
                        v[1] /= (double)wgrid[idx];
                    }
                }
            }
        }
    }
    return BL_OK;
}

int bl_grid_simple(double *g, int N, double theta, const bl_set *s) {
    for (int64_t b = 0; b < s->nbl; b++)
        for (int t = 0; t < s->nt[b]; t++)
            for (int f = 0; f < s->nf[b]; f++) {
                int64_t idx;
                int rc = coord(N, theta, uvw_lambda(s, b, t, f, 0),
                               uvw_lambda(s, b, t, f, 1), &idx);
                if (rc) return rc;
                const double *v = s->vis + 2 * (s->v0[b] + (int64_t)t * s->nf[b] + f);
                g[2 * idx] += v[0];
                g[2 * idx + 1] += v[1];
            }
    return BL_OK;
}

int bl_degrid_simple(const double *g, int N, double theta, const bl_set *s) {
    for (int64_t b = 0; b < s->nbl; b++)
        for (int t = 0; t < s->nt[b]; t++)
            for (int f = 0; f < s->nf[b]; f++) {
                int64_t idx;
                int rc = coord(N, theta, uvw_lambda(s, b, t, f, 0),
                               uvw_lambda(s, b, t, f, 1), &idx);
                if (rc) return rc;
                double *v = s->vis + 2 * (s->v0[b] + (int64_t)t * s->nf[b] + f);
                v[0] = g[2 * idx];
                v[1] = g[2 * idx + 1];
            }
    return BL_OK;
}
