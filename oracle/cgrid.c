/*
 * FP64 C restatement of crocodile's convolutional gridding core -- TEST
 * INFRASTRUCTURE (oracle).  Used by tests and by bench.py's cpu_baseline leg
 * only; the product library never links it.
 *
 * Follows crocodile/synthesis.py:
 *   frac_coord   :419-435  (x = N//2 + p*N; fx = round-half-even(-x*Q);
 *                           ix = -floor(fx/Q); xf = fx mod Q)
 *   convgrid     :464-481  (window rows iy - S//2 .. iy + (S+1)//2 - 1,
 *                           a[win] += k(yf, xf) * v, no conjugation)
 *   convdegrid   :484-503  (v = sum(a[win] * k(yf, xf)))
 *   make_gcf     :451-461  (separable: k = outer(g[yf], g[xf]); 2D: g[yf,xf])
 * and the per-visibility w-bin selection of w_cache_imaging :792-795.
 *
 * Build with -ffp-contract=off so products and sums round exactly as NumPy's
 * complex arithmetic does.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#define CG_OK 0
#define CG_EDOMAIN -1   /* p outside [-0.5, 0.5)  (AssertionError)        */
#define CG_EBOUNDS -2   /* kernel window leaves the grid (ValueError)      */
#define CG_EARG -3      /* inconsistent arguments                          */

static int64_t floordiv(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

static int frac1(int N, int Q, double p, int64_t *ix, int64_t *xf) {
    if (!(p >= -0.5 && p < 0.5)) return CG_EDOMAIN;
    volatile double prod = p * (double)N;          /* round p*N first   */
    volatile double x = (double)(N / 2) + prod;    /* then the add      */
    volatile double neg = -x * (double)Q;
    double fx = nearbyint(neg);                    /* default mode: half-even */
    int64_t f = (int64_t)fx;
    *ix = -floordiv(f, Q);
    *xf = f - floordiv(f, Q) * Q;
    return CG_OK;
}

int cg_frac_coord(int N, int Q, const double *p, int64_t n, int64_t stride,
                  int64_t *ix, int64_t *xf) {
    for (int64_t i = 0; i < n; i++) {
        int rc = frac1(N, Q, p[i * stride], &ix[i], &xf[i]);
        if (rc) return rc;
    }
    return CG_OK;
}

/* Kernel bank: plane k holds a table at complex offset plane_off[k] with
 * support plane_s[k] (square).  Layout [Q][Q][S][S] (2D) or [Q][S]
 * (separable).  vplane[i] selects the plane of visibility i (NULL: plane 0). */
typedef struct {
    const double *kern;          /* complex128 interleaved */
    const int64_t *plane_off;
    const int32_t *plane_s;
    int Q;
    int separable;
} cg_bank;

static inline void tap(const cg_bank *b, int64_t off, int S, int64_t yf,
                       int64_t xf, int i, int j, double *kr, double *ki) {
    if (b->separable) {
        const double *gy = b->kern + 2 * (off + yf * S + i);
        const double *gx = b->kern + 2 * (off + xf * S + j);
        *kr = gy[0] * gx[0] - gy[1] * gx[1];
        *ki = gy[0] * gx[1] + gy[1] * gx[0];
    } else {
        const double *g = b->kern + 2 * (off + ((yf * b->Q + xf) * S + i) * S + j);
        *kr = g[0];
        *ki = g[1];
    }
}

static int locate(int N, const cg_bank *b, const double *pi, int S,
                  int64_t *x0, int64_t *xf, int64_t *y0, int64_t *yf) {
    int64_t ix, iy;
    int rc = frac1(N, b->Q, pi[0], &ix, xf);
    if (rc) return rc;
    rc = frac1(N, b->Q, pi[1], &iy, yf);
    if (rc) return rc;
    *x0 = ix - S / 2;
    *y0 = iy - S / 2;
    if (*x0 < 0 || *y0 < 0 || *x0 + S > N || *y0 + S > N) return CG_EBOUNDS;
    return CG_OK;
}

/* grid: complex128 [N][N] updated in place; p: [nv][pstride] fraction of
 * the grid (uvw / lam); vis: complex128 [nv]. */
int cg_convgrid(double *grid, int N, const double *p, int64_t nv,
                int64_t pstride, const double *vis, const double *kern, int Q,
                int separable, const int64_t *plane_off,
                const int32_t *plane_s, const int32_t *vplane) {
    cg_bank b = {kern, plane_off, plane_s, Q, separable};
    for (int64_t v = 0; v < nv; v++) {
        int k = vplane ? vplane[v] : 0;
        int S = plane_s[k];
        int64_t x0, xf, y0, yf;
        int rc = locate(N, &b, p + v * pstride, S, &x0, &xf, &y0, &yf);
        if (rc) return rc;
        double vr = vis[2 * v], vi = vis[2 * v + 1];
        for (int i = 0; i < S; i++) {
            double *row = grid + 2 * ((y0 + i) * (int64_t)N + x0);
            for (int j = 0; j < S; j++) {
                double kr, ki;
                tap(&b, plane_off[k], S, yf, xf, i, j, &kr, &ki);
                row[2 * j] += kr * vr - ki * vi;
                row[2 * j + 1] += kr * vi + ki * vr;
            }
        }
    }
    return CG_OK;
}

int cg_convdegrid(const double *grid, int N, const double *p, int64_t nv,
                  int64_t pstride, double *vis, const double *kern, int Q,
                  int separable, const int64_t *plane_off,
                  const int32_t *plane_s, const int32_t *vplane) {
    cg_bank b = {kern, plane_off, plane_s, Q, separable};
    for (int64_t v = 0; v < nv; v++) {
        int k = vplane ? vplane[v] : 0;
        int S = plane_s[k];
        int64_t x0, xf, y0, yf;
        int rc = locate(N, &b, p + v * pstride, S, &x0, &xf, &y0, &yf);
        if (rc) return rc;
        double sr = 0, si = 0;
        for (int i = 0; i < S; i++) {
            const double *row = grid + 2 * ((y0 + i) * (int64_t)N + x0);
            for (int j = 0; j < S; j++) {
                double kr, ki;
                tap(&b, plane_off[k], S, yf, xf, i, j, &kr, &ki);
                sr += row[2 * j] * kr - row[2 * j + 1] * ki;
                si += row[2 * j] * ki + row[2 * j + 1] * kr;
            }
        }
        vis[2 * v] = sr;
        vis[2 * v + 1] = si;
    }
    return CG_OK;
}

/* w-bin index of each visibility: round-half-even(w / wstep), the
 * numpy.round of w_cache_imaging (synthesis.py:793). */
int cg_wbin(const double *w, int64_t n, int64_t stride, double wstep,
            int64_t *out) {
    for (int64_t i = 0; i < n; i++) {
        volatile double q = w[i * stride] / wstep;
        out[i] = (int64_t)nearbyint(q);
    }
    return CG_OK;
}
