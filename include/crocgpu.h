/*
 * crocgpu.h -- C ABI of the MI355X (gfx950) gridding / degridding library
 * libcrocgpu.so.
 *
 * Every entry point takes plain device pointers and sizes, returns an int
 * status (CROC_OK or a negative CROC_E* code; croc_last_error() gives the
 * message for the calling thread) and runs on the context's HIP stream.
 * Entry points that report per-visibility errors (coordinate domain, window
 * bounds) synchronise the stream before returning; on error the output
 * arrays are left unmodified.
 *
 * Array conventions (all device memory):
 *   cf32  complex64 stored as interleaved float pairs (re, im)
 *   cf64  complex128 stored as interleaved double pairs
 *   uvw   double [nv][stride], columns u, v, w (wavelengths); the grid
 *         coordinate is p = uvw / lam computed in IEEE FP64 as NumPy does
 *   grid  cf32 [N][N] row-major, rows = v, columns = u, centre at N/2
 *   kernel bank: kernel tables in one allocation of at least
 *         max_k(plane_off[k] + table size) complex elements; plane k starts
 *         at element plane_off[k] and has support plane_s[k] (square);
 *         2D table layout [Q][Q][S][S] = [yf][xf][y][x]; separable table
 *         layout [Q][S] (the kernel is the outer product of rows yf and xf)
 *
 * Plane selection (which table a visibility uses):
 *   CROC_PLANE_SINGLE  every visibility uses plane 0
 *   CROC_PLANE_INDEX   plane = vplane[i]            (src-indexed tables)
 *   CROC_PLANE_WBIN    plane = rint(w_i / wstep) - wbin0   (w-cache bins,
 *                      rint is round-half-even like numpy.round)
 */
#ifndef CROCGPU_H
#define CROCGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CROC_OK 0
#define CROC_EINVAL (-1)   /* invalid argument                                   */
#define CROC_EDOMAIN (-2)  /* grid coordinate outside [-0.5, 0.5): AssertionError */
#define CROC_EBOUNDS (-3)  /* kernel window leaves the grid: ValueError          */
#define CROC_EHIP (-4)     /* HIP runtime failure                                 */
#define CROC_ENOMEM (-5)   /* device allocation failed                            */
#define CROC_EPLANE (-6)   /* w-bin / plane index outside the kernel bank         */

#define CROC_PLANE_SINGLE 0
#define CROC_PLANE_INDEX 1
#define CROC_PLANE_WBIN 2

typedef struct croc_ctx croc_ctx;

/* Library / context management (no reference counterpart: the reference is
 * a set of pure NumPy functions; state it keeps implicitly -- the pylru
 * kernel cache of synthesis.py:788 -- lives in a context here). */
int croc_version(void);
const char *croc_last_error(void);
int croc_ctx_create(int device, croc_ctx **out);
int croc_ctx_destroy(croc_ctx *ctx);
int croc_ctx_set_stream(croc_ctx *ctx, void *hip_stream);
/* Tuning knobs of convgrid / convdegrid: tile edge in grid cells (0 =
 * automatic, else a multiple of 16 in [16, 80]), the most visibilities one
 * work item of a tile holds (>= 64), and the kernel-bank KiB per plane-group
 * (0 = automatic, the default: eight groups when the bank can exceed 4 MiB,
 * else one).  Degridding splits the groups into one contiguous range per
 * XCD, so that each XCD's L2 serves its own slice of a bank far larger than
 * one L2; gridding visits each tile's visibilities group by group.  Results
 * do not depend on them. */
int croc_ctx_set_tuning(croc_ctx *ctx, int tile, int chunk, int group_kib);
/* Device time in milliseconds of the last gridding / degridding kernel
 * (HIP events around its launch on the context stream). */
int croc_last_kernel_ms(croc_ctx *ctx, float *grid_ms, float *degrid_ms);
/* Device time in milliseconds summed over the launches of the tap kernels
 * proper (the kernels that spread / gather kernel taps: k_grid_tiles,
 * k_tw_grid_jobs, k_degrid_tiles, k_tw_degrid_jobs, ...) of the last
 * gridding and degridding call on this context, and their launch counts
 * (HIP events around each launch, on the context stream). */
int croc_last_tap_kernel_ms(croc_ctx *ctx, float *grid_ms, int *grid_launches,
                            float *degrid_ms, int *degrid_launches);

/* frac_coords: replaces crocodile/synthesis.py:438-449 (frac_coords) and
 * :419-435 (frac_coord) -- x = N//2 + p*N, fx = round-half-even(-x*Q),
 * ix = -floor(fx/Q), xf = fx mod Q; p = uvw/lam in FP64.  Outputs int32
 * [nv].  Returns CROC_EDOMAIN if any p lies outside [-0.5, 0.5).
 * Native counterpart: examples/grid/grid.h:177-178 (frac_coord). */
int croc_frac_coords(croc_ctx *ctx, const double *uvw, int64_t nv,
                     int64_t stride, double lam, int Nh, int Nw, int Q,
                     int32_t *x, int32_t *xf, int32_t *y, int32_t *yf);

/* convgrid: replaces crocodile/synthesis.py:464-481 (convgrid, in-place
 * accumulation a[window] += k(s, xf, yf) * v, no conjugation) with the
 * kernel selection of make_gcf :451-461 and of w_cache_imaging :792-795
 * (plane_mode CROC_PLANE_WBIN).  grid is updated in place.
 * Native counterparts: examples/grid/grid.h:192-193 (grid_wprojection),
 * grid.c:750-792 (grid_awprojection). */
int croc_convgrid(croc_ctx *ctx, void *grid, int N, const double *uvw,
                  int64_t nv, int64_t stride, double lam, const void *vis,
                  const void *kern, int Q, int separable, int nplanes,
                  const int64_t *plane_off, const int32_t *plane_s,
                  int max_support, int plane_mode, const int32_t *vplane,
                  double wstep, int wbin0);

/* convdegrid: replaces crocodile/synthesis.py:484-503 (convdegrid,
 * v = sum(a[window] * k(s, xf, yf))) and the per-visibility loop of
 * w_cache_predict :799-828.  vis_out cf32 [nv].
 * Native counterpart: examples/grid/grid.h:183-186 (degrid_conv_uv). */
int croc_convdegrid(croc_ctx *ctx, const void *grid, int N, const double *uvw,
                    int64_t nv, int64_t stride, double lam, void *vis_out,
                    const void *kern, int Q, int separable, int nplanes,
                    const int64_t *plane_off, const int32_t *plane_s,
                    int max_support, int plane_mode, const int32_t *vplane,
                    double wstep, int wbin0);

/* The plain per-visibility GPU form of convgrid / convdegrid
 * (crocodile/synthesis.py:464-503, same arguments and errors as
 * croc_convgrid / croc_convdegrid): one wavefront per visibility, FP32
 * atomic adds straight into the grid, no binning and no LDS.  It is the
 * baseline the w-towers pipeline is compared with; its gridding rounding
 * depends on the order of the atomics. */
int croc_convgrid_pervis(croc_ctx *ctx, void *grid, int N, const double *uvw,
                         int64_t nv, int64_t stride, double lam, const void *vis,
                         const void *kern, int Q, int separable, int nplanes,
                         const int64_t *plane_off, const int32_t *plane_s,
                         int max_support, int plane_mode, const int32_t *vplane,
                         double wstep, int wbin0);
int croc_convdegrid_pervis(croc_ctx *ctx, const void *grid, int N, const double *uvw,
                           int64_t nv, int64_t stride, double lam, void *vis_out,
                           const void *kern, int Q, int separable, int nplanes,
                           const int64_t *plane_off, const int32_t *plane_s,
                           int max_support, int plane_mode, const int32_t *vplane,
                           double wstep, int wbin0);

/* Plans (no reference counterpart: crocodile bins nothing, each convgrid
 * call walks its visibilities one by one).  A plan holds the binning of one
 * set of coordinates for one kernel-bank layout, for gridding and
 * degridding the same visibilities repeatedly, as imaging followed by
 * prediction in an iterative reconstruction does.  croc_plan_bin takes the
 * arguments of croc_convgrid without grid / vis / kern, validates every
 * coordinate with the same error codes and synchronises; after a
 * successful binning croc_convgrid_plan and croc_convdegrid_plan only
 * enqueue work on the context stream and report no per-visibility errors.
 * uvw may change after binning; plane_off and plane_s (device arrays) must
 * stay valid and unchanged while the plan is used, and kern must keep the
 * bank layout they describe (its values may change).  Calls with one plan
 * must not overlap on two streams. */
typedef struct croc_plan croc_plan;
int croc_plan_create(croc_ctx *ctx, croc_plan **out);
int croc_plan_destroy(croc_ctx *ctx, croc_plan *plan);
int croc_plan_bin(croc_ctx *ctx, croc_plan *plan, int N, const double *uvw,
                  int64_t nv, int64_t stride, double lam, int Q, int separable,
                  int nplanes, const int64_t *plane_off, const int32_t *plane_s,
                  int max_support, int plane_mode, const int32_t *vplane,
                  double wstep, int wbin0);
int croc_convgrid_plan(croc_ctx *ctx, croc_plan *plan, void *grid,
                       const void *vis, const void *kern);
int croc_convdegrid_plan(croc_ctx *ctx, croc_plan *plan, const void *grid,
                         void *vis_out, const void *kern);

/* convdegrid without device allocation.  No counterpart in the reference
 * (whose convdegrid allocates NumPy temporaries freely): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same arguments, results
 * and errors as croc_convdegrid (crocodile/synthesis.py:484-503), through
 * the same binning and tile kernels; every scratch buffer (records, bin
 * counters, scan space, work items, status) is carved from one static
 * device array of CROC_NOALLOC_SCRATCH_MIB MiB per GPU, held from library
 * load until the process ends.  Limits: at most CROC_NOALLOC_MAX_NV
 * visibilities, and the binning of the grid's tiles must fit the array
 * (grids up to 8192^2 do at supports >= 16); beyond either, CROC_EINVAL
 * with a message, outputs untouched.  All calls on one GPU share the
 * array: calls that overlap on one GPU (two streams or threads) race and
 * must be serialised by the caller.  Allocates nothing, ever. */
#define CROC_NOALLOC_SCRATCH_MIB 32
#define CROC_NOALLOC_MAX_NV 524288
int croc_convdegrid_noalloc(croc_ctx *ctx, const void *grid, int N,
                            const double *uvw, int64_t nv, int64_t stride,
                            double lam, void *vis_out, const void *kern, int Q,
                            int separable, int nplanes, const int64_t *plane_off,
                            const int32_t *plane_s, int max_support,
                            int plane_mode, const int32_t *vplane, double wstep,
                            int wbin0);

/* Non-square grids [Nh][Nw] (synthesis.py's convgrid / convdegrid / grid /
 * degrid on any 2D array): each visibility's cell and sub-offset by the
 * reference's rule for that shape -- frac_coords (:438-449, x from Nw, y
 * from Nh) when nearest == 0, with the window rows y - gh//2 ..
 * y + (gh+1)//2 - 1 and columns x - gw//2 .. checked inside the grid
 * (CROC_EBOUNDS); the nearest cell of grid / degrid (:389-416: N = Nh for
 * both axes, NumPy's negative-index wrap) when nearest == 1 -- written as
 * the coordinate p_out (double [nv][2]) of a square [Nb][Nb] grid (Nb even)
 * that holds the [Nh][Nw] grid at rows oy.., columns ox..: the square-grid
 * entry points given p_out and lam = 1 then land every tap on the same cell
 * of that embedding (exactly: the cell and sub-offset are recovered bit for
 * bit).  No separate counterpart in the reference, whose NumPy functions
 * take any shape. */
int croc_remap_coords(croc_ctx *ctx, const double *uvw, int64_t nv, int64_t stride,
                      double lam, int Nh, int Nw, int Q, int gh, int gw, int nearest,
                      int Nb, int oy, int ox, double *p_out);

/* grid / degrid (nearest cell): replace crocodile/synthesis.py:389-401 and
 * :404-416, cell N//2 + floor(0.5 + N*p) (round-half-up). */
int croc_grid_nearest(croc_ctx *ctx, void *grid, int N, const double *uvw,
                      int64_t nv, int64_t stride, double lam, const void *vis);
int croc_degrid_nearest(croc_ctx *ctx, const void *grid, int N,
                        const double *uvw, int64_t nv, int64_t stride,
                        double lam, void *vis_out);

/* Full-grid w-stacking, the per-plane loop of
 * examples/notebooks/wstacking.ipynb cell 10 (JSON lines 191-240:
 * simple_imaging per w-plane, ifft, divide by w_kernel_function(l, m, mean w),
 * sum), split into two batched steps around a batched inverse FFT:
 *
 * croc_wstack_grid adds each visibility whose plane
 * t = round-half-even(w / wstep) - plane0 has slot_of[t] >= 0 to
 * stack[slot_of[t]] ([P][N][N] cf32) at its nearest cell (as
 * croc_grid_nearest; stride >= 3).  A plane outside [0, ntab) is
 * CROC_EPLANE.  With centred (even N only) each value is multiplied by
 * (-1)^(x+y), the input half of the centred transform's shifts.
 *
 * croc_wstack_add adds sum_p stack[p] * conj(w_kernel_function(l, m, midw[p]))
 * (midw: device [P] FP64; l, m = theta * coordinates2(N)) to image ([N][N]
 * cf64), times (-1)^(x+y) when centred (the output half of the shifts).
 * CROC_EDOMAIN when the image corner lies outside the unit circle
 * (synthesis.py:290). */
int croc_wstack_grid(croc_ctx *ctx, void *stack, int P, int N, const double *uvw,
                     int64_t nv, int64_t stride, double lam, double wstep,
                     int64_t plane0, int ntab, const int32_t *slot_of, const void *vis,
                     int centred);
int croc_wstack_add(croc_ctx *ctx, void *image, const void *stack, int P, int N,
                    double theta, const double *midw, int centred);

/* AW-towers imaging, examples/notebooks/awtowers.ipynb cells 17
 * (pregrid_tower_plane), 19 (grid_tower_plane: A-screen ratio per antenna
 * pair) and 22 (grid_tower, grid_towers), as three batched steps around
 * batched FFTs.  Visibilities are sorted by group = (chunk, w-plane, antenna
 * pair); gid[v] is visibility v's group, rel its [nv][3] coordinates
 * relative to its group's (u_mid, v_mid, wp * wstep).
 *
 * croc_awt_pregrid grids visibilities [v0, v1) onto stack[gid - g0]
 * ([G][M][M] cf32, M = 2 * subgrid_size) with the separable cf32 table gcf
 * [Q][K] at p = rel / lam_sub (synthesis.py:451-481's convgrid, frac_coord
 * on the M grid), each tap times (-1)^(x+y).  check_only: validate only
 * (CROC_EDOMAIN for p outside [-0.5, 0.5), CROC_EBOUNDS for a window that
 * leaves the M grid).
 *
 * croc_awt_apply adds to bins[c0 + k] ([C][S][S] cf32), k < nc, the sum
 * over groups g in [start[k], start[k+1]) of the cut-out middle pixel of
 * the (inverse transformed) stack[g] times ascreen[ga2[g]] / ascreen[ga1[g]]
 * ([nant][S][S] cf32) times wkern^-gwp[g] (wkern = 2^lmag exp(2 pi i turns)
 * per pixel), all times win ([S][S]).
 *
 * croc_awt_place adds (-1)^(x+y) bins[c] (forward transformed) into grid
 * ([N][N] cf32) at corner[c] = (x0, y0), clipped at the grid's edges. */
int croc_awt_pregrid(croc_ctx *ctx, void *stack, int M, const double *rel, int64_t v0,
                     int64_t v1, double lam_sub, int Q, int K, const void *gcf,
                     const void *vis, const int32_t *gid, int32_t g0, int check_only);
int croc_awt_apply(croc_ctx *ctx, void *bins, int c0, int nc, const void *stack, int S,
                   const int32_t *start, const int32_t *gwp, const int32_t *ga1,
                   const int32_t *ga2, const void *ascreen, const float *win,
                   const double *turns, const float *lmag);
int croc_awt_place(croc_ctx *ctx, void *grid, int N, const void *bins, int nc, int S,
                   const int32_t *corner);

/* w-kernel bank generation: replaces crocodile/synthesis.py:357-377
 * (w_kernel) = :250-270 (kernel_coordinates) + :273-293
 * (w_kernel_function, with 1/k when invert) + :330-354 (kernel_oversample)
 * for nw planes in one launch.  w: host array [nw]; S: host array [nw]
 * (NpixKern per plane); T: host 2x2 matrix row-major or NULL.  Output
 * written at out + plane_off[k] (host offsets, complex elements), layout
 * [Q][Q][S][S]; out_double selects cf64 (1) or cf32 (0) output.
 * aterm: NULL, or a device array double [npix_ff][npix_ff] multiplying
 * every plane's far field (an anti-aliasing taper times an A-term), which
 * makes AW kernels: the image-plane product whose uv-plane form
 * aw_kernel_fn (synthesis.py:701-734) builds by convolution.
 * Computed in FP64 (phase and the oversampled DFT). */
int croc_w_kernels(croc_ctx *ctx, double theta, const double *w, int nw,
                   int npix_ff, const int32_t *S, int Q, int invert,
                   double dl, double dm, const double *T,
                   const int64_t *plane_off, void *out, int out_double,
                   const double *aterm);

/* kernel_oversample: replaces crocodile/synthesis.py:330-354 for a device
 * far field ff (cf64, [n][n] when ndim == 2, [n] when ndim == 1):
 * pad to n*Q, centred FFT, extract -- computed as the equivalent direct
 * zoom DFT in FP64.  out cf64 [Q][Q][s][s] or [Q][s]. */
int croc_kernel_oversample(croc_ctx *ctx, const void *ff, int n, int ndim,
                           int Q, int s, void *out);

/* Batched centred 2D DFT of n x n complex64 sub-grids (n = 128):
 * inverse = 0: numpy.fft.fftshift(fft2(ifftshift(x))) -- crocodile fft,
 * synthesis.py:83-93; inverse = 1: the ifft counterpart with 1/n^2,
 * synthesis.py:96-106.  Matrix-core (fp16x3 split) evaluation; in and out
 * may alias.  The sub-grid transform of the w-towers drivers. */
int croc_dft2d_batched(croc_ctx *ctx, const void *in, void *out, int n,
                       int64_t batch, int inverse);

/* The same centred transforms as croc_dft2d_batched (n = 128 only), as a
 * radix 16 x 8 FFT in FP32 on the vector ALUs with the sub-grid held in LDS:
 * ~30x fewer operations than the dense products, the transform the w-towers
 * drivers use. */
int croc_fft2d_batched(croc_ctx *ctx, const void *in, void *out, int n,
                       int64_t batch, int inverse);

/* W-towers imaging: replaces the w-stacking-by-sub-grid drivers of
 * examples/notebooks/wtowers.ipynb (imaging cell) and grid.c:482-710
 * (grid_wtowers).  Visibilities are binned into (chunk, w-plane) jobs, chunk
 * = subgrid - margin cells, w-plane wp = rint(w / wstep); each job is
 * gridded on a subgrid x subgrid sub-grid (kern == NULL: nearest cell as
 * simple_imaging; else the Q-oversampled kernel, separable [Q][S] or 2D
 * [Q][Q][S][S]), inverse-transformed, multiplied by
 * exp(-2 pi i wp wstep (1 - n)), summed per chunk, transformed back and
 * added into grid (cf32 [N][N], in place).  subgrid must be 128, margin
 * even, S <= margin <= subgrid/2.  uvw in wavelengths, stride >= 3.
 * kern may hold a w-kernel set (grid.c's w_kernel_data, grid.h:114-122,
 * used by w_project grid.c:277-305): nwk tables of the shape above, table k
 * for the residual w - wp wstep = wk_min + k wk_step, chosen per visibility
 * as floor((w - wp wstep - wk_min) / wk_step + 0.5); the set must cover
 * [-wstep/2, wstep/2] (grid.c:493-496).  nwk = 1 uses the one table for
 * every visibility (the notebook's form; wk_min, wk_step ignored). */
int croc_wtowers_grid(croc_ctx *ctx, void *grid, int N, const double *uvw,
                      int64_t nv, int64_t stride, double theta, const void *vis,
                      double wstep, int subgrid, int margin, const void *kern,
                      int Q, int S, int separable, int nwk, double wk_min,
                      double wk_step);

/* W-towers predict: replaces examples/notebooks/wtowers-predict.ipynb
 * (predict cell): per chunk, inverse transform of the grid window; per
 * w-plane, multiply by exp(+2 pi i wp wstep (1 - n)), transform, degrid.
 * vis_out cf32 [nv]. */
int croc_wtowers_degrid(croc_ctx *ctx, const void *grid, int N,
                        const double *uvw, int64_t nv, int64_t stride,
                        double theta, void *vis_out, double wstep, int subgrid,
                        int margin, const void *kern, int Q, int S,
                        int separable, int nwk, double wk_min, double wk_step);

/* W-towers plans: the (chunk, w-plane) binning of one set of coordinates,
 * reused by any number of croc_wtowers_grid_plan / croc_wtowers_degrid_plan
 * calls (the one-shot calls above bin on every call).  No counterpart in the
 * reference, whose notebooks bin inside each imaging / predict cell: an
 * extension for callers that image and predict the same visibilities
 * repeatedly.  croc_wtowers_plan_bin takes the coordinate and geometry
 * arguments of the one-shot calls, the kernel kind has_kernel (0: none, the
 * one-shot calls' kern == NULL; 1: separable [Q][S] tables; 2:
 * two-dimensional [Q][Q][S][S] tables) and the kernel shape Q, S; it
 * validates every coordinate and makes the job schedules of that kind.  The
 * plan calls then take kernel values of that kind and shape (kern NULL
 * exactly when has_kernel was 0, separable exactly when it was 1).  uvw may
 * change after binning.  Calls with one plan must not overlap. */
typedef struct croc_tw_plan croc_tw_plan;
int croc_wtowers_plan_create(croc_ctx *ctx, croc_tw_plan **out);
int croc_wtowers_plan_destroy(croc_ctx *ctx, croc_tw_plan *plan);
int croc_wtowers_plan_bin(croc_ctx *ctx, croc_tw_plan *plan, int N,
                          const double *uvw, int64_t nv, int64_t stride,
                          double theta, double wstep, int subgrid, int margin,
                          int has_kernel, int Q, int S, int nwk, double wk_min,
                          double wk_step);
int croc_wtowers_grid_plan(croc_ctx *ctx, croc_tw_plan *plan, void *grid,
                           const void *vis, const void *kern, int separable);
int croc_wtowers_degrid_plan(croc_ctx *ctx, croc_tw_plan *plan,
                             const void *grid, void *vis_out,
                             const void *kern, int separable);

/* 2D convolution, scipy.signal.convolve2d(a, b[n], mode='same') for nb
 * kernels b sharing one a (cf64, [ah][aw] and [nb][bh][bw]; out
 * [nb][ah][aw]), FP64 direct sums.  The building block of aw_kernel_fn,
 * crocodile/synthesis.py:701-734 (:724 A1*A2, :729-731 per w sub-kernel). */
int croc_convolve_same(croc_ctx *ctx, const void *a, int ah, int aw,
                       const void *b, int bh, int bw, int64_t nb, void *out);

/* make_grid_hermitian: replaces crocodile/synthesis.py:855-880 (in place,
 * [N][N], cf64 when is_double else cf32).
 * Native counterpart: examples/grid/grid.c:794-817. */
int croc_make_hermitian(croc_ctx *ctx, void *grid, int N, int is_double);

/* ------------------------------------------------------------------------
 * Baseline / time / frequency layout: the native counterparts of
 * examples/grid/grid.h:14-40 (struct bl_data, struct vis_data) and the
 * functions of grid.c / streamer.c over it.
 *
 * A set of baselines in device memory, packed: baseline b has
 * time_count[b] dumps and freq_count[b] channels; its uvw (metres) are rows
 * uvw_first[b] .. + time_count[b] - 1 of uvw_m [uvw_rows][3]; its channel
 * frequencies (Hz) freq[freq_first[b] ..]; its visibilities
 * vis[vis_first[b] + time * freq_count[b] + freq] (cf32 [nvis]).  A
 * visibility's coordinate in wavelengths is uvw_m * f / c (grid.h:53-60,
 * c = 299792458 m/s, FP64, rounded as written).  The struct itself is host
 * memory; every array it points to is device memory.  Ranges are checked
 * on the device (CROC_EINVAL when a baseline's arrays or a time / frequency
 * window leave their arrays).
 *
 * Coordinates follow grid.c's frac_coord (:35-53): ox = rint((N/2 - x) Q)
 * (round-half-even: the fesetround(3) call there is rejected by glibc),
 * cell N - ox/Q, sub-offset ox % Q, where x = theta * u.  A kernel window
 * that leaves the grid is CROC_EBOUNDS (grid.c reads or writes the
 * neighbouring row, or outside the array, there).  Separable kernels here
 * are real FP64 tables [Q][S] (grid.h:101-106 sep_kernel_data; sub-pixel
 * offset outer).  All of these validate every input they will read before
 * writing anything (one stream synchronisation) and leave their outputs
 * unmodified on error; the degridding functions then return without waiting
 * for the device unless flops is requested.  flops (host, may be NULL)
 * receives the reference's flop count of the call (4 (1 + S) S per
 * degridded visibility, grid.c:222; 8 S^2 per gridded one, :304).
 * ------------------------------------------------------------------------ */
typedef struct croc_bl_set {
  int64_t nbl;                /* baselines (vis_data.bl_count)            */
  int64_t uvw_rows;           /* rows of uvw_m                            */
  int64_t freq_len;           /* entries of freq                          */
  int64_t nvis;               /* entries of vis                           */
  const int32_t *time_count;  /* [nbl]                                    */
  const int32_t *freq_count;  /* [nbl]                                    */
  const int64_t *uvw_first;   /* [nbl]                                    */
  const int64_t *freq_first;  /* [nbl]                                    */
  const int64_t *vis_first;   /* [nbl]                                    */
  const double *uvw_m;        /* [uvw_rows][3], metres                    */
  const double *freq;         /* [freq_len], Hz                           */
  void *vis;                  /* cf32 [nvis]                              */
} croc_bl_set;

/* weight (grid.c:119-141): counts the visibilities of every nearest cell
 * (floor(theta * u + .5) + N/2, grid.c:22-32) into wgrid (uint32 [N][N],
 * overwritten), then divides each visibility by its cell's count. */
int croc_bl_weight(croc_ctx *ctx, uint32_t *wgrid, int N, double theta,
                   const croc_bl_set *vis);

/* grid_simple (grid.c:161-177): uvgrid (cf32 [N][N]) += each visibility at
 * its nearest cell, accumulated exactly (int64 fixed point) in any order.
 * degrid_simple (grid.c:143-159): each visibility = uvgrid at its cell. */
int croc_bl_grid_simple(croc_ctx *ctx, void *uvgrid, int N, double theta,
                        const croc_bl_set *vis, uint64_t *flops);
int croc_bl_degrid_simple(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                          const croc_bl_set *vis, uint64_t *flops);

/* degrid_conv_uv (grid.c:179-232) for n coordinates uv (double [n][2],
 * wavelengths): vis_out[i] (cf32) = sum_y k[fy][y] sum_x k[fx][x]
 * uvgrid[iy - S/2 + y][ix - S/2 + x], kernel k real double [Q][S]. */
int croc_degrid_conv_uv(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                        const double *uv, int64_t n, const double *kern, int S,
                        int Q, void *vis_out, uint64_t *flops);

/* degrid_conv_bl (grid.c:234-259), for every baseline of the set: the
 * visibilities (time, freq) of [time0, time1) x [freq0, freq1) whose
 * (u, v) lie in [min_u, max_u) x [min_v, max_v) are degridded at
 * (u - d_u, v - d_v) into vis[vis_first[b] + (time - time0) * (freq1 -
 * freq0) + freq - freq0] (the reference's chunk-local index); the others
 * are left untouched. */
int croc_degrid_conv_bl(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                        double d_u, double d_v, double min_u, double max_u,
                        double min_v, double max_v, const croc_bl_set *bl,
                        int time0, int time1, int freq0, int freq1,
                        const double *kern, int S, int Q, uint64_t *flops);

/* degrid_conv (grid.c:261-275): degrid_conv_bl of every baseline over all
 * its dumps and channels with the grid's own window
 * [d - N/theta/2, d + N/theta/2). */
int croc_degrid_conv(croc_ctx *ctx, const void *uvgrid, int N, double theta,
                     double d_u, double d_v, const croc_bl_set *vis,
                     const double *kern, int S, int Q, uint64_t *flops);

/* grid_wprojection (grid.c:307-322 over w_project :277-305): every
 * visibility adds v * conj(k) over its S x S window, k the table
 * (yf, xf) of w-plane floor((w - w_min) / w_step + .5) of wkern (cf32
 * [plane_count][Q][Q][S][S], w_kernel_data grid.h:114-122), into uvgrid
 * (cf32 [N][N], in place), accumulated exactly (int64 fixed point).  A
 * plane outside [0, plane_count) is CROC_EPLANE. */
int croc_grid_wprojection(croc_ctx *ctx, void *uvgrid, int N, double theta,
                          const croc_bl_set *vis, const void *wkern,
                          int plane_count, double w_min, double w_step, int S,
                          int Q, uint64_t *flops);

/* streamer_degrid_worker (streamer.c:366-486) for nchunks chunks of one
 * subgrid (cf32 [subgrid_size][subgrid_size], centred at (mid_u, mid_v)):
 * chunk c = chunks[5c .. 5c + 4] = (baseline, it0, it1, if0, if1), its
 * visibility (time, freq) written to vis_queue[c * time_chunk * freq_chunk
 * + (time - it0) * freq_chunk + freq - if0] (cf32): degridded at
 * (u - mid_u, v - mid_v) when u >= max(min_u, 0), u < max_u, v in
 * [min_v, max_v); else the conjugate of the degrid at (-u - mid_u,
 * -v - mid_v) when -u >= max(min_u, DBL_MIN) and (-u, -v) lies in the
 * window; else 0 (streamer.c:418-438).  Slot entries beyond a chunk's
 * extent are not written.  ndegrid (device int64 [nchunks], may be NULL)
 * receives each chunk's count of degridded visibilities (the reference's
 * "flops == 0" test for an empty chunk). */
int croc_streamer_degrid(croc_ctx *ctx, const void *subgrid, int subgrid_size,
                         double theta, double mid_u, double mid_v, double min_u,
                         double max_u, double min_v, double max_v,
                         const croc_bl_set *bl, const int32_t *chunks,
                         int64_t nchunks, int time_chunk, int freq_chunk,
                         const double *kern, int S, int Q, void *vis_queue,
                         int64_t *ndegrid, uint64_t *flops);

/* croc_streamer_degrid for many subgrids in one launch.  No counterpart in
 * the reference, whose streamer degrids one subgrid per call as the
 * recombination delivers it: an extension for callers that hold a work
 * assignment's subgrids in device memory together.  subgrids cf32
 * [nsub][subgrid_size][subgrid_size]; windows host double [nsub][6] =
 * (mid_u, mid_v, min_u, max_u, min_v, max_v) per subgrid; chunks int32
 * [nchunks][6] = (subgrid, baseline, it0, it1, if0, if1); the queue slots,
 * ndegrid and flops as for croc_streamer_degrid (slot c for chunk c). */
int croc_streamer_degrid_batch(croc_ctx *ctx, const void *subgrids, int nsub,
                               int subgrid_size, double theta, const double *windows,
                               const croc_bl_set *bl, const int32_t *chunks,
                               int64_t nchunks, int time_chunk, int freq_chunk,
                               const double *kern, int S, int Q, void *vis_queue,
                               int64_t *ndegrid, uint64_t *flops);

#ifdef __cplusplus
}
#endif
#endif /* CROCGPU_H */
