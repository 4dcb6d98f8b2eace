"""FP64 NumPy restatement of crocodile's synthesis module (the oracle).

TEST INFRASTRUCTURE: imported only by tests, smoke() and bench.py's
cpu_baseline leg.  Each function cites the reference line range it follows
(``synthesis.py`` = crocodile/synthesis.py).  Behaviour, argument meaning,
rounding conventions and error cases follow the reference; where the
reference is broken as written (``w_slice_*`` and ``w_cache_predict`` raise
NameError) the intended semantics are implemented and the deviation is noted
in the docstring.

Pinning (what checks each function independently of this file's own
formulas) is listed in the tests that exercise it:

* coordinates / pad / extract: closed forms from tests/test_synthesis.py:15-44
* frac_coord(s): known-answer vectors of examples/grid/test_recombine.c:380-401
  and the tie property of tests/test_synthesis.py:393-409
* kernel_oversample / w_kernel: brute-force direct DFT, the oversampling
  consistency test (test_synthesis.py:77-100), the perfect-sampling
  degrid-equals-DFT test (test_synthesis.py:242-281)
* convgrid / convdegrid: the exact delta-image test (test_synthesis.py:109-138),
  adjointness, brute-force dense summation on tiny grids
* anti_aliasing_function: scipy.special.pro_ang1 (library routine), shape and
  centre value (test_synthesis.py:46-50); PSWF values beyond that are
  "parity unpinned" but identical tables feed both oracle and GPU.
"""

import numpy
import scipy.special
import scipy.signal


# --------------------------------------------------------------------------
# L0: coordinates, FFT helpers, padding            (synthesis.py:38-197)
# --------------------------------------------------------------------------

def coordinateBounds(N):
    """Lowest / highest coordinate of an N-point axis, step 1/N, zero at
    index N//2.  synthesis.py:38-55."""
    if N % 2 == 0:
        return -0.5, 0.5 * (N - 2) / N
    return -0.5 * (N - 1) / N, 0.5 * (N - 1) / N


def coordinates(N):
    """1D axis ``(i - N//2)/N``, i in [0,N).  synthesis.py:57-63."""
    return (numpy.arange(N) - N // 2) / N


def coordinates2(N):
    """Pair (cx, cy) of 2D axes: cx varies along axis 1, cy along axis 0,
    both zero at index N//2.  synthesis.py:66-80."""
    c = coordinates(N)
    cx = numpy.broadcast_to(c[numpy.newaxis, :], (N, N)).copy()
    cy = numpy.broadcast_to(c[:, numpy.newaxis], (N, N)).copy()
    return numpy.array([cx, cy])


def fft(a):
    """Centred forward FFT fftshift(fft(ifftshift(a))), unnormalised.
    synthesis.py:83-93."""
    if a.ndim == 1:
        return numpy.fft.fftshift(numpy.fft.fft(numpy.fft.ifftshift(a)))
    if a.ndim == 2:
        return numpy.fft.fftshift(numpy.fft.fft2(numpy.fft.ifftshift(a)))
    raise AssertionError("Unsupported image shape for FFT!")


def ifft(a):
    """Centred inverse FFT (1/N^2 normalised).  synthesis.py:96-106."""
    if a.ndim == 1:
        return numpy.fft.fftshift(numpy.fft.ifft(numpy.fft.ifftshift(a)))
    if a.ndim == 2:
        return numpy.fft.fftshift(numpy.fft.ifft2(numpy.fft.ifftshift(a)))
    raise AssertionError("Unsupported grid shape for iFFT!")


def pad_mid(ff, N):
    """Zero-pad around index N//2 to size N (asymmetric split
    ``(N//2 - N0//2, (N+1)//2 - (N0+1)//2)``).  synthesis.py:109-129."""
    N0 = ff.shape[0]
    if N == N0:
        return ff
    assert N > N0
    lo, hi = N // 2 - N0 // 2, (N + 1) // 2 - (N0 + 1) // 2
    if ff.ndim == 2:
        assert N0 == ff.shape[1]
        return numpy.pad(ff, [(lo, hi), (lo, hi)], mode='constant')
    return numpy.pad(ff, [(lo, hi)], mode='constant')


def extract_mid(a, N):
    """Centre section of size N (inverse of pad_mid).  synthesis.py:131-158."""
    assert N <= a.shape[0]
    h = N // 2
    odd = N % 2
    c0 = a.shape[0] // 2
    if a.ndim == 2:
        assert N <= a.shape[1]
        c1 = a.shape[1] // 2
        return a[c0 - h:c0 + h + odd, c1 - h:c1 + h + odd]
    if a.ndim == 1:
        return a[c0 - h:c0 + h + odd]
    raise AssertionError("Unsupported grid shape for extract_mid!")


def extract_oversampled(a, Qpx, N):
    """Sub-kernel (yf, xf) = every Qpx-th sample starting at
    ``Na//2 - Qpx*(N//2) + f``.  Returns [Qpx,Qpx,N,N] (2D) or [Qpx,N] (1D).
    synthesis.py:160-197."""
    Na = a.shape[0]
    start = Na // 2 - Qpx * (N // 2)
    if a.ndim == 1:
        out = numpy.empty((Qpx, N), dtype=complex)
        for f in range(Qpx):
            out[f] = a[start + f: start + f + Qpx * N: Qpx]
        return out
    if a.ndim == 2:
        out = numpy.empty((Qpx, Qpx, N, N), dtype=complex)
        for fy in range(Qpx):
            for fx in range(Qpx):
                out[fy, fx] = a[start + fy: start + fy + Qpx * N: Qpx,
                                start + fx: start + fx + Qpx * N: Qpx]
        return out
    raise AssertionError("Unsupported grid shape for extract_oversampled!")


# --------------------------------------------------------------------------
# L1: kernels                                       (synthesis.py:199-387)
# --------------------------------------------------------------------------

def anti_aliasing_function(shape, m, c):
    """Prolate-spheroidal AA function; 2D is the outer product.
    synthesis.py:199-218."""
    if numpy.ndim(shape) == 0:
        mult = 2 - 1 / shape / 4
        return scipy.special.pro_ang1(m, m, c, mult * coordinates(shape))[0]
    return numpy.outer(anti_aliasing_function(shape[0], m, c),
                       anti_aliasing_function(shape[1], m, c))


def kernel_transform(dl, dm):
    """2x2 matrix keeping w-kernel coordinates steady under a (dl,dm) shift.
    synthesis.py:220-248."""
    dn = numpy.sqrt(1 - dl ** 2 - dm ** 2)
    f = (dn - 1) / (dl ** 2 + dm ** 2) if (dl != 0 or dm != 0) else 0
    return numpy.array([[dn - dm * dm * f, dl * dm * f],
                        [dl * dm * f, dn - dl * dl * f]])


def kernel_coordinates(N, theta, dl=0, dm=0, T=None):
    """(l, m) far-field coordinates, optionally transformed and shifted.
    synthesis.py:250-270."""
    l, m = coordinates2(N) * theta
    if T is not None:
        l, m = T[0, 0] * l + T[1, 0] * m, T[0, 1] * l + T[1, 1] * m
    return l + dl, m + dm


def w_kernel_function(l, m, w):
    """exp(2 pi i w (1 - sqrt(1 - l^2 - m^2))).  synthesis.py:273-293."""
    r2 = l ** 2 + m ** 2
    assert numpy.all(r2 < 1.0), "Error in image coordinate system"
    return numpy.exp(2j * numpy.pi * w * (1 - numpy.sqrt(1.0 - r2)))


def kernel_recentre(cp, theta, w, dl, dm):
    """Image-space phase ramp exp(-2 pi i w (dl l + dm m)).
    synthesis.py:296-311."""
    l, m = coordinates2(cp.shape[0]) * theta
    return cp * numpy.exp(-2j * numpy.pi * w * (dl * l + dm * m))


def visibility_recentre(uvw, dl, dm):
    """u -= w dl, v -= w dm.  synthesis.py:314-327."""
    out = numpy.array(uvw, dtype=float)
    out[:, 0] = uvw[:, 0] - uvw[:, 2] * dl
    out[:, 1] = uvw[:, 1] - uvw[:, 2] * dm
    return out


def kernel_oversample(ff, Qpx, s=None):
    """pad far field to N*Qpx, centred FFT, extract oversampled kernels.
    Result layout [yf, xf, y, x] (sub-pixel offsets outer).
    synthesis.py:330-354."""
    N = ff.shape[0]
    if s is None:
        s = N
    return extract_oversampled(fft(pad_mid(ff, N * Qpx)), Qpx, s)


def w_kernel(theta, w, NpixFF, NpixKern, Qpx, invert=True, **kwargs):
    """Oversampled w-kernel [Qpx,Qpx,NpixKern,NpixKern], unnormalised.
    synthesis.py:357-377."""
    assert NpixFF > NpixKern or (NpixFF == NpixKern and Qpx == 1)
    l, m = kernel_coordinates(NpixFF, theta, **kwargs)
    kern = w_kernel_function(l, m, w)
    if invert:
        kern = 1 / kern
    return kernel_oversample(kern, Qpx, NpixKern)


def w_conj_kernel_fn(kernel_fn):
    """Fold k(-w) = conj(k(w)).  synthesis.py:682-699."""
    def fn(theta, w, *args, **kw):
        if w < 0:
            return numpy.conj(kernel_fn(theta, -w, *args, **kw))
        return kernel_fn(theta, w, *args, **kw)
    return fn


def aw_kernel_fn(a_kernel_fn, w_kernel_fn=w_kernel):
    """AW kernel = conv2d(conv2d(a1, a2, same), wk[yf,xf], same) for every
    sub-kernel.  synthesis.py:701-734."""
    def fn(theta, w, a1, a2, *src):
        akern = scipy.signal.convolve2d(a_kernel_fn(theta, a1, *src),
                                        a_kernel_fn(theta, a2, *src),
                                        mode='same')
        wkern = w_kernel_fn(theta, w)
        return numpy.array([[scipy.signal.convolve2d(akern, wk, mode='same')
                             for wk in row] for row in wkern])
    return fn


# --------------------------------------------------------------------------
# L2: gridding core                                 (synthesis.py:389-503)
# --------------------------------------------------------------------------

def grid(a, p, v):
    """Nearest-cell gridding, cell N//2 + floor(0.5 + N p) (half-up).
    synthesis.py:389-401."""
    assert numpy.max(p) < 0.5
    N = a.shape[0]
    xy = N // 2 + numpy.floor(0.5 + N * p[:, 0:2]).astype(int)
    for (x, y), vv in zip(xy, v):
        a[y, x] += vv


def degrid(a, p):
    """Nearest-cell degridding.  synthesis.py:404-416."""
    assert numpy.max(p) < 0.5
    N = a.shape[0]
    xy = N // 2 + numpy.floor(0.5 + p[:, 0:2] * N).astype(int)
    return numpy.array([a[y, x] for x, y in xy])


def frac_coord(N, Qpx, p):
    """Whole cell ``ix`` and sub-offset ``xf`` with ``ix - xf/Qpx`` the
    nearest 1/Qpx approximation of ``N//2 + p N``.  numpy.around is
    half-to-even.  synthesis.py:419-435."""
    assert (p >= -0.5).all() and (p < 0.5).all()
    x = N // 2 + p * N
    fx = numpy.around(-x * Qpx).astype(int)
    return -(fx // Qpx), fx % Qpx


def frac_coords(shape, Qpx, p):
    """(x, xf, y, yf) for p[:, 0] (width) and p[:, 1] (height).
    synthesis.py:438-449."""
    h, w = shape
    x, xf = frac_coord(w, Qpx, p[:, 0])
    y, yf = frac_coord(h, Qpx, p[:, 1])
    return x, xf, y, yf


def make_gcf(src, gcf):
    """Kernel accessor.  Separable tables ([*src, Q, S]) give the outer
    product gcf[s,yf] x gcf[s,xf]; 2D tables ([*src, Q, Q, gh, gw]) give
    gcf[s, yf, xf].  Returned extents follow the reference order
    (Q, gw, gh) for 2D tables.  synthesis.py:451-461."""
    src = numpy.array(src)
    k = src.shape[1]
    if gcf.ndim - k == 2:
        Qpx, gw = gcf.shape[-2:]
        return Qpx, gw, gw, lambda s, xf, yf: numpy.outer(gcf[(*s, yf)], gcf[(*s, xf)])
    if gcf.ndim - k == 4:
        Qpx, gh, gw = gcf.shape[-3:]
        return Qpx, gw, gh, lambda s, xf, yf: gcf[(*s, yf, xf)]
    raise AssertionError("unsupported kernel table rank")


def convgrid(gcf, a, p, src, v):
    """a[y-gh//2 : y+(gh+1)//2, x-gw//2 : x+(gw+1)//2] += k(s,xf,yf) v per
    visibility; no conjugation, in place.  Extents unpacked as (Q, gh, gw)
    from make_gcf's (Q, gw, gh) exactly as the reference does.
    synthesis.py:464-481."""
    if src is None:
        src = numpy.zeros((p.shape[0], 0))
    Qpx, gh, gw, mk = make_gcf(src, gcf)
    x, xf, y, yf = frac_coords(a.shape, Qpx, p)
    for vv, s, xi, xfi, yi, yfi in zip(v, src, x, xf, y, yf):
        a[yi - gh // 2: yi + (gh + 1) // 2,
          xi - gw // 2: xi + (gw + 1) // 2] += mk(s, xfi, yfi) * vv


def convdegrid(gcf, a, p, src=None):
    """v_i = sum(a[window_i] * k(s,xf,yf)); no conjugation.
    synthesis.py:484-503."""
    if src is None:
        src = numpy.zeros((p.shape[0], 0))
    Qpx, gw, gh, mk = make_gcf(src, gcf)
    x, xf, y, yf = frac_coords(a.shape, Qpx, p)
    return numpy.array([
        numpy.sum(a[yi - gh // 2: yi + (gh + 1) // 2,
                    xi - gw // 2: xi + (gw + 1) // 2] * mk(s, xfi, yfi))
        for s, xi, xfi, yi, yfi in zip(src, x, xf, y, yf)])


# --------------------------------------------------------------------------
# L3: imaging drivers                               (synthesis.py:506-927)
# --------------------------------------------------------------------------

def sort_vis_w(p, v=None):
    """Order by w.  The reference uses numpy's default (unstable) argsort;
    a stable sort is used here so ties are deterministic.
    synthesis.py:506-515."""
    zs = numpy.argsort(p[:, 2], kind='stable')
    if v is not None:
        return p[zs], v[zs]
    return p[zs]


def slice_vis(step, uvw, v=None):
    """Chunks of ``step`` visibilities.  synthesis.py:518-532."""
    starts = range(0, len(uvw), step)
    if v is None:
        return [uvw[i:i + step] for i in starts]
    return [(uvw[i:i + step], v[i:i + step]) for i in starts]


def doweight(theta, lam, p, v):
    """Uniform re-weighting by visibility count per cell (Qpx=1 frac_coords;
    note the reference indexes the weight grid as gw[x, y]).
    synthesis.py:535-549."""
    N = int(round(theta * lam))
    assert N > 1
    gw = numpy.zeros([N, N])
    x, xf, y, yf = frac_coords(gw.shape, 1, p / lam)
    numpy.add.at(gw, (x, y), 1)
    return v / gw[x, y]


def simple_imaging(theta, lam, uvw, src, vis):
    """synthesis.py:552-562."""
    N = int(round(theta * lam))
    assert N > 1
    guv = numpy.zeros([N, N], dtype=complex)
    grid(guv, uvw / lam, vis)
    return guv


def simple_predict(theta, lam, uvw, src, guv):
    """synthesis.py:565-580."""
    N = int(round(theta * lam))
    assert N > 1
    return degrid(guv, uvw / lam)


def conv_imaging(theta, lam, uvw, src, vis, kv):
    """synthesis.py:583-598."""
    N = int(round(theta * lam))
    assert N > 1
    guv = numpy.zeros([N, N], dtype=complex)
    convgrid(kv, guv, uvw / lam, src, vis)
    return guv


def conv_predict(theta, lam, uvw, src, guv, kv):
    """synthesis.py:600-613."""
    if src is None:
        src = numpy.ndarray((len(uvw), 0))
    return convdegrid(kv, guv, uvw / lam, src)


def w_slice_imaging(theta, lam, uvw, src, vis, wstep=2000, kernel_fn=w_kernel,
                    **kwargs):
    """Sort by w, slices of ``wstep`` visibilities, kernel at the slice's
    mean w.  The reference passes undefined ``p, v`` (NameError at
    synthesis.py:638); the intended ``uvw, vis`` are used, and ``src`` is
    sorted and sliced with them.  synthesis.py:616-644."""
    N = int(round(theta * lam))
    assert N > 1
    if src is None:
        src = numpy.zeros((len(uvw), 0))
    order = numpy.argsort(uvw[:, 2], kind='stable')
    guv = numpy.zeros([N, N], dtype=complex)
    for i in range(0, len(uvw), wstep):
        sel = order[i:i + wstep]
        ps = uvw[sel]
        wg = kernel_fn(theta, numpy.mean(ps[:, 2]), **kwargs)
        convgrid(wg, guv, ps / lam, src[sel], vis[sel])
    return guv


def w_slice_predict(theta, lam, uvw, src, guv, wstep=2000, kernel_fn=w_kernel,
                    **kwargs):
    """Predict counterpart of w_slice_imaging; undoes the sort.  The
    reference uses undefined ``p`` (NameError at synthesis.py:672).
    synthesis.py:647-679."""
    N = int(round(theta * lam))
    assert N > 1
    if src is None:
        src = numpy.zeros((len(uvw), 0))
    order = numpy.argsort(uvw[:, 2], kind='stable')
    v = numpy.empty(len(uvw), dtype=complex)
    for i in range(0, len(uvw), wstep):
        sel = order[i:i + wstep]
        ps = uvw[sel]
        wg = kernel_fn(theta, numpy.mean(ps[:, 2]), **kwargs)
        v[sel] = convdegrid(wg, guv, ps / lam, src[sel])
    return v


def w_cache_imaging(theta, lam, uvw, src, vis, wstep=2000, kernel_cache=None,
                    kernel_fn=w_kernel, **kwargs):
    """Per visibility: kernel at wstep*round(w/wstep) (half-to-even),
    then a one-visibility convgrid.  ``kernel_cache`` replaces ``kernel_fn``
    when given (the reference's default cache is an LRU of size 1, which
    does not change results).  synthesis.py:757-796."""
    fn = kernel_cache if kernel_cache is not None else kernel_fn
    if src is None:
        src = numpy.zeros((len(uvw), 0))
    N = int(round(theta * lam))
    guv = numpy.zeros([N, N], dtype=complex)
    for p, s, v in zip(uvw, src, vis):
        wbin = wstep * numpy.round(p[2] / wstep)
        wg = fn(theta, wbin, *s, **kwargs)
        convgrid(wg, guv, numpy.array([p / lam]), [()], numpy.array([v]))
    return guv


def w_cache_predict(theta, lam, uvw, src, guv, wstep=2000, kernel_cache=None,
                    kernel_fn=w_kernel, **kwargs):
    """Predict counterpart of w_cache_imaging.  The reference uses undefined
    ``nv`` and ``ixs`` (NameError at synthesis.py:823/827); the intended
    per-visibility result is returned in input order.
    synthesis.py:799-828."""
    fn = kernel_cache if kernel_cache is not None else kernel_fn
    if src is None:
        src = numpy.zeros((len(uvw), 0))
    v = numpy.empty(len(uvw), dtype=complex)
    for i, (p, s) in enumerate(zip(uvw, src)):
        wbin = wstep * numpy.round(p[2] / wstep)
        wg = fn(theta, wbin, *s, **kwargs)
        v[i] = convdegrid(wg, guv, numpy.array([p / lam]))[0]
    return v


def mirror_uvw(uvw, vis):
    """Flip baselines with v < 0, conjugating their visibilities.
    synthesis.py:831-852."""
    neg = uvw[:, 1] < 0
    uvw = numpy.array(uvw, dtype=float, copy=True)
    vis = numpy.array(vis, copy=True)
    uvw[neg] = -uvw[neg]
    vis[neg] = numpy.conj(vis[neg])
    return uvw, vis


def make_grid_hermitian(guv):
    """Add the conjugated point-mirror about cell N//2 (even N: rows/cols
    1..N-1 only).  In place, returns guv.  synthesis.py:855-880."""
    if guv.shape[0] % 2 == 0:
        guv[1:, 1:] += numpy.conj(guv[:0:-1, :0:-1])
    else:
        guv += numpy.conj(guv[::-1, ::-1])
    return guv


def do_imaging(theta, lam, uvw, src, vis, imgfn, **kwargs):
    """Mirror, weight, image, hermitian, ifft; psf likewise; normalise.
    synthesis.py:883-911."""
    if src is None:
        src = numpy.ndarray((len(vis), 0))
    uvw, vis = mirror_uvw(uvw, vis)
    wt = doweight(theta, lam, uvw, numpy.ones(len(uvw)))
    cdrt = imgfn(theta, lam, uvw, src, wt * vis, **kwargs)
    drt = numpy.real(ifft(make_grid_hermitian(cdrt)))
    c = imgfn(theta, lam, uvw, src, wt, **kwargs)
    psf = numpy.real(ifft(make_grid_hermitian(c)))
    pmax = psf.max()
    assert pmax > 0.0
    return drt / pmax, psf / pmax, pmax


def do_predict(theta, lam, uvw, src, modelimage, predfn, **kwargs):
    """fft(model image) then predfn.  synthesis.py:914-927."""
    return predfn(theta, lam, uvw, src, fft(modelimage.astype(complex)), **kwargs)


# --------------------------------------------------------------------------
# Input synthesis (crocodile/simulate.py restated; off the hot path)
# --------------------------------------------------------------------------

def simulate_point(dist_uvw, l, m):
    """Unit point source at (l, m), phase-tracked: exp(-2 pi i uvw.(l, m,
    n-1)).  simulate.py:210-230."""
    s = numpy.array([l, m, numpy.sqrt(1 - l ** 2 - m ** 2) - 1.0])
    return numpy.exp(-2j * numpy.pi * numpy.dot(dist_uvw, s))


def xyz_to_uvw(xyz, ha, dec):
    """Rotate earth (x,y,z) to (u,v,w) for hour angle ha and declination
    dec.  simulate.py:74-100."""
    x, y, z = numpy.hsplit(xyz, 3)
    u = x * numpy.cos(ha) - y * numpy.sin(ha)
    v0 = x * numpy.sin(ha) + y * numpy.cos(ha)
    w = z * numpy.sin(dec) - v0 * numpy.cos(dec)
    v = z * numpy.cos(dec) + v0 * numpy.sin(dec)
    return numpy.hstack([u, v, w])


def visibility_shift(uvw, vis, dl, dm):
    """Image-space shift by (dl, dm): vis * exp(-2 pi i (u dl + v dm)).
    simulate.py:259-275."""
    return vis * numpy.exp(-2j * numpy.pi * (uvw[:, 0] * dl + uvw[:, 1] * dm))


def uvw_transform(uvw, T):
    """uv' = uv (T^T)^-1 so the image transforms by T; w kept.
    simulate.py:278-300."""
    uv1 = numpy.dot(uvw[:, 0:2], numpy.linalg.inv(numpy.transpose(T)))
    return numpy.hstack([uv1, uvw[:, 2:3]])
