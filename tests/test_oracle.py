"""Pinning of the CPU oracle (oracle/synthesis_ref.py, oracle/cgrid.c).

Each check is independent of the oracle's own formulas: golden vectors of the
reference's C unit test, the closed-form tests of the reference's
tests/test_synthesis.py (transcribed), brute-force Fourier sums written from
the mathematics, and invariants (adjointness, linearity, hermitian symmetry).
"""
import itertools
import json
import os

import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import cgrid
from oracle import synthesis_ref as R

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def pattern(N):
    return R.coordinates2(N)[1] + R.coordinates2(N)[0] * 1j


def uvw_grid(N, uw=0, vw=0):
    u, v = R.coordinates2(N)
    u = numpy.hstack(u)
    v = numpy.hstack(v)
    return numpy.transpose([u, v, uw * u + vw * v])


# ---------------------------------------------------------------- frac_coord

def test_frac_coord_t05_golden():
    """Known answers of examples/grid/test_recombine.c:378-412."""
    g = json.load(open(os.path.join(GOLDEN, 't05_frac_coord.json')))
    N, Q = g['grid_size'], g['oversample']
    p = numpy.array(g['u']) / N
    x, xf = R.frac_coord(N, Q, p)
    assert x.tolist() == g['x']
    assert xf.tolist() == g['fx']
    cx, cxf = cgrid.frac_coord(N, Q, p)
    assert cx.tolist() == g['x'] and cxf.tolist() == g['fx']


@pytest.mark.parametrize('Qpx', range(1, 6))
def test_frac_coord_ties(Qpx):
    """tests/test_synthesis.py:393-409: exact ties round half to even."""
    p = numpy.arange(-2 * Qpx, 2 * Qpx) / 4 / Qpx
    flx, fracx = R.frac_coord(2, Qpx, p)
    assert (fracx >= 0).all() and (fracx < Qpx).all()
    assert_allclose(flx - fracx / Qpx, numpy.around(numpy.arange(0, 4 * Qpx) / 2) / Qpx)
    cx, cxf = cgrid.frac_coord(2, Qpx, p)
    assert (cx == flx).all() and (cxf == fracx).all()


def test_frac_coord_bruteforce():
    """Nearest multiple of 1/Q by exhaustive search, on random coordinates."""
    rng = numpy.random.default_rng(5)
    for N, Q in [(16, 8), (17, 3), (64, 16)]:
        p = rng.uniform(-0.5, 0.5, 500)
        x, xf = R.frac_coord(N, Q, p)
        target = N // 2 + p * N
        cand = numpy.arange(-1, N * Q + 2) / Q
        best = cand[numpy.argmin(abs(cand[None, :] - target[:, None]), axis=1)]
        assert_allclose(x - xf / Q, best, atol=1e-12)
        assert ((xf >= 0) & (xf < Q)).all()


def test_frac_coord_domain():
    with pytest.raises(AssertionError):
        R.frac_coord(8, 2, numpy.array([0.5]))
    with pytest.raises(AssertionError):
        cgrid.frac_coord(8, 2, numpy.array([-0.51]))


# ------------------------------------------------------------- L0 helpers

def test_coordinates():
    """tests/test_synthesis.py:15-28."""
    for N in [4, 5, 6, 7, 8, 9, 1000, 1001, 1002, 1003]:
        low, high = R.coordinateBounds(N)
        c = R.coordinates(N)
        cx, cy = R.coordinates2(N)
        for a in (c, cx, cy):
            assert abs(a.min() - low) < 1e-12 and abs(a.max() - high) < 1e-12
        assert c[N // 2] == 0
        assert (cx[:, N // 2] == 0).all()
        assert (cy[N // 2, :] == 0).all()


def test_pad_extract():
    """tests/test_synthesis.py:33-44."""
    for N, N2 in [(1, 1), (1, 2), (2, 3), (3, 4), (2, 5), (4, 6)]:
        cs = 1 + pattern(N)
        cs_pad = R.pad_mid(cs, N2)
        cs2 = 1 + pattern(N2) * N2 / N
        equal = numpy.abs(cs_pad - cs2) < 1e-15
        zero = numpy.abs(cs_pad) < 1e-15
        assert numpy.all(equal + zero)
        assert_allclose(R.extract_mid(cs_pad, N), cs)


def test_anti_aliasing():
    """tests/test_synthesis.py:46-50."""
    for shape in [(4, 4), (5, 5), (4, 6), (7, 3)]:
        aaf = R.anti_aliasing_function(shape, 0, 10)
        assert aaf.shape == shape
        assert abs(aaf[shape[0] // 2, shape[1] // 2] - 1) < 1e-7


# ------------------------------------------------------------------ kernels

def test_w_kernel_function():
    """tests/test_synthesis.py:52-58."""
    l, m = R.kernel_coordinates(5, 0.1)
    assert_allclose(R.w_kernel_function(l, m, 0), 1)
    assert abs(R.w_kernel_function(l, m, 100)[2, 2] - 1) < 1e-7
    l, m = R.kernel_coordinates(10, 0.1)
    assert abs(R.w_kernel_function(l, m, 100)[5, 5] - 1) < 1e-7
    assert abs(R.w_kernel_function(l, m, 1000)[5, 5] - 1) < 1e-7


def test_w_kernel_function_closed_form():
    """Phase 2 pi w (1 - n) from the stable form r^2 / (1 + n)."""
    l, m = R.kernel_coordinates(33, 0.3, dl=0.01, dm=-0.02)
    w = 1234.5
    r2 = l ** 2 + m ** 2
    ph = 2 * numpy.pi * w * r2 / (1 + numpy.sqrt(1 - r2))
    assert_allclose(R.w_kernel_function(l, m, w), numpy.exp(1j * ph), atol=1e-10)


def test_w_kernel_function_recentre():
    """tests/test_synthesis.py:60-75."""
    for lam in [4, 1000]:
        for N in range(2, 6):
            for w in [0.1, 1, 10]:
                for dl, dm in [(1, 1), (2, 0), (1, -3)]:
                    l, m = R.kernel_coordinates(N, N / lam)
                    cp = R.w_kernel_function(l, m, w)
                    kern = R.ifft(cp)
                    kerns = R.ifft(R.kernel_recentre(cp, N / lam, w, dl * lam / N / w,
                                                     dm * lam / N / w))
                    assert_allclose(numpy.roll(numpy.roll(kern, dl, axis=1), dm, axis=0),
                                    kerns, atol=1e-15)


def test_kernel_oversampled_subgrid():
    """tests/test_synthesis.py:77-91."""
    for N in range(3, 30):
        pat = pattern(N)
        kern = R.kernel_oversample(pat, 1, N - 2)
        kern2 = R.kernel_oversample(pat, 2, N - 2)
        assert_allclose(kern[0, 0], kern2[0, 0], atol=1e-13)
        kern3 = R.kernel_oversample(pat, 3, N - 2)
        assert_allclose(kern[0, 0], kern3[0, 0], atol=1e-13)
        kern4 = R.kernel_oversample(pat, 4, N - 2)
        for ux, uy in itertools.product(range(2), range(2)):
            assert_allclose(kern2[uy, ux], kern4[2 * uy, 2 * ux], atol=1e-13)
        kern8 = R.kernel_oversample(pat, 8, N - 2)
        for ux, uy in itertools.product(range(3), range(3)):
            assert_allclose(kern4[uy, ux], kern8[2 * uy, 2 * ux], atol=1e-13)


def test_kernel_scale():
    """tests/test_synthesis.py:93-100."""
    N = 10
    wff = numpy.zeros((N, N))
    wff[N // 2, N // 2] = 1
    assert_allclose(R.kernel_oversample(wff, 1, N),
                    R.kernel_oversample(R.pad_mid(wff, N * 2), 1, N))


def direct_oversampled(ff, Q, S):
    """Brute-force Fourier sum: K[yf,xf,y,x] = sum_{m,l} ff[m,l]
    exp(-2 pi i (u l + v m)) with the kernel sample at u = (x - S//2 + xf/Q)
    cells and far-field pixel l = (l_idx - F//2)/F of the field."""
    F = ff.shape[0]
    n = numpy.arange(F) - F // 2
    out = numpy.zeros((Q, Q, S, S), dtype=complex)
    for yf in range(Q):
        for xf in range(Q):
            ku = (numpy.arange(S) - S // 2) + xf / Q
            kv = (numpy.arange(S) - S // 2) + yf / Q
            Ex = numpy.exp(-2j * numpy.pi * numpy.outer(ku, n) / F)
            Ey = numpy.exp(-2j * numpy.pi * numpy.outer(kv, n) / F)
            out[yf, xf] = Ey @ ff @ Ex.T
    return out


@pytest.mark.parametrize('F,Q,S', [(8, 4, 6), (9, 3, 7), (16, 8, 12), (7, 1, 7)])
def test_kernel_oversample_bruteforce(F, Q, S):
    rng = numpy.random.default_rng(F * 100 + Q)
    ff = rng.standard_normal((F, F)) + 1j * rng.standard_normal((F, F))
    assert_allclose(R.kernel_oversample(ff, Q, S), direct_oversampled(ff, Q, S),
                    atol=1e-11 * F * F)


def test_w_kernel_sign_convention():
    """A w-kernel made with invert=False and used for degridding reproduces
    the phase-tracked DFT of a point source (physics: V = sum I G_w e^{...});
    tests/test_synthesis.py:242-281 in its perfect-sampling form."""
    lam = 1000
    for uw, vw in [(.5, 0), (0, .5), (-1, 0), (0, -1)]:
        for N in range(1, 6):
            uvw_all = uvw_grid(N, uw, vw) * lam
            slices = R.slice_vis(N, R.sort_vis_w(uvw_all))
            gcfs = [(R.w_kernel(N / lam, numpy.mean(u[:, 2]), N, N, 1) / N ** 2,
                     R.w_kernel(N / lam, numpy.mean(u[:, 2]), N, N, 1, invert=False) / N ** 2)
                    for u in slices]
            for x, y in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
                img_ref = numpy.zeros((N, N))
                img_ref[N // 2 + y, N // 2 + x] = 1
                a_ref = numpy.fft.ifftshift(R.fft(img_ref))
                a_ref = numpy.vstack(2 * [numpy.hstack(2 * [a_ref])])
                a = numpy.zeros((2 * N, 2 * N), dtype=complex)
                for u, (gcf, gcf_p) in zip(slices, gcfs):
                    vis = R.simulate_point(u, x / lam, y / lam)
                    assert_allclose(vis, R.convdegrid(gcf_p, a_ref, u / lam / 2))
                    R.convgrid(gcf, a, u / lam / 2, None, vis)
                a = numpy.fft.fftshift(a[:N, :N] + a[:N, N:] + a[N:, :N] + a[N:, N:])
                img = numpy.real(R.ifft(a))
                assert_allclose(img[N // 2 + y, N // 2 + x], 1)
                assert_allclose(img, img_ref, atol=2e-3)


# ------------------------------------------------------------ gridding core

def test_grid_degrid_exact():
    """tests/test_synthesis.py:109-138: points on grid nodes give an exact
    delta image and an exact round trip."""
    for lam in [6, 1e15]:
        gcf = numpy.conj(R.w_kernel(1 / lam, 0, 1, 1, 1))
        for N in range(2, 6):
            uvw = uvw_grid(N) * lam
            for x, y in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
                vis = R.simulate_point(uvw, x / lam, y / lam)
                a = numpy.zeros((N, N), dtype=complex)
                R.grid(a, uvw / lam, vis)
                R.convgrid(gcf, a, uvw / lam, None, vis)
                a /= 2
                img = numpy.real(R.ifft(a))
                assert abs(img[N // 2 + y, N // 2 + x] - 1) < 1e-7
                img[N // 2 + y, N // 2 + x] = 0
                assert_allclose(img, 0, atol=1e-14)
                assert_allclose(vis, R.degrid(a, uvw / lam))
                assert_allclose(vis, R.convdegrid(gcf, a, uvw / lam))


def random_case(rng, N, S, Q, nv, planes=1, separable=False):
    if separable:
        tables = [rng.standard_normal((Q, S)) + 1j * rng.standard_normal((Q, S))
                  for _ in range(planes)]
    else:
        tables = [rng.standard_normal((Q, Q, S, S)) + 1j * rng.standard_normal((Q, Q, S, S))
                  for _ in range(planes)]
    lim = 0.5 - (S + 2) / N
    p = numpy.stack([rng.uniform(-lim, lim, nv), rng.uniform(-lim, lim, nv),
                     rng.uniform(-1, 1, nv)], axis=1)
    v = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    return tables, p, v


@pytest.mark.parametrize('separable', [False, True])
def test_grid_degrid_adjoint(separable):
    """convdegrid is the transpose of convgrid: sum_i (B a)_i w_i equals
    sum_c a(c) (A w)(c) for any grid a and visibilities w.  Fails if the
    two disagree on window placement or kernel orientation."""
    rng = numpy.random.default_rng(7)
    (gcf,), p, w = random_case(rng, 24, 5, 3, 40, separable=separable)
    a = rng.standard_normal((24, 24)) + 1j * rng.standard_normal((24, 24))
    Aw = numpy.zeros((24, 24), dtype=complex)
    R.convgrid(gcf, Aw, p, None, w)
    Ba = R.convdegrid(gcf, a, p)
    assert_allclose(numpy.sum(Ba * w), numpy.sum(a * Aw), rtol=1e-12)


def test_convgrid_orientation():
    """A kernel with a single non-zero tap (yf, xf, i, j) moves one unit of
    visibility to cell (iy - S//2 + i, ix - S//2 + j): rows are v, columns u,
    table layout [yf][xf][y][x] (synthesis.py:341-342, :480-481)."""
    N, S, Q = 16, 4, 4
    p = numpy.array([[0.1234, -0.2111, 0.0]])
    ix, xf = R.frac_coord(N, Q, p[:, 0])
    iy, yf = R.frac_coord(N, Q, p[:, 1])
    for i, j in [(0, 0), (1, 3), (3, 2)]:
        gcf = numpy.zeros((Q, Q, S, S), dtype=complex)
        gcf[yf[0], xf[0], i, j] = 1
        a = numpy.zeros((N, N), dtype=complex)
        R.convgrid(gcf, a, p, None, numpy.array([2 + 1j]))
        nz = numpy.argwhere(a != 0)
        assert nz.tolist() == [[iy[0] - S // 2 + i, ix[0] - S // 2 + j]]


def test_convgrid_bounds_error():
    """A window leaving the grid makes NumPy's slice-add fail (ValueError)."""
    gcf = numpy.ones((2, 2, 4, 4), dtype=complex)
    a = numpy.zeros((16, 16), dtype=complex)
    with pytest.raises(ValueError):
        R.convgrid(gcf, a, numpy.array([[0.49, 0.0]]), None, numpy.array([1.0 + 0j]))


def test_hermitian():
    """tests/test_synthesis.py:358-391."""
    lam = 100
    _or, _and = numpy.logical_or, numpy.logical_and
    for N in range(2, 20):
        uvw = numpy.round(uvw_grid(N) * lam)
        non_zero = _or(uvw[:, 0] != 0, uvw[:, 1] != 0)
        non_red = _or(uvw[:, 0] < 0, _and(uvw[:, 0] == 0, uvw[:, 1] < 0))
        if N % 2 == 0:
            non_red = _or(non_red, _or(uvw[:, 0] == -lam / 2, uvw[:, 1] == -lam / 2))
        for x, y in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
            vis = R.simulate_point(uvw, x / lam, y / lam)
            a = numpy.zeros((N, N), dtype=complex)
            R.grid(a, uvw[non_zero] / lam, vis[non_zero])
            a_bl = numpy.zeros((N, N), dtype=complex)
            R.grid(a_bl, uvw[non_red] / lam, vis[non_red])
            assert_allclose(a, R.make_grid_hermitian(a_bl), rtol=1e-5, atol=1e-5)
            assert_allclose(R.ifft(a).imag, 0, atol=1e-14)


# ----------------------------------------------------- C oracle = NumPy oracle

@pytest.mark.parametrize('separable', [False, True])
def test_c_oracle_matches_numpy(separable):
    rng = numpy.random.default_rng(11)
    tables, p, v = random_case(rng, 48, 6, 4, 300, planes=3, separable=separable)
    vplane = rng.integers(0, 3, 300)
    src = vplane[:, None]
    gcf = numpy.array(tables)
    a = numpy.zeros((48, 48), dtype=complex)
    R.convgrid(gcf, a, p, src, v)
    b = numpy.zeros((48, 48), dtype=complex)
    cgrid.convgrid(cgrid.Bank(tables, separable), b, p, v, vplane)
    assert_allclose(b, a, rtol=1e-13, atol=1e-12)
    vd = R.convdegrid(gcf, a, p, src)
    vc = cgrid.convdegrid(cgrid.Bank(tables, separable), a, p, vplane)
    assert_allclose(vc, vd, rtol=1e-11, atol=1e-11)


def test_c_oracle_wcache_matches_numpy():
    theta, lam, wstep = 0.1, 400.0, 25.0
    rng = numpy.random.default_rng(3)
    nv = 60
    uvw = numpy.stack([rng.uniform(-150, 150, nv), rng.uniform(-150, 150, nv),
                       rng.uniform(-100, 100, nv)], axis=1)
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)

    def kfn(theta, w):
        return R.w_kernel(theta, w, 12, 6, 2) / 144

    ref = R.w_cache_imaging(theta, lam, uvw, None, vis, wstep=wstep, kernel_fn=kfn)
    bins = cgrid.wbin(uvw[:, 2], wstep)
    b0 = bins.min()
    tables = [kfn(theta, wstep * b) for b in range(b0, bins.max() + 1)]
    g = numpy.zeros((40, 40), dtype=complex)
    cgrid.convgrid(cgrid.Bank(tables), g, uvw / lam, vis, (bins - b0).astype(numpy.int32))
    assert_allclose(g, ref, rtol=1e-12, atol=1e-12)
    vref = R.w_cache_predict(theta, lam, uvw, None, ref, wstep=wstep, kernel_fn=kfn)
    vc = cgrid.convdegrid(cgrid.Bank(tables), ref, uvw / lam, (bins - b0).astype(numpy.int32))
    assert_allclose(vc, vref, rtol=1e-11, atol=1e-11)


def test_w_slice_composition():
    """w_slice_imaging (intended semantics) equals explicit per-slice convgrid
    with the kernel at the slice mean w."""
    theta, lam = 0.1, 300.0
    rng = numpy.random.default_rng(4)
    uvw = numpy.stack([rng.uniform(-100, 100, 50), rng.uniform(-100, 100, 50),
                       rng.uniform(-50, 50, 50)], axis=1)
    vis = rng.standard_normal(50) + 0j

    def kfn(theta, w):
        return R.w_kernel(theta, w, 8, 4, 2)

    g = R.w_slice_imaging(theta, lam, uvw, None, vis, wstep=7, kernel_fn=kfn)
    order = numpy.argsort(uvw[:, 2], kind='stable')
    ref = numpy.zeros((30, 30), dtype=complex)
    for i in range(0, 50, 7):
        sel = order[i:i + 7]
        R.convgrid(kfn(theta, uvw[sel, 2].mean()), ref, uvw[sel] / lam, None, vis[sel])
    assert_allclose(g, ref, rtol=1e-13)


# ------------------------------------- shifted / transformed w-kernels
# tests/test_synthesis.py:283-356 transcribed: a point source simulated at a
# shifted (and linearly transformed) position must image back to its grid
# cell when the visibilities are phase-shifted and gridded with w-kernels
# built on the shifted / transformed kernel coordinates.

SHIFT_TS = [
    numpy.array([[1, 0], [0, 1]]),
    numpy.array([[2, 0], [0, 3]]),
    numpy.array([[0, 1], [-1, 0]]),
    numpy.array([[1 / numpy.sqrt(2), 1 / numpy.sqrt(2)], [-1 / numpy.sqrt(2), 1 / numpy.sqrt(2)]]),
    numpy.array([[1e5, -5e3], [6e4, -1e6]]),
    numpy.array([[0, .5], [-.5, 0]]),
]


def shift_w_peaks(M, lam=10, Ns=range(1, 6)):
    """Peaks of test_grid_shift_w (tests/test_synthesis.py:283-310) with the
    oracle or product module M: every value must be 1."""
    peaks = []
    for uw, vw in [(.5, 0), (0, .5), (-1, 0), (0, -1)]:
        for N in Ns:
            for dl, dm in [(1 / lam, 1 / lam), (-1 / lam, 2 / lam), (5 / lam, 0)]:
                theta = N / lam
                uvw_all = uvw_grid(N, uw, vw) * lam
                slices = R.slice_vis(N, R.sort_vis_w(uvw_all))
                gcfs = [M.w_kernel(theta, numpy.mean(u[:, 2]), N, N, 1, dl=-dl, dm=-dm) / N ** 2
                        for u in slices]
                for x, y in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
                    a = numpy.zeros((2 * N, 2 * N), dtype=complex)
                    for u, gcf in zip(slices, gcfs):
                        vis = R.simulate_point(u, x / lam - dl, y / lam - dm)
                        viss = R.visibility_shift(u, vis, dl, dm)
                        assert_allclose(R.visibility_shift(u, viss, -dl, -dm), vis)
                        M.convgrid(gcf, a, u / lam / 2, None, viss)
                    a2 = numpy.fft.fftshift(a[:N, :N] + a[:N, N:] + a[N:, :N] + a[N:, N:])
                    img = numpy.real(R.ifft(a2))
                    peaks.append(img[N // 2 + y, N // 2 + x])
    return numpy.array(peaks)


def transform_shift_w_peaks(M, lam=10, Ns=range(1, 6)):
    """Peaks of test_grid_transform_shift_w (tests/test_synthesis.py:312-356)
    with module M: every value must be 1."""
    peaks = []
    for T in SHIFT_TS:
        Ti = numpy.linalg.inv(T)
        for uw, vw in [(.5, 0), (0, .5), (0, -1)]:
            for dl, dm in [(0, 0), (1 / lam, 1 / lam), (-1 / lam, 4 / lam)]:
                for N in Ns:
                    uvwt_all = uvw_grid(N, uw, vw) * lam
                    uvw_all = R.uvw_transform(uvwt_all, Ti)
                    slices = R.slice_vis(N, *R.sort_vis_w(uvwt_all, uvw_all))
                    gcfs = [M.w_kernel(N / lam, numpy.mean(ut[:, 2]), N, N, 1, T=Ti, dl=-dl,
                                       dm=-dm) / N ** 2 for ut, _ in slices]
                    for xt, yt in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
                        x, y = numpy.dot([xt, yt], Ti)
                        assert_allclose(numpy.dot([x, y], T), [xt, yt], atol=1e-14)
                        a = numpy.zeros((2 * N, 2 * N), dtype=complex)
                        for (ut, u), gcf in zip(slices, gcfs):
                            vis = R.simulate_point(u, x / lam - dl, y / lam - dm)
                            vis = R.visibility_shift(u, vis, dl, dm)
                            M.convgrid(gcf, a, ut / lam / 2, None, vis)
                        a = numpy.fft.fftshift(a[:N, :N] + a[:N, N:] + a[N:, :N] + a[N:, N:])
                        img = numpy.real(R.ifft(a))
                        peaks.append(img[N // 2 + yt, N // 2 + xt])
    return numpy.array(peaks)


def test_grid_shift_w():
    """tests/test_synthesis.py:283-310 (w_kernel with dl, dm; visibility_shift)."""
    assert_allclose(shift_w_peaks(R), 1)


def test_grid_transform_shift_w():
    """tests/test_synthesis.py:312-356 (w_kernel with T, dl, dm; uvw_transform)."""
    assert_allclose(transform_shift_w_peaks(R), 1)


def test_shift_w_fails_without_kernel_shift():
    """The shift test is sensitive: kernels built without the (dl, dm) offset
    leave the peak off 1 (so a dropped or sign-flipped shift is caught)."""
    class NoShift:
        w_kernel = staticmethod(lambda theta, w, F, S, Q, dl=0, dm=0: R.w_kernel(theta, w, F, S, Q))
        convgrid = staticmethod(R.convgrid)

    class SignFlip:
        w_kernel = staticmethod(lambda theta, w, F, S, Q, dl=0, dm=0:
                                R.w_kernel(theta, w, F, S, Q, dl=-dl, dm=-dm))
        convgrid = staticmethod(R.convgrid)
    for M in (NoShift, SignFlip):
        peaks = shift_w_peaks(M, Ns=[3, 4])
        assert numpy.abs(peaks - 1).max() > 1e-3


def test_uvw_transform_and_shift_invert():
    """simulate.py:259-300: uvw_transform by T then by T^-1 is the identity,
    and it maps the image by T in the reference's row-vector convention (a
    point at x images at x T, test_synthesis.py:218-223: the uv phase is
    preserved, u' . (x T) = u . x); visibility_shift by (dl, dm)
    then (-dl, -dm) is the identity."""
    rng = numpy.random.default_rng(7)
    uvw = rng.standard_normal((20, 3)) * 100
    for T in SHIFT_TS:
        Ti = numpy.linalg.inv(T)
        assert_allclose(R.uvw_transform(R.uvw_transform(uvw, T), Ti), uvw, rtol=1e-9,
                        atol=1e-9 * numpy.abs(uvw).max())
        x = rng.standard_normal(2)
        ut = R.uvw_transform(uvw, T)
        assert_allclose(ut[:, :2] @ (x @ T), uvw[:, :2] @ x, rtol=1e-9, atol=1e-6)
        assert_allclose(ut[:, 2], uvw[:, 2])
    vis = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    assert_allclose(R.visibility_shift(uvw, R.visibility_shift(uvw, vis, .01, -.02), -.01, .02),
                    vis)


def test_kernel_transform_eigenstructure():
    """synthesis.py:220-248: T is symmetric, maps the shift direction d to
    dn d (foreshortening by the cosine dn = sqrt(1 - |d|^2)) and leaves the
    perpendicular direction unchanged; the identity at d = 0."""
    assert_allclose(R.kernel_transform(0, 0), numpy.eye(2))
    for dl, dm in [(0.1, 0.0), (0.0, -0.3), (0.2, 0.5), (-0.6, 0.7), (1e-4, 2e-4)]:
        T = R.kernel_transform(dl, dm)
        dn = numpy.sqrt(1 - dl ** 2 - dm ** 2)
        assert_allclose(T, T.T)
        assert_allclose(T @ [dl, dm], dn * numpy.array([dl, dm]), atol=1e-15)
        assert_allclose(T @ [-dm, dl], [-dm, dl], atol=1e-15)
        assert_allclose(numpy.linalg.det(T), dn, rtol=1e-12)


def test_w_conj_kernel_fn_dispatch():
    """synthesis.py:682-699: negative w is served by conj(k(theta, -w));
    w >= 0 passes through; extra arguments are forwarded."""
    calls = []

    def kfn(theta, w, *a, **kw):
        calls.append((theta, w, a, kw))
        return numpy.array([[1 + 2j, w + 1j]])
    fn = R.w_conj_kernel_fn(kfn)
    assert_allclose(fn(0.1, -3.0, 7, k=1), numpy.conj(kfn(0.1, 3.0)))
    assert calls[0] == (0.1, 3.0, (7,), {'k': 1})
    assert_allclose(fn(0.1, 2.0), kfn(0.1, 2.0))
    assert calls[2][1] == 2.0


def test_w_kernel_conjugate_symmetry():
    """k(-w) = conj(k(w)) holds exactly for odd far-field sizes at Q = 1
    (symmetric coordinates), the identity w_conj_kernel_fn folds on."""
    for F, w in [(5, 30.0), (9, 250.0), (15, 1000.0)]:
        assert_allclose(R.w_kernel(0.1, -w, F, F, 1), numpy.conj(R.w_kernel(0.1, w, F, F, 1)),
                        atol=1e-12)


def test_doweight_hand_built():
    """synthesis.py:533-549 on a hand-built case: theta 0.1, lam 100 (N =
    10, cells of 10 wavelengths); the weight cell is round-half-even(N//2 +
    u theta) as frac_coord with Qpx = 1 gives, so u = 5 (5.5 -> 6) shares
    cell 6 with u = 15 (6.5 -> 6), and u = 25, 26 share cell 8."""
    uvw = numpy.array([[0, 0, 0], [1, 1, 0], [-2, 3, 0], [15, 0, 0], [5, 0, 0],
                       [25, 0, 0], [26, 0, 0], [0, 31, 0]], dtype=float)
    want = numpy.array([1 / 3, 1 / 3, 1 / 3, 1 / 2, 1 / 2, 1 / 2, 1 / 2, 1.0])
    assert_allclose(R.doweight(0.1, 100.0, uvw, numpy.ones(8)), want)
    v = numpy.arange(1, 9) * (1 + 1j)
    assert_allclose(R.doweight(0.1, 100.0, uvw, v), v * want)


def test_mirror_uvw_hand_built():
    """synthesis.py:831-852: rows with v < 0 are negated (u, v and w) and
    their visibilities conjugated; v >= 0 (including -0.0) is kept."""
    uvw = numpy.array([[1., -2., 3.], [4., 5., -6.], [-7., 0., 8.], [9., -0.0, 1.],
                       [-1., -1e-9, -2.]])
    vis = numpy.array([1 + 2j, 3 - 4j, 5 + 6j, 7 + 8j, -1 - 1j])
    u2, v2 = R.mirror_uvw(uvw, vis)
    assert_allclose(u2, [[-1., 2., -3.], [4., 5., -6.], [-7., 0., 8.], [9., 0., 1.],
                         [1., 1e-9, 2.]])
    assert_allclose(v2, [1 - 2j, 3 - 4j, 5 + 6j, 7 + 8j, -1 + 1j])
    assert uvw[0, 1] == -2.0 and vis[0] == 1 + 2j  # inputs untouched
