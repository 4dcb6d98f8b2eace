"""Pinning of the w-towers oracle (oracle/wtowers_ref.py).

* With every w = 0 the towers must reproduce conv_imaging / conv_predict
  exactly (the chunk geometry, sub-grid coordinates and clipped adds are
  then the only moving parts).
* For w != 0 the w-stacking must move the result towards the direct
  Fourier transform: a wrong Fresnel sign or plane index makes the error
  larger than doing no w-correction at all.
"""
import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import synthesis_ref as R
from oracle import wtowers_ref as WT

THETA, LAM = 0.1, 2560.0      # N = 256
SG, MARGIN = 64, 16


def aa_kernel(N, S=8, Q=32):
    aaf = R.anti_aliasing_function(N, 0, S)
    kv = R.kernel_oversample(aaf, Q, S)
    return aaf, kv / kv[0].sum()


def random_uvw(rng, nv, wmax):
    r = 0.42 * LAM * numpy.sqrt(rng.uniform(0, 1, nv))
    ph = rng.uniform(0, 2 * numpy.pi, nv)
    return numpy.stack([r * numpy.cos(ph), r * numpy.sin(ph), rng.uniform(-wmax, wmax, nv)], 1)


@pytest.mark.parametrize('separable', [True, False])
def test_towers_equal_conv_at_w0(separable):
    rng = numpy.random.default_rng(1)
    uvw = random_uvw(rng, 3000, 0.0)
    vis = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
    if separable:
        _, kv = aa_kernel(256)
    else:
        kv = rng.standard_normal((4, 4, 6, 6)) + 1j * rng.standard_normal((4, 4, 6, 6))
    g = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 10.0, SG, MARGIN, kv)
    assert_allclose(g, R.conv_imaging(THETA, LAM, uvw, None, vis, kv), atol=1e-9)
    grid = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    v = WT.w_towers_predict(THETA, LAM, uvw, None, grid, 10.0, SG, MARGIN, kv)
    assert_allclose(v, R.conv_predict(THETA, LAM, uvw, None, grid, kv), atol=1e-9)


def test_towers_nearest_equal_simple_at_w0():
    rng = numpy.random.default_rng(2)
    uvw = random_uvw(rng, 2000, 0.0)
    vis = rng.standard_normal(2000) + 0j
    g = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 10.0, SG, MARGIN, None)
    assert_allclose(g, R.simple_imaging(THETA, LAM, uvw, None, vis), atol=1e-9)



def test_towers_predict_physics():
    """Predict from a model image of pixel-centred sources, compared with
    the phase-tracked DFT including w.  w-stacking must beat no correction
    by a wide margin and reach the anti-aliasing accuracy."""
    rng = numpy.random.default_rng(3)
    N = 256
    aaf, kv = aa_kernel(N)
    pix = rng.integers(-40, 41, (6, 2))
    flux = rng.uniform(0.5, 2.0, 6)
    image = numpy.zeros((N, N))
    for (px, py), f in zip(pix, flux):
        image[N // 2 + py, N // 2 + px] += f
    grid = R.fft(image / numpy.outer(aaf, aaf))
    uvw = random_uvw(rng, 2000, 800.0)
    ref = sum(f * R.simulate_point(uvw, px / LAM, py / LAM) for (px, py), f in zip(pix, flux))
    towers = WT.w_towers_predict(THETA, LAM, uvw, None, grid, 10.0, SG, MARGIN, kv)
    flat = WT.w_towers_predict(THETA, LAM, uvw, None, grid, 1e9, SG, MARGIN, kv)
    e_towers = numpy.abs(towers - ref).mean() / numpy.abs(ref).mean()
    e_flat = numpy.abs(flat - ref).mean() / numpy.abs(ref).mean()
    assert e_towers < 1e-2, e_towers  # the 8-tap PSWF floor is 7.7e-3 at w = 0
    assert e_flat > 20 * e_towers, (e_flat, e_towers)


def test_towers_imaging_physics():
    """Dirty image of one point source at large w: the towers recover the
    peak; without w-correction it is decorrelated."""
    rng = numpy.random.default_rng(4)
    N = 256
    aaf, kv = aa_kernel(N)
    px, py = 80, -65
    uvw = random_uvw(rng, 4000, 1500.0)
    vis = R.simulate_point(uvw, px / LAM, py / LAM)
    peak = {}
    for name, wstep in (('towers', 10.0), ('flat', 1e9)):
        g = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, wstep, SG, MARGIN, kv)
        img = numpy.real(R.ifft(g)) / numpy.outer(aaf, aaf) * N * N / len(vis)
        peak[name] = img[N // 2 + py, N // 2 + px]
    assert abs(peak['towers'] - 1) < 0.01, peak
    assert peak['flat'] < 0.8 * peak['towers'], peak


def test_towers_linearity():
    rng = numpy.random.default_rng(5)
    uvw = random_uvw(rng, 800, 300.0)
    a = rng.standard_normal(800) + 1j * rng.standard_normal(800)
    b = rng.standard_normal(800) + 1j * rng.standard_normal(800)
    _, kv = aa_kernel(256)
    ga = WT.w_towers_imaging(THETA, LAM, uvw, None, a, 25.0, SG, MARGIN, kv)
    gb = WT.w_towers_imaging(THETA, LAM, uvw, None, b, 25.0, SG, MARGIN, kv)
    gab = WT.w_towers_imaging(THETA, LAM, uvw, None, 2 * a - 3j * b, 25.0, SG, MARGIN, kv)
    assert_allclose(gab, 2 * ga - 3j * gb, atol=1e-10)


@pytest.mark.parametrize('with_kernel', [False, True])
def test_imaging_and_predict_are_adjoint(with_kernel):
    """<A v, G> = <v, B' G>: w-towers imaging A and predict B' with the
    conjugated kernel tables are adjoint (F^H = n^2 iF, the Fresnel factors
    conjugate, cutting is the transpose of adding).  Holds for any data."""
    rng = numpy.random.default_rng(3)
    N, theta = 256, 0.1
    lam = N / theta
    nv = 3000
    uvw = numpy.stack([rng.uniform(-0.45, 0.45, nv) * lam, rng.uniform(-0.45, 0.45, nv) * lam,
                       rng.uniform(-300, 300, nv)], 1)
    v = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    G = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    kv = None
    if with_kernel:
        kv = rng.standard_normal((4, 4, 8, 8)) + 1j * rng.standard_normal((4, 4, 8, 8))
    A = WT.w_towers_imaging(theta, lam, uvw, None, v, 20.0, 64, 16, kv)
    B = WT.w_towers_predict(theta, lam, uvw, None, G, 20.0, 64, 16,
                           None if kv is None else numpy.conj(kv))
    lhs, rhs = numpy.vdot(A, G), numpy.vdot(v, B)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def residual_set(N, wincrement, step, invert, S=8, Q=32):
    """Separable w-kernel set over [-wincrement/2, wincrement/2]: table k is
    kernel_oversample of aaf(l) exp(-+ i pi w_k l^2) (the paraxial residual
    w-term per axis), normalised like aa_kernel's w = 0 kernel."""
    aaf = R.anti_aliasing_function(N, 0, S)
    norm = R.kernel_oversample(aaf, Q, S)[0].sum()
    n = int(numpy.ceil(wincrement / step - 1e-9)) + 1
    w_min = -(n - 1) * step / 2
    l = R.coordinates(N) * THETA
    sign = -1 if invert else 1
    tabs = numpy.array([R.kernel_oversample(aaf * numpy.exp(sign * 1j * numpy.pi *
                                                            (w_min + k * step) * l ** 2), Q, S)
                        for k in range(n)]) / norm
    return aaf, (tabs, w_min, step)


def test_kernel_set_of_equal_tables_is_single_table():
    """A w-kernel set whose tables are all the same kernel gives the single
    kernel's result exactly (only the per-visibility table choice moves)."""
    rng = numpy.random.default_rng(6)
    uvw = random_uvw(rng, 1500, 300.0)
    vis = rng.standard_normal(1500) + 1j * rng.standard_normal(1500)
    _, kv = aa_kernel(256)
    kset = (numpy.array([kv] * 5), -20.0, 10.0)
    assert_allclose(WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kset),
                    WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv),
                    atol=1e-12)
    grid = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    assert_allclose(WT.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kset),
                    WT.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kv),
                    atol=1e-12)


def test_kernel_set_selects_residual_plane():
    """Tables marked by their index: a visibility of plane wp = round(w /
    wstep) uses table floor((w - wp wstep - w_min) / w_step + 0.5), the
    w_project rule (grid.c:286-289); checked on single-visibility gridding
    where the table's mark (squared: the separable kernel is the outer
    product of two table rows) scales the grid."""
    _, kv = aa_kernel(256)
    wstep, w_min, w_step = 40.0, -20.0, 10.0
    tabs = numpy.array([kv * (k + 1) for k in range(5)])
    for w, want in [(0.0, 2), (4.9, 2), (5.1, 3), (19.0, 4), (-19.0, 0), (-14.9, 1),
                    (41.0, 2), (-61.0, 4), (58.0, 4), (-75.0, 3)]:
        uvw = numpy.array([[100.0, -50.0, w]])
        g = WT.w_towers_imaging(THETA, LAM, uvw, None, numpy.array([1.0 + 0j]), wstep, SG,
                                MARGIN, (tabs, w_min, w_step))
        g1 = WT.w_towers_imaging(THETA, LAM, uvw, None, numpy.array([1.0 + 0j]), wstep, SG,
                                 MARGIN, kv)
        assert_allclose(g, (want + 1) ** 2 * g1, atol=1e-12), (w, want)


def test_kernel_set_predict_physics():
    """Coarse w-planes (wstep 80, residual |dw| <= 40, ~0.6 rad at the field
    corner): with the residual w-kernel set the towers predict reaches the
    anti-aliasing floor against the DFT; with one kernel for all residuals
    (the notebooks' form) they do not."""
    rng = numpy.random.default_rng(8)
    N = 256
    aaf, kset = residual_set(N, 80.0, 1.0, invert=False)
    _, kv = aa_kernel(N)
    pix = rng.integers(-90, 91, (6, 2))
    flux = rng.uniform(0.5, 2.0, 6)
    image = numpy.zeros((N, N))
    for (px, py), f in zip(pix, flux):
        image[N // 2 + py, N // 2 + px] += f
    grid = R.fft(image / numpy.outer(aaf, aaf))
    uvw = random_uvw(rng, 2000, 800.0)
    ref = sum(f * R.simulate_point(uvw, px / LAM, py / LAM) for (px, py), f in zip(pix, flux))
    e = {}
    for name, k, wstep in (('set', kset, 80.0), ('single', kv, 80.0), ('fine', kv, 4.0)):
        v = WT.w_towers_predict(THETA, LAM, uvw, None, grid, wstep, SG, MARGIN, k)
        e[name] = numpy.abs(v - ref).mean() / numpy.abs(ref).mean()
    # 'fine': 20x finer planes with one kernel, residual phase <= 0.03 rad:
    # the anti-aliasing floor of the 8-tap PSWF at these source positions
    assert e['set'] < 1.2 * e['fine'], e
    assert e['single'] > 3 * e['set'], e
