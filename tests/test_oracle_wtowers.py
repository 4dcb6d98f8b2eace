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
