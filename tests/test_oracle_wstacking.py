"""Pinning of the w-stacking restatement (oracle/wstacking_ref.py:
wstacking.ipynb cells 10 and 12) by closed forms and physics: at w = 0 both
variants reduce to simple imaging exactly, and for point sources far from
the phase centre with |w| large enough to decorrelate them, the w-stacked
image recovers each source's full peak where the uncorrected image loses
it."""
import numpy
import pytest
from numpy.testing import assert_allclose

from oracle import synthesis_ref as R
from oracle import wstacking_ref as WS


def test_full_grid_at_w0_is_simple_imaging():
    rng = numpy.random.default_rng(1)
    N, theta = 64, 0.1
    lam = N / theta
    uvw = numpy.zeros((500, 3))
    uvw[:, :2] = rng.uniform(-0.45, 0.45, (500, 2)) * lam
    vis = rng.normal(size=500) + 1j * rng.normal(size=500)
    got = WS.w_stacking_imaging(theta, lam, uvw, None, vis, 10.0)
    ref = R.ifft(R.simple_imaging(theta, lam, uvw, None, vis))
    assert_allclose(got, ref, rtol=0, atol=0)


def test_subgrid_within_half_step_is_simple_imaging():
    rng = numpy.random.default_rng(2)
    N, theta = 64, 0.1
    lam = N / theta
    uvw = numpy.zeros((500, 3))
    uvw[:, :2] = rng.uniform(-0.45, 0.45, (500, 2)) * lam
    uvw[:, 2] = rng.uniform(-4.9, 4.9, 500)
    vis = rng.normal(size=500) + 1j * rng.normal(size=500)
    got = WS.w_stacking_subgrid_imaging(theta, lam, uvw, None, vis, 10.0, 8)
    assert_allclose(got, R.simple_imaging(theta, lam, uvw, None, vis), rtol=0, atol=0)


def sources_setup(seed=3, N=256, theta=0.5, nv=4000, wmax=240.0):
    # wide field, |w| < lam / 2 (grid's assertion covers the w column)
    rng = numpy.random.default_rng(seed)
    lam = N / theta
    r = 0.35 * lam * numpy.sqrt(rng.uniform(0, 1, nv))
    ph = rng.uniform(0, 2 * numpy.pi, nv)
    uvw = numpy.stack([r * numpy.cos(ph), r * numpy.sin(ph), rng.uniform(-wmax, wmax, nv)], 1)
    pix = [(60, 45), (-70, 30), (50, -65)]
    vis = numpy.zeros(nv, dtype=complex)
    for px, py in pix:
        vis += R.simulate_point(uvw, px * theta / N, py * theta / N)
    return N, theta, lam, uvw, vis, pix


@pytest.mark.parametrize('variant', ['full', 'subgrid'])
def test_w_stacking_recovers_sources(variant):
    """Each source's image peak is nv / N^2 (the ifft normalisation) times
    nearest-cell gridding's image taper sinc(px / N) sinc(py / N) once the
    Fresnel pattern of every plane is divided out, less the residual w
    within a plane (|w - mean w| <= 5, 2 pi 5 (1 - n) <= 0.34 rad: a few
    per cent); without the correction the w-term spreads the source
    (|w| (1 - n) reaches ~4 turns here)."""
    N, theta, lam, uvw, vis, pix = sources_setup()
    nv = len(vis)
    if variant == 'full':
        img = WS.w_stacking_imaging(theta, lam, uvw, None, vis, 10.0)
    else:
        img = R.ifft(WS.w_stacking_subgrid_imaging(theta, lam, uvw, None, vis, 10.0, 48))
    plain = R.ifft(R.simple_imaging(theta, lam, uvw, None, vis))
    for px, py in pix:
        want = nv / N ** 2 * numpy.sinc(px / N) * numpy.sinc(py / N)
        got = img[N // 2 + py, N // 2 + px]
        assert abs(got - want) < 0.06 * want, (variant, px, py, got, want)
        assert abs(plain[N // 2 + py, N // 2 + px]) < 0.5 * want
