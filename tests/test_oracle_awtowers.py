"""Pinning of the AW-towers restatement (oracle/awtowers_ref.py:
awtowers.ipynb cells 17, 19, 22) by physics: for point sources seen through
per-antenna direction-dependent phase screens (the notebook's ionosphere,
awtowers.ipynb cell 13's add_point), the AW-towers image equals the direct
Fourier dirty image with the A-correction applied pixel by pixel, up to the
anti-aliasing kernel's aliasing and the w residual inside a plane.  Without
the A-screens, or without the Fresnel stepping, it does not."""
import itertools

import numpy
import pytest

from oracle import awtowers_ref as AWT
from oracle import synthesis_ref as R


def scenario(seed=0, nant=6, ndump=24, N=256, theta=0.1, S=64, margin=16, wvl=4.0,
             aa_over=1024):
    """Baseline tracks of nant antennas over ndump hour angles, five point
    sources, antenna phase screens exp(2 pi i phi_a(l, m)); visibilities
    flipped to v >= 0 with conjugated values and swapped antennas (the
    notebook's cell 15)."""
    rng = numpy.random.default_rng(seed)
    lam = N / theta
    xyz = numpy.stack([rng.uniform(-1400, 1400, nant), rng.uniform(-1400, 1400, nant),
                       rng.uniform(-30, 30, nant)], 1)
    pairs = numpy.array(list(itertools.combinations(range(nant), 2)))
    uvw = numpy.concatenate([R.xyz_to_uvw(xyz[pairs[:, 0]] - xyz[pairs[:, 1]], h,
                                          numpy.radians(45)) / wvl
                             for h in numpy.linspace(-0.7, 0.7, ndump)])
    src = numpy.tile(pairs, (ndump, 1))
    coef = rng.uniform(-0.5, 0.5, (nant, 4))

    def ascreen(a, l, m):
        return numpy.exp(2j * numpy.pi * (coef[a, 0] + 3 * coef[a, 1] * l + 3 * coef[a, 2] * m
                                          + 60 * coef[a, 3] * (l * l + m * m)))

    sources = [(30, 20), (-50, 10), (40, -60), (0, 0), (-20, -35)]
    vis = numpy.zeros(len(uvw), dtype=complex)
    for px, py in sources:
        l, m = px * theta / N, py * theta / N
        ratio = numpy.array([ascreen(s[0], l, m) / ascreen(s[1], l, m) for s in src])
        vis += ratio * R.simulate_point(uvw, l, m)
    flip = uvw[:, 1] < 0
    uvw[flip] = -uvw[flip]
    vis[flip] = numpy.conj(vis[flip])
    src[flip] = src[flip, ::-1]
    aa = R.anti_aliasing_function(N, 0, 6)
    l, m = theta * R.coordinates2(S)
    return dict(N=N, theta=theta, lam=lam, S=S, margin=margin, uvw=uvw, src=src, vis=vis,
                ascreen=ascreen, sources=sources, aa=aa,
                aa_gcf=R.kernel_oversample(aa, aa_over, 6),
                aa_sub=R.anti_aliasing_function(S, 0, 6),
                aa_sub2=R.anti_aliasing_function(2 * S, 0, 6),
                aimage=numpy.array([ascreen(a, l, m) for a in range(nant)]),
                l=l, m=m)


def image(sc, grid_sum):
    return R.ifft(grid_sum) / numpy.outer(sc['aa'], sc['aa'])


def run(sc, wstep, aimage=None, wkern=None):
    wk = R.w_kernel_function(sc['l'], sc['m'], wstep) if wkern is None else wkern
    return AWT.aw_towers_imaging(sc['theta'], sc['lam'], sc['uvw'], sc['src'], sc['vis'],
                                 wstep, sc['S'], sc['margin'], sc['aa_gcf'], sc['aa_sub'],
                                 sc['aa_sub2'],
                                 sc['aimage'] if aimage is None else aimage, wk)


def compare(sc, img, npix=100):
    """(scale, residual): the least-squares scale of img onto the DFT image
    at the sources and npix random pixels, and the relative residual."""
    N = sc['N']
    rng = numpy.random.default_rng(1)
    pix = [(N // 2 + px, N // 2 + py) for px, py in sc['sources']]
    pix += [tuple(x) for x in rng.integers(N // 2 - 100, N // 2 + 100, (npix, 2))]
    d = AWT.dft_image(sc['theta'], N, sc['uvw'], sc['src'], sc['vis'], sc['ascreen'], pix)
    got = numpy.array([img[y, x] for x, y in pix])
    c = numpy.vdot(got, d) / numpy.vdot(got, got)
    return c, numpy.linalg.norm(got * c - d) / numpy.linalg.norm(d)


def test_matches_corrected_dft_image():
    """Scale 4 = (2 S)^2 / S^2: the pre-grid's inverse transform over the
    twice-larger grid against the sub-grid's forward transform."""
    sc = scenario()
    c, res = compare(sc, image(sc, run(sc, 2.0)))
    assert abs(c - 4) < 0.01 and res < 0.01, (c, res)


def test_without_a_screens_or_w_stepping_it_fails():
    """Residuals 0.94 without screens, 0.34 without the Fresnel stepping and
    0.51 with the stepping's sign reversed, against 0.0024."""
    sc = scenario()
    _, res_a = compare(sc, image(sc, run(sc, 2.0, aimage=numpy.ones_like(sc['aimage']))))
    _, res_w = compare(sc, image(sc, run(sc, 2.0, wkern=numpy.ones((sc['S'], sc['S'])))))
    wrong = numpy.conj(R.w_kernel_function(sc['l'], sc['m'], 2.0))
    _, res_s = compare(sc, image(sc, run(sc, 2.0, wkern=wrong)))
    assert res_a > 0.5 and res_w > 0.2 and res_s > 0.2, (res_a, res_w, res_s)


def test_swapped_screen_ratio_fails():
    """aimage[a1] / aimage[a2] instead of aimage[a2] / aimage[a1]: with
    unit-modulus screens that is the conjugate screen."""
    sc = scenario()
    _, res = compare(sc, image(sc, run(sc, 2.0, aimage=numpy.conj(sc['aimage']))))
    assert res > 0.5, res


@pytest.mark.parametrize('S,margin', [(64, 16), (64, 15), (32, 5)])
def test_chunk_geometry_covers_its_visibilities(S, margin):
    """Every visibility's cell, with the anti-aliasing kernel's half width,
    lies inside its chunk's sub-grid at (corner, corner + S): odd chunk
    widths take the notebook's truncating float corners."""
    sc = scenario(S=S, margin=margin)
    theta, lam, N = sc['theta'], sc['lam'], sc['N']
    chunk = S - margin
    for u, v in sc['uvw'][:, :2]:
        ub = int(numpy.floor(u * theta / chunk + 0.5))
        vb = int(numpy.floor(v * theta / chunk + 0.5))
        uv_mid, x0, y0 = AWT.chunk_geometry(theta, lam, S, margin, ub, vb)
        cx = int(lam * theta) // 2 + u * theta
        cy = int(lam * theta) // 2 + v * theta
        assert x0 + 1 <= cx <= x0 + S - 1 and y0 + 1 <= cy <= y0 + S - 1
        assert abs(u - uv_mid[0]) * theta <= chunk / 2 + 1
