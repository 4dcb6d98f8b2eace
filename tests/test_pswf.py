"""anti_aliasing_function (synthesis.py:199-217) pinned by the mathematics
of its values: the zeroth prolate spheroidal angular function S_00(c, x) is
the eigenfunction of the band-limited Fourier transform,
    integral_{-1}^{1} exp(i c x t) S_00(c, t) dt = lambda S_00(c, x),
so the transform of the function, evaluated by Gauss-Legendre quadrature
at the very points the grid samples (x = (2 - 1/(4N)) (i - N//2) / N),
divided by the function's values there, is one constant.  A wrong spheroidal
parameter, mode, or sample mapping breaks the constancy.  Both the oracle's
and the product's host function are checked; they are the same scipy call."""
import numpy
import pytest
import scipy.special

import crocodile_amd.synthesis as G
from oracle import synthesis_ref as R


def eigen_ratio(aa, N, c, mult=None):
    mult = 2 - 1 / N / 4 if mult is None else mult
    x = mult * (numpy.arange(N) - N // 2) / N
    t, w = numpy.polynomial.legendre.leggauss(200)
    s = scipy.special.pro_ang1(0, 0, c, t)[0]
    F = numpy.array([numpy.sum(w * s * numpy.exp(1j * c * xi * t)) for xi in x])
    inner = numpy.abs(x) < 0.95          # the function is small and ill-conditioned at |x| -> 1
    return F[inner] / aa[inner]


@pytest.mark.parametrize('fn', [R.anti_aliasing_function, G.anti_aliasing_function])
@pytest.mark.parametrize('N,c', [(64, 6), (256, 6), (128, 13.0), (96, 20.8)])
def test_pswf_eigen_equation(fn, N, c):
    aa = fn(N, 0, c)
    r = eigen_ratio(aa, N, c)
    assert numpy.abs(r.imag).max() < 1e-8 * numpy.abs(r.real).max()
    assert numpy.ptp(r.real) < 1e-7 * numpy.abs(r.real).mean(), numpy.ptp(r.real)
    # eigenvalue of S_00: 2 R_00^(1)(c, 1) times S_00's normalisation, the
    # same for every N at this c
    r2 = eigen_ratio(fn(N + 10, 0, c), N + 10, c)
    assert abs(r.real.mean() - r2.real.mean()) < 1e-7 * abs(r.real.mean())


def test_pswf_mistakes_fail():
    """The check rejects a wrong c, a wrong mode and a wrong sample map."""
    N, c = 128, 8
    good = numpy.ptp(eigen_ratio(R.anti_aliasing_function(N, 0, c), N, c).real)
    wrong_c = numpy.ptp(eigen_ratio(R.anti_aliasing_function(N, 0, c + 0.5), N, c).real)
    wrong_map = numpy.ptp(eigen_ratio(R.anti_aliasing_function(N, 0, c), N, c, mult=2.0 - 0.1).real)
    x = (2 - 1 / N / 4) * (numpy.arange(N) - N // 2) / N
    mode2 = scipy.special.pro_ang1(0, 2, c, x)[0]
    wrong_mode = numpy.ptp(eigen_ratio(mode2, N, c).real)
    scale = abs(eigen_ratio(R.anti_aliasing_function(N, 0, c), N, c).real.mean())
    assert good < 1e-7 * scale
    assert min(wrong_c, wrong_map, wrong_mode) > 1e-3 * scale


def test_pswf_shape_two_dimensional():
    """The 2D form is the outer product of the two 1D ones."""
    a = G.anti_aliasing_function((32, 48), 0, 6)
    numpy.testing.assert_allclose(a, numpy.outer(G.anti_aliasing_function(32, 0, 6),
                                                 G.anti_aliasing_function(48, 0, 6)),
                                  rtol=1e-15, atol=0)
