"""AW-kernel generation (aw_kernel_fn, crocodile/synthesis.py:701-734) on the
device against scipy.signal.convolve2d and the oracle's aw_kernel_fn."""
import numpy
import pytest
import scipy.signal
from numpy.testing import assert_allclose

import crocodile_amd.synthesis as G
from crocodile_amd.awkernel import convolve_same
from oracle import synthesis_ref as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('sa,sb', [((5, 5), (3, 3)), ((8, 8), (8, 8)), ((7, 4), (2, 5)),
                                   ((16, 16), (9, 9)), ((3, 3), (6, 6))])
def test_convolve_same_matches_scipy(gpu, sa, sb):
    rng = numpy.random.default_rng(sum(sa) * 10 + sum(sb))
    a = rng.standard_normal(sa) + 1j * rng.standard_normal(sa)
    b = rng.standard_normal(sb) + 1j * rng.standard_normal(sb)
    assert_allclose(convolve_same(a, b), scipy.signal.convolve2d(a, b, mode='same'),
                    rtol=1e-13, atol=1e-13)


def test_aw_kernel_fn_matches_oracle(gpu):
    theta = 0.08
    rng = numpy.random.default_rng(3)
    akerns = {a: (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))) / 8
              for a in range(4)}

    def a_fn(theta, ant, t, f):
        return akerns[ant] * (1 + 0.01 * t) * (1 - 0.02 * f)

    def w_fn_gpu(theta, w):
        return G.w_kernel(theta, w, 32, 8, 4)

    def w_fn_ref(theta, w):
        return R.w_kernel(theta, w, 32, 8, 4)

    got = G.aw_kernel_fn(a_fn, w_fn_gpu)(theta, 250.0, 1, 3, 2.0, 1.0)
    ref = R.aw_kernel_fn(a_fn, w_fn_ref)(theta, 250.0, 1, 3, 2.0, 1.0)
    assert got.shape == ref.shape == (4, 4, 8, 8)
    assert_allclose(got, ref, rtol=0, atol=1e-11 * numpy.abs(ref).max())
