"""Batched centred 2D transforms of 128 x 128 sub-grids against NumPy's FFT
(crocodile fft / ifft, synthesis.py:83-106): the dense DFT on the matrix
cores (fp16x3 split products, croc_dft2d_batched) and the radix 16 x 8 FFT
on the vector ALUs (croc_fft2d_batched)."""
import numpy
import pytest
import torch

from crocodile_amd import _lib
from oracle import synthesis_ref as R

pytestmark = pytest.mark.gpu


def dft(x, inverse, entry='croc_dft2d_batched'):
    t = torch.as_tensor(x, device='cuda').to(torch.complex64).contiguous()
    out = torch.empty_like(t)
    _lib.check(getattr(_lib.load(), entry)(_lib.ctx(), _lib.ptr(t), _lib.ptr(out),
                                           t.shape[-1], t.shape[0], int(inverse)))
    return out.cpu().numpy()


def fft(x, inverse):
    return dft(x, inverse, 'croc_fft2d_batched')


def rel(a, b):
    return numpy.linalg.norm(a - b) / numpy.linalg.norm(b)


@pytest.mark.parametrize('inverse', [False, True])
def test_dft_matches_numpy(gpu, inverse):
    rng = numpy.random.default_rng(1 + inverse)
    x = rng.standard_normal((5, 128, 128)) + 1j * rng.standard_normal((5, 128, 128))
    got = dft(x, inverse)
    f = R.ifft if inverse else R.fft
    for b in range(5):
        e = rel(got[b], f(x[b]))
        print('batch %d rms_rel %.3e' % (b, e))
        assert e < 2e-6, e


def test_dft_structured_inputs(gpu):
    """Delta, plane wave and a wide-dynamic-range field."""
    n = 128
    x = numpy.zeros((3, n, n), dtype=complex)
    x[0, 70, 33] = 1.0
    yy, xx = numpy.mgrid[0:n, 0:n]
    x[1] = numpy.exp(2j * numpy.pi * (5 * xx + 17 * yy) / n)
    rng = numpy.random.default_rng(3)
    x[2] = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * \
        10.0 ** rng.uniform(-6, 3, (n, n))
    got = dft(x, False)
    for b in range(3):
        e = rel(got[b], R.fft(x[b]))
        print('case %d rms_rel %.3e' % (b, e))
        assert e < 2e-6, e


def test_dft_round_trip(gpu):
    rng = numpy.random.default_rng(4)
    x = rng.standard_normal((2, 128, 128)) + 1j * rng.standard_normal((2, 128, 128))
    y = dft(dft(x, False), True)
    assert rel(y, x) < 3e-6


@pytest.mark.parametrize('inverse', [False, True])
def test_fft_matches_numpy(gpu, inverse):
    rng = numpy.random.default_rng(11 + inverse)
    x = rng.standard_normal((7, 128, 128)) + 1j * rng.standard_normal((7, 128, 128))
    got = fft(x, inverse)
    f = R.ifft if inverse else R.fft
    for b in range(7):
        e = rel(got[b], f(x[b]))
        print('batch %d rms_rel %.3e' % (b, e))
        assert e < 1e-6, e


def test_fft_structured_inputs(gpu):
    """Delta, plane wave, wide dynamic range; in place."""
    n = 128
    x = numpy.zeros((3, n, n), dtype=complex)
    x[0, 70, 33] = 1.0
    yy, xx = numpy.mgrid[0:n, 0:n]
    x[1] = numpy.exp(2j * numpy.pi * (5 * xx + 17 * yy) / n)
    rng = numpy.random.default_rng(13)
    x[2] = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * \
        10.0 ** rng.uniform(-6, 3, (n, n))
    t = torch.as_tensor(x, device='cuda').to(torch.complex64).contiguous()
    _lib.check(_lib.load().croc_fft2d_batched(_lib.ctx(), _lib.ptr(t), _lib.ptr(t), n, 3, 0))
    got = t.cpu().numpy()
    for b in range(3):
        e = rel(got[b], R.fft(x[b]))
        print('case %d rms_rel %.3e' % (b, e))
        assert e < 1e-6, e


def test_fft_agrees_with_dft_and_rejects_other_sizes(gpu):
    rng = numpy.random.default_rng(14)
    x = rng.standard_normal((3, 128, 128)) + 1j * rng.standard_normal((3, 128, 128))
    assert rel(fft(x, False), dft(x, False)) < 3e-6
    assert rel(fft(fft(x, False), True), x) < 1e-6
    with pytest.raises(ValueError):
        fft(numpy.zeros((1, 64, 64), complex), False)


@pytest.mark.parametrize('in_place', [False, True])
def test_fft_batch_larger_than_the_device(gpu, in_place):
    """More sub-grids than compute units: each workgroup transforms several,
    loading the next one's rows while it works on the current one."""
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    nb = 2 * ncu + 37
    rng = numpy.random.default_rng(15 + in_place)
    x = (rng.standard_normal((nb, 128, 128)) +
         1j * rng.standard_normal((nb, 128, 128))).astype(numpy.complex64)
    t = torch.as_tensor(x, device='cuda')
    out = t if in_place else torch.empty_like(t)
    _lib.check(_lib.load().croc_fft2d_batched(_lib.ctx(), _lib.ptr(t), _lib.ptr(out), 128, nb,
                                              0))
    got = out.cpu().numpy()
    worst = 0.0
    for b in range(nb):
        worst = max(worst, rel(got[b], R.fft(x[b].astype(complex))))
    print('batch %d worst rms_rel %.3e' % (nb, worst))
    assert worst < 1e-6, worst
