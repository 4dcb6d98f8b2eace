"""W-stacking on the device (crocodile_amd.synthesis.w_stacking_imaging and
w_stacking_subgrid_imaging) against the FP64 oracle (oracle/wstacking_ref.py,
wstacking.ipynb cells 10 and 12), which test_oracle_wstacking.py pins.  The
device grids planes in complex64 and transforms them in complex64 FFTs, so
the bound is a few 1e-7 of the image's largest value."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from oracle import synthesis_ref as R
from oracle import wstacking_ref as WS

pytestmark = pytest.mark.gpu

TOL = 2e-6


def case(seed, N, theta, nv, wmax, wstep, fill=0.35, ties=True):
    rng = numpy.random.default_rng(seed)
    lam = N / theta
    r = fill * lam * numpy.sqrt(rng.uniform(0, 1, nv))
    ph = rng.uniform(0, 2 * numpy.pi, nv)
    w = rng.uniform(-wmax, wmax, nv)
    if ties:
        # exact half-way w values: numpy.around's half-even rule decides the plane
        k = rng.integers(-int(wmax / wstep), int(wmax / wstep), nv // 20)
        w[:nv // 20] = (k + 0.5) * wstep
    uvw = numpy.stack([r * numpy.cos(ph), r * numpy.sin(ph), w], 1)
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    return lam, uvw, vis


def rel(got, ref):
    return numpy.abs(got - ref).max() / numpy.abs(ref).max()


@pytest.mark.parametrize('N,theta,wstep,budget', [(256, 0.5, 10.0, 4 << 30),
                                                   (256, 0.5, 10.0, 5 * 256 * 256 * 8),
                                                   (255, 0.5, 10.0, 4 << 30),
                                                   (130, 0.25, 7.0, 3 * 130 * 130 * 8)])
def test_full_grid_matches_oracle(gpu, N, theta, wstep, budget):
    """Even N (centring folded into the kernels), odd N (shifted FFTs), and
    stacks split into batches of 5 and 3 planes."""
    print('case wstack full N=%d budget=%d' % (N, budget), flush=True)
    lam, uvw, vis = case(N, N, theta, 6000, 0.45 * N / theta * 0.5, wstep)
    ref = WS.w_stacking_imaging(theta, lam, uvw, None, vis, wstep)
    got = G.w_stacking_imaging(theta, lam, uvw, None, vis, wstep, max_stack_bytes=budget)
    assert got.dtype == numpy.complex128 and got.shape == (N, N)
    assert rel(got, ref) < TOL, rel(got, ref)


def test_full_grid_torch_in_torch_out(gpu):
    lam, uvw, vis = case(5, 128, 0.3, 3000, 100.0, 8.0)
    ref = WS.w_stacking_imaging(0.3, lam, uvw, None, vis, 8.0)
    got = G.w_stacking_imaging(0.3, lam, torch.as_tensor(uvw, device='cuda'),
                               None, torch.as_tensor(vis, device='cuda'), 8.0)
    assert isinstance(got, torch.Tensor) and got.is_cuda
    assert rel(got.cpu().numpy(), ref) < TOL


def test_full_grid_recovers_sources(gpu):
    """The physics check of the oracle test on the device result: each point
    source's peak is nv / N^2 sinc(px / N) sinc(py / N) within 6 %."""
    rng = numpy.random.default_rng(3)
    N, theta, nv = 256, 0.5, 4000
    lam = N / theta
    r = 0.35 * lam * numpy.sqrt(rng.uniform(0, 1, nv))
    ph = rng.uniform(0, 2 * numpy.pi, nv)
    uvw = numpy.stack([r * numpy.cos(ph), r * numpy.sin(ph), rng.uniform(-240, 240, nv)], 1)
    pix = [(60, 45), (-70, 30), (50, -65)]
    vis = sum(R.simulate_point(uvw, px * theta / N, py * theta / N) for px, py in pix)
    img = G.w_stacking_imaging(theta, lam, uvw, None, vis, 10.0)
    for px, py in pix:
        want = nv / N ** 2 * numpy.sinc(px / N) * numpy.sinc(py / N)
        assert abs(img[N // 2 + py, N // 2 + px] - want) < 0.06 * want


@pytest.mark.parametrize('N,theta,wstep,npk', [(256, 0.5, 10.0, 8), (256, 0.5, 25.0, 48),
                                                (200, 0.4, 9.0, 11)])
def test_subgrid_matches_oracle(gpu, N, theta, wstep, npk):
    print('case wstack subgrid N=%d wstep=%g npixkern=%d' % (N, wstep, npk), flush=True)
    lam, uvw, vis = case(N + 1, N, theta, 5000, 0.45 * N / theta * 0.5, wstep, fill=0.3)
    ref = WS.w_stacking_subgrid_imaging(theta, lam, uvw, None, vis, wstep, npk)
    got = G.w_stacking_subgrid_imaging(theta, lam, uvw, None, vis, wstep, npk)
    assert got.dtype == numpy.complex128 and got.shape == (N, N)
    assert rel(got, ref) < TOL, rel(got, ref)


def test_subgrid_and_full_grid_images_agree(gpu):
    """Both variants image the same sky, up to the sub-grid's truncation of
    each plane's Fresnel spread (|w| theta^2 / 2 = 25 cells here): 19 % of
    the image norm at npixkern = 16, 3 % at 48."""
    lam, uvw, vis = case(9, 256, 0.5, 5000, 200.0, 10.0, fill=0.25, ties=False)
    full = G.w_stacking_imaging(0.5, lam, uvw, None, vis, 10.0)
    err = {}
    for npk in (16, 48):
        sub = G.ifft(G.w_stacking_subgrid_imaging(0.5, lam, uvw, None, vis, 10.0, npk))
        err[npk] = numpy.linalg.norm(sub - full) / numpy.linalg.norm(full)
    assert err[48] < 0.04 and err[16] > 3 * err[48], err


def test_errors_and_empty(gpu):
    N, theta = 64, 0.5
    lam = N / theta
    uvw = numpy.array([[0.0, 0.0, 0.6 * lam]])
    with pytest.raises(AssertionError):
        G.w_stacking_imaging(theta, lam, uvw, None, numpy.ones(1), 10.0)
    wide = numpy.array([[1.0, 2.0, 0.0]])
    with pytest.raises(AssertionError):
        G.w_stacking_imaging(1.5, 64 / 1.5, wide, None, numpy.ones(1), 10.0)
    edge = numpy.array([[0.49 * lam, 0.0, 30.0], [-0.2 * lam, 0.1 * lam, 30.0]])
    with pytest.raises(ValueError):
        G.w_stacking_subgrid_imaging(theta, lam, edge, None, numpy.ones(2), 10.0, 16)
    with pytest.raises(ValueError):
        G.w_stacking_imaging(theta, lam, uvw[:, :2], None, numpy.ones(1), 10.0)
    empty = numpy.zeros((0, 3))
    assert not G.w_stacking_imaging(theta, lam, empty, None, numpy.zeros(0), 10.0).any()
    assert not G.w_stacking_subgrid_imaging(theta, lam, empty, None, numpy.zeros(0), 10.0,
                                            8).any()
