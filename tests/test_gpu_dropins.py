"""Device forms of the reference's shifted / transformed w-kernel tests and
of the host drop-ins for device tensors.

test_grid_shift_w / test_grid_transform_shift_w (tests/test_synthesis.py:
283-356) run with the product's w_kernel (FP64 device generation with dl,
dm, T) and convgrid (complex64 device gridding): every imaged peak must be
1 within 1e-5 (the complex64 tolerance of the path), as the oracle's are
within 1e-7.
"""
import numpy
import pytest
import torch
from numpy.testing import assert_allclose

import crocodile_amd.synthesis as G
from oracle import synthesis_ref as R
from test_oracle import shift_w_peaks, transform_shift_w_peaks

pytestmark = pytest.mark.gpu


def test_grid_shift_w_device(gpu):
    peaks = shift_w_peaks(G)
    print('shift_w: max |peak - 1| = %.3e over %d peaks' % (numpy.abs(peaks - 1).max(),
                                                            len(peaks)), flush=True)
    assert_allclose(peaks, 1, atol=1e-5)


def test_grid_transform_shift_w_device(gpu):
    peaks = transform_shift_w_peaks(G)
    print('transform_shift_w: max |peak - 1| = %.3e over %d peaks'
          % (numpy.abs(peaks - 1).max(), len(peaks)), flush=True)
    assert_allclose(peaks, 1, atol=1e-5)


def test_w_kernel_shift_transform_matches_oracle(gpu):
    """The device w_kernel with dl, dm and T against the oracle's NumPy FFT
    route, at sizes where the kernel is far from a delta."""
    T = numpy.array([[0.9, 0.2], [-0.1, 1.1]])
    for kw in [dict(dl=0.01, dm=-0.02), dict(T=T), dict(T=T, dl=-0.03, dm=0.015)]:
        got = G.w_kernel(0.1, 800.0, 32, 12, 4, **kw)
        ref = R.w_kernel(0.1, 800.0, 32, 12, 4, **kw)
        assert_allclose(got, ref, rtol=0, atol=1e-11 * numpy.abs(ref).max())


def test_doweight_device_hand_built(gpu):
    """The hand-built weights of test_oracle.test_doweight_hand_built through
    the product's doweight (device frac_coords + bincount)."""
    uvw = numpy.array([[0, 0, 0], [1, 1, 0], [-2, 3, 0], [15, 0, 0], [5, 0, 0],
                       [25, 0, 0], [26, 0, 0], [0, 31, 0]], dtype=float)
    want = numpy.array([1 / 3, 1 / 3, 1 / 3, 1 / 2, 1 / 2, 1 / 2, 1 / 2, 1.0])
    assert_allclose(G.doweight(0.1, 100.0, uvw, numpy.ones(8)), want)
    t = G.doweight(0.1, 100.0, torch.as_tensor(uvw, device=gpu),
                   torch.ones(8, dtype=torch.complex128, device=gpu))
    assert_allclose(t.cpu().numpy(), want)
    rng = numpy.random.default_rng(3)
    p = rng.uniform(-40, 40, (3000, 3))
    v = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
    assert_allclose(G.doweight(0.1, 800.0, p, v), R.doweight(0.1, 800.0, p, v), rtol=1e-15)


def test_device_tensor_dropins(gpu):
    """pad_mid / extract_mid / extract_oversampled / kernel_recentre /
    visibility_recentre / sort_vis_w / slice_vis / mirror_uvw keep device
    tensors on the device and equal the oracle."""
    rng = numpy.random.default_rng(9)
    a = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    t = torch.as_tensor(a, device=gpu)
    p = G.pad_mid(t, 11)
    assert p.is_cuda
    assert_allclose(p.cpu().numpy(), R.pad_mid(a, 11))
    assert_allclose(G.extract_mid(p, 6).cpu().numpy(), a)
    big = rng.standard_normal((24, 24)) + 0j
    assert_allclose(G.extract_oversampled(torch.as_tensor(big, device=gpu), 4, 5).cpu().numpy(),
                    R.extract_oversampled(big, 4, 5))
    assert_allclose(G.kernel_recentre(t, 0.1, 30.0, 0.01, -0.02).cpu().numpy(),
                    R.kernel_recentre(a, 0.1, 30.0, 0.01, -0.02), rtol=1e-14)
    uvw = rng.standard_normal((50, 3)) * 100
    ut = torch.as_tensor(uvw, device=gpu)
    assert_allclose(G.visibility_recentre(ut, 0.01, 0.02).cpu().numpy(),
                    R.visibility_recentre(uvw, 0.01, 0.02))
    vis = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    vt = torch.as_tensor(vis, device=gpu)
    ps, vs = G.sort_vis_w(ut, vt)
    rp, rv = R.sort_vis_w(uvw, vis)
    assert_allclose(ps.cpu().numpy(), rp)
    assert_allclose(vs.cpu().numpy(), rv)
    assert_allclose(torch.cat(G.slice_vis(7, ps)).cpu().numpy(), rp)
    mu, mv = G.mirror_uvw(ut, vt)
    ru, rv = R.mirror_uvw(uvw, vis)
    assert mu.is_cuda and mv.is_cuda
    assert_allclose(mu.cpu().numpy(), ru)
    assert_allclose(mv.cpu().numpy(), rv)
