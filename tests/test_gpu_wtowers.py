"""GPU w-towers imaging / predict against the FP64 oracle
(oracle/wtowers_ref.py), nearest-cell and convolutional sub-grid kernels."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import _lib
from oracle import synthesis_ref as R
from oracle import wtowers_ref as WT

from test_gpu_core import assert_close

pytestmark = pytest.mark.gpu

THETA, LAM = 0.1, 5120.0   # N = 512
SG, MARGIN = 128, 32


def case(seed, nv, wmax):
    rng = numpy.random.default_rng(seed)
    r = 0.45 * LAM * numpy.sqrt(rng.uniform(0, 1, nv))
    ph = rng.uniform(0, 2 * numpy.pi, nv)
    uvw = numpy.stack([r * numpy.cos(ph), r * numpy.sin(ph), rng.uniform(-wmax, wmax, nv)], 1)
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    return rng, uvw, vis


def kernels(rng):
    aaf = R.anti_aliasing_function(512, 0, 8)
    sep = R.kernel_oversample(aaf, 16, 8)
    sep /= sep[0].sum()
    k2 = rng.standard_normal((4, 4, 6, 6)) + 1j * rng.standard_normal((4, 4, 6, 6))
    # S = 16, the support of the AW workloads (its own degridding walk)
    k16 = rng.standard_normal((4, 4, 16, 16)) + 1j * rng.standard_normal((4, 4, 16, 16))
    return {'nearest': None, 'separable': sep, '2d': k2, '2d16': k16}


@pytest.mark.parametrize('kind', ['nearest', 'separable', '2d', '2d16'])
def test_wtowers_imaging(gpu, kind):
    rng, uvw, vis = case(1, 6000, 400.0)
    kv = kernels(rng)[kind]
    g = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    ref = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    assert_close(g, ref, what='w_towers_imaging ' + kind)


@pytest.mark.parametrize('kind', ['nearest', 'separable', '2d', '2d16'])
def test_wtowers_predict(gpu, kind):
    rng, uvw, vis = case(2, 6000, 400.0)
    kv = kernels(rng)[kind]
    grid = rng.standard_normal((512, 512)) + 1j * rng.standard_normal((512, 512))
    v = G.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kv)
    ref = WT.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kv)
    assert_close(v, ref, what='w_towers_predict ' + kind)


def test_wtowers_deterministic(gpu):
    rng, uvw, vis = case(3, 4000, 300.0)
    kv = kernels(rng)['separable']
    a = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 30.0, SG, MARGIN, kv)
    b = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 30.0, SG, MARGIN, kv)
    assert (a == b).all()


def test_wtowers_outside_grid(gpu):
    rng, uvw, vis = case(4, 100, 10.0)
    uvw[5, 0] = 2.0 * LAM  # beyond the last chunk
    with pytest.raises((ValueError, AssertionError)):
        G.w_towers_imaging(THETA, LAM, uvw, None, vis, 30.0, SG, MARGIN, None)


def test_wtowers_clips_last_chunk_like_oracle(gpu):
    """A visibility past the grid edge but inside the last chunk is
    gridded on its sub-grid and clipped on the add, as the oracle does."""
    rng, uvw, vis = case(5, 500, 50.0)
    uvw[5, 0] = 0.6 * LAM
    g = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 30.0, SG, MARGIN, None)
    ref = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 30.0, SG, MARGIN, None)
    assert_close(g, ref, what='clipped chunk')


def test_wtowers_job_longer_than_a_part_window(gpu):
    """One (chunk, w-plane) job of 2.2e6 visibilities.  Gridding cuts a job
    into at most 512 parts, so each part here holds ~4300 visibilities, more
    than the 4096-record window the job kernel lists per pass: the kernel
    must walk a part in several windows.  With every w = 0 the towers
    reproduce conv_imaging (the oracle equality of the notebook), here
    against the device conv_imaging, whose parity is tested separately."""
    rng = numpy.random.default_rng(7)
    nv = 2_200_000
    uvw = numpy.stack([rng.uniform(-300, 300, nv), rng.uniform(-300, 300, nv),
                       numpy.zeros(nv)], 1)   # cells 226..286: one 96-cell chunk
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    kv = kernels(rng)['separable']
    g = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    ref = G.conv_imaging(THETA, LAM, uvw, None, vis, kv)
    assert_close(g, ref, what='w-towers long job vs conv_imaging')


@pytest.mark.parametrize('kind', ['nearest', 'separable', '2d'])
def test_wtowers_plan_matches_one_shot(gpu, kind):
    """A plan bins once; its imaging and predict calls equal the one-shot
    calls bit for bit (the same records, in the same order), on repeated
    calls, after re-binning other coordinates and with new kernel values of
    the planned shape."""
    rng, uvw, vis = case(6, 5000, 400.0)
    kv = kernels(rng)[kind]
    grid = rng.standard_normal((512, 512)) + 1j * rng.standard_normal((512, 512))
    plan = G.WTowersPlan(THETA, LAM, uvw, 40.0, SG, MARGIN, kv)
    one_g = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    one_v = G.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kv)
    for _ in range(2):
        assert (plan.imaging(vis, kv) == one_g).all()
        assert (plan.predict(grid, kv) == one_v).all()
    assert_close(plan.imaging(vis, kv),
                 WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv),
                 what='planned w_towers_imaging ' + kind)
    _, uvw2, vis2 = case(7, 3000, 300.0)
    kv2 = None if kv is None else kv * (0.5 + 0.25j)
    plan.bin(uvw2, kv2)
    assert (plan.imaging(vis2, kv2) ==
            G.w_towers_imaging(THETA, LAM, uvw2, None, vis2, 40.0, SG, MARGIN, kv2)).all()
    assert (plan.predict(grid, kv2) ==
            G.w_towers_predict(THETA, LAM, uvw2, None, grid, 40.0, SG, MARGIN, kv2)).all()


def test_wtowers_plan_errors_and_empty(gpu):
    rng, uvw, vis = case(8, 1000, 100.0)
    kv = kernels(rng)['2d']
    plan = G.WTowersPlan(THETA, LAM, uvw, 40.0, SG, MARGIN, kv)
    with pytest.raises(ValueError):
        plan.imaging(vis, None)            # binned for a kernel
    with pytest.raises(ValueError):
        plan.imaging(vis[:10], kv)         # visibility count
    with pytest.raises(ValueError):        # the C ABI refuses a missing kernel too
        _lib.check(_lib.load().croc_wtowers_grid_plan(
            _lib.ctx(), plan._h, _lib.ptr(torch.zeros((512, 512), dtype=torch.complex64,
                                                      device='cuda')),
            _lib.ptr(torch.zeros(1000, dtype=torch.complex64, device='cuda')), None, 0))
    with pytest.raises(ValueError):        # binned for a two-dimensional kernel
        plan.imaging(vis, kernels(rng)['separable'])
    k2 = torch.as_tensor(kv.astype(numpy.complex64), device='cuda')
    with pytest.raises(ValueError):        # the C ABI checks the kernel kind as well
        _lib.check(_lib.load().croc_wtowers_grid_plan(
            _lib.ctx(), plan._h, _lib.ptr(torch.zeros((512, 512), dtype=torch.complex64,
                                                      device='cuda')),
            _lib.ptr(torch.zeros(1000, dtype=torch.complex64, device='cuda')), _lib.ptr(k2), 1))
    bad = uvw.copy()
    bad[3, 0] = 2.0 * LAM                  # beyond the last chunk
    with pytest.raises((ValueError, AssertionError)):
        plan.bin(bad, kv)
    with pytest.raises(ValueError):        # a failed binning leaves nothing to use
        plan.imaging(vis, kv)
    empty = G.WTowersPlan(THETA, LAM, numpy.zeros((0, 3)), 40.0, SG, MARGIN, None)
    assert not empty.imaging(numpy.zeros(0, complex)).any()
    assert empty.predict(numpy.ones((512, 512), complex)).shape == (0,)


def kernel_sets(rng, wstep):
    """A separable residual w-kernel set (PSWF taper times the paraxial
    residual w-term, G.w_towers_kernel_set) and a random 2D set, both
    covering [-wstep/2, wstep/2]; each as the product's WKernelSet and the
    oracle's (tables, w_min, w_step)."""
    aaf = R.anti_aliasing_function(512, 0, 8)
    sep = G.w_towers_kernel_set(THETA, wstep, 512, 8, 16, 2.0, taper=aaf, invert=True)
    sep_ref = (sep.tables.cpu().numpy().astype(complex), sep.w_min, sep.w_step)
    n = 7
    t2 = rng.standard_normal((n, 4, 4, 6, 6)) + 1j * rng.standard_normal((n, 4, 4, 6, 6))
    w_min = -(n - 1) * (wstep / (n - 1)) / 2
    two = G.WKernelSet(t2, w_min, wstep / (n - 1))
    two_ref = (two.tables.cpu().numpy().astype(complex), two.w_min, two.w_step)
    return {'separable-set': (sep, sep_ref), '2d-set': (two, two_ref)}


def test_wtowers_kernel_set_generation(gpu):
    """w_towers_kernel_set: table k is kernel_oversample of taper(l)
    exp(-i pi w_k l^2) (1D, FP64 on the device), against the oracle's
    NumPy FFT route."""
    aaf = R.anti_aliasing_function(256, 0, 8)
    ks = G.w_towers_kernel_set(0.1, 40.0, 256, 8, 8, 4.0, taper=aaf, invert=False)
    assert ks.nplanes == 11 and ks.w_min == -20.0 and ks.w_max == 20.0
    l = R.coordinates(256) * 0.1
    for k in (0, 3, 10):
        w = ks.w_min + k * ks.w_step
        ref = R.kernel_oversample(aaf * numpy.exp(1j * numpy.pi * w * l ** 2), 8, 8)
        got = ks.tables[k].cpu().numpy()
        assert numpy.abs(got - ref).max() <= 1e-6 * numpy.abs(ref).max()


@pytest.mark.parametrize('kind', ['separable-set', '2d-set'])
def test_wtowers_kernel_set_parity(gpu, kind):
    """Imaging and predict with a residual w-kernel set (w_project's table
    choice per visibility) against the FP64 oracle, one-shot and planned."""
    rng, uvw, vis = case(9, 6000, 400.0)
    kset, kref = kernel_sets(rng, 40.0)[kind]
    g = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kset)
    ref = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kref)
    assert_close(g, ref, what='w_towers_imaging ' + kind)
    grid = rng.standard_normal((512, 512)) + 1j * rng.standard_normal((512, 512))
    v = G.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kset)
    vref = WT.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kref)
    assert_close(v, vref, what='w_towers_predict ' + kind)
    plan = G.WTowersPlan(THETA, LAM, uvw, 40.0, SG, MARGIN, kset)
    assert (plan.imaging(vis, kset) == g).all()
    assert (plan.predict(grid, kset) == v).all()
    with pytest.raises(ValueError):  # another set layout than the planned one
        plan.imaging(vis, G.WKernelSet(kset.tables, kset.w_min - 1.0, kset.w_step))


def test_wtowers_kernel_set_must_cover_increment(gpu):
    rng, uvw, vis = case(10, 500, 100.0)
    kset, _ = kernel_sets(rng, 40.0)['2d-set']
    narrow = G.WKernelSet(kset.tables, kset.w_min / 2, kset.w_step / 2)
    with pytest.raises(ValueError):
        G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, narrow)


def boundary_jobs(seed):
    """One chunk (cells around 240 of 512: chunk 2 of 96) whose w-planes
    hold jobs of exactly 64, 65, 128, 129, 1025 and 4097 visibilities: the
    small-job staging (64 records at a time), the small / matrix-core split
    (128), the gridding window (1024) and the part size (4096) on both
    sides of each bound; plus a one-visibility job and a second chunk."""
    rng = numpy.random.default_rng(seed)
    sizes = [64, 65, 128, 129, 1025, 4097, 1]
    cells, ws = [], []
    for plane, n in zip(range(-3, 4), sizes):
        cells.append(rng.uniform(-30, 30, (n, 2)) + 240.0)
        ws.append(plane * 40.0 + rng.uniform(-15, 15, n))
    cells.append(rng.uniform(-10, 10, (3, 2)) + numpy.array([330.0, 240.0]))  # chunk 3
    ws.append(rng.uniform(-15, 15, 3))
    cell = numpy.concatenate(cells)
    uvw = numpy.concatenate([(cell - 256.0) / THETA, numpy.concatenate(ws)[:, None]], 1)
    vis = rng.standard_normal(len(uvw)) + 1j * rng.standard_normal(len(uvw))
    return rng, uvw, vis


@pytest.mark.parametrize('support', [8, 16])
def test_wtowers_job_size_bounds(gpu, support):
    """Imaging and predict at the job-size bounds of the gridding and summing
    kernels and the predict's run splitting (this set's runs cut every
    job), against the oracle, with separable kernels of support 8 (the
    generic-support code) and 16 (the fixed-support code)."""
    print('case job size bounds S=%d' % support, flush=True)
    rng, uvw, vis = boundary_jobs(30 + support)
    aaf = R.anti_aliasing_function(512, 0, support)
    kv = R.kernel_oversample(aaf, 8, support)
    kv /= kv[0].sum()
    g = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    ref = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    assert_close(g, ref, what='w_towers_imaging job size bounds S=%d' % support)
    grid = rng.standard_normal((512, 512)) + 1j * rng.standard_normal((512, 512))
    v = G.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kv)
    vref = WT.w_towers_predict(THETA, LAM, uvw, None, grid, 40.0, SG, MARGIN, kv)
    assert_close(v, vref, what='w_towers_predict job size bounds S=%d' % support)
    plan = G.WTowersPlan(THETA, LAM, uvw, 40.0, SG, MARGIN, kv)
    assert (plan.imaging(vis, kv) == g).all() and (plan.predict(grid, kv) == v).all()
