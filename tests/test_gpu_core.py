"""GPU parity of the core path against the FP64 oracle.

Tolerances: integer outputs (frac_coord) bit-exact; FP64 kernel generation
within 1e-11 relative of the NumPy FFT route; complex64 gridding and
degridding within 1e-5 RMS relative (the stated target) and 1e-4 max
relative of the complex128 oracle.
"""
import itertools
import json
import os

import numpy
import pytest
import torch
from numpy.testing import assert_allclose

import crocodile_amd.synthesis as G
from crocodile_amd import _lib
from oracle import cgrid
from oracle import synthesis_ref as R

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')

RMS_TOL = 1e-5
MAX_TOL = 1e-4


def rel_err(x, ref):
    x = numpy.asarray(x)
    ref = numpy.asarray(ref)
    d = x - ref
    rms = numpy.linalg.norm(d) / max(numpy.linalg.norm(ref), 1e-300)
    mx = numpy.abs(d).max() / max(numpy.abs(ref).max(), 1e-300)
    return rms, mx


def assert_close(x, ref, rms_tol=RMS_TOL, max_tol=MAX_TOL, what=''):
    rms, mx = rel_err(x, ref)
    print('%s rms_rel=%.3e max_rel=%.3e' % (what, rms, mx), flush=True)
    assert rms <= rms_tol and mx <= max_tol, (what, rms, mx)


# ---------------------------------------------------------------- frac_coord

def test_frac_coord_golden(gpu):
    g = json.load(open(os.path.join(GOLDEN, 't05_frac_coord.json')))
    p = numpy.array(g['u']) / g['grid_size']
    x, xf = G.frac_coord(g['grid_size'], g['oversample'], p)
    assert x.tolist() == g['x'] and xf.tolist() == g['fx']


@pytest.mark.parametrize('N,Q', [(2, 1), (2, 3), (512, 8), (2048, 8), (4095, 16), (8192, 32)])
def test_frac_coord_bit_exact(gpu, N, Q):
    rng = numpy.random.default_rng(N + Q)
    p = rng.uniform(-0.5, 0.5, 20000)
    # exact ties of -x*Q at every half-integer the grid reaches
    ties = ((numpy.arange(-N * Q, N * Q) + 0.5) / Q - N // 2) / N
    ties = ties[(ties >= -0.5) & (ties < 0.5)]
    p = numpy.concatenate([p, ties, [-0.5, 0.0, numpy.nextafter(0.5, 0)]])
    x, xf = G.frac_coord(N, Q, p)
    rx, rxf = R.frac_coord(N, Q, p)
    assert (x == rx).all() and (xf == rxf).all()


def test_frac_coords_shape_order(gpu):
    rng = numpy.random.default_rng(1)
    p = rng.uniform(-0.5, 0.5, (1000, 3))
    got = G.frac_coords((64, 32), 4, p)
    ref = R.frac_coords((64, 32), 4, p)
    for a, b in zip(got, ref):
        assert (a == b).all()


def test_frac_coord_domain_error(gpu):
    with pytest.raises(AssertionError):
        G.frac_coord(8, 2, numpy.array([0.1, 0.5]))


# ------------------------------------------------------------------ kernels

@pytest.mark.parametrize('theta,w,F,S,Q,kw', [
    (0.1, 0.0, 16, 8, 4, {}),
    (0.1, 137.0, 64, 32, 8, {}),
    (0.1, -1500.0, 64, 32, 8, {}),
    (0.08, 620.0, 33, 16, 16, {}),
    (0.2, 55.5, 20, 9, 3, {'dl': 0.01, 'dm': -0.02}),
    (0.05, 300.0, 24, 12, 2, {'T': numpy.array([[1.0, 0.02], [-0.01, 0.98]])}),
])
@pytest.mark.parametrize('invert', [True, False])
def test_w_kernel_matches_oracle(gpu, theta, w, F, S, Q, kw, invert):
    got = G.w_kernel(theta, w, F, S, Q, invert=invert, **kw)
    ref = R.w_kernel(theta, w, F, S, Q, invert=invert, **kw)
    assert got.shape == ref.shape and got.dtype == numpy.complex128
    assert_allclose(got, ref, rtol=0, atol=1e-11 * numpy.abs(ref).max())


def test_w_kernel_perfect_sampling(gpu):
    for N in range(1, 6):
        for w in [0.0, 3.0, -7.5]:
            assert_allclose(G.w_kernel(N / 1000, w, N, N, 1),
                            R.w_kernel(N / 1000, w, N, N, 1), atol=1e-12)


def test_w_kernel_coordinate_error(gpu):
    with pytest.raises(AssertionError):
        G.w_kernel(2.5, 10.0, 16, 8, 2)


@pytest.mark.parametrize('N,Q,s', [(5, 1, 3), (16, 4, 14), (31, 8, 29), (12, 3, 10)])
def test_kernel_oversample_2d(gpu, N, Q, s):
    pat = R.coordinates2(N)[1] + R.coordinates2(N)[0] * 1j
    assert_allclose(G.kernel_oversample(pat, Q, s), R.kernel_oversample(pat, Q, s),
                    atol=1e-12 * N * N)


def test_kernel_oversample_1d_pswf(gpu):
    aaf = R.anti_aliasing_function(512, 0, 8)
    got = G.kernel_oversample(aaf, 8, 8)
    ref = R.kernel_oversample(aaf, 8, 8)
    assert_allclose(got, ref, atol=1e-12 * numpy.abs(ref).max())


def test_kernel_oversample_subgrid_property(gpu):
    """tests/test_synthesis.py:77-91 on the device implementation."""
    for N in range(3, 20):
        pat = R.coordinates2(N)[1] + R.coordinates2(N)[0] * 1j
        k1 = G.kernel_oversample(pat, 1, N - 2)
        k2 = G.kernel_oversample(pat, 2, N - 2)
        k4 = G.kernel_oversample(pat, 4, N - 2)
        assert_allclose(k1[0, 0], k2[0, 0], atol=1e-12)
        for ux, uy in itertools.product(range(2), range(2)):
            assert_allclose(k2[uy, ux], k4[2 * uy, 2 * ux], atol=1e-12)


# ------------------------------------------------------------ gridding core

def random_case(rng, N, S, Q, nv, planes=1, separable=False, margin=None):
    shape = (Q, S) if separable else (Q, Q, S, S)
    tables = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
              for _ in range(planes)]
    lim = 0.5 - (S // 2 + 2) / N if margin is None else margin
    p = numpy.stack([rng.uniform(-lim, lim, nv), rng.uniform(-lim, lim, nv),
                     rng.uniform(-1, 1, nv)], axis=1)
    v = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    return tables, p, v


CASES = [
    # N, S, Q, nv, planes, separable
    (64, 4, 2, 3000, 1, False),
    (128, 8, 8, 20000, 1, True),
    (256, 16, 4, 20000, 3, False),
    (512, 32, 8, 30000, 2, False),
    (300, 7, 5, 10000, 1, False),
    (200, 24, 3, 15000, 4, True),
    (96, 64, 1, 500, 1, False),
    (2048, 20, 8, 50000, 5, False),
]


@pytest.mark.parametrize('N,S,Q,nv,planes,separable', CASES)
def test_convgrid_convdegrid_parity(gpu, N, S, Q, nv, planes, separable):
    print('case N=%d S=%d Q=%d nv=%d planes=%d sep=%d' % (N, S, Q, nv, planes, separable),
          flush=True)
    rng = numpy.random.default_rng(N * 7 + S)
    tables, p, v = random_case(rng, N, S, Q, nv, planes, separable)
    vplane = rng.integers(0, planes, nv)
    gcf = numpy.array(tables) if planes > 1 else tables[0]
    src = vplane[:, None] if planes > 1 else None
    bank = cgrid.Bank(tables, separable)
    ref = numpy.zeros((N, N), dtype=complex)
    cgrid.convgrid(bank, ref, p, v, vplane.astype(numpy.int32) if planes > 1 else None)
    a = numpy.zeros((N, N), dtype=complex)
    G.convgrid(gcf, a, p, src, v)
    assert_close(a, ref, what='grid')
    vref = cgrid.convdegrid(bank, ref, p, vplane.astype(numpy.int32) if planes > 1 else None)
    vd = G.convdegrid(gcf, ref, p, src)
    assert_close(vd, vref, what='degrid')


def test_every_support_size(gpu):
    """Supports 1..64: the window walk splits S*S taps into unmasked blocks
    of 1024 / 512 / 256 taps and one masked remainder, and finds each tap's
    row with a 24-bit reciprocal multiply; every split and every S is hit."""
    for S in range(1, 65):
        for separable in (False, True):
            rng = numpy.random.default_rng(1000 + 2 * S + separable)
            N, Q, nv = 2 * S + 40, 3, 150
            tables, p, v = random_case(rng, N, S, Q, nv, 1, separable)
            bank = cgrid.Bank(tables, separable)
            ref = numpy.zeros((N, N), dtype=complex)
            cgrid.convgrid(bank, ref, p, v, None)
            a = numpy.zeros((N, N), dtype=complex)
            G.convgrid(tables[0], a, p, None, v)
            assert_close(a, ref, what='grid S=%d sep=%d' % (S, separable))
            vd = G.convdegrid(tables[0], ref, p, None)
            assert_close(vd, cgrid.convdegrid(bank, ref, p, None),
                         what='degrid S=%d sep=%d' % (S, separable))


def test_convgrid_accumulates_in_place(gpu):
    rng = numpy.random.default_rng(2)
    (gcf,), p, v = random_case(rng, 64, 6, 4, 500)
    a = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
    ref = a.copy()
    R.convgrid(gcf, ref, p, None, v)
    G.convgrid(gcf, a, p, None, v)
    assert_close(a, ref, what='in-place')


def test_torch_device_inputs(gpu):
    rng = numpy.random.default_rng(3)
    (gcf,), p, v = random_case(rng, 128, 8, 4, 4000)
    ref = numpy.zeros((128, 128), dtype=complex)
    R.convgrid(gcf, ref, p, None, v)
    a = torch.zeros((128, 128), dtype=torch.complex64, device=gpu)
    G.convgrid(torch.as_tensor(gcf, device=gpu), a, torch.as_tensor(p, device=gpu), None,
               torch.as_tensor(v, device=gpu))
    assert isinstance(a, torch.Tensor)
    assert_close(a.cpu().numpy(), ref, what='torch grid')
    out = G.convdegrid(gcf, a, torch.as_tensor(p, device=gpu))
    assert isinstance(out, torch.Tensor) and out.is_cuda


def test_reference_exact_delta_image(gpu):
    """tests/test_synthesis.py:109-138 through the device path."""
    for lam in [6, 1e15]:
        gcf = numpy.conj(G.w_kernel(1 / lam, 0, 1, 1, 1))
        for N in range(2, 6):
            u, v = R.coordinates2(N)
            uvw = numpy.transpose([numpy.hstack(u), numpy.hstack(v), 0 * numpy.hstack(u)]) * lam
            for x, y in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
                vis = R.simulate_point(uvw, x / lam, y / lam)
                a = numpy.zeros((N, N), dtype=complex)
                G.grid(a, uvw / lam, vis)
                G.convgrid(gcf, a, uvw / lam, None, vis)
                a /= 2
                img = numpy.real(R.ifft(a))
                assert abs(img[N // 2 + y, N // 2 + x] - 1) < 1e-6
                img[N // 2 + y, N // 2 + x] = 0
                assert_allclose(img, 0, atol=1e-6)
                assert_allclose(G.degrid(a, uvw / lam), vis, rtol=1e-6, atol=1e-6)
                assert_allclose(G.convdegrid(gcf, a, uvw / lam), vis, rtol=1e-6, atol=1e-6)


def test_reference_grid_degrid_w(gpu):
    """tests/test_synthesis.py:242-281 through the device path (kernels of
    size N on a 2N grid, perfect sampling)."""
    lam = 1000
    for uw, vw in [(.5, 0), (0, .5), (-1, 0), (0, -1)]:
        for N in range(1, 6):
            u, v = R.coordinates2(N)
            u, v = numpy.hstack(u), numpy.hstack(v)
            uvw_all = numpy.transpose([u, v, uw * u + vw * v]) * lam
            slices = R.slice_vis(N, R.sort_vis_w(uvw_all))
            gcfs = [(G.w_kernel(N / lam, numpy.mean(s[:, 2]), N, N, 1) / N ** 2,
                     G.w_kernel(N / lam, numpy.mean(s[:, 2]), N, N, 1, invert=False) / N ** 2)
                    for s in slices]
            for x, y in itertools.product(range(-(N // 2), (N + 1) // 2), repeat=2):
                img_ref = numpy.zeros((N, N))
                img_ref[N // 2 + y, N // 2 + x] = 1
                a_ref = numpy.fft.ifftshift(R.fft(img_ref))
                a_ref = numpy.vstack(2 * [numpy.hstack(2 * [a_ref])])
                a = numpy.zeros((2 * N, 2 * N), dtype=complex)
                for s, (gcf, gcf_p) in zip(slices, gcfs):
                    vis = R.simulate_point(s, x / lam, y / lam)
                    assert_allclose(G.convdegrid(gcf_p, a_ref, s / lam / 2), vis,
                                    rtol=2e-6, atol=2e-6)
                    G.convgrid(gcf, a, s / lam / 2, None, vis)
                a = numpy.fft.fftshift(a[:N, :N] + a[:N, N:] + a[N:, :N] + a[N:, N:])
                img = numpy.real(R.ifft(a))
                assert_allclose(img[N // 2 + y, N // 2 + x], 1, rtol=1e-5)
                assert_allclose(img, img_ref, atol=2e-3)


def test_window_outside_grid_raises_and_leaves_grid(gpu):
    gcf = numpy.ones((2, 2, 4, 4), dtype=complex)
    a = numpy.zeros((16, 16), dtype=complex)
    p = numpy.array([[0.0, 0.0], [0.49, 0.0]])
    with pytest.raises(ValueError):
        G.convgrid(gcf, a, p, None, numpy.array([1.0 + 0j, 1.0 + 0j]))
    assert (a == 0).all()
    with pytest.raises(ValueError):
        G.convdegrid(gcf, a, p)


def test_domain_error(gpu):
    gcf = numpy.ones((2, 2, 4, 4), dtype=complex)
    a = numpy.zeros((16, 16), dtype=complex)
    with pytest.raises(AssertionError):
        G.convgrid(gcf, a, numpy.array([[0.0, -0.6]]), None, numpy.array([1.0 + 0j]))


def test_non_square_kernel_grid_error(gpu):
    gcf = numpy.ones((2, 2, 4, 6), dtype=complex)
    with pytest.raises(ValueError):
        G.convgrid(gcf, numpy.zeros((32, 32), dtype=complex), numpy.zeros((1, 2)), None,
                   numpy.ones(1, dtype=complex))


def test_empty_input(gpu):
    gcf = numpy.ones((2, 2, 4, 4), dtype=complex)
    a = numpy.zeros((16, 16), dtype=complex)
    G.convgrid(gcf, a, numpy.zeros((0, 3)), None, numpy.zeros(0, dtype=complex))
    assert (a == 0).all()
    assert G.convdegrid(gcf, a, numpy.zeros((0, 3))).shape == (0,)


def test_nearest_grid_degrid(gpu):
    rng = numpy.random.default_rng(9)
    N = 100
    p = rng.uniform(-0.5, 0.49, (5000, 3))
    v = rng.standard_normal(5000) + 1j * rng.standard_normal(5000)
    a = numpy.zeros((N, N), dtype=complex)
    ref = numpy.zeros((N, N), dtype=complex)
    G.grid(a, p, v)
    R.grid(ref, p, v)
    assert_close(a, ref, what='nearest grid')
    assert_close(G.degrid(ref, p), R.degrid(ref, p), what='nearest degrid')


def test_make_grid_hermitian(gpu):
    rng = numpy.random.default_rng(10)
    for N in [2, 5, 8, 33, 64]:
        g = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        ref = R.make_grid_hermitian(g.copy())
        got = G.make_grid_hermitian(g.copy())
        assert_allclose(got, ref, rtol=1e-15, atol=1e-15)
        t = torch.as_tensor(g, device=gpu).to(torch.complex64)
        G.make_grid_hermitian(t)
        assert_close(t.cpu().numpy(), ref, what='hermitian c64')


def test_tiling_knobs_do_not_change_results(gpu):
    rng = numpy.random.default_rng(12)
    (gcf,), p, v = random_case(rng, 256, 12, 4, 20000)
    ref = numpy.zeros((256, 256), dtype=complex)
    R.convgrid(gcf, ref, p, None, v)
    try:
        for tile, chunk in [(16, 64), (48, 1000), (80, 100000)]:
            _lib.set_tuning(tile, chunk)
            a = numpy.zeros((256, 256), dtype=complex)
            G.convgrid(gcf, a, p, None, v)
            assert_close(a, ref, what='tile=%d chunk=%d' % (tile, chunk))
            assert_close(G.convdegrid(gcf, ref, p), R.convdegrid(gcf, ref, p),
                         what='degrid tile=%d' % tile)
    finally:
        _lib.set_tuning()


def test_plane_groups_do_not_change_results(gpu):
    """Binning by (tile, kernel-bank slice): every group size, from one
    group to a group per fraction of a plane, gives the oracle's degridded
    visibilities and bit-identical grids."""
    rng = numpy.random.default_rng(14)
    N, S, Q, nv, planes = 256, 12, 4, 30000, 12
    tables, p, v = random_case(rng, N, S, Q, nv, planes)
    vplane = rng.integers(0, planes, nv).astype(numpy.int32)
    bank = cgrid.Bank(tables, False)
    ref = numpy.zeros((N, N), dtype=complex)
    cgrid.convgrid(bank, ref, p, v, vplane)
    vref = cgrid.convdegrid(bank, ref, p, vplane)
    gcf, src = numpy.array(tables), vplane[:, None]
    grids = []
    try:
        # one table is 144 complex64 (1.1 KiB); a plane is 18 KiB
        for group_kib in (0, 1, 2, 8, 64, 2048):
            _lib.set_tuning(0, 1024, group_kib)
            a = numpy.zeros((N, N), dtype=complex)
            G.convgrid(gcf, a, p, src, v)
            assert_close(a, ref, what='grid group_kib=%d' % group_kib)
            grids.append(a)
            assert_close(G.convdegrid(gcf, ref, p, src), vref,
                         what='degrid group_kib=%d' % group_kib)
    finally:
        _lib.set_tuning()
    for g in grids[1:]:
        assert (g == grids[0]).all()


def test_dense_tile_split_is_exact(gpu):
    """Tiles holding at least 1.5 chunks of visibilities are gridded per
    plane-group range from per-XCD queues, the others whole from a shared
    queue.  With a 64-visibility chunk nearly every tile takes the split
    path, with 4096 none does; with one plane group nothing is split.  Every
    grid is bit-identical, and equal to the oracle's."""
    rng = numpy.random.default_rng(21)
    N, S, Q, nv, planes = 256, 12, 4, 30000, 12
    tables, p, v = random_case(rng, N, S, Q, nv, planes)
    vplane = rng.integers(0, planes, nv).astype(numpy.int32)
    bank = cgrid.Bank(tables, False)
    ref = numpy.zeros((N, N), dtype=complex)
    cgrid.convgrid(bank, ref, p, v, vplane)
    gcf, src = numpy.array(tables), vplane[:, None]
    grids = []
    try:
        for chunk, group_kib in ((4096, 0), (64, 2), (256, 2), (64, 8), (4096, 2)):
            _lib.set_tuning(0, chunk, group_kib)
            a = numpy.zeros((N, N), dtype=complex)
            G.convgrid(gcf, a, p, src, v)
            assert_close(a, ref, what='grid chunk=%d group_kib=%d' % (chunk, group_kib))
            grids.append(a)
    finally:
        _lib.set_tuning()
    for g in grids[1:]:
        assert (g == grids[0]).all()


def test_gridding_is_bit_reproducible(gpu):
    """Fixed-point accumulation is order independent: different tilings,
    chunkings and repeated runs give bit-identical grids."""
    rng = numpy.random.default_rng(13)
    (gcf,), p, v = random_case(rng, 512, 24, 8, 60000)
    grids = []
    try:
        for tile, chunk in [(0, 4096), (16, 64), (48, 999), (0, 4096)]:
            _lib.set_tuning(tile, chunk)
            a = numpy.zeros((512, 512), dtype=complex)
            G.convgrid(gcf, a, p, None, v)
            grids.append(a)
    finally:
        _lib.set_tuning()
    for g in grids[1:]:
        assert (g == grids[0]).all()


def test_gridding_dynamic_range(gpu):
    """Visibilities spanning 12 orders of magnitude keep FP32-level relative
    accuracy on the grid (fixed-point scale chosen from the largest)."""
    rng = numpy.random.default_rng(14)
    (gcf,), p, v = random_case(rng, 256, 8, 4, 5000)
    v = v * 10.0 ** rng.uniform(-12, 0, len(v))
    ref = numpy.zeros((256, 256), dtype=complex)
    cgrid.convgrid(cgrid.Bank([gcf]), ref, p, v)
    a = numpy.zeros((256, 256), dtype=complex)
    G.convgrid(gcf, a, p, None, v)
    assert_close(a, ref, what='dynamic range')


def test_plan_matches_one_shot_calls(gpu):
    """A plan bins once; its gridding is bit-identical to convgrid's and its
    degridding identical to convdegrid's, on repeated calls with new
    visibility values and new kernel values of the same layout."""
    rng = numpy.random.default_rng(21)
    N, S, Q, nv, planes = 512, 16, 8, 40000, 5
    tables, p, v = random_case(rng, N, S, Q, nv, planes)
    bank = G.KernelBank.from_tables(tables)
    bank2 = G.KernelBank.from_tables([t * (1 + 0.5j) for t in tables])
    uvw = torch.as_tensor(p, device=gpu).contiguous()
    vp = torch.as_tensor(rng.integers(0, planes, nv).astype(numpy.int32), device=gpu)
    plan = G.GridPlan(bank, N, uvw, 1.0, _lib.PLANE_INDEX, vp)
    for it, bk in enumerate((bank, bank2, bank)):
        vis = torch.as_tensor(rng.standard_normal(nv) + 1j * rng.standard_normal(nv),
                              device=gpu).to(torch.complex64)
        a = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
        b = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
        G.grid_device(bk, a, uvw, 1.0, vis, _lib.PLANE_INDEX, vp)
        plan.grid(b, vis, bk)
        assert torch.equal(a, b), 'grid differs on call %d' % it
        o1 = torch.empty(nv, dtype=torch.complex64, device=gpu)
        o2 = torch.empty(nv, dtype=torch.complex64, device=gpu)
        G.degrid_device(bk, a, uvw, 1.0, o1, _lib.PLANE_INDEX, vp)
        plan.degrid(a, o2, bk)
        assert torch.equal(o1, o2), 'degrid differs on call %d' % it


def test_plan_errors(gpu):
    rng = numpy.random.default_rng(22)
    (gcf,), p, v = random_case(rng, 128, 8, 4, 1000)
    bank = G.KernelBank.from_tables([gcf])
    bad = p.copy()
    bad[7, 0] = 0.5
    plan = G.GridPlan(bank, 128)
    with pytest.raises(AssertionError):
        plan.bin(torch.as_tensor(bad, device=gpu).contiguous(), 1.0)
    g = torch.zeros((128, 128), dtype=torch.complex64, device=gpu)
    vis = torch.zeros(1000, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):
        plan.grid(g, vis)  # no successful binning
    edge = p.copy()
    edge[3, 1] = 0.5 - 1.0 / 128  # window leaves the grid
    with pytest.raises(ValueError):
        plan.bin(torch.as_tensor(edge, device=gpu).contiguous(), 1.0)
    plan.bin(torch.as_tensor(p, device=gpu).contiguous(), 1.0)
    plan.grid(g, vis)
    other = G.KernelBank.from_tables([numpy.zeros((4, 4, 6, 6), complex)])
    with pytest.raises(ValueError):
        plan.grid(g, vis, other)


@pytest.mark.parametrize('N,S,Q,nv,planes,separable', [CASES[2], CASES[5], CASES[7]])
def test_pervis_kernels_parity(gpu, N, S, Q, nv, planes, separable):
    """The plain per-visibility kernels (croc_convgrid_pervis /
    croc_convdegrid_pervis, the w-towers comparison baseline) against the
    FP64 oracle, with the same kernel selection as the binned path."""
    rng = numpy.random.default_rng(N * 11 + S)
    tables, p, v = random_case(rng, N, S, Q, nv, planes, separable)
    vplane = rng.integers(0, planes, nv).astype(numpy.int32)
    bank_ref = cgrid.Bank(tables, separable)
    ref = numpy.zeros((N, N), dtype=complex)
    cgrid.convgrid(bank_ref, ref, p, v, vplane)
    bank = G.KernelBank.from_tables(tables, separable)
    uvw = torch.as_tensor(p, device=gpu).contiguous()
    vp = torch.as_tensor(vplane, device=gpu)
    vis = torch.as_tensor(v, device=gpu).to(torch.complex64)
    a = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    G.grid_device_pervis(bank, a, uvw, 1.0, vis, _lib.PLANE_INDEX, vp)
    assert_close(a.cpu().numpy(), ref, what='pervis grid')
    g = torch.as_tensor(ref, device=gpu).to(torch.complex64)
    out = torch.empty(nv, dtype=torch.complex64, device=gpu)
    G.degrid_device_pervis(bank, g, uvw, 1.0, out, _lib.PLANE_INDEX, vp)
    assert_close(out.cpu().numpy(), cgrid.convdegrid(bank_ref, ref, p, vplane),
                 what='pervis degrid')
    # errors as convgrid: a window outside the grid leaves the grid unmodified
    bad = uvw.clone()
    bad[3, 0] = 0.5 - 1.0 / N
    before = a.clone()
    with pytest.raises(ValueError):
        G.grid_device_pervis(bank, a, bad, 1.0, vis, _lib.PLANE_INDEX, vp)
    assert torch.equal(a, before)


@pytest.mark.parametrize('gh,gw,src', [(6, 10, False), (11, 4, False), (7, 8, True)])
def test_convdegrid_non_square_kernel(gpu, gh, gw, src):
    """convdegrid with a non-square [Q, Q, gh, gw] table (its window is
    gh rows x gw columns, synthesis.py:497-503), optionally indexed by a
    source column, against the oracle; windows kept off the grid edge."""
    rng = numpy.random.default_rng(gh * 31 + gw)
    N, Q, nv = 128, 4, 3000
    shape = ((3,) if src else ()) + (Q, Q, gh, gw)
    gcf = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    lim = 0.5 - (max(gh, gw) // 2 + 2) / N
    p = rng.uniform(-lim, lim, (nv, 2))
    s = rng.integers(-3, 3, (nv, 1)) if src else None
    a = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    got = G.convdegrid(gcf, a, p, s)
    ref = R.convdegrid(gcf, a, p, s)
    assert_close(got, ref, what='convdegrid %dx%d' % (gh, gw))
    with pytest.raises(ValueError):  # convgrid transposes the window (synthesis.py:477)
        G.convgrid(gcf, numpy.zeros((N, N), complex), p, s, numpy.ones(nv))


@pytest.mark.parametrize('Nh,Nw', [(96, 160), (150, 64)])
@pytest.mark.parametrize('kind', ['separable', '2d', '2d_nonsquare_kernel', 'src'])
def test_non_square_grids(gpu, Nh, Nw, kind):
    """convgrid / convdegrid on an [h, w] grid (frac_coords with per-axis
    sizes, synthesis.py:438-449) against the oracle, with separable, 2D,
    non-square 2D (degrid only) and source-indexed tables; the edge rule
    for windows; grid / degrid (nearest cell, N = h for both axes as
    synthesis.py:398, :413)."""
    rng = numpy.random.default_rng(Nh + Nw + len(kind))
    Q, nv = 8, 4000
    if kind == 'separable':
        gcf, s, gh, gw = rng.standard_normal((Q, 7)) + 0j, None, 7, 7
    elif kind == '2d':
        gcf, s, gh, gw = rng.standard_normal((Q, Q, 8, 8)) + 1j, None, 8, 8
    elif kind == '2d_nonsquare_kernel':
        gcf, s, gh, gw = rng.standard_normal((Q, Q, 5, 9)) + 1j, None, 5, 9
    else:
        gcf = rng.standard_normal((3, Q, Q, 6, 6)) + 1j * rng.standard_normal((3, Q, Q, 6, 6))
        s, gh, gw = rng.integers(0, 3, (nv, 1)), 6, 6
    p = numpy.stack([rng.uniform(-0.5 + (gw // 2 + 2) / Nw, 0.5 - (gw // 2 + 2) / Nw, nv),
                     rng.uniform(-0.5 + (gh // 2 + 2) / Nh, 0.5 - (gh // 2 + 2) / Nh, nv)], 1)
    print('case non-square grid %dx%d %s' % (Nh, Nw, kind), flush=True)
    a = rng.standard_normal((Nh, Nw)) + 1j * rng.standard_normal((Nh, Nw))
    got = G.convdegrid(gcf, a, p, s)
    assert_close(got, R.convdegrid(gcf, a, p, s), what='convdegrid %dx%d %s' % (Nh, Nw, kind))
    v = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    if kind != '2d_nonsquare_kernel':
        g = numpy.zeros((Nh, Nw), complex)
        G.convgrid(gcf, g, p, s, v)
        ref = numpy.zeros((Nh, Nw), complex)
        R.convgrid(gcf, ref, p, s, v)
        assert_close(g, ref, what='convgrid %dx%d %s' % (Nh, Nw, kind))
        # a window over the bottom edge raises and leaves the grid as it was
        bad = p.copy()
        bad[5, 1] = 0.5 - 1.0 / Nh
        before = g.copy()
        with pytest.raises(ValueError):
            G.convgrid(gcf, g, bad, s, v)
        assert (g == before).all()
    # nearest cell
    # x = h//2 + floor(0.5 + h p_x) must index the w columns
    q = numpy.stack([rng.uniform(-0.5, min(0.49, (Nw - Nh // 2 - 1) / Nh), nv),
                     rng.uniform(-0.5, 0.49, nv)], 1)
    gn = numpy.zeros((Nh, Nw), complex)
    G.grid(gn, q, v)
    refn = numpy.zeros((Nh, Nw), complex)
    R.grid(refn, q, v)
    assert_close(gn, refn, what='grid %dx%d' % (Nh, Nw))
    assert_close(G.degrid(a, q), R.degrid(a, q), what='degrid %dx%d' % (Nh, Nw))
