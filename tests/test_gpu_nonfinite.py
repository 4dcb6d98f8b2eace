"""NaN and infinite inputs through the fixed-point gridders, against the
FP64 oracle's propagation (NumPy arithmetic as synthesis.py:479-480 and the
w-towers notebook do it): the cells the reference leaves non-finite are NaN
here, every other cell keeps its value.  Exact fixed point cannot hold
them, so the device excludes them from the scale and marks their cells
afterwards (tile path: the window of a non-finite visibility, the cell of a
non-finite tap; w-towers: the chunk's sub-grid region)."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import _lib
from oracle import synthesis_ref as R
from oracle import wtowers_ref as WT

pytestmark = pytest.mark.gpu


def compare(got, ref):
    bad_ref = ~numpy.isfinite(ref)
    assert bad_ref.any()
    assert numpy.array_equal(numpy.isnan(got), bad_ref)
    assert numpy.isfinite(got[~bad_ref]).all()
    fin = ref[~bad_ref]
    err = numpy.abs(got[~bad_ref] - fin).max() / numpy.abs(fin).max()
    assert err < 1e-6, err


def points(rng, nv, N, S):
    lim = 0.5 - (S // 2 + 2) / N
    return numpy.stack([rng.uniform(-lim, lim, nv), rng.uniform(-lim, lim, nv),
                        numpy.zeros(nv)], 1)


@pytest.mark.parametrize('separable', [False, True])
def test_tile_path_nonfinite_visibilities(gpu, separable):
    print('case nonfinite vis separable=%s' % separable, flush=True)
    rng = numpy.random.default_rng(11)
    N, S, Q, nv = 128, 8, 4, 3000
    shape = (Q, S) if separable else (Q, Q, S, S)
    kern = rng.standard_normal(shape) + 1j * rng.standard_normal(shape) + 0.1
    p = points(rng, nv, N, S)
    v = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    v[3] = numpy.nan
    v[7] = numpy.inf
    v[11] = complex(0.0, -numpy.inf)
    ref = numpy.zeros((N, N), complex)
    R.convgrid(kern, ref, p, None, v)
    got = numpy.zeros((N, N), complex)
    G.convgrid(kern, got, p, None, v)
    compare(got, ref)
    # the planned path (one device kernel bank, binned once)
    bank = G.KernelBank.from_tables([kern], separable=separable)
    plan = G.GridPlan(bank, N, torch.as_tensor(p, device=gpu), 1.0)
    g = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    plan.grid(g, torch.as_tensor(v, device=gpu).to(torch.complex64))
    compare(g.cpu().numpy().astype(complex), ref)


def test_tile_path_nonfinite_tap(gpu):
    """A NaN tap in one (yf, xf) table: the cells of exactly the visibilities
    using that table, at that tap."""
    rng = numpy.random.default_rng(12)
    N, S, Q, nv = 128, 6, 4, 4000
    kern = rng.standard_normal((Q, Q, S, S)) + 1j * rng.standard_normal((Q, Q, S, S)) + 0.1
    kern[1, 2, 3, 4] = numpy.nan
    p = points(rng, nv, N, S)
    v = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    ref = numpy.zeros((N, N), complex)
    R.convgrid(kern, ref, p, None, v)
    got = numpy.zeros((N, N), complex)
    G.convgrid(kern, got, p, None, v)
    compare(got, ref)


@pytest.mark.parametrize('kind', ['nearest', 'separable'])
def test_wtowers_nonfinite_visibility(gpu, kind):
    """W-towers: the chunk of a NaN or infinite visibility turns NaN over its
    whole sub-grid region, as the reference's transforms make it."""
    print('case nonfinite w-towers %s' % kind, flush=True)
    from test_gpu_wtowers import LAM, MARGIN, SG, THETA, case, kernels
    rng, uvw, vis = case(13, 4000, 300.0)
    kv = kernels(rng)[kind]
    vis[10] = numpy.nan
    vis[20] = numpy.inf
    ref = WT.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    got = G.w_towers_imaging(THETA, LAM, uvw, None, vis, 40.0, SG, MARGIN, kv)
    bad_ref = ~numpy.isfinite(ref)
    assert bad_ref.any() and not bad_ref.all()
    assert numpy.array_equal(numpy.isnan(got), bad_ref)
    fin = ref[~bad_ref]
    assert numpy.abs(got[~bad_ref] - fin).max() < 1e-5 * numpy.abs(fin).max()
