"""AW-towers imaging on the device (crocodile_amd.synthesis.aw_towers_imaging)
against the FP64 oracle (oracle/awtowers_ref.py, awtowers.ipynb cells 17,
19, 22), which test_oracle_awtowers.py pins against the A-corrected direct
Fourier image.  The device pre-grids in complex64 with FP32 atomics and
transforms in complex64 FFTs: the bound is 2e-6 of the grid's largest
value."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from oracle import awtowers_ref as AWT
from oracle import synthesis_ref as R
from test_oracle_awtowers import compare, scenario

pytestmark = pytest.mark.gpu

TOL = 2e-6


def both(sc, wstep, dl=0.0, dm=0.0, **kw):
    wk = R.w_kernel_function(sc['l'], sc['m'], wstep)
    args = (sc['theta'], sc['lam'], sc['uvw'], sc['src'], sc['vis'], wstep, sc['S'],
            sc['margin'], sc['aa_gcf'], sc['aa_sub'], sc['aa_sub2'], sc['aimage'], wk)
    ref = AWT.aw_towers_imaging(*args, dl=dl, dm=dm)
    got = G.aw_towers_imaging(*args, dl=dl, dm=dm, **kw)
    return got, ref


def rel(got, ref):
    return numpy.abs(got - ref).max() / numpy.abs(ref).max()


@pytest.mark.parametrize('S,margin,wstep,budget', [(64, 16, 2.0, 4 << 30),
                                                    (64, 15, 5.0, 4 << 30),
                                                    (64, 16, 2.0, 37 * 128 * 128 * 8),
                                                    (32, 6, 3.0, 4 << 30)])
def test_matches_oracle(gpu, S, margin, wstep, budget):
    """Even and odd chunk widths; groups split into batches of 37 (batches
    that cut through chunks)."""
    print('case awtowers S=%d margin=%d wstep=%g budget=%d' % (S, margin, wstep, budget),
          flush=True)
    sc = scenario(S=S, margin=margin)
    got, ref = both(sc, wstep, max_stack_bytes=budget)
    assert got.shape == ref.shape and got.dtype == numpy.complex128
    assert rel(got, ref) < TOL, rel(got, ref)


def test_shifted_field_matches_oracle(gpu):
    print('case awtowers dl dm', flush=True)
    sc = scenario(seed=2)
    got, ref = both(sc, 2.0, dl=0.02, dm=-0.015)
    assert rel(got, ref) < TOL, rel(got, ref)


def test_physics_on_device(gpu):
    """The device grid's image against the A-corrected DFT image: scale 4
    and residual < 1 %."""
    sc = scenario(seed=4, nant=8, ndump=30)
    wk = R.w_kernel_function(sc['l'], sc['m'], 2.0)
    g = G.aw_towers_imaging(sc['theta'], sc['lam'], sc['uvw'], sc['src'], sc['vis'], 2.0,
                            sc['S'], sc['margin'], sc['aa_gcf'], sc['aa_sub'], sc['aa_sub2'],
                            sc['aimage'], wk)
    c, res = compare(sc, R.ifft(g) / numpy.outer(sc['aa'], sc['aa']))
    assert abs(c - 4) < 0.01 and res < 0.01, (c, res)


def test_torch_in_torch_out(gpu):
    sc = scenario(seed=1)
    wk = R.w_kernel_function(sc['l'], sc['m'], 2.0)
    ref = AWT.aw_towers_imaging(sc['theta'], sc['lam'], sc['uvw'], sc['src'], sc['vis'], 2.0,
                                sc['S'], sc['margin'], sc['aa_gcf'], sc['aa_sub'],
                                sc['aa_sub2'], sc['aimage'], wk)
    d = lambda a: torch.as_tensor(a, device='cuda')  # noqa: E731
    got = G.aw_towers_imaging(sc['theta'], sc['lam'], d(sc['uvw']), d(sc['src']), d(sc['vis']),
                              2.0, sc['S'], sc['margin'], d(sc['aa_gcf']), sc['aa_sub'],
                              sc['aa_sub2'], d(sc['aimage']), d(wk))
    assert isinstance(got, torch.Tensor) and got.is_cuda
    assert rel(got.cpu().numpy(), ref) < TOL


def test_errors(gpu):
    sc = scenario()
    wk = R.w_kernel_function(sc['l'], sc['m'], 2.0)
    args = [sc['theta'], sc['lam'], sc['uvw'], sc['src'], sc['vis'], 2.0, sc['S'],
            sc['margin'], sc['aa_gcf'], sc['aa_sub'], sc['aa_sub2'], sc['aimage'], wk]
    bad = list(args)
    bad[3] = sc['src'] + 10
    with pytest.raises(IndexError):
        G.aw_towers_imaging(*bad)
    bad = list(args)
    bad[6] = 63          # odd sub-grid
    with pytest.raises(ValueError):
        G.aw_towers_imaging(*bad)
    empty = G.aw_towers_imaging(sc['theta'], sc['lam'], numpy.zeros((0, 3)),
                                numpy.zeros((0, 2), dtype=int), numpy.zeros(0), *args[5:])
    assert empty.shape == (sc['N'], sc['N']) and not empty.any()
