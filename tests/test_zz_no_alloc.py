"""croc_convdegrid_noalloc (the allocation-free degridding entry point)
through the C ABI against croc_convdegrid: the same binning and tile
kernels with scratch from a static device array, so the visibilities must
be bit-identical, on two calls in a row (the second smaller, so stale
scratch contents of the first call would show), and sizes beyond the static
maximum must fail with an error that leaves the output untouched."""
import ctypes

import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import _lib
from test_oracle import random_case

pytestmark = pytest.mark.gpu

MAX_NV = 524288  # CROC_NOALLOC_MAX_NV of include/crocgpu.h


def _call(entry, bank, grid, uvw, out, vp):
    L = _lib.load()
    return getattr(L, entry)(
        _lib.ctx(), _lib.ptr(grid), grid.shape[0], _lib.ptr(uvw), uvw.shape[0], uvw.shape[1],
        1.0, _lib.ptr(out), _lib.ptr(bank.kern), bank.Q, int(bank.separable), bank.nplanes,
        _lib.ptr(bank.plane_off), _lib.ptr(bank.plane_s), bank.max_support,
        _lib.PLANE_INDEX, _lib.ptr(vp), 1.0, 0)


def _case(gpu, seed, N, S, Q, nv, planes, separable):
    rng = numpy.random.default_rng(seed)
    tables, p, v = random_case(rng, N, S, Q, nv, planes, separable)
    bank = G.KernelBank.from_tables(tables, separable)
    g = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    grid = torch.as_tensor(g, device=gpu).to(torch.complex64).contiguous()
    uvw = torch.as_tensor(p, device=gpu).contiguous()
    vp = torch.as_tensor(rng.integers(0, planes, nv).astype(numpy.int32), device=gpu)
    return bank, grid, uvw, vp


@pytest.mark.parametrize('separable', [False, True])
def test_noalloc_matches_convdegrid_two_calls(gpu, separable):
    cases = [(41, 1024, 16, 8, 300000, 4), (42, 512, 8, 4, 20000, 3)]  # second smaller
    for seed, N, S, Q, nv, planes in cases:
        print('noalloc case seed %d N %d S %d nv %d separable %d' % (seed, N, S, nv, separable),
              flush=True)
        bank, grid, uvw, vp = _case(gpu, seed, N, S, Q, nv, planes, separable)
        ref = torch.empty(nv, dtype=torch.complex64, device=gpu)
        got = torch.full((nv,), float('nan'), dtype=torch.complex64, device=gpu)
        _lib.check(_call('croc_convdegrid', bank, grid, uvw, ref, vp))
        _lib.check(_call('croc_convdegrid_noalloc', bank, grid, uvw, got, vp))
        torch.cuda.synchronize()
        assert torch.equal(got, ref), 'noalloc degrid differs (seed %d)' % seed


def test_noalloc_limits_and_errors(gpu):
    bank, grid, uvw, vp = _case(gpu, 43, 256, 8, 4, 1000, 2, False)
    out = torch.full((1000,), 3.0, dtype=torch.complex64, device=gpu)
    # beyond the visibility maximum: a clear error, output untouched
    big = torch.zeros((MAX_NV + 1, 3), dtype=torch.float64, device=gpu)
    bigvp = torch.zeros(MAX_NV + 1, dtype=torch.int32, device=gpu)
    bigout = torch.full((MAX_NV + 1,), 3.0, dtype=torch.complex64, device=gpu)
    rc = _call('croc_convdegrid_noalloc', bank, grid, big, bigout, bigvp)
    assert rc == _lib.EINVAL
    assert b'at most' in _lib.load().croc_last_error()
    assert bool((bigout == 3.0).all())
    # a coordinate outside [-0.5, 0.5): the reference's AssertionError,
    # output untouched
    bad = uvw.clone()
    bad[5, 0] = 0.75
    with pytest.raises(AssertionError):
        _lib.check(_call('croc_convdegrid_noalloc', bank, grid, bad, out, vp))
    assert bool((out == 3.0).all())
    # and the next good call is still right
    ref = torch.empty(1000, dtype=torch.complex64, device=gpu)
    _lib.check(_call('croc_convdegrid', bank, grid, uvw, ref, vp))
    _lib.check(_call('croc_convdegrid_noalloc', bank, grid, uvw, out, vp))
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
