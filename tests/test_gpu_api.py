"""Argument checking of the device entry points and per-stream contexts.

The C ABI takes raw device pointers and trusts the sizes it is given, so
the Python layer checks dtypes, devices and lengths before each call
(ValueError otherwise).  Library contexts are kept per (device, stream):
calls queued on two torch streams do not share scratch buffers.
"""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import _lib
from test_oracle import random_case

pytestmark = pytest.mark.gpu


def _case(gpu, N=256, S=8, Q=4, nv=5000, seed=31):
    rng = numpy.random.default_rng(seed)
    tables, p, v = random_case(rng, N, S, Q, nv, planes=2)
    bank = G.KernelBank.from_tables(tables)
    uvw = torch.as_tensor(p, device=gpu).contiguous()
    vis = torch.as_tensor(v, device=gpu).to(torch.complex64)
    vp = torch.as_tensor(rng.integers(0, 2, nv).astype(numpy.int32), device=gpu)
    return bank, uvw, vis, vp


def test_wrong_dtypes_and_shapes_rejected(gpu):
    N = 256
    bank, uvw, vis, vp = _case(gpu, N)
    g = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    nv = uvw.shape[0]
    bad_calls = [
        dict(guv=g.to(torch.complex128)),                        # grid dtype
        dict(guv=torch.zeros((N, N // 2), dtype=torch.complex64, device=gpu)),
        dict(guv=g.cpu()),                                       # grid on the host
        dict(uvw=uvw.to(torch.float32)),                         # uvw dtype
        dict(uvw=uvw[:, :1].contiguous()),                       # no v column
        dict(vis=vis[:-1]),                                      # too short
        dict(vis=vis.to(torch.complex128)),                      # vis dtype
        dict(vplane=vp.to(torch.int64)),                         # vplane dtype
        dict(vplane=vp[:10]),                                    # vplane length
        dict(vplane=None),                                       # missing vplane
    ]
    for fn in (G.grid_device, G.degrid_device):
        for bad in bad_calls:
            kw = dict(guv=g, uvw=uvw, vis=vis if fn is G.grid_device else
                      torch.empty(nv, dtype=torch.complex64, device=gpu), vplane=vp)
            if 'vis' in bad and fn is G.degrid_device:
                bad = dict(vis=bad['vis'])
            kw.update(bad)
            with pytest.raises(ValueError):
                fn(bank, kw['guv'], kw['uvw'], 1.0, kw['vis'], _lib.PLANE_INDEX, kw['vplane'])
    # a complex128 bank is converted, not misread
    b128 = G.KernelBank(bank.kern.to(torch.complex128), bank.plane_off_h, bank.plane_s_h,
                        bank.Q, bank.separable)
    g1 = torch.zeros_like(g)
    g2 = torch.zeros_like(g)
    G.grid_device(bank, g1, uvw, 1.0, vis, _lib.PLANE_INDEX, vp)
    G.grid_device(b128, g2, uvw, 1.0, vis, _lib.PLANE_INDEX, vp)
    assert torch.equal(g1, g2)
    plan = G.GridPlan(bank, N, uvw, 1.0, _lib.PLANE_INDEX, vp)
    with pytest.raises(ValueError):
        plan.grid(g, vis[:-1])
    with pytest.raises(ValueError):
        plan.degrid(g, torch.empty(nv, dtype=torch.complex128, device=gpu))
    with pytest.raises(ValueError):
        plan.bin(uvw.to(torch.float32), 1.0, _lib.PLANE_INDEX, vp)


def test_two_streams_match_one_stream(gpu):
    """Gridding two visibility sets on two torch streams at once gives the
    grids of running them one after the other on the default stream."""
    N = 256
    bank, uvw, vis, vp = _case(gpu, N)
    vis2 = (vis * (0.5 - 2j)).contiguous()
    ref1 = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    ref2 = torch.zeros_like(ref1)
    G.grid_device(bank, ref1, uvw, 1.0, vis, _lib.PLANE_INDEX, vp)
    G.grid_device(bank, ref2, uvw, 1.0, vis2, _lib.PLANE_INDEX, vp)
    torch.cuda.synchronize()
    # planned calls only enqueue work (no host synchronisation), so the two
    # streams' gridding kernels really overlap
    # (one plan per stream: a plan's queue counters are its own)
    plan1 = G.GridPlan(bank, N, uvw, 1.0, _lib.PLANE_INDEX, vp)
    plan2 = G.GridPlan(bank, N, uvw, 1.0, _lib.PLANE_INDEX, vp)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        a1 = torch.zeros_like(ref1)
        a2 = torch.zeros_like(ref1)
        torch.cuda.synchronize()
        with torch.cuda.stream(s1):
            plan1.grid(a1, vis)
        with torch.cuda.stream(s2):
            plan2.grid(a2, vis2)
        torch.cuda.synchronize()
        assert torch.equal(a1, ref1) and torch.equal(a2, ref2)
    with torch.cuda.stream(s1):
        c1 = _lib.ctx()
    with torch.cuda.stream(s2):
        c2 = _lib.ctx()
    assert c1.value != c2.value
