"""The planned gridding / degridding calls enqueue only kernels and event
records on the stream (no host synchronisation, no memset nodes), so they can
be captured in a HIP graph and replayed.  Every replay, with eager work of
the program between replays, must give the eager result bit for bit; and the
tuner's choice must leave results unchanged."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import _lib, tuning

pytestmark = pytest.mark.gpu


def case(gpu, seed=31, N=512, S=16, Q=8, nv=50000, planes=6):
    rng = numpy.random.default_rng(seed)
    tables = [rng.standard_normal((Q, Q, S, S)) + 1j * rng.standard_normal((Q, Q, S, S))
              for _ in range(planes)]
    lim = 0.5 - (S // 2 + 2) / N
    p = numpy.stack([rng.uniform(-lim, lim, nv), rng.uniform(-lim, lim, nv),
                     rng.uniform(-1, 1, nv)], axis=1)
    uvw = torch.as_tensor(p, device=gpu).contiguous()
    vis = torch.as_tensor(rng.standard_normal(nv) + 1j * rng.standard_normal(nv),
                          device=gpu).to(torch.complex64)
    vp = torch.as_tensor(rng.integers(0, planes, nv).astype(numpy.int32), device=gpu)
    return G.KernelBank.from_tables(tables), N, uvw, vis, vp


def test_plan_calls_replay_in_a_graph(gpu):
    bank, N, uvw, vis, vp = case(gpu)
    plan = G.GridPlan(bank, N, uvw, 1.0, _lib.PLANE_INDEX, vp)
    guv = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    out = torch.empty(uvw.shape[0], dtype=torch.complex64, device=gpu)

    def body():
        guv.zero_()
        plan.grid(guv, vis)
        plan.degrid(guv, out)

    body()  # eager: sizes every workspace buffer before capture
    torch.cuda.synchronize()
    ref_g, ref_v = guv.clone(), out.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        body()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    # captured on the warmed-up stream: library contexts are per stream, and
    # the warm-up sized that stream's context workspace
    with torch.cuda.graph(g, stream=s):
        body()
    for rep in range(4):
        guv.fill_(1.0)
        out.fill_(2.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(guv, ref_g), 'grid differs on replay %d' % rep
        assert torch.equal(out, ref_v), 'visibilities differ on replay %d' % rep
        # eager work of the program between replays
        G.grid_device(bank, torch.zeros_like(guv), uvw, 1.0, vis, _lib.PLANE_INDEX, vp)
        (vis * 2).sum().item()


def test_autotune_keeps_results(gpu):
    bank, N, uvw, vis, vp = case(gpu, seed=32, nv=30000)
    plan = G.GridPlan(bank, N, uvw, 1.0, _lib.PLANE_INDEX, vp)
    a = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    plan.grid(a, vis)
    try:
        choice, times = tuning.autotune(bank, N, uvw, 1.0, vis, _lib.PLANE_INDEX, vp,
                                        repeats=1)
        print('autotune choice', choice, {k: round(v, 3) for k, v in times.items()})
        assert choice in times and len(times) >= 3
        plan.bin(uvw, 1.0, _lib.PLANE_INDEX, vp)
        b = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
        plan.grid(b, vis)
        assert torch.equal(a, b)
    finally:
        _lib.set_tuning()


@pytest.mark.parametrize('kind', ['separable-set', 'nearest', '2d'])
def test_wtowers_plan_calls_replay_in_a_graph(gpu, kind):
    """W-towers planned imaging and predict keep their job schedule on the
    device from binning on, so a call enqueues kernels and event records
    only: captured in a graph, every replay equals the eager result bit for
    bit (the matrix-core separable path, nearest cell and 2D kernels)."""
    from test_gpu_wtowers import LAM, MARGIN, SG, THETA, case, kernel_sets, kernels
    rng, uvw, vis = case(21, 20000, 400.0)
    if kind == 'separable-set':
        kv = kernel_sets(rng, 40.0)['separable-set'][0]
    else:
        kv = kernels(rng)[kind]
        kv = None if kv is None else torch.as_tensor(kv, device=gpu).to(torch.complex64)
    uvw_d = torch.as_tensor(uvw, device=gpu)
    vis_d = torch.as_tensor(vis, device=gpu).to(torch.complex64)
    grid_in = torch.as_tensor(rng.standard_normal((512, 512)) +
                              1j * rng.standard_normal((512, 512)), device=gpu).to(torch.complex64)
    plan = G.WTowersPlan(THETA, LAM, uvw_d, 40.0, SG, MARGIN, kv)
    out = {}

    def body():
        out['g'] = plan.imaging(vis_d, kv)
        out['v'] = plan.predict(grid_in, kv)

    body()
    torch.cuda.synchronize()
    ref_g, ref_v = out['g'].clone(), out['v'].clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        body()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    assert torch.equal(out['g'], ref_g) and torch.equal(out['v'], ref_v)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        body()
    for rep in range(3):
        out['g'].fill_(1.0)
        out['v'].fill_(2.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out['g'], ref_g), 'grid differs on replay %d' % rep
        assert torch.equal(out['v'], ref_v), 'visibilities differ on replay %d' % rep
        # eager work of the program between replays (another context)
        G.w_towers_imaging(THETA, LAM, uvw_d, None, vis_d, 40.0, SG, MARGIN, kv)
        (vis_d * 2).sum().item()
