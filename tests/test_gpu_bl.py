"""GPU parity of the baseline-layout entry points (crocodile_amd.bl over
croc_bl_* / croc_degrid_conv* / croc_grid_wprojection / croc_streamer_degrid)
against the C restatement of grid.c / streamer.c (oracle/blgrid.c).

Tolerances: selections, counts, untouched entries and zero slots exact;
nearest-cell gathers exact; complex64 values within 1e-5 RMS relative and
1e-4 max relative of the FP64 oracle fed the same complex64 inputs.
"""
import numpy
import pytest
import torch
from numpy.testing import assert_array_equal

from crocodile_amd import bl as BL
from oracle import blgrid as B
from oracle import synthesis_ref as R

from test_gpu_core import assert_close
from test_oracle_bl import bl_tracks, cut, lam, tiles

pytestmark = pytest.mark.gpu


def c64(a):
    return numpy.asarray(a).astype(numpy.complex64).astype(numpy.complex128)


def make_set(rng, nbl, nt, nf, umax, ragged=False, with_vis=True):
    uvw, freq = bl_tracks(rng, nbl, nt, nf, umax, ragged=ragged)
    vis = None
    if with_vis:
        vis = [c64(rng.normal(size=len(a) * len(f)) + 1j * rng.normal(size=len(a) * len(f)))
               for a, f in zip(uvw, freq)]
    host = B.BlSet(uvw, freq, vis)
    dev = BL.VisData([BL.BlData(0, b + 1, numpy.arange(len(u)), f, u, host.baseline_vis(b))
                      for b, (u, f) in enumerate(zip(uvw, freq))])
    return host, dev


def rand_grid(rng, N):
    g = c64(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    return g, torch.as_tensor(g.astype(numpy.complex64)).cuda()


@pytest.mark.parametrize('S,Q', [(4, 16), (7, 8), (8, 1024), (12, 32), (16, 64), (30, 4),
                                 (64, 2)])
def test_degrid_conv_uv(gpu, S, Q):
    print('case degrid_conv_uv S=%d Q=%d' % (S, Q), flush=True)
    rng = numpy.random.default_rng(S + Q)
    N, theta = 256, 0.02
    g, gd = rand_grid(rng, N)
    kern = rng.normal(size=(Q, S))
    lim = (N / 2 - S) / theta
    uv = rng.uniform(-lim, lim, (5000, 2))
    fl = []
    got = BL.degrid_conv_uv(gd, N, theta, uv[:, 0], uv[:, 1], BL.SepKernelData(kern), flops=fl)
    ref = B.degrid_conv_uv(g, theta, uv, kern)
    assert_close(got.cpu().numpy(), ref, what='degrid_conv_uv S=%d Q=%d' % (S, Q))
    assert fl == [5000 * 4 * (1 + S) * S]


@pytest.mark.parametrize('ragged', [False, True])
def test_degrid_conv_bl(gpu, ragged):
    print('case degrid_conv_bl ragged=%s' % ragged, flush=True)
    rng = numpy.random.default_rng(21 + ragged)
    N, theta, S, Q = 128, 0.01, 8, 128
    g, gd = rand_grid(rng, N)
    kern = rng.normal(size=(Q, S))
    host, dev = make_set(rng, 40, 24, 40, 9000.0, ragged=False)
    if ragged:  # per-baseline counts differ; the window fits them all
        uvw, freq = bl_tracks(rng, 40, 30, 50, 9000.0, ragged=True)
        uvw = [numpy.vstack([u, u[-1] + numpy.zeros((24, 3))])[:max(24, len(u))] for u in uvw]
        freq = [numpy.concatenate([f, f[-1] + 1e6 * numpy.arange(1, 41)])[:max(40, len(f))]
                for f in freq]
        host = B.BlSet(uvw, freq)
        dev = BL.VisData([BL.BlData(0, 1, numpy.arange(len(u)), f, u) for u, f in zip(uvw, freq)])
    sentinel = 3 - 2j
    host.vis[:] = sentinel
    dev.vis[:] = sentinel
    box = (-1500.0, 1800.0, -1700.0, 1600.0)
    n = B.degrid_conv_bl(g, theta, 120.0, -80.0, *box, host, 3, 21, 5, 37, kern)
    fl = BL.degrid_conv_bl(gd, N, theta, 120.0, -80.0, *box, dev, 3, 21, 5, 37,
                           BL.SepKernelData(kern))
    got = dev.vis.cpu().numpy()
    untouched = host.vis == sentinel
    assert 0 < n < len(host.vis) and untouched.sum() > 0
    assert fl == n * 4 * (1 + S) * S
    assert_array_equal(got[untouched], numpy.complex64(sentinel))
    assert_array_equal(got == numpy.complex64(sentinel), untouched)
    assert_close(got[~untouched], host.vis[~untouched], what='degrid_conv_bl')


def test_degrid_conv(gpu):
    print('case degrid_conv', flush=True)
    rng = numpy.random.default_rng(5)
    N, theta, S, Q = 128, 0.01, 10, 64
    g, gd = rand_grid(rng, N)
    kern = rng.normal(size=(Q, S))
    uvw, freq = bl_tracks(rng, 30, 20, 30, 5000.0, ragged=True)
    # one baseline wholly beyond the grid's window (left untouched)
    uvw.append(numpy.array([[20000.0, 100.0, 0.0]] * 7))
    freq.append(freq[0])
    host = B.BlSet(uvw, freq)
    dev = BL.VisData([BL.BlData(0, 1, numpy.arange(len(u)), f, u) for u, f in zip(uvw, freq)])
    host.vis[:] = 9.0
    dev.vis[:] = 9.0
    h = N / theta / 2
    n = B.degrid_conv_bl(g, theta, 0.0, 0.0, -h, h, -h, h, host, 0, 0, 0, 0, kern)
    assert n == 0
    # degrid_conv walks every baseline's full extent: restate per baseline
    for b in range(host.nbl):
        one = B.BlSet([host.uvw_m[host.uvw0[b]:host.uvw0[b] + host.nt[b]]],
                      [host.freq[host.f0[b]:host.f0[b] + host.nf[b]]], [host.baseline_vis(b)])
        B.degrid_conv_bl(g, theta, 0.0, 0.0, -h, h, -h, h, one, 0, int(host.nt[b]), 0,
                         int(host.nf[b]), kern)
        host.baseline_vis(b)[:] = one.vis
    BL.degrid_conv(gd, N, theta, 0.0, 0.0, dev, BL.SepKernelData(kern))
    got = dev.vis.cpu().numpy()
    out = host.vis == 9.0
    assert 0 < out.sum() < len(out)
    assert_array_equal(got == numpy.complex64(9.0), out)
    assert_close(got[~out], host.vis[~out], what='degrid_conv')


def test_streamer_degrid_worker(gpu):
    print('case streamer_degrid_worker', flush=True)
    rng = numpy.random.default_rng(12)
    N, theta, S, Q, xA, xM = 512, 0.05, 8, 256, 64, 128
    g, _ = rand_grid(rng, N)
    kern = rng.normal(size=(Q, S))
    host, dev = make_set(rng, 60, 16, 48, 12000.0, with_vis=False)
    tc, fc = 8, 16
    chunks = numpy.array([[b, t, t + tc, f, f + fc] for b in range(host.nbl)
                          for t in range(0, 16, tc) for f in range(0, 48, fc)], dtype=numpy.int32)
    total = 0
    for off_u, off_v in [(0, 0), (64, -128), (-64, 64), (128, 0)]:
        sub = cut(g, off_u, off_v, xM)
        box = ((off_u - xA // 2) / theta, (off_u + xA // 2) / theta,
               (off_v - xA // 2) / theta, (off_v + xA // 2) / theta)
        ref, nd_ref = B.streamer_degrid(sub, theta, off_u / theta, off_v / theta, *box, host,
                                        chunks, tc, fc, kern)
        subd = torch.as_tensor(numpy.ascontiguousarray(sub).astype(numpy.complex64)).cuda()
        q, nd, fl = BL.streamer_degrid_worker(dev, subd, xM, theta, off_u / theta,
                                              off_v / theta, chunks, *box,
                                              BL.SepKernelData(kern), tc, fc)
        q = q.cpu().numpy()
        assert_array_equal(nd.cpu().numpy(), nd_ref)
        assert fl == int(nd_ref.sum()) * 4 * (1 + S) * S
        zero = ref == 0
        assert_array_equal(q == 0, zero)
        total += int(nd_ref.sum())
        if nd_ref.sum():
            assert_close(q[~zero], ref[~zero], what='streamer slots off=(%d,%d)' % (off_u, off_v))
    assert total > 200


def test_streamer_tiling_matches_full_grid(gpu):
    """Streamer over a tiling of subgrids cut from one grid, chunk selection
    by bounding boxes, slots merged by the writer: every visibility inside
    the tiled area equals degrid_conv_uv on the full grid (conjugated at
    (-u, -v) for u < 0), all others stay zero."""
    print('case streamer tiling', flush=True)
    rng = numpy.random.default_rng(13)
    N, theta, S, Q, xA, xM = 512, 0.05, 8, 64, 64, 128
    g, _ = rand_grid(rng, N)
    kern = rng.normal(size=(Q, S))
    host, dev = make_set(rng, 80, 32, 64, 14000.0, with_vis=False)
    st = BL.Streamer(dev, theta, xA, xM, 16, 32, BL.SepKernelData(kern))
    nsel = 0
    for off_u, off_v in tiles(N, xA, xM):
        sub = torch.as_tensor(numpy.ascontiguousarray(cut(g, off_u, off_v, xM))
                              .astype(numpy.complex64)).cuda()
        ch, q, nd, _ = st.degrid(sub, off_u, off_v)
        nsel += ch.shape[0]
        st.write(ch, q)
    print('chunks degridded: %d of %d (subgrid, chunk) pairs' % (
        nsel, len(tiles(N, xA, xM)) * st._chunks.shape[0]), flush=True)
    assert nsel < len(tiles(N, xA, xM)) * st._chunks.shape[0] // 2  # boxes prune
    got = dev.vis.cpu().numpy()
    u, v = lam(host, 0), lam(host, 1)
    k = (N - xM) // 2 // xA
    edge = (k * xA + xA // 2) / theta
    inside = (numpy.abs(u) < edge - 1) & (numpy.abs(v) < edge - 1) & (u != 0)
    outside = (numpy.abs(u) > edge + 1) | (numpy.abs(v) > edge + 1)
    assert inside.sum() > 1000 and outside.sum() > 1000
    uv = numpy.stack([u, v], axis=1)
    want = numpy.zeros(len(u), dtype=complex)
    pos, neg = inside & (u > 0), inside & (u < 0)
    want[pos] = B.degrid_conv_uv(g, theta, uv[pos], kern)
    want[neg] = numpy.conj(B.degrid_conv_uv(g, theta, -uv[neg], kern))
    assert (got[outside] == 0).all()
    assert_close(got[inside], want[inside], what='streamer tiling')


def test_streamer_predicts_point_sources(gpu):
    """The streamer's own spot check (streamer.c:444-464): subgrids of the
    grid fft(image / aa x aa) degridded with the PSWF table of aa give the
    direct Fourier sum of the image's point sources."""
    print('case streamer DFT check', flush=True)
    N, theta, S, Q, xA, xM = 512, 0.04, 10, 1024, 64, 128
    aa = R.anti_aliasing_function(N, 0, S)
    k = R.kernel_oversample(aa, Q, S).real
    kern = k / k[0].sum()
    rng = numpy.random.default_rng(4)
    pix = rng.integers(-N // 8, N // 8, (6, 2))
    img = numpy.zeros((N, N))
    for il, im in pix:
        img[N // 2 + im, N // 2 + il] += 1.0
    g = R.fft(img / numpy.outer(aa, aa))
    host, dev = make_set(rng, 40, 16, 32, 9000.0, with_vis=False)
    st = BL.Streamer(dev, theta, xA, xM, 8, 16, BL.SepKernelData(kern))
    for off_u, off_v in tiles(N, xA, xM):
        sub = torch.as_tensor(numpy.ascontiguousarray(cut(g, off_u, off_v, xM))
                              .astype(numpy.complex64)).cuda()
        ch, q, _, _ = st.degrid(sub, off_u, off_v)
        st.write(ch, q)
    got = dev.vis.cpu().numpy()
    u, v = lam(host, 0), lam(host, 1)
    k_ = (N - xM) // 2 // xA
    edge = (k_ * xA + xA // 2) / theta
    inside = (numpy.abs(u) < edge - 1) & (numpy.abs(v) < edge - 1) & (u != 0)
    l = pix[:, 0] * theta / N
    m = pix[:, 1] * theta / N
    dft = numpy.exp(-2j * numpy.pi * (u[:, None] * l + v[:, None] * m)).sum(axis=1)
    err = numpy.abs(got[inside] - dft[inside]).max()
    print('streamer vs DFT: max error %.2e over %d visibilities' % (err, inside.sum()), flush=True)
    assert inside.sum() > 1000 and err < 1e-3 * len(pix)


def test_grid_wprojection(gpu):
    print('case grid_wprojection', flush=True)
    rng = numpy.random.default_rng(8)
    N, theta, S, Q, P = 256, 0.02, 8, 8, 7
    wk = c64(rng.normal(size=(P, Q, Q, S, S)) + 1j * rng.normal(size=(P, Q, Q, S, S)))
    host, dev = make_set(rng, 50, 20, 24, 9000.0, ragged=True)
    w = lam(host, 2)
    w_min, w_step = w.min() + 5.0, (w.max() - w.min()) / (P - 1.5)
    ref = numpy.zeros((N, N), dtype=complex)
    B.grid_wprojection(ref, theta, host, wk, w_min, w_step)
    wkd = BL.WKernelData(wk, w_min, w_step)
    gd = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    fl = BL.grid_wprojection(gd, N, theta, dev, wkd)
    assert fl == len(host.vis) * 8 * S * S
    assert_close(gd.cpu().numpy(), ref, what='grid_wprojection')
    g2 = torch.zeros_like(gd)
    BL.grid_wprojection(g2, N, theta, dev, wkd)
    assert torch.equal(gd, g2)
    # adds into what the grid holds
    g3 = torch.ones_like(gd)
    BL.grid_wprojection(g3, N, theta, dev, wkd)
    assert_close(g3.cpu().numpy(), ref + 1, what='grid_wprojection into ones')


def test_simple_functions(gpu):
    print('case weight / grid_simple / degrid_simple', flush=True)
    rng = numpy.random.default_rng(9)
    N, theta = 128, 0.05
    host, dev = make_set(rng, 30, 16, 12, 2000.0, ragged=True)
    ref = numpy.zeros((N, N), dtype=complex)
    B.grid_simple(ref, theta, host)
    gd = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    assert BL.grid_simple(gd, N, theta, dev) == 2 * len(host.vis)
    assert_close(gd.cpu().numpy(), ref, what='grid_simple')
    g = c64(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    h2, d2 = make_set(numpy.random.default_rng(9), 30, 16, 12, 2000.0, ragged=True,
                      with_vis=False)
    B.degrid_simple(g, theta, h2)
    BL.degrid_simple(torch.as_tensor(g.astype(numpy.complex64)).cuda(), N, theta, d2)
    assert_array_equal(d2.vis.cpu().numpy(), h2.vis.astype(numpy.complex64))
    wg_ref = B.weight(N, theta, host)
    wg = BL.weight(None, N, theta, dev)
    assert_array_equal(wg.cpu().numpy().astype(numpy.uint32), wg_ref)
    assert_close(dev.vis.cpu().numpy(), host.vis, what='weight')


def test_errors_leave_outputs_unmodified(gpu):
    print('case bl errors', flush=True)
    rng = numpy.random.default_rng(3)
    N, theta, S, Q = 64, 0.05, 8, 16
    kern = BL.SepKernelData(rng.normal(size=(Q, S)))
    gd = torch.ones((N, N), dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):  # window over the grid's edge
        BL.degrid_conv_uv(gd, N, theta, [(N / 2 - 2) / theta], [0.0], kern)
    host, dev = make_set(rng, 4, 6, 5, 500.0)
    before = dev.vis.clone()
    with pytest.raises(ValueError):  # window beyond a baseline's dumps
        BL.degrid_conv_bl(gd, N, theta, 0, 0, -1e9, 1e9, -1e9, 1e9, dev, 0, 7, 0, 5, kern)
    assert torch.equal(dev.vis, before)
    wk = BL.WKernelData(numpy.ones((2, 2, 2, 4, 4), complex), 1e6, 1.0)
    with pytest.raises(ValueError):  # w-plane outside the set
        BL.grid_wprojection(gd, N, theta, dev, wk)
    assert torch.equal(gd, torch.ones_like(gd))
    with pytest.raises(ValueError):  # chunk longer than a queue slot
        BL.streamer_degrid_worker(dev, gd, N, theta, 0, 0, [[0, 0, 6, 0, 5]], -1e9, 1e9,
                                  -1e9, 1e9, kern, 4, 5)


def test_streamer_batch_matches_per_subgrid_calls(gpu):
    """croc_streamer_degrid_batch over every subgrid of a tiling in one
    launch gives the per-subgrid croc_streamer_degrid slots bit for bit."""
    print('case streamer batch', flush=True)
    rng = numpy.random.default_rng(31)
    N, theta, S, Q, xA, xM = 512, 0.05, 8, 64, 64, 128
    g, _ = rand_grid(rng, N)
    kern = BL.SepKernelData(rng.normal(size=(Q, S)))
    host, dev = make_set(rng, 50, 32, 64, 14000.0, with_vis=False)
    st = BL.Streamer(dev, theta, xA, xM, 16, 32, kern)
    offs = tiles(N, xA, xM)
    subs = torch.stack([torch.as_tensor(numpy.ascontiguousarray(cut(g, ou, ov, xM))
                                        .astype(numpy.complex64)) for ou, ov in offs]).cuda()
    ch6, qb, ndb, flb = st.degrid_batch(subs, offs)
    singles, nds, fls = [], [], 0
    for i, (ou, ov) in enumerate(offs):
        ch, q, nd, fl = st.degrid(subs[i].contiguous(), ou, ov)
        singles.append(q)
        nds.append(nd)
        fls += fl
    assert torch.equal(qb, torch.cat(singles))
    assert torch.equal(ndb, torch.cat(nds))
    assert flb == fls and flb > 0


def test_streamer_windows_near_subgrid_edge(gpu):
    """Windows that reach near the subgrid's edge take the per-visibility
    bounds pass: still equal to the oracle when every window fits, and a
    window over the edge raises with the queue untouched."""
    print('case streamer near edge', flush=True)
    rng = numpy.random.default_rng(32)
    N, theta, S, Q, xM = 256, 0.05, 8, 32, 128
    g, _ = rand_grid(rng, N)
    kern = rng.normal(size=(Q, S))
    host, dev = make_set(rng, 60, 16, 32, 4000.0, with_vis=False)
    tc, fc = 8, 16
    chunks = numpy.array([[b, t, t + tc, f, f + fc] for b in range(host.nbl)
                          for t in range(0, 16, tc) for f in range(0, 32, fc)], dtype=numpy.int32)
    sub = cut(g, 0, 0, xM)
    subd = torch.as_tensor(numpy.ascontiguousarray(sub).astype(numpy.complex64)).cuda()
    xA = 110  # |x| <= 55 cells: fits the 128 subgrid, beyond the provable margin
    box = (-xA / 2 / theta, xA / 2 / theta, -xA / 2 / theta, xA / 2 / theta)
    ref, nd_ref = B.streamer_degrid(sub, theta, 0.0, 0.0, *box, host, chunks, tc, fc, kern)
    q, nd, _ = BL.streamer_degrid_worker(dev, subd, xM, theta, 0.0, 0.0, chunks, *box,
                                         BL.SepKernelData(kern), tc, fc)
    assert_array_equal(nd.cpu().numpy(), nd_ref)
    assert nd_ref.sum() > 100
    q = q.cpu().numpy()
    zero = ref == 0
    assert_array_equal(q == 0, zero)
    assert_close(q[~zero], ref[~zero], what='streamer near edge')
    xA = 124  # windows at |x| near 62 leave the subgrid
    box = (-xA / 2 / theta, xA / 2 / theta, -xA / 2 / theta, xA / 2 / theta)
    with pytest.raises(ValueError):
        B.streamer_degrid(sub, theta, 0.0, 0.0, *box, host, chunks, tc, fc, kern)
    queue = torch.full((len(chunks), tc * fc), 5.0, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):
        BL.streamer_degrid_worker(dev, subd, xM, theta, 0.0, 0.0, chunks, *box,
                                  BL.SepKernelData(kern), tc, fc, vis_queue=queue)
    assert bool((queue == 5.0).all())
