"""GPU parity of the imaging / predict drivers (imgfn / predfn protocol of
crocodile/synthesis.py:552-927) against the FP64 oracle, including the
full-size w-projection workload (2e6 visibilities, 151 w-planes)."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import workloads as W
from oracle import cgrid
from oracle import synthesis_ref as R

from test_gpu_core import assert_close

pytestmark = pytest.mark.gpu


def small_wproj(nv, seed=21):
    c = W.SPEC['wproj2048']
    return W.wproj2048(nv=nv, seed=seed), c


def wproj_kernel_fn(c, invert):
    def fn(theta, w):
        S = W.wproj_support(w, c['wmax'], c['S_min'], c['S_max'])
        return R.w_kernel(theta, w, c['NpixFF'], S, c['Q'], invert=invert) / c['NpixFF'] ** 2
    return fn


def gpu_wproj_kernel_fn(c, invert):
    def fn(theta, w):
        S = W.wproj_support(w, c['wmax'], c['S_min'], c['S_max'])
        return G.w_kernel(theta, w, c['NpixFF'], S, c['Q'], invert=invert) / c['NpixFF'] ** 2
    return fn


def test_simple_imaging_predict(gpu):
    rng = numpy.random.default_rng(1)
    theta, lam = 0.1, 1000.0
    uvw = numpy.stack([rng.uniform(-450, 450, 4000), rng.uniform(-450, 450, 4000),
                       rng.uniform(-10, 10, 4000)], axis=1)
    vis = rng.standard_normal(4000) + 1j * rng.standard_normal(4000)
    g = G.simple_imaging(theta, lam, uvw, None, vis)
    assert_close(g, R.simple_imaging(theta, lam, uvw, None, vis), what='simple_imaging')
    assert_close(G.simple_predict(theta, lam, uvw, None, g),
                 R.simple_predict(theta, lam, uvw, None, g), what='simple_predict')


def test_conv_imaging_predict_plain512(gpu):
    """plain512 workload (separable PSWF kernel, S=8, Q=8) on 20k of its
    visibilities."""
    c = W.SPEC['plain512']
    uvw, vis = W.plain512(nv=20000)
    aaf = R.anti_aliasing_function(c['N'], 0, c['S'])
    kv = R.kernel_oversample(aaf, c['Q'], c['S'])
    kv /= kv[0].sum()
    g = G.conv_imaging(c['theta'], c['lam'], uvw, None, vis, kv)
    ref = R.conv_imaging(c['theta'], c['lam'], uvw, None, vis, kv)
    assert_close(g, ref, what='conv_imaging')
    assert_close(G.conv_predict(c['theta'], c['lam'], uvw, None, ref, kv),
                 R.conv_predict(c['theta'], c['lam'], uvw, None, ref, kv), what='conv_predict')


def test_w_cache_native_kernels(gpu):
    """kernel_fn = w_kernel with its kwargs: kernels are generated on the
    device for every w-bin."""
    (uvw, vis), c = small_wproj(3000)
    kw = dict(NpixFF=32, NpixKern=16, Qpx=4)
    theta, lam, wstep = c['theta'], c['lam'], 100.0
    g = G.w_cache_imaging(theta, lam, uvw, None, vis, wstep=wstep, **kw)
    ref = R.w_cache_imaging(theta, lam, uvw, None, vis, wstep=wstep, **kw)
    assert_close(g, ref, what='w_cache_imaging')
    kwp = dict(kw, invert=False)
    v = G.w_cache_predict(theta, lam, uvw, None, ref, wstep=wstep, **kwp)
    vref = R.w_cache_predict(theta, lam, uvw, None, ref, wstep=wstep, **kwp)
    assert_close(v, vref, what='w_cache_predict')


def test_w_cache_user_kernel_fn_and_cache(gpu):
    """A host kernel_fn (variable support per w-bin) and a kernel_cache
    callable, as scripts/image_dataset.py plugs them."""
    (uvw, vis), c = small_wproj(2500, seed=22)
    theta, lam, wstep = c['theta'], c['lam'], c['wstep']
    calls = []

    def cache(theta, w):
        calls.append(w)
        return wproj_kernel_fn(c, True)(theta, w)

    g = G.w_cache_imaging(theta, lam, uvw, None, vis, wstep=wstep, kernel_cache=cache)
    assert len(calls) == len(set(calls))  # one evaluation per distinct bin
    ref = R.w_cache_imaging(theta, lam, uvw, None, vis, wstep=wstep,
                            kernel_fn=wproj_kernel_fn(c, True))
    assert_close(g, ref, what='w_cache_imaging(kernel_cache)')
    v = G.w_cache_predict(theta, lam, uvw, None, ref, wstep=wstep,
                          kernel_fn=wproj_kernel_fn(c, False))
    vref = R.w_cache_predict(theta, lam, uvw, None, ref, wstep=wstep,
                             kernel_fn=wproj_kernel_fn(c, False))
    assert_close(v, vref, what='w_cache_predict(kernel_fn)')


def test_w_cache_sparse_bins(gpu):
    """Widely separated w-bins use the indexed plane path."""
    rng = numpy.random.default_rng(5)
    theta, lam = 0.1, 2000.0
    nv = 400
    uvw = numpy.stack([rng.uniform(-800, 800, nv), rng.uniform(-800, 800, nv),
                       rng.choice([-5000.0, 0.0, 7000.0], nv)], axis=1)
    vis = rng.standard_normal(nv) + 0j
    kw = dict(NpixFF=16, NpixKern=8, Qpx=2)
    g = G.w_cache_imaging(theta, lam, uvw, None, vis, wstep=10.0, **kw)
    ref = R.w_cache_imaging(theta, lam, uvw, None, vis, wstep=10.0, **kw)
    assert_close(g, ref, what='sparse bins')


def test_w_cache_with_source_columns(gpu):
    """kernel_cache(theta, w, *src_row): one kernel per (w-bin, src) key,
    the AW-kernel protocol of scripts/image_dataset.py:153-171."""
    rng = numpy.random.default_rng(6)
    theta, lam, wstep = 0.1, 1000.0, 50.0
    nv = 1500
    uvw = numpy.stack([rng.uniform(-400, 400, nv), rng.uniform(-400, 400, nv),
                       rng.uniform(-200, 200, nv)], axis=1)
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    src = rng.integers(0, 3, (nv, 2))

    def kfn(theta, w, a1, a2):
        k = R.w_kernel(theta, w, 16, 8, 2) / 256
        return k * (1 + 0.1 * a1 - 0.05j * a2)

    g = G.w_cache_imaging(theta, lam, uvw, src, vis, wstep=wstep, kernel_cache=kfn)
    ref = R.w_cache_imaging(theta, lam, uvw, src, vis, wstep=wstep, kernel_cache=kfn)
    assert_close(g, ref, what='w_cache src columns')


def test_w_slice_imaging_predict(gpu):
    (uvw, vis), c = small_wproj(3000, seed=23)
    theta, lam = c['theta'], c['lam']
    kw = dict(NpixFF=32, NpixKern=16, Qpx=4)
    g = G.w_slice_imaging(theta, lam, uvw, None, vis, wstep=500, **kw)
    ref = R.w_slice_imaging(theta, lam, uvw, None, vis, wstep=500, **kw)
    assert_close(g, ref, what='w_slice_imaging')
    v = G.w_slice_predict(theta, lam, uvw, None, ref, wstep=500, invert=False, **kw)
    vref = R.w_slice_predict(theta, lam, uvw, None, ref, wstep=500, invert=False, **kw)
    assert_close(v, vref, what='w_slice_predict')


def test_do_imaging_do_predict(gpu):
    c = W.SPEC['plain512']
    uvw, vis = W.plain512(nv=5000, seed=31)
    aaf = R.anti_aliasing_function(c['N'], 0, c['S'])
    kv = R.kernel_oversample(aaf, c['Q'], c['S'])
    kv /= kv[0].sum()
    d, p, pm = G.do_imaging(c['theta'], c['lam'], uvw, None, vis, G.conv_imaging, kv=kv)
    dr, pr, pmr = R.do_imaging(c['theta'], c['lam'], uvw, None, vis, R.conv_imaging, kv=kv)
    assert abs(pm - pmr) / pmr < 1e-5
    assert_close(d, dr, what='dirty image')
    assert_close(p, pr, what='psf')
    model = numpy.zeros((c['N'], c['N']))
    model[200, 300] = 1.0
    model[256, 256] = 2.0
    v = G.do_predict(c['theta'], c['lam'], uvw, None, model, G.conv_predict, kv=kv)
    vr = R.do_predict(c['theta'], c['lam'], uvw, None, model, R.conv_predict, kv=kv)
    assert_close(v, vr, what='do_predict')


def test_doweight_and_mirror(gpu):
    rng = numpy.random.default_rng(8)
    uvw = numpy.stack([rng.uniform(-40, 40, 3000), rng.uniform(-40, 40, 3000),
                       rng.uniform(-5, 5, 3000)], axis=1)
    vis = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
    u1, v1 = G.mirror_uvw(uvw, vis)
    u2, v2 = R.mirror_uvw(uvw, vis)
    assert (u1 == u2).all() and (v1 == v2).all()
    w1 = G.doweight(0.1, 1000.0, u1, numpy.ones(3000))
    w2 = R.doweight(0.1, 1000.0, u2, numpy.ones(3000))
    assert numpy.allclose(w1, w2, rtol=0, atol=0)


def test_fft_helpers(gpu):
    rng = numpy.random.default_rng(9)
    for N in [7, 8, 64]:
        a = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        assert numpy.allclose(G.fft(a), R.fft(a), atol=1e-10)
        assert numpy.allclose(G.ifft(a), R.ifft(a), atol=1e-12)


def test_wproj2048_full_size(gpu):
    """The benchmark workload at full size: 2e6 visibilities, 151 w-planes,
    support 16..32, Q=8, N=2048.  Grid and visibilities within 1e-5 RMS
    relative of the FP64 C oracle (same kernels, FP64 accumulation)."""
    c = W.SPEC['wproj2048']
    uvw, vis = W.wproj2048()
    wb = numpy.round(uvw[:, 2] / c['wstep']).astype(int)
    lo, hi = wb.min(), wb.max()
    ws = [c['wstep'] * b for b in range(lo, hi + 1)]
    S = [W.wproj_support(w, c['wmax'], c['S_min'], c['S_max']) for w in ws]
    scale = 1.0 / c['NpixFF'] ** 2
    dev = gpu
    bank_i = G.w_kernel_bank(c['theta'], ws, c['NpixFF'], S, c['Q'], True, scale, double=True)
    bank_p = G.w_kernel_bank(c['theta'], ws, c['NpixFF'], S, c['Q'], False, scale, double=True)
    # the oracle uses the device-generated FP64 tables (checked against the
    # NumPy FFT route in test_gpu_core); its accumulation is FP64
    ti = bank_i.kern.cpu().numpy()
    tp = bank_p.kern.cpu().numpy()
    ob_i, ob_p = cgrid.Bank([ti[:1]]), cgrid.Bank([tp[:1]])
    for ob, t in ((ob_i, ti), (ob_p, tp)):
        ob.kern = t
        ob.plane_off = bank_i.plane_off_h.copy()
        ob.plane_s = bank_i.plane_s_h.copy()
        ob.Q = c['Q']
    vplane = (wb - lo).astype(numpy.int32)
    p = uvw / c['lam']
    ref = numpy.zeros((c['N'], c['N']), dtype=complex)
    cgrid.convgrid(ob_i, ref, p, vis, vplane)
    uvw_d = torch.as_tensor(uvw, device=dev)
    vis_d = torch.as_tensor(vis, device=dev).to(torch.complex64)
    guv = torch.zeros((c['N'], c['N']), dtype=torch.complex64, device=dev)
    G.grid_device(bank_i.as_complex64(), guv, uvw_d, c['lam'], vis_d, 2, None, c['wstep'], lo)
    assert_close(guv.cpu().numpy(), ref, what='wproj2048 grid')
    vref = cgrid.convdegrid(ob_p, ref, p, vplane)
    out = torch.empty(len(vis), dtype=torch.complex64, device=dev)
    G.degrid_device(bank_p.as_complex64(), torch.as_tensor(ref, device=dev).to(torch.complex64),
                    uvw_d, c['lam'], out, 2, None, c['wstep'], lo)
    assert_close(out.cpu().numpy(), vref, what='wproj2048 degrid')


@pytest.mark.parametrize('separable', [False, True])
def test_w_slice_with_source_columns(gpu, separable):
    """w_slice_* with source columns: each slice's table carries leading
    dimensions indexed by the visibility's source row (make_gcf,
    synthesis.py:451-461; negative rows wrap), against the oracle."""
    theta, lam = 0.1, 3000.0
    rng = numpy.random.default_rng(12)
    nv = 4000
    uvw = numpy.stack([rng.uniform(-1000, 1000, nv), rng.uniform(-1000, 1000, nv),
                       rng.uniform(-300, 300, nv)], axis=1)
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    src = numpy.stack([rng.integers(-2, 2, nv), rng.integers(0, 3, nv)], axis=1)
    f = rng.standard_normal((2, 3)) + 1j * rng.standard_normal((2, 3))

    def kfn(theta, w, invert=True):
        k = R.w_kernel(theta, w, 16, 8, 4, invert=invert) / 256
        if separable:
            k = k[:, 0, 4, :]  # a [Q, S] table, made up but fixed per w
        return numpy.array([[k * f[i, j] for j in range(3)] for i in range(2)])
    g = G.w_slice_imaging(theta, lam, uvw, src, vis, wstep=700, kernel_fn=kfn)
    ref = R.w_slice_imaging(theta, lam, uvw, src, vis, wstep=700, kernel_fn=kfn)
    assert_close(g, ref, what='w_slice_imaging src sep=%d' % separable)
    v = G.w_slice_predict(theta, lam, uvw, src, ref, wstep=700, kernel_fn=kfn, invert=False)
    vref = R.w_slice_predict(theta, lam, uvw, src, ref, wstep=700, kernel_fn=kfn, invert=False)
    assert_close(v, vref, what='w_slice_predict src sep=%d' % separable)
