"""Files read through crocodile_amd.formats feed the native entry points as
the same data built in memory does (examples/grid/hdf5.c load_vis /
load_wkern / load_sep_kern feeding grid.c): gridding and degridding from a
visibility file and kernel files equal the in-memory results bit for bit,
and the FP64 oracle to the usual bound; an image_dataset.py-style kernel
file drives w_cache_imaging as its kernel function does."""
import numpy
import pytest
import torch

import crocodile_amd.synthesis as G
from crocodile_amd import bl as BL
from crocodile_amd import formats
from oracle import blgrid as B

from test_gpu_bl import lam, make_set
from test_gpu_core import assert_close

pytestmark = pytest.mark.gpu


def per_baseline(host):
    """[(uvw [nt, 3], freq [nf])] of the oracle's packed baseline set."""
    return [(host.uvw_m[host.uvw0[b]:host.uvw0[b] + host.nt[b]],
             host.freq[host.f0[b]:host.f0[b] + host.nf[b]]) for b in range(host.nbl)]


def test_wprojection_from_files(gpu, tmp_path):
    print('case grid_wprojection from files', flush=True)
    rng = numpy.random.default_rng(21)
    N, theta, S, Q, P = 256, 0.02, 8, 8, 7
    # a complete baseline set of 9 antennas: hdf5.c:330-362 reads pairs
    # a1 < a2 of (number of a1 groups) + 1 antennas
    a1, a2 = numpy.triu_indices(9, 1)
    host, _ = make_set(rng, len(a1), 16, 12, 9000.0, ragged=True)
    w = lam(host, 2)
    w_min, w_step = w.min() + 5.0, (w.max() - w.min()) / (P - 1.5)
    wk = rng.normal(size=(P, Q, Q, S, S)) + 1j * rng.normal(size=(P, Q, Q, S, S))
    wk = wk.astype(numpy.complex64).astype(numpy.complex128)
    vis_file, kern_file = str(tmp_path / 'vis.h5'), str(tmp_path / 'wkern.h5')
    bls = [BL.BlData(a1[b], a2[b], numpy.arange(len(u)), f, u, host.baseline_vis(b))
           for b, (u, f) in enumerate(per_baseline(host))]
    dev = BL.VisData(bls, 9)
    formats.write_vis(vis_file, bls)
    formats.write_kernels(kern_file, wkern={(theta, w_min + k * w_step): wk[k] for k in range(P)})
    loaded = formats.load_vis(vis_file)
    wkd = formats.load_wkern(kern_file, theta)
    assert loaded.antenna_count == 9 and loaded.antennas == dev.antennas and wkd.plane_count == P
    assert abs(wkd.w_min - w_min) < 1e-9 and abs(wkd.w_step - w_step) < 1e-9
    g_file = torch.zeros((N, N), dtype=torch.complex64, device=gpu)
    BL.grid_wprojection(g_file, N, theta, loaded, wkd)
    g_mem = torch.zeros_like(g_file)
    # the file's w-step is recomputed from the plane names: use it throughout
    BL.grid_wprojection(g_mem, N, theta, dev, BL.WKernelData(wk, wkd.w_min, wkd.w_step))
    assert torch.equal(g_file, g_mem)
    ref = numpy.zeros((N, N), dtype=complex)
    B.grid_wprojection(ref, theta, host, wk, wkd.w_min, wkd.w_step)
    assert_close(g_file.cpu().numpy(), ref, what='grid_wprojection from files')


def test_degrid_from_files(gpu, tmp_path):
    print('case degrid_conv_bl from files', flush=True)
    rng = numpy.random.default_rng(22)
    N, theta, S, Q = 256, 0.02, 8, 16
    a1, a2 = numpy.triu_indices(8, 1)
    host, _ = make_set(rng, len(a1), 12, 10, 9000.0, with_vis=False)
    kern = rng.normal(size=(Q, S))
    grid = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    gd = torch.as_tensor(grid.astype(numpy.complex64), device=gpu)
    vis_file, kern_file = str(tmp_path / 'vis.h5'), str(tmp_path / 'sep.h5')
    bls = [BL.BlData(a1[b], a2[b], numpy.arange(len(u)), f, u)
           for b, (u, f) in enumerate(per_baseline(host))]
    dev = BL.VisData(bls, 8)
    formats.write_vis(vis_file, bls)
    formats.write_kernels(kern_file, sepkern=kern)
    loaded = formats.load_vis(vis_file)
    sk = formats.load_sepkern(kern_file)
    BL.degrid_conv(gd, N, theta, 0.0, 0.0, loaded, sk)
    BL.degrid_conv(gd, N, theta, 0.0, 0.0, dev, BL.SepKernelData(kern))
    assert torch.equal(loaded.vis, dev.vis)


def test_kernel_file_as_kernel_cache(gpu, tmp_path):
    """scripts/image_dataset.py:114-137: w-kernels from the file (closest
    w), used as w_cache_imaging's kernel_cache."""
    theta, wstep = 0.05, 50.0
    kw = dict(NpixFF=32, NpixKern=8, Qpx=4)
    ws = numpy.arange(-6, 7) * wstep
    kern_file = str(tmp_path / 'wkern.h5')
    formats.write_kernels(kern_file, wkern={(theta, w): G.w_kernel(theta, w, **kw) for w in ws})
    cache = formats.H5KernelFile(kern_file)
    assert cache.wstep(theta) == wstep
    rng = numpy.random.default_rng(23)
    N, lam_ = 256, 256 / theta
    nv = 20000
    uvw = numpy.stack([rng.uniform(-0.4, 0.4, nv) * lam_, rng.uniform(-0.4, 0.4, nv) * lam_,
                       rng.uniform(-290.0, 290.0, nv)], axis=1)
    vis = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    g1 = G.w_cache_imaging(theta, lam_, uvw, None, vis, wstep=wstep, kernel_cache=cache)
    g2 = G.w_cache_imaging(theta, lam_, uvw, None, vis, wstep=wstep, **kw)
    assert numpy.abs(g1 - g2).max() <= 1e-6 * numpy.abs(g2).max()


def test_generate_w_kernels_h5(gpu, tmp_path):
    """generate_wkernels.py's HDF5 output: every w of the range, equal to
    w_kernel, extended without recomputing what the file holds."""
    from crocodile_amd import kernels as K
    from oracle import synthesis_ref as R
    p = str(tmp_path / 'wk.h5')
    written = K.generate_w_kernels_h5(p, 0.05, 40.0, 2, 32, 8, 4)
    assert written == [-80.0, -40.0, 0.0, 40.0, 80.0]
    assert K.generate_w_kernels_h5(p, 0.05, 40.0, 2, 32, 8, 4) == []
    assert K.generate_w_kernels_h5(p, 0.05, 40.0, 3, 32, 8, 4) == [-120.0, 120.0]
    ks = formats.read_wkernels(p, 0.05)
    assert sorted(ks) == [-120.0, -80.0, -40.0, 0.0, 40.0, 80.0, 120.0]
    ref = R.w_kernel(0.05, 80.0, 32, 8, 4)
    assert numpy.abs(ks[80.0] - ref).max() <= 1e-12 * numpy.abs(ref).max()
