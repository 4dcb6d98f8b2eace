"""The product's host-side drop-ins (crocodile_amd.synthesis), against the
reference's own closed-form tests and the pinned oracle.

These functions hold no device arithmetic (index layout, scipy's PSWF, small
2x2 algebra, sorting), so they run on the CPU, for NumPy and for torch CPU
tensors; the device-tensor forms are covered by tests/test_gpu_dropins.py.
"""
import itertools

import numpy
import pytest
import torch
from numpy.testing import assert_allclose

import crocodile_amd.synthesis as G
from oracle import synthesis_ref as R


def pattern(N):
    return R.coordinates2(N)[1] + R.coordinates2(N)[0] * 1j


def test_coordinates_product():
    """tests/test_synthesis.py:15-28 on the product functions."""
    for N in [4, 5, 6, 7, 8, 9, 1000, 1001]:
        low, high = G.coordinateBounds(N)
        c = G.coordinates(N)
        cx, cy = G.coordinates2(N)
        for a in (c, cx, cy):
            assert abs(a.min() - low) < 1e-12 and abs(a.max() - high) < 1e-12
        assert c[N // 2] == 0 and (cx[:, N // 2] == 0).all() and (cy[N // 2, :] == 0).all()
        assert_allclose(cx, R.coordinates2(N)[0], atol=0)
        assert_allclose(cy, R.coordinates2(N)[1], atol=0)


@pytest.mark.parametrize('torch_in', [False, True])
def test_pad_extract_product(torch_in):
    """tests/test_synthesis.py:33-44 on pad_mid / extract_mid (1D and 2D)."""
    for N, N2 in [(1, 1), (1, 2), (2, 3), (3, 4), (2, 5), (4, 6)]:
        cs = 1 + pattern(N)
        x = torch.as_tensor(cs) if torch_in else cs
        cs_pad = numpy.asarray(G.pad_mid(x, N2))
        cs2 = 1 + pattern(N2) * N2 / N
        assert numpy.all((numpy.abs(cs_pad - cs2) < 1e-15) | (numpy.abs(cs_pad) < 1e-15))
        back = G.extract_mid(torch.as_tensor(cs_pad) if torch_in else cs_pad, N)
        assert_allclose(numpy.asarray(back), cs)
        row = cs[0]
        p1 = G.pad_mid(torch.as_tensor(row) if torch_in else row, N2)
        assert_allclose(numpy.asarray(p1), R.pad_mid(row, N2))
        assert_allclose(numpy.asarray(G.extract_mid(p1, N)), row)


@pytest.mark.parametrize('torch_in', [False, True])
def test_extract_oversampled_product(torch_in):
    """synthesis.py:160-197: sub-kernel (yf, xf) holds a[m + Q*i] with m =
    Na//2 - Q*(N//2) + f, written out element by element here."""
    rng = numpy.random.default_rng(5)
    for Na, Q, N in [(24, 4, 5), (21, 3, 6), (16, 1, 16), (30, 5, 4)]:
        a = rng.standard_normal((Na, Na)) + 1j * rng.standard_normal((Na, Na))
        got = numpy.asarray(G.extract_oversampled(torch.as_tensor(a) if torch_in else a, Q, N))
        assert got.shape == (Q, Q, N, N)
        for yf, xf, y, x in itertools.product(range(Q), range(Q), range(N), range(N)):
            m0 = Na // 2 - Q * (N // 2)
            assert got[yf, xf, y, x] == a[m0 + yf + Q * y, m0 + xf + Q * x]
        a1 = a[0]
        got1 = numpy.asarray(G.extract_oversampled(torch.as_tensor(a1) if torch_in else a1, Q, N))
        assert_allclose(got1, R.extract_oversampled(a1, Q, N), atol=0)


def test_anti_aliasing_product():
    """tests/test_synthesis.py:46-50 (shape, centre value 1) and equality with
    the oracle's scipy pro_ang1 call (the same library routine)."""
    for shape in [(4, 4), (5, 5), (4, 6), (7, 3)]:
        aaf = G.anti_aliasing_function(shape, 0, 10)
        assert aaf.shape == shape
        assert abs(aaf[shape[0] // 2, shape[1] // 2] - 1) < 1e-7
        assert_allclose(aaf, R.anti_aliasing_function(shape, 0, 10), rtol=1e-14)
    a1 = G.anti_aliasing_function(64, 0, 8)
    assert a1.shape == (64,) and numpy.all(a1 > 0) and a1.argmax() == 32
    assert_allclose(a1, a1[::-1][numpy.r_[-1, 0:63]], rtol=1e-12)  # even about N//2


def test_kernel_transform_product():
    """Eigenstructure of kernel_transform (see test_oracle): symmetric, d ->
    dn d, d-perpendicular unchanged, identity at 0; equal to the oracle."""
    assert_allclose(G.kernel_transform(0, 0), numpy.eye(2))
    for dl, dm in [(0.1, 0.0), (0.0, -0.3), (0.2, 0.5), (-0.6, 0.7)]:
        T = G.kernel_transform(dl, dm)
        dn = numpy.sqrt(1 - dl ** 2 - dm ** 2)
        assert_allclose(T, T.T)
        assert_allclose(T @ [dl, dm], dn * numpy.array([dl, dm]), atol=1e-15)
        assert_allclose(T @ [-dm, dl], [-dm, dl], atol=1e-15)
        assert_allclose(T, R.kernel_transform(dl, dm), rtol=1e-15)


def test_kernel_coordinates_product():
    """(l, m) = theta * coordinates2 transformed by T (row-vector convention
    of synthesis.py:265-268) plus (dl, dm)."""
    T = numpy.array([[2., .5], [-1., 3.]])
    l, m = G.kernel_coordinates(6, 0.2, 0.01, -0.02, T)
    cx, cy = R.coordinates2(6)
    for i, j in itertools.product(range(6), repeat=2):
        x = numpy.array([cx[i, j], cy[i, j]]) * 0.2
        lm = x @ T + [0.01, -0.02]
        assert_allclose([l[i, j], m[i, j]], lm, rtol=1e-14, atol=1e-16)


def test_recentre_product():
    """tests/test_synthesis.py:60-75 on the product's kernel_recentre (an
    image-plane ramp moves the kernel by whole cells), and
    visibility_recentre's u -= w dl, v -= w dm, by hand."""
    for lam in [4, 1000]:
        for N in range(2, 6):
            for w in [0.1, 1, 10]:
                for dl, dm in [(1, 1), (2, 0), (1, -3)]:
                    l, m = R.kernel_coordinates(N, N / lam)
                    cp = R.w_kernel_function(l, m, w)
                    kern = R.ifft(cp)
                    kerns = R.ifft(G.kernel_recentre(cp, N / lam, w, dl * lam / N / w,
                                                     dm * lam / N / w))
                    assert_allclose(numpy.roll(numpy.roll(kern, dl, axis=1), dm, axis=0),
                                    kerns, atol=1e-15)
    uvw = numpy.array([[1., 2., 10.], [-3., 4., -20.]])
    for x in (uvw, torch.as_tensor(uvw)):
        out = numpy.asarray(G.visibility_recentre(x, 0.1, -0.05))
        assert_allclose(out, [[0., 2.5, 10.], [-1., 3., -20.]])
    assert uvw[0, 0] == 1.0


@pytest.mark.parametrize('torch_in', [False, True])
def test_sort_slice_product(torch_in):
    """sort_vis_w orders by w (stable, ties in input order), slice_vis cuts
    runs of `step` visibilities; tests/test_synthesis.py:235-240."""
    uvw = numpy.array([[0, 0, 3.], [1, 0, -1.], [2, 0, 3.], [3, 0, 0.], [4, 0, -1.]])
    vis = numpy.arange(5) + 0j
    x = torch.as_tensor(uvw) if torch_in else uvw
    v = torch.as_tensor(vis) if torch_in else vis
    p, s = G.sort_vis_w(x, v)
    assert numpy.asarray(p)[:, 0].tolist() == [1, 4, 3, 0, 2]
    assert numpy.asarray(s).real.tolist() == [1, 4, 3, 0, 2]
    assert numpy.asarray(G.sort_vis_w(x))[:, 0].tolist() == [1, 4, 3, 0, 2]
    for N in range(2, 6):
        cs = numpy.transpose([*map(numpy.ravel, R.coordinates2(N)), numpy.zeros(N * N)])
        for step in range(1, 10):
            sl = G.slice_vis(step, torch.as_tensor(cs) if torch_in else cs)
            assert_allclose(numpy.vstack([numpy.asarray(a) for a in sl]), cs)
            assert all(len(a) == step for a in sl[:-1])
        pairs = G.slice_vis(3, cs, numpy.arange(N * N))
        assert_allclose(numpy.concatenate([b for _, b in pairs]), numpy.arange(N * N))


def test_w_conj_kernel_fn_product():
    """synthesis.py:682-699 dispatch: w < 0 -> conj(k(theta, -w, ...))."""
    calls = []

    def kfn(theta, w, *a, **kw):
        calls.append((theta, w, a, kw))
        return numpy.array([[1 + 2j, w + 1j]])
    fn = G.w_conj_kernel_fn(kfn)
    assert_allclose(fn(0.1, -3.0, 7, k=1), numpy.conj(kfn(0.1, 3.0)))
    assert calls[0] == (0.1, 3.0, (7,), {'k': 1})
    assert_allclose(fn(0.1, 2.0), kfn(0.1, 2.0))

    def tfn(theta, w):
        return torch.tensor([[1 + 2j, w + 1j]])
    assert_allclose(G.w_conj_kernel_fn(tfn)(0.1, -2.0).numpy(), [[1 - 2j, -(-2.0) - 1j]])


@pytest.mark.parametrize('torch_in', [False, True])
def test_mirror_uvw_product(torch_in):
    """synthesis.py:831-852 on the hand-built case of test_oracle."""
    uvw = numpy.array([[1., -2., 3.], [4., 5., -6.], [-7., 0., 8.], [9., -0.0, 1.],
                       [-1., -1e-9, -2.]])
    vis = numpy.array([1 + 2j, 3 - 4j, 5 + 6j, 7 + 8j, -1 - 1j])
    x = torch.as_tensor(uvw) if torch_in else uvw
    v = torch.as_tensor(vis) if torch_in else vis
    u2, v2 = G.mirror_uvw(x, v)
    assert_allclose(numpy.asarray(u2), [[-1., 2., -3.], [4., 5., -6.], [-7., 0., 8.],
                                        [9., 0., 1.], [1., 1e-9, 2.]])
    assert_allclose(numpy.asarray(v2), [1 - 2j, 3 - 4j, 5 + 6j, 7 + 8j, -1 + 1j])
    assert uvw[0, 1] == -2.0 and vis[0] == 1 + 2j


def test_make_gcf_product():
    """make_gcf's accessor (synthesis.py:451-461): separable tables give the
    outer product of rows yf and xf; 2D tables index [src..., yf, xf]."""
    rng = numpy.random.default_rng(2)
    sep = rng.standard_normal((3, 4, 5)) + 0j  # [src][Q][S]
    Q, gw, gh, fn = G.make_gcf(numpy.zeros((1, 1), int), sep)
    assert (Q, gw, gh) == (4, 5, 5)
    assert_allclose(fn((2,), 1, 3), numpy.outer(sep[2, 3], sep[2, 1]))
    k2 = rng.standard_normal((2, 3, 3, 4, 6)) + 0j  # [src][Q][Q][gh][gw]
    Q, gw, gh, fn = G.make_gcf(numpy.zeros((1, 1), int), k2)
    assert (Q, gw, gh) == (3, 6, 4)
    assert_allclose(fn((1,), 2, 0), k2[1, 0, 2])
    with pytest.raises(AssertionError):
        G.make_gcf(numpy.zeros((1, 0), int), numpy.zeros((3, 3, 3)))
