"""Pinning of the C restatement of the reference's visibility-layout functions
(oracle/blgrid.c: examples/grid/grid.c:22-322 and the streamer's degrid
worker, streamer.c:366-486).

Each check rests on something other than the restatement's own formulas:
the T05 frac_coord golden vectors of the reference's C unit test, the
Python restatement (oracle/synthesis_ref.py, pinned by the reference's own
tests) for the same arithmetic in the other coordinate convention, brute
force in NumPy on small inputs, the direct Fourier sum, and invariants of the
subgrid decomposition (each visibility degridded by exactly one subgrid, the
value equal to degridding the full grid there).
"""
import json
import os

import numpy
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import blgrid as B
from oracle import synthesis_ref as R

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
C = 299792458.0


def bl_tracks(rng, nbl, nt, nf, umax, f0=1.0e8, df=2.0e6, ragged=False):
    """Random straight uvw tracks in metres, nf channels from f0."""
    uvw, freq = [], []
    for b in range(nbl):
        t = nt if not ragged else int(rng.integers(1, nt + 1))
        f = nf if not ragged else int(rng.integers(1, nf + 1))
        start = rng.uniform(-umax, umax, 3)
        step = rng.uniform(-umax, umax, 3) / max(nt, 1) / 4
        uvw.append(start + numpy.arange(t)[:, None] * step)
        freq.append(f0 + df * numpy.arange(f) + rng.uniform(0, df))
    return uvw, freq


def lam(s, k):
    """uvw_lambda of every visibility, in layout order (u * f / c)."""
    out = numpy.empty(s.vis.shape[0])
    for b in range(s.nbl):
        nt, nf = int(s.nt[b]), int(s.nf[b])
        u = s.uvw_m[int(s.uvw0[b]):int(s.uvw0[b]) + nt, k]
        f = s.freq[int(s.f0[b]):int(s.f0[b]) + nf]
        o = int(s.v0[b])
        out[o:o + nt * nf] = (u[:, None] * f[None, :] / C).ravel()
    return out


def pswf_sep(S, Q, N=256):
    aa = R.anti_aliasing_function(N, 0, S)
    k = R.kernel_oversample(aa, Q, S).real
    return k / k[0].sum(), aa


# ---------------------------------------------------------------- frac_coord

def test_frac_coord_t05_golden():
    """Known answers of examples/grid/test_recombine.c:378-412 (C
    convention: u in cells)."""
    g = json.load(open(os.path.join(GOLDEN, 't05_frac_coord.json')))
    x, fx = B.frac_coord(g['grid_size'], g['oversample'], numpy.array(g['u']))
    assert_array_equal(x, g['x'])
    assert_array_equal(fx, g['fx'])


@pytest.mark.parametrize('N,Q', [(64, 8), (256, 16), (1024, 32), (8192, 1024)])
def test_frac_coord_matches_python_convention(N, Q):
    """grid.c's frac_coord on u cells equals synthesis.py's frac_coord on
    p = u / N: random positions and every half-step tie."""
    rng = numpy.random.default_rng(N + Q)
    ties = rng.choice(numpy.arange(-N * Q // 2, N * Q // 2), min(N * Q, 20000), replace=False)
    u = numpy.concatenate([rng.uniform(-N / 2 + 1, N / 2 - 1, 2000), (ties + 0.5) / Q])
    u = u[(u >= -N / 2) & (u < N / 2)]
    x, fx = B.frac_coord(N, Q, u)
    xp, fxp = R.frac_coord(N, Q, u / N)
    assert_array_equal(x, xp)
    assert_array_equal(fx, fxp)


# ------------------------------------------------------------ degrid_conv_uv

@pytest.mark.parametrize('S,Q', [(8, 16), (7, 8), (12, 32)])
def test_degrid_conv_uv_is_separable_convdegrid(S, Q):
    """degrid_conv_uv (grid.c:179-205) = convdegrid with the separable table
    (synthesis.py:484-503, make_gcf's outer product) at p = theta u / N."""
    rng = numpy.random.default_rng(S * Q)
    N, theta = 64, 0.02
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    kern = rng.normal(size=(Q, S))
    lim = (N / 2 - S) / theta
    uv = rng.uniform(-lim, lim, (300, 2))
    got = B.degrid_conv_uv(g, theta, uv, kern)
    ref = R.convdegrid(kern, g, theta * uv / N)
    assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_degrid_conv_uv_window_outside_grid_raises():
    N, theta = 32, 0.1
    kern = numpy.ones((4, 8))
    g = numpy.zeros((N, N), dtype=complex)
    with pytest.raises(ValueError):
        B.degrid_conv_uv(g, theta, numpy.array([[(N / 2 - 2) / theta, 0.0]]), kern)


def test_degrid_conv_uv_predicts_point_sources():
    """Direct Fourier sum closed form: degridding fft(image / aa x aa) with
    the PSWF table of aa predicts the sky's visibilities."""
    N, theta, S, Q = 256, 0.04, 10, 1024
    kern, aa = pswf_sep(S, Q, N)
    rng = numpy.random.default_rng(3)
    img = numpy.zeros((N, N))
    pix = rng.integers(-N // 4, N // 4, (5, 2))
    for il, im in pix:
        img[N // 2 + im, N // 2 + il] += 1.0
    g = R.fft(img / numpy.outer(aa, aa))
    uv = rng.uniform(-(N / 2 - S) / theta, (N / 2 - S) / theta, (400, 2))
    got = B.degrid_conv_uv(g, theta, uv, kern)
    l = pix[:, 0] * theta / N
    m = pix[:, 1] * theta / N
    dft = numpy.exp(-2j * numpy.pi * (uv[:, :1] * l + uv[:, 1:] * m)).sum(axis=1)
    # floor: positions rounded to 1/Q cell, phase error 2 pi (1/4) / 2Q
    # per source at N/4 from the centre
    assert numpy.abs(got - dft).max() < 1e-3 * len(pix)


# ------------------------------------------------------------ degrid_conv_bl

@pytest.mark.parametrize('ragged', [False, True])
def test_degrid_conv_bl_window_brute_force(ragged):
    """Window rule, chunk-local index and untouched entries of
    degrid_conv_bl (grid.c:234-259), enumerated in NumPy."""
    rng = numpy.random.default_rng(11 + ragged)
    N, theta, S, Q = 64, 0.02, 8, 16
    kern = rng.normal(size=(Q, S))
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    uvw, freq = bl_tracks(rng, 6, 12, 9, 300.0, ragged=ragged)
    if ragged:  # the window must fit every baseline
        for i in range(len(uvw)):
            uvw[i] = numpy.vstack([uvw[i], uvw[i][-1:] + numpy.zeros((12, 3))])[:12]
            freq[i] = numpy.concatenate([freq[i], freq[i][-1] + 1e6 * numpy.arange(1, 10)])[:9]
    s = B.BlSet(uvw, freq)
    sentinel = 7 + 7j
    s.vis[:] = sentinel
    t0, t1, f0, f1 = 2, 10, 1, 8
    d_u, d_v = 30.0, -20.0
    box = (-60.0, 70.0, -50.0, 65.0)
    n = B.degrid_conv_bl(g, theta, d_u, d_v, *box, s, t0, t1, f0, f1, kern)
    want = numpy.full(s.vis.shape[0], sentinel)
    count = 0
    for b in range(s.nbl):
        for t in range(t0, t1):
            for f in range(f0, f1):
                fr = s.freq[int(s.f0[b]) + f]
                u = s.uvw_m[int(s.uvw0[b]) + t, 0] * fr / C
                v = s.uvw_m[int(s.uvw0[b]) + t, 1] * fr / C
                if not (box[0] <= u < box[1] and box[2] <= v < box[3]):
                    continue
                o = int(s.v0[b]) + (t - t0) * (f1 - f0) + (f - f0)
                want[o] = B.degrid_conv_uv(g, theta, numpy.array([[u - d_u, v - d_v]]), kern)[0]
                count += 1
    assert n == count and 0 < count < s.nbl * (t1 - t0) * (f1 - f0)
    assert_allclose(s.vis, want, rtol=0, atol=0)


# ---------------------------------------------------------- streamer degrid

def tiles(N, xA, xM):
    """Subgrid offsets (cells from the grid centre) tiling the grid with
    windows xA apart; every subgrid lies inside the grid."""
    k = (N - xM) // 2 // xA
    return [(iu * xA, iv * xA) for iv in range(-k, k + 1) for iu in range(-k, k + 1)]


def cut(g, off_u, off_v, xM):
    N = g.shape[0]
    y0 = N // 2 + off_v - xM // 2
    x0 = N // 2 + off_u - xM // 2
    return g[y0:y0 + xM, x0:x0 + xM]


def test_streamer_subgrids_reproduce_full_grid_degrid():
    """streamer_degrid_worker over a tiling of subgrids (windows xA wide,
    subgrids xM wide, cut from one grid): every visibility inside the tiled
    area is degridded by exactly one subgrid, and its value equals
    degrid_conv_uv on the full grid at (u, v), or the conjugate at (-u, -v)
    for the mirrored half; all other slot entries are zero."""
    rng = numpy.random.default_rng(5)
    N, theta, S, Q, xA, xM = 256, 0.05, 8, 16, 32, 64
    kern = rng.normal(size=(Q, S))
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    tc, fc = 4, 8
    uvw, freq = bl_tracks(rng, 10, 8, 16, 6000.0)
    s = B.BlSet(uvw, freq)
    u, v = lam(s, 0), lam(s, 1)
    chunks = numpy.array([[b, t, t + tc, f, f + fc] for b in range(s.nbl)
                          for t in range(0, 8, tc) for f in range(0, 16, fc)], dtype=numpy.int32)
    merged = numpy.zeros(s.vis.shape[0], dtype=complex)
    times = numpy.zeros(s.vis.shape[0], dtype=int)
    for off_u, off_v in tiles(N, xA, xM):
        sub = cut(g, off_u, off_v, xM)
        box = ((off_u - xA // 2) / theta, (off_u + xA // 2) / theta,
               (off_v - xA // 2) / theta, (off_v + xA // 2) / theta)
        out, nd = B.streamer_degrid(sub, theta, off_u / theta, off_v / theta, *box, s, chunks,
                                    tc, fc, kern)
        for c, (b, t0, t1, f0, f1) in enumerate(chunks):
            slot = out[c].reshape(tc, fc)
            idx = (int(s.v0[b]) + numpy.arange(t0, t1)[:, None] * int(s.nf[b])
                   + numpy.arange(f0, f1)[None, :])
            hit = slot != 0
            assert hit.sum() == nd[c]
            times[idx[hit]] += 1
            merged[idx[hit]] = slot[hit]
    k = (N - xM) // 2 // xA
    edge = (k * xA + xA // 2) / theta
    inside = (numpy.abs(u) < edge - 1) & (numpy.abs(v) < edge - 1) & (u != 0)
    assert inside.sum() > 100 and (~inside).sum() > 100
    assert (times[inside] == 1).all()
    assert (times <= 1).all()
    pos = inside & (u > 0)
    neg = inside & (u < 0)
    uv = numpy.stack([u, v], axis=1)
    assert_allclose(merged[pos], B.degrid_conv_uv(g, theta, uv[pos], kern), rtol=1e-11, atol=1e-11)
    assert_allclose(merged[neg], numpy.conj(B.degrid_conv_uv(g, theta, -uv[neg], kern)),
                    rtol=1e-11, atol=1e-11)


def test_streamer_u_zero_is_degridded_once():
    """u = 0 counts as the normal half only (streamer.c:418-433: -u must be
    at least DBL_MIN)."""
    N, theta, S, Q, xA, xM = 128, 0.1, 4, 4, 32, 64
    kern = numpy.ones((Q, S))
    g = numpy.arange(N * N).reshape(N, N).astype(complex)
    s = B.BlSet([numpy.array([[0.0, 30.0, 0.0], [0.0, -30.0, 0.0]])], [numpy.array([C])])
    chunks = numpy.array([[0, 0, 2, 0, 1]], dtype=numpy.int32)
    hits = numpy.zeros(2, dtype=int)
    for off_u, off_v in tiles(N, xA, xM):
        sub = cut(g, off_u, off_v, xM)
        box = ((off_u - xA // 2) / theta, (off_u + xA // 2) / theta,
               (off_v - xA // 2) / theta, (off_v + xA // 2) / theta)
        out, _ = B.streamer_degrid(sub, theta, off_u / theta, off_v / theta, *box, s, chunks,
                                   2, 1, kern)
        hits += out[0] != 0
    assert_array_equal(hits, [1, 1])


# ------------------------------------------------------------ w-projection

def test_grid_wprojection_is_conj_convgrid_per_plane():
    """w_project (grid.c:277-305): plane floor((w - w_min)/w_step + .5),
    grid += v conj(k): equal to synthesis.py's convgrid with the conjugated
    table of each plane at p = theta u / N."""
    rng = numpy.random.default_rng(8)
    N, theta, S, Q, P = 64, 0.02, 6, 4, 5
    wk = rng.normal(size=(P, Q, Q, S, S)) + 1j * rng.normal(size=(P, Q, Q, S, S))
    uvw, freq = bl_tracks(rng, 5, 6, 5, 400.0)
    vis = [rng.normal(size=len(a) * len(f)) + 1j * rng.normal(size=len(a) * len(f))
           for a, f in zip(uvw, freq)]
    s = B.BlSet(uvw, freq, vis)
    w = lam(s, 2)
    w_min, w_step = w.min() + 3.0, (w.max() - w.min()) / (P - 1)
    g = numpy.zeros((N, N), dtype=complex)
    B.grid_wprojection(g, theta, s, wk, w_min, w_step)
    ref = numpy.zeros((N, N), dtype=complex)
    plane = numpy.floor((w - w_min) / w_step + .5).astype(int)
    p = numpy.stack([lam(s, 0), lam(s, 1)], axis=1) * theta / N
    for k in range(P):
        sel = plane == k
        R.convgrid(numpy.conj(wk[k]), ref, p[sel], None, s.vis[sel])
    assert_allclose(g, ref, rtol=1e-12, atol=1e-12)


# -------------------------------------------------------- nearest-cell layer

def test_simple_grid_degrid_weight():
    """grid_simple / degrid_simple / weight (grid.c:119-177) against
    synthesis.py's nearest-cell grid / degrid at p = theta u / N and a NumPy
    count of visibilities per cell."""
    rng = numpy.random.default_rng(9)
    N, theta = 64, 0.05
    uvw, freq = bl_tracks(rng, 7, 9, 6, 250.0, ragged=True)
    vis = [rng.normal(size=len(a) * len(f)) + 1j * rng.normal(size=len(a) * len(f))
           for a, f in zip(uvw, freq)]
    s = B.BlSet(uvw, freq, vis)
    p = numpy.stack([lam(s, 0), lam(s, 1)], axis=1) * theta / N
    g = numpy.zeros((N, N), dtype=complex)
    B.grid_simple(g, theta, s)
    ref = numpy.zeros((N, N), dtype=complex)
    R.grid(ref, p, s.vis)
    assert_allclose(g, ref, rtol=1e-14, atol=1e-14)
    s2 = B.BlSet(uvw, freq)
    B.degrid_simple(ref, theta, s2)
    assert_allclose(s2.vis, R.degrid(ref, p), rtol=0, atol=0)
    v0 = s.vis.copy()
    wg = B.weight(N, theta, s)
    cell = (N // 2 + numpy.floor(0.5 + N * p).astype(int))
    flat = cell[:, 1] * N + cell[:, 0]
    counts = numpy.bincount(flat, minlength=N * N)
    assert_array_equal(wg.ravel(), counts)
    assert_allclose(s.vis, v0 / counts[flat], rtol=1e-15)
