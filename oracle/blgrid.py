"""ctypes binding of the C restatement of the reference's visibility-layout
functions (blgrid.c) -- TEST INFRASTRUCTURE.

A baseline set holds, per baseline, uvw rows in metres [nt, 3], channel
frequencies in Hz [nf] and visibilities [nt * nf] (time-major), as the
reference's struct bl_data (examples/grid/grid.h:14-34); here they are packed
into flat arrays with per-baseline offsets.
"""

import ctypes

import numpy

from .cgrid import lib as _cglib, _ptr

_EBOUNDS, _EARG = -2, -3


class _BlSetC(ctypes.Structure):
    _fields_ = [('nbl', ctypes.c_int64), ('nt', ctypes.c_void_p), ('nf', ctypes.c_void_p),
                ('uvw0', ctypes.c_void_p), ('f0', ctypes.c_void_p), ('v0', ctypes.c_void_p),
                ('uvw_m', ctypes.c_void_p), ('freq', ctypes.c_void_p), ('vis', ctypes.c_void_p)]


_SIGS = None


def lib():
    global _SIGS
    L = _cglib()
    if _SIGS is None:
        P, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        B = ctypes.POINTER(_BlSetC)
        L.bl_frac_coords.argtypes = [I, I, P, I64, P, P]
        L.bl_degrid_conv_uv.argtypes = [P, I, D, P, I64, P, I, I, P]
        L.bl_degrid_conv_bl.argtypes = [P, I, D, D, D, D, D, D, D, B, I, I, I, I, P, I, I, P]
        L.bl_streamer_degrid.argtypes = [P, I, D, D, D, D, D, D, D, B, P, I64, I, I, P, I, I,
                                         P, P]
        L.bl_grid_wprojection.argtypes = [P, I, D, B, P, I, D, D, I, I]
        L.bl_weight.argtypes = [P, I, D, B]
        L.bl_grid_simple.argtypes = [P, I, D, B]
        L.bl_degrid_simple.argtypes = [P, I, D, B]
        _SIGS = True
    return L


def _check(rc):
    if rc == _EBOUNDS:
        raise ValueError('kernel window / cell outside the grid')
    if rc == _EARG:
        raise ValueError('time / frequency window outside a baseline')
    if rc != 0:
        raise RuntimeError('blgrid error %d' % rc)


class BlSet:
    """Packed baselines: ``uvw_m`` list of [nt, 3] metres, ``freq`` list of
    [nf] Hz, ``vis`` list of [nt * nf] complex (zeros when omitted)."""

    def __init__(self, uvw_m, freq, vis=None):
        self.nbl = len(uvw_m)
        self.nt = numpy.array([len(u) for u in uvw_m], dtype=numpy.int32)
        self.nf = numpy.array([len(f) for f in freq], dtype=numpy.int32)
        self.uvw0 = numpy.concatenate([[0], numpy.cumsum(self.nt)[:-1]]).astype(numpy.int64)
        self.f0 = numpy.concatenate([[0], numpy.cumsum(self.nf)[:-1]]).astype(numpy.int64)
        nvis = self.nt.astype(numpy.int64) * self.nf
        self.v0 = numpy.concatenate([[0], numpy.cumsum(nvis)[:-1]]).astype(numpy.int64)
        self.uvw_m = numpy.ascontiguousarray(numpy.concatenate(uvw_m), dtype=numpy.float64)
        self.freq = numpy.ascontiguousarray(numpy.concatenate(freq), dtype=numpy.float64)
        if vis is None:
            self.vis = numpy.zeros(int(nvis.sum()), dtype=numpy.complex128)
        else:
            self.vis = numpy.ascontiguousarray(numpy.concatenate(vis), dtype=numpy.complex128)
        assert self.vis.shape[0] == int(nvis.sum())

    @classmethod
    def uniform(cls, uvw_m, freq, with_vis=True):
        """Baselines sharing one channel axis: ``uvw_m`` [nbl, nt, 3],
        ``freq`` [nf]; visibilities zero (none held without ``with_vis``:
        for streamer_degrid, which writes queue slots only)."""
        self = cls.__new__(cls)
        uvw_m = numpy.ascontiguousarray(uvw_m, dtype=numpy.float64)
        nbl, nt = uvw_m.shape[:2]
        nf = len(freq)
        self.nbl = nbl
        self.nt = numpy.full(nbl, nt, dtype=numpy.int32)
        self.nf = numpy.full(nbl, nf, dtype=numpy.int32)
        self.uvw0 = numpy.arange(nbl, dtype=numpy.int64) * nt
        self.f0 = numpy.zeros(nbl, dtype=numpy.int64)
        self.v0 = numpy.arange(nbl, dtype=numpy.int64) * nt * nf
        self.uvw_m = uvw_m.reshape(nbl * nt, 3)
        self.freq = numpy.ascontiguousarray(freq, dtype=numpy.float64)
        self.vis = numpy.zeros(nbl * nt * nf if with_vis else 1, dtype=numpy.complex128)
        return self

    def c(self):
        return _BlSetC(self.nbl, _ptr(self.nt).value, _ptr(self.nf).value,
                       _ptr(self.uvw0).value, _ptr(self.f0).value, _ptr(self.v0).value,
                       _ptr(self.uvw_m).value, _ptr(self.freq).value, _ptr(self.vis).value)

    def baseline_vis(self, b):
        o = int(self.v0[b])
        return self.vis[o:o + int(self.nt[b]) * int(self.nf[b])]


def _grid(g):
    g = numpy.ascontiguousarray(g, dtype=numpy.complex128)
    assert g.ndim == 2 and g.shape[0] == g.shape[1]
    return g


def _kern(k):
    k = numpy.ascontiguousarray(k, dtype=numpy.float64)
    assert k.ndim == 2
    return k


def frac_coord(grid_size, oversample, u):
    u = numpy.ascontiguousarray(u, dtype=numpy.float64).ravel()
    x = numpy.empty(u.shape[0], dtype=numpy.int64)
    fx = numpy.empty(u.shape[0], dtype=numpy.int64)
    _check(lib().bl_frac_coords(grid_size, oversample, _ptr(u), u.shape[0], _ptr(x), _ptr(fx)))
    return x, fx


def degrid_conv_uv(g, theta, uv, kern):
    g, kern = _grid(g), _kern(kern)
    uv = numpy.ascontiguousarray(uv, dtype=numpy.float64)
    out = numpy.empty(uv.shape[0], dtype=numpy.complex128)
    _check(lib().bl_degrid_conv_uv(_ptr(g), g.shape[0], theta, _ptr(uv), uv.shape[0],
                                   _ptr(kern), kern.shape[1], kern.shape[0], _ptr(out)))
    return out


def degrid_conv_bl(g, theta, d_u, d_v, min_u, max_u, min_v, max_v, s, time0, time1,
                   freq0, freq1, kern):
    """Writes into ``s.vis``; returns the number of degridded visibilities."""
    g, kern = _grid(g), _kern(kern)
    n = ctypes.c_int64()
    c = s.c()
    _check(lib().bl_degrid_conv_bl(_ptr(g), g.shape[0], theta, d_u, d_v, min_u, max_u,
                                   min_v, max_v, ctypes.byref(c), time0, time1, freq0, freq1,
                                   _ptr(kern), kern.shape[1], kern.shape[0], ctypes.byref(n)))
    return n.value


def streamer_degrid(g, theta, mid_u, mid_v, min_u, max_u, min_v, max_v, s, chunks,
                    time_chunk, freq_chunk, kern):
    """(slots [nchunks, time_chunk * freq_chunk], degridded count per chunk)."""
    g, kern = _grid(g), _kern(kern)
    chunks = numpy.ascontiguousarray(chunks, dtype=numpy.int32).reshape(-1, 5)
    out = numpy.empty((chunks.shape[0], time_chunk * freq_chunk), dtype=numpy.complex128)
    nd = numpy.empty(chunks.shape[0], dtype=numpy.int64)
    c = s.c()
    _check(lib().bl_streamer_degrid(_ptr(g), g.shape[0], theta, mid_u, mid_v, min_u, max_u,
                                    min_v, max_v, ctypes.byref(c), _ptr(chunks),
                                    chunks.shape[0], time_chunk, freq_chunk, _ptr(kern),
                                    kern.shape[1], kern.shape[0], _ptr(out), _ptr(nd)))
    return out, nd


def grid_wprojection(g, theta, s, wkern, w_min, w_step):
    """In-place into complex128 ``g``; ``wkern`` [planes, Q, Q, S, S]."""
    assert g.dtype == numpy.complex128 and g.flags.c_contiguous
    wkern = numpy.ascontiguousarray(wkern, dtype=numpy.complex128)
    P, Q, _, S, _ = wkern.shape
    c = s.c()
    _check(lib().bl_grid_wprojection(_ptr(g), g.shape[0], theta, ctypes.byref(c), _ptr(wkern),
                                     P, w_min, w_step, S, Q))


def weight(N, theta, s):
    """Returns the count grid; divides ``s.vis`` in place."""
    wg = numpy.empty((N, N), dtype=numpy.uint32)
    c = s.c()
    _check(lib().bl_weight(_ptr(wg), N, theta, ctypes.byref(c)))
    return wg


def grid_simple(g, theta, s):
    assert g.dtype == numpy.complex128 and g.flags.c_contiguous
    c = s.c()
    _check(lib().bl_grid_simple(_ptr(g), g.shape[0], theta, ctypes.byref(c)))


def degrid_simple(g, theta, s):
    g = _grid(g)
    c = s.c()
    _check(lib().bl_degrid_simple(_ptr(g), g.shape[0], theta, ctypes.byref(c)))
