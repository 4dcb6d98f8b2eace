"""The reference's native visibility layout and its gridding functions on the
GPU.

Mirrors examples/grid/grid.h -- ``struct bl_data`` (:14-34), ``struct
vis_data`` (:35-40), ``struct sep_kernel_data`` (:101-106), ``struct
w_kernel_data`` (:114-122) -- and the functions of grid.c and streamer.c over
them, with the reference's names and argument meaning:

* ``weight``, ``grid_simple``, ``degrid_simple`` (grid.c:119-177),
* ``degrid_conv_uv``, ``degrid_conv_bl``, ``degrid_conv`` (grid.c:179-275),
* ``grid_wprojection`` (grid.c:307-322 over ``w_project`` :277-305),
* ``streamer_degrid_worker`` (streamer.c:366-486) and the chunk selection of
  ``streamer_degrid_chunk`` (streamer.c:488-600) as ``Streamer``.

``VisData`` packs its baselines into flat device arrays once (the C ABI's
``croc_bl_set``); every function then runs in libcrocgpu on the device.
Visibilities are complex64 on the device, coordinates and kernel tables FP64.
Each function returns the reference's flop count.  Nothing falls back to the
CPU.
"""

import ctypes

import numpy
import torch

from . import _lib

SPEED_OF_LIGHT = 299792458.0  # grid.h:51


def _dev():
    return _lib.device()


def _t(x, dtype, device=None):
    device = device or _dev()
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).resolve_conj().resolve_neg().contiguous()
    return torch.as_tensor(numpy.ascontiguousarray(x)).to(device=device, dtype=dtype).contiguous()


def uvw_m_to_l(u, f):
    """grid.h:53-55: metres to wavelengths at frequency f (Hz)."""
    return u * f / SPEED_OF_LIGHT


def uvw_l_to_m(u, f):
    """grid.h:56-58."""
    return u / f * SPEED_OF_LIGHT


class BlData:
    """One baseline (grid.h:14-34): ``time`` [nt], ``freq`` [nf] (Hz),
    ``uvw_m`` [nt, 3] (metres), ``vis`` [nt * nf] (time-major)."""

    def __init__(self, antenna1, antenna2, time, freq, uvw_m, vis=None):
        self.antenna1, self.antenna2 = int(antenna1), int(antenna2)
        self.time = numpy.asarray(time, dtype=numpy.float64)
        self.freq = numpy.asarray(freq, dtype=numpy.float64)
        self.uvw_m = numpy.asarray(uvw_m, dtype=numpy.float64).reshape(-1, 3)
        self.time_count, self.freq_count = self.uvw_m.shape[0], self.freq.shape[0]
        n = self.time_count * self.freq_count
        self.vis = (numpy.zeros(n, dtype=numpy.complex128) if vis is None
                    else numpy.asarray(vis, dtype=numpy.complex128).reshape(n))

    def uvw_lambda(self, time, freq, uvw):
        """grid.h:59-62."""
        return uvw_m_to_l(self.uvw_m[time, uvw], self.freq[freq])


class VisData:
    """A set of baselines (grid.h:35-40) packed in device memory: per
    baseline counts and offsets, ``uvw_m`` [rows, 3], ``freq`` [entries],
    ``vis`` complex64 [nvis]."""

    def __init__(self, bls=(), antenna_count=0, device=None):
        bls = list(bls)
        self.antenna_count = int(antenna_count)
        self.antennas = [(b.antenna1, b.antenna2) for b in bls]
        nt = numpy.array([b.time_count for b in bls], dtype=numpy.int32)
        nf = numpy.array([b.freq_count for b in bls], dtype=numpy.int32)
        uvw = (numpy.concatenate([b.uvw_m for b in bls]) if bls
               else numpy.zeros((0, 3)))
        freq = numpy.concatenate([b.freq for b in bls]) if bls else numpy.zeros(0)
        vis = numpy.concatenate([b.vis for b in bls]) if bls else numpy.zeros(0, complex)
        self._pack(nt, nf, uvw, freq, vis, device)
        self.time = [b.time for b in bls]

    @classmethod
    def uniform(cls, uvw_m, freq, vis=None, device=None):
        """Baselines sharing one channel axis (the streamer's vis_spec,
        config.c:835-861): ``uvw_m`` [nbl, nt, 3] metres, ``freq`` [nf],
        ``vis`` [nbl, nt, nf] or None.  Arrays may be device tensors."""
        self = cls.__new__(cls)
        device = device or _dev()
        uvw_m = _t(uvw_m, torch.float64, device)
        nbl, nt = uvw_m.shape[0], uvw_m.shape[1]
        freq = _t(freq, torch.float64, device)
        nf = freq.shape[0]
        self.antenna_count = 0
        self.antennas = [(0, 0)] * nbl
        self.time = None
        self.device = device
        self.bl_count = nbl
        self.time_count = torch.full((nbl,), nt, dtype=torch.int32, device=device)
        self.freq_count = torch.full((nbl,), nf, dtype=torch.int32, device=device)
        ar = torch.arange(nbl, dtype=torch.int64, device=device)
        self.uvw_first = ar * nt
        self.freq_first = torch.zeros(nbl, dtype=torch.int64, device=device)
        self.vis_first = ar * (nt * nf)
        self.uvw_m = uvw_m.reshape(nbl * nt, 3).contiguous()
        self.freq = freq
        self.vis = (torch.zeros(nbl * nt * nf, dtype=torch.complex64, device=device)
                    if vis is None else _t(vis, torch.complex64, device).reshape(-1))
        self._nt_h = numpy.full(nbl, nt, dtype=numpy.int64)
        self._nf_h = numpy.full(nbl, nf, dtype=numpy.int64)
        self._v0_h = numpy.arange(nbl, dtype=numpy.int64) * nt * nf
        return self

    def _pack(self, nt, nf, uvw, freq, vis, device):
        device = device or _dev()
        self.device = device
        self.bl_count = len(nt)
        self._nt_h = nt.astype(numpy.int64)
        self._nf_h = nf.astype(numpy.int64)
        uvw0 = numpy.concatenate([[0], numpy.cumsum(self._nt_h)[:-1]]).astype(numpy.int64)
        f0 = numpy.concatenate([[0], numpy.cumsum(self._nf_h)[:-1]]).astype(numpy.int64)
        self._v0_h = numpy.concatenate(
            [[0], numpy.cumsum(self._nt_h * self._nf_h)[:-1]]).astype(numpy.int64)
        self.time_count = _t(nt, torch.int32, device)
        self.freq_count = _t(nf, torch.int32, device)
        self.uvw_first = _t(uvw0[:len(nt)], torch.int64, device)
        self.freq_first = _t(f0[:len(nt)], torch.int64, device)
        self.vis_first = _t(self._v0_h[:len(nt)], torch.int64, device)
        self.uvw_m = _t(uvw.reshape(-1, 3), torch.float64, device)
        self.freq = _t(freq, torch.float64, device)
        self.vis = _t(vis, torch.complex64, device)

    def c_struct(self):
        """The croc_bl_set of this set (device pointers; keep self alive)."""
        for name in ('time_count', 'freq_count', 'uvw_first', 'freq_first', 'vis_first',
                     'uvw_m', 'freq', 'vis'):
            t = getattr(self, name)
            if not t.is_contiguous():
                raise ValueError('VisData.%s must be contiguous' % name)
        return _lib.BlSetC(self.bl_count, self.uvw_m.shape[0], self.freq.shape[0],
                           self.vis.shape[0], self.time_count.data_ptr(),
                           self.freq_count.data_ptr(), self.uvw_first.data_ptr(),
                           self.freq_first.data_ptr(), self.vis_first.data_ptr(),
                           self.uvw_m.data_ptr(), self.freq.data_ptr(), self.vis.data_ptr())

    def baseline_vis(self, b):
        """Device view of baseline b's visibilities, [nt, nf]."""
        o = int(self._v0_h[b])
        nt, nf = int(self._nt_h[b]), int(self._nf_h[b])
        return self.vis[o:o + nt * nf].view(nt, nf)

    def to_bl_data(self):
        """Host BlData list with the device visibilities."""
        vis = self.vis.cpu().numpy().astype(numpy.complex128)
        uvw = self.uvw_m.cpu().numpy()
        freq = self.freq.cpu().numpy()
        u0 = self.uvw_first.cpu().numpy()
        f0 = self.freq_first.cpu().numpy()
        out = []
        for b in range(self.bl_count):
            nt, nf, v0 = int(self._nt_h[b]), int(self._nf_h[b]), int(self._v0_h[b])
            a1, a2 = self.antennas[b]
            t = self.time[b] if self.time else numpy.arange(nt, dtype=float)
            out.append(BlData(a1, a2, t, freq[f0[b]:f0[b] + nf], uvw[u0[b]:u0[b] + nt],
                              vis[v0:v0 + nt * nf]))
        return out


class SepKernelData:
    """Separable real kernel (grid.h:101-106): ``data`` [oversampling, size]
    FP64, row = sub-pixel offset."""

    def __init__(self, data, device=None):
        d = data if isinstance(data, torch.Tensor) else numpy.asarray(data)
        if numpy.iscomplexobj(d.cpu().numpy() if isinstance(d, torch.Tensor) else d):
            raise ValueError('separable kernels are real (sep_kernel_data holds doubles)')
        self.data = _t(d, torch.float64, device)
        if self.data.ndim != 2:
            raise ValueError('separable kernel data must be [oversampling, size]')
        self.oversampling, self.size = self.data.shape


class WKernelData:
    """Equal-step w-kernel set (grid.h:114-122): ``data`` [plane_count,
    oversampling, oversampling, size, size] complex, plane k at
    w = w_min + k w_step."""

    def __init__(self, data, w_min, w_step, device=None):
        self.data = _t(data, torch.complex64, device)
        if self.data.ndim != 5 or self.data.shape[1] != self.data.shape[2] or \
                self.data.shape[3] != self.data.shape[4]:
            raise ValueError('w-kernel data must be [planes, Q, Q, S, S]')
        self.plane_count, self.oversampling, _, self.size_y, self.size_x = self.data.shape
        self.w_min, self.w_step = float(w_min), float(w_step)
        self.w_max = self.w_min + (self.plane_count - 1) * self.w_step


def _grid_arg(uvgrid, grid_size, dtype=torch.complex64):
    if not isinstance(uvgrid, torch.Tensor) or uvgrid.dtype != dtype or \
            uvgrid.device.type != 'cuda' or not uvgrid.is_contiguous():
        raise ValueError('uvgrid must be a contiguous %s device tensor' % dtype)
    if tuple(uvgrid.shape) != (grid_size, grid_size):
        raise ValueError('uvgrid must be [grid_size, grid_size]')
    return _lib.ptr(uvgrid)


def _flops_ref():
    return ctypes.c_uint64(0)


def weight(wgrid, grid_size, theta, vis):
    """grid.c:119-141: uniform weighting in place; ``wgrid`` an int32/uint32
    [grid_size, grid_size] device tensor (overwritten with the counts) or
    None (allocated and returned)."""
    if wgrid is None:
        wgrid = torch.empty((grid_size, grid_size), dtype=torch.int32, device=vis.device)
    if wgrid.dtype not in (torch.int32, torch.uint32) or tuple(wgrid.shape) != \
            (grid_size, grid_size) or not wgrid.is_contiguous():
        raise ValueError('wgrid must be a contiguous 32-bit [grid_size, grid_size] tensor')
    s = vis.c_struct()
    _lib.check(_lib.load().croc_bl_weight(_lib.ctx(), _lib.ptr(wgrid), grid_size, theta,
                                          ctypes.byref(s)))
    return wgrid


def grid_simple(uvgrid, grid_size, theta, vis):
    """grid.c:161-177; returns flops."""
    f = _flops_ref()
    s = vis.c_struct()
    _lib.check(_lib.load().croc_bl_grid_simple(_lib.ctx(), _grid_arg(uvgrid, grid_size),
                                               grid_size, theta, ctypes.byref(s),
                                               ctypes.byref(f)))
    return f.value


def degrid_simple(uvgrid, grid_size, theta, vis):
    """grid.c:143-159; returns flops."""
    f = _flops_ref()
    s = vis.c_struct()
    _lib.check(_lib.load().croc_bl_degrid_simple(_lib.ctx(), _grid_arg(uvgrid, grid_size),
                                                 grid_size, theta, ctypes.byref(s),
                                                 ctypes.byref(f)))
    return f.value


def degrid_conv_uv(uvgrid, grid_size, theta, u, v, kernel, flops=None):
    """grid.c:179-232 for arrays u, v (wavelengths): complex64 device tensor
    of visibilities.  ``flops``: a one-element list that receives the count."""
    uv = torch.stack([_t(u, torch.float64).reshape(-1), _t(v, torch.float64).reshape(-1)],
                     dim=1).contiguous()
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    f = _flops_ref()
    _lib.check(_lib.load().croc_degrid_conv_uv(
        _lib.ctx(), _grid_arg(uvgrid, grid_size), grid_size, theta, _lib.ptr(uv), uv.shape[0],
        _lib.ptr(kernel.data), kernel.size, kernel.oversampling, _lib.ptr(out),
        ctypes.byref(f)))
    if flops is not None:
        flops.append(f.value)
    return out


def degrid_conv_bl(uvgrid, grid_size, theta, d_u, d_v, min_u, max_u, min_v, max_v, bl,
                   time0, time1, freq0, freq1, kernel):
    """grid.c:234-259 for every baseline of ``bl`` (a VisData; one baseline
    is the reference's call); returns flops."""
    f = _flops_ref()
    s = bl.c_struct()
    _lib.check(_lib.load().croc_degrid_conv_bl(
        _lib.ctx(), _grid_arg(uvgrid, grid_size), grid_size, theta, d_u, d_v, min_u, max_u,
        min_v, max_v, ctypes.byref(s), time0, time1, freq0, freq1, _lib.ptr(kernel.data),
        kernel.size, kernel.oversampling, ctypes.byref(f)))
    return f.value


def degrid_conv(uvgrid, grid_size, theta, d_u, d_v, vis, kernel):
    """grid.c:261-275; returns flops."""
    f = _flops_ref()
    s = vis.c_struct()
    _lib.check(_lib.load().croc_degrid_conv(
        _lib.ctx(), _grid_arg(uvgrid, grid_size), grid_size, theta, d_u, d_v, ctypes.byref(s),
        _lib.ptr(kernel.data), kernel.size, kernel.oversampling, ctypes.byref(f)))
    return f.value


def grid_wprojection(uvgrid, grid_size, theta, vis, wkern):
    """grid.c:307-322 (w_project :277-305: v * conj(k) of the w-plane
    floor((w - w_min) / w_step + .5)); returns flops."""
    if wkern.size_x != wkern.size_y:
        raise ValueError('w_project needs square kernels (grid.c:296)')
    f = _flops_ref()
    s = vis.c_struct()
    _lib.check(_lib.load().croc_grid_wprojection(
        _lib.ctx(), _grid_arg(uvgrid, grid_size), grid_size, theta, ctypes.byref(s),
        _lib.ptr(wkern.data), wkern.plane_count, wkern.w_min, wkern.w_step, wkern.size_x,
        wkern.oversampling, ctypes.byref(f)))
    return f.value


def streamer_degrid_worker(vis, subgrid, subgrid_size, theta, mid_u, mid_v, chunks,
                           min_u, max_u, min_v, max_v, kernel, time_chunk, freq_chunk,
                           vis_queue=None, ndegrid=None, want_flops=True):
    """streamer.c:366-486 for a list of chunks (int32 [n, 5]: baseline, it0,
    it1, if0, if1) of one subgrid: returns (vis_queue complex64 [n,
    time_chunk * freq_chunk], ndegrid int64 [n], flops).  Slot entries beyond
    a chunk's extent are left as they were (zeros when allocated here).
    With ``want_flops`` False the call returns as soon as its inputs are
    validated (flops None; ndegrid and the slots fill in on the stream)."""
    chunks = _t(chunks, torch.int32, vis.device).reshape(-1, 5)
    n = chunks.shape[0]
    vis_queue, ndegrid = _queue(vis, n, time_chunk, freq_chunk, vis_queue, ndegrid)
    f = _flops_ref()
    s = vis.c_struct()
    _lib.check(_lib.load().croc_streamer_degrid(
        _lib.ctx(), _grid_arg(subgrid, subgrid_size), subgrid_size, theta, mid_u, mid_v,
        min_u, max_u, min_v, max_v, ctypes.byref(s), _lib.ptr(chunks), n, time_chunk,
        freq_chunk, _lib.ptr(kernel.data), kernel.size, kernel.oversampling,
        _lib.ptr(vis_queue), _lib.ptr(ndegrid), ctypes.byref(f) if want_flops else None))
    return vis_queue, ndegrid, (f.value if want_flops else None)


def _queue(vis, n, time_chunk, freq_chunk, vis_queue, ndegrid):
    if vis_queue is None:
        vis_queue = torch.zeros((n, time_chunk * freq_chunk), dtype=torch.complex64,
                                device=vis.device)
    if ndegrid is None:
        ndegrid = torch.empty(n, dtype=torch.int64, device=vis.device)
    if vis_queue.dtype != torch.complex64 or vis_queue.numel() < n * time_chunk * freq_chunk \
            or not vis_queue.is_contiguous():
        raise ValueError('vis_queue must hold n * time_chunk * freq_chunk complex64')
    if ndegrid.dtype != torch.int64 or ndegrid.numel() < n or not ndegrid.is_contiguous():
        raise ValueError('ndegrid must hold n int64')
    return vis_queue, ndegrid


def streamer_degrid_batch(vis, subgrids, theta, windows, chunks, kernel, time_chunk,
                          freq_chunk, vis_queue=None, ndegrid=None, want_flops=True):
    """streamer_degrid_worker for many subgrids in one launch (an extension:
    the reference degrids one subgrid per call).  ``subgrids`` complex64
    [nsub, xM, xM] device tensor; ``windows`` [nsub, 6] = (mid_u, mid_v,
    min_u, max_u, min_v, max_v); ``chunks`` int32 [n, 6] = (subgrid,
    baseline, it0, it1, if0, if1).  Returns (vis_queue, ndegrid, flops)."""
    if not isinstance(subgrids, torch.Tensor) or subgrids.dtype != torch.complex64 or \
            subgrids.ndim != 3 or subgrids.shape[1] != subgrids.shape[2] or \
            not subgrids.is_contiguous() or subgrids.device.type != 'cuda':
        raise ValueError('subgrids must be a contiguous complex64 [nsub, xM, xM] device tensor')
    win = numpy.ascontiguousarray(windows, dtype=numpy.float64).reshape(-1, 6)
    if win.shape[0] != subgrids.shape[0]:
        raise ValueError('one window per subgrid')
    chunks = _t(chunks, torch.int32, vis.device).reshape(-1, 6)
    n = chunks.shape[0]
    vis_queue, ndegrid = _queue(vis, n, time_chunk, freq_chunk, vis_queue, ndegrid)
    f = _flops_ref()
    s = vis.c_struct()
    _lib.check(_lib.load().croc_streamer_degrid_batch(
        _lib.ctx(), _lib.ptr(subgrids), subgrids.shape[0], subgrids.shape[1], theta,
        win.ctypes.data_as(ctypes.c_void_p), ctypes.byref(s), _lib.ptr(chunks), n,
        time_chunk, freq_chunk, _lib.ptr(kernel.data), kernel.size, kernel.oversampling,
        _lib.ptr(vis_queue), _lib.ptr(ndegrid), ctypes.byref(f) if want_flops else None))
    return vis_queue, ndegrid, (f.value if want_flops else None)


class Streamer:
    """The streamer's degridding side (streamer.c:488-600) on one GPU: for
    each subgrid (a cf32 [xM, xM] device array centred at cell offset
    (subgrid_off_u, subgrid_off_v) from the grid centre, as the facet
    recombination delivers it), select the (baseline, time chunk, frequency
    chunk) triples whose uv bounding box meets the subgrid's xA window or
    its mirror image (streamer_degrid_chunk :503-535), degrid them into
    queue slots (streamer_degrid_worker) and merge the slots into the
    visibility set (the writer, :322-337: nonzero slot entries are copied,
    and an entry is never written twice)."""

    def __init__(self, vis, theta, xA_size, xM_size, time_chunk, freq_chunk, kernel,
                 image_size=None):
        self.vis, self.theta = vis, float(theta)
        self.xA, self.xM = int(xA_size), int(xM_size)
        self.time_chunk, self.freq_chunk = int(time_chunk), int(freq_chunk)
        self.kernel = kernel
        self.image_size = image_size
        uvw = vis.uvw_m.cpu().numpy()
        freq = vis.freq.cpu().numpy()
        u0 = vis.uvw_first.cpu().numpy()
        f0 = vis.freq_first.cpu().numpy()
        nt, nf = vis._nt_h, vis._nf_h
        tc, fc = self.time_chunk, self.freq_chunk
        ntc, nfc = (nt + tc - 1) // tc, (nf + fc - 1) // fc
        per = ntc * nfc
        b = numpy.repeat(numpy.arange(vis.bl_count), per)
        j = numpy.arange(int(per.sum())) - numpy.repeat(numpy.cumsum(per) - per, per)
        it0 = (j // nfc[b]) * tc
        if0 = (j % nfc[b]) * fc
        it1 = numpy.minimum(nt[b], it0 + tc)
        if1 = numpy.minimum(nf[b], if0 + fc)
        self._chunks = numpy.stack([b, it0, it1, if0, if1], axis=1).astype(numpy.int32)
        # uv bounding box of each chunk from its corner dumps and channels
        # (streamer.c:522-531)
        fa = uvw_m_to_l(1, freq[f0[b] + if0])
        fb = uvw_m_to_l(1, freq[f0[b] + if1 - 1])
        ua, va = uvw[u0[b] + it0, 0], uvw[u0[b] + it0, 1]
        ub, vb = uvw[u0[b] + it1 - 1, 0], uvw[u0[b] + it1 - 1, 1]
        cu = numpy.stack([ua * fa, ua * fb, ub * fa, ub * fb])
        cv = numpy.stack([va * fa, va * fb, vb * fa, vb * fb])
        self._box = (cu.min(axis=0), cu.max(axis=0), cv.min(axis=0), cv.max(axis=0))

    def window(self, subgrid_off_u, subgrid_off_v):
        """(mid_u, mid_v, min_u, max_u, min_v, max_v) of a subgrid
        (streamer.c:502-511)."""
        th = self.theta
        mid_u, mid_v = subgrid_off_u / th, subgrid_off_v / th
        min_u = (subgrid_off_u - self.xA / 2) / th
        min_v = (subgrid_off_v - self.xA / 2) / th
        max_u = (subgrid_off_u + self.xA / 2) / th
        max_v = (subgrid_off_v + self.xA / 2) / th
        if self.image_size is not None and min_v > self.image_size / th / 2:
            min_v -= self.image_size / th / 2
            max_v -= self.image_size / th / 2
        return mid_u, mid_v, min_u, max_u, min_v, max_v

    def chunks_for(self, subgrid_off_u, subgrid_off_v):
        """Chunks (int32 [n, 5]) whose bounding box overlaps the subgrid's
        window or its mirror (streamer.c:527-535)."""
        _, _, smin_u, smax_u, smin_v, smax_v = self.window(subgrid_off_u, subgrid_off_v)
        min_u, max_u, min_v, max_v = self._box
        overlap = (min_u < smax_u) & (max_u > smin_u) & (min_v < smax_v) & (max_v > smin_v)
        inv = (-max_u < smax_u) & (-min_u > smin_u) & (-max_v < smax_v) & (-min_v > smin_v)
        return self._chunks[overlap | inv]

    def degrid(self, subgrid, subgrid_off_u, subgrid_off_v, chunks=None):
        """Degrid one subgrid into queue slots: (chunks, vis_queue,
        ndegrid, flops)."""
        if chunks is None:
            chunks = self.chunks_for(subgrid_off_u, subgrid_off_v)
        mid_u, mid_v, min_u, max_u, min_v, max_v = self.window(subgrid_off_u, subgrid_off_v)
        ch = _t(chunks, torch.int32, self.vis.device).reshape(-1, 5)
        q, nd, flops = streamer_degrid_worker(
            self.vis, subgrid, self.xM, self.theta, mid_u, mid_v, ch, min_u, max_u, min_v,
            max_v, self.kernel, self.time_chunk, self.freq_chunk)
        return ch, q, nd, flops

    def batch(self, offsets):
        """Work of many subgrids at once: (windows [nsub, 6], chunks int32
        [n, 6] device tensor with the subgrid index first)."""
        wins, parts = [], []
        for i, (ou, ov) in enumerate(offsets):
            wins.append(self.window(ou, ov))
            ch = self.chunks_for(ou, ov)
            parts.append(numpy.concatenate(
                [numpy.full((len(ch), 1), i, dtype=numpy.int32), ch], axis=1))
        ch = numpy.concatenate(parts) if parts else numpy.zeros((0, 6), numpy.int32)
        return numpy.array(wins).reshape(-1, 6), _t(ch, torch.int32, self.vis.device)

    def degrid_batch(self, subgrids, offsets, work=None):
        """Degrid subgrids [nsub, xM, xM] at cell offsets in one launch:
        (chunks [n, 6], vis_queue, ndegrid, flops)."""
        wins, ch = work if work is not None else self.batch(offsets)
        q, nd, flops = streamer_degrid_batch(self.vis, subgrids, self.theta, wins, ch,
                                             self.kernel, self.time_chunk, self.freq_chunk)
        return ch, q, nd, flops

    def write(self, chunks, vis_queue):
        """The writer (streamer.c:322-337): copy every nonzero slot entry into
        the visibility set; raises if one lands on an entry written before."""
        if chunks.shape[0] == 0:
            return
        tc, fc = self.time_chunk, self.freq_chunk
        ch = chunks.to(torch.int64)
        if ch.shape[1] == 6:  # batched: drop the subgrid column
            ch = ch[:, 1:]
        b, t0, t1, f0, f1 = ch.unbind(1)
        tt = torch.arange(tc, device=ch.device)
        ff = torch.arange(fc, device=ch.device)
        nf = self.vis.freq_count.to(torch.int64)[b]
        base = self.vis.vis_first[b]
        idx = (base[:, None, None] + (t0[:, None, None] + tt[None, :, None]) * nf[:, None, None]
               + f0[:, None, None] + ff[None, None, :])
        valid = ((t0[:, None, None] + tt[None, :, None]) < t1[:, None, None]) & \
                ((f0[:, None, None] + ff[None, None, :]) < f1[:, None, None])
        q = vis_queue.reshape(-1, tc, fc)
        take = valid & (q != 0)
        sel = idx[take]
        if bool((self.vis.vis[sel] != 0).any()):
            raise RuntimeError('a visibility was degridded by two subgrids')
        self.vis.vis[sel] = q[take]
