"""GPU (MI355X / gfx950) drop-in for crocodile's gridding path.

Same function names, positional order, defaults and exceptions as
crocodile/synthesis.py for the convolutional gridding / degridding path:
coordinate helpers, kernel generation (``kernel_oversample``, ``w_kernel``),
``frac_coord(s)``, ``convgrid`` / ``convdegrid``, nearest-cell
``grid`` / ``degrid`` and the imaging / predict drivers (imgfn / predfn
protocol of synthesis.py:883-927).  Every function accepts NumPy arrays
(copied to the device; ``convgrid`` still updates the caller's array in
place) or torch tensors, and returns NumPy for NumPy input and torch device
tensors for torch input.

All arithmetic on the path runs in libcrocgpu (hand-written HIP):
coordinates and w-kernel phases in FP64, tap products in complex FP32,
gridding accumulated exactly in 64-bit fixed point (bit-reproducible),
degridding in FP32.  Nothing here falls back to the CPU; a missing library
or GPU raises.
"""

import ctypes
import math
import numbers

import numpy
import scipy.special
import torch

from . import _lib

# --------------------------------------------------------------------------
# helpers: array plumbing between NumPy / torch and device buffers
# --------------------------------------------------------------------------


def _is_torch(x):
    return isinstance(x, torch.Tensor)


def _dev(x, dtype):
    """Contiguous device tensor of ``dtype`` holding ``x``."""
    dev = _lib.device()
    if _is_torch(x):
        return x.to(device=dev, dtype=dtype).resolve_conj().resolve_neg().contiguous()
    arr = numpy.asarray(x)
    if arr.dtype == object:
        arr = arr.astype(numpy.complex128 if dtype.is_complex else numpy.float64)
    return torch.as_tensor(numpy.ascontiguousarray(arr)).to(device=dev, dtype=dtype).contiguous()


def _uvw(p):
    """[nv, k>=2] float64 device tensor of coordinates."""
    t = _dev(p, torch.float64)
    if t.ndim != 2 or t.shape[1] < 2:
        raise ValueError('coordinates must have shape [nv, >=2]')
    return t


def _out(t, like_torch, dtype=numpy.complex128):
    """Return device tensor ``t`` as torch (torch callers) or NumPy."""
    if like_torch:
        return t
    a = t.cpu().numpy()
    return a.astype(dtype) if dtype is not None else a


def _torch_in(*xs):
    return any(_is_torch(x) for x in xs)


# --------------------------------------------------------------------------
# L0: coordinates, FFT, padding                      (synthesis.py:38-197)
# --------------------------------------------------------------------------


def coordinateBounds(N):
    """Lowest and highest coordinate, step 1/N, zero at N//2.
    synthesis.py:38-55."""
    if N % 2 == 0:
        return -0.5, 0.5 * (N - 2) / N
    return -0.5 * (N - 1) / N, 0.5 * (N - 1) / N


def coordinates(N):
    """``(i - N//2)/N`` for i in [0, N).  synthesis.py:57-63."""
    return (numpy.arange(N) - N // 2) / N


def coordinates2(N):
    """(cx, cy): cx varies along axis 1, cy along axis 0.
    synthesis.py:66-80."""
    c = coordinates(N)
    return numpy.array([numpy.tile(c, (N, 1)), numpy.tile(c[:, None], (1, N))])


def _fft2c(a, inverse):
    was_torch = _is_torch(a)
    t = _dev(a, torch.complex128) if not (was_torch and a.dtype == torch.complex64) \
        else a.to(_lib.device()).contiguous()
    dims = tuple(range(t.ndim))
    if t.ndim not in (1, 2):
        raise AssertionError('Unsupported shape for FFT!')
    t = torch.fft.ifftshift(t, dim=dims)
    t = torch.fft.ifftn(t, dim=dims) if inverse else torch.fft.fftn(t, dim=dims)
    t = torch.fft.fftshift(t, dim=dims)
    return t if was_torch else t.cpu().numpy()


def fft(a):
    """Centred forward FFT (unnormalised), on the device (rocFFT through
    torch.fft).  synthesis.py:83-93."""
    return _fft2c(a, False)


def ifft(a):
    """Centred inverse FFT (1/N^2), on the device.  synthesis.py:96-106."""
    return _fft2c(a, True)


def pad_mid(ff, N):
    """Zero-pad around N//2.  synthesis.py:109-129."""
    N0 = ff.shape[0]
    if N == N0:
        return ff
    assert N > N0
    lo, hi = N // 2 - N0 // 2, (N + 1) // 2 - (N0 + 1) // 2
    if _is_torch(ff):
        pads = [lo, hi] * ff.ndim
        if ff.ndim == 2:
            assert N0 == ff.shape[1]
        return torch.nn.functional.pad(ff, pads)
    if ff.ndim == 2:
        assert N0 == ff.shape[1]
        return numpy.pad(ff, [(lo, hi), (lo, hi)], mode='constant')
    return numpy.pad(ff, [(lo, hi)], mode='constant')


def extract_mid(a, N):
    """Centre N (x N) section.  synthesis.py:131-158."""
    assert N <= a.shape[0]
    h, odd, c0 = N // 2, N % 2, a.shape[0] // 2
    if a.ndim == 2:
        assert N <= a.shape[1]
        c1 = a.shape[1] // 2
        return a[c0 - h:c0 + h + odd, c1 - h:c1 + h + odd]
    if a.ndim == 1:
        return a[c0 - h:c0 + h + odd]
    raise AssertionError('Unsupported grid shape for extract_mid!')


def extract_oversampled(a, Qpx, N):
    """[Qpx,Qpx,N,N] (2D) / [Qpx,N] (1D) sub-sampled kernels.
    synthesis.py:160-197."""
    Na = a.shape[0]
    start = Na // 2 - Qpx * (N // 2)
    lib = torch if _is_torch(a) else numpy
    if a.ndim == 1:
        return lib.stack([a[start + f:start + f + Qpx * N:Qpx] for f in range(Qpx)])
    if a.ndim == 2:
        return lib.stack([lib.stack([a[start + fy:start + fy + Qpx * N:Qpx,
                                       start + fx:start + fx + Qpx * N:Qpx]
                                     for fx in range(Qpx)]) for fy in range(Qpx)])
    raise AssertionError('Unsupported grid shape for extract_oversampled!')


# --------------------------------------------------------------------------
# L1: kernels                                         (synthesis.py:199-387)
# --------------------------------------------------------------------------


def anti_aliasing_function(shape, m, c):
    """Prolate-spheroidal anti-aliasing function, evaluated once on the host
    with scipy.special.pro_ang1 exactly as the reference does (there is no
    device routine for PSWFs; the result only seeds kernel tables).
    synthesis.py:199-218."""
    if numpy.ndim(shape) == 0:
        mult = 2 - 1 / shape / 4
        return scipy.special.pro_ang1(m, m, c, mult * coordinates(shape))[0]
    return numpy.outer(anti_aliasing_function(shape[0], m, c),
                       anti_aliasing_function(shape[1], m, c))


def kernel_transform(dl, dm):
    """synthesis.py:220-248."""
    dn = numpy.sqrt(1 - dl ** 2 - dm ** 2)
    f = (dn - 1) / (dl ** 2 + dm ** 2) if (dl != 0 or dm != 0) else 0
    return numpy.array([[dn - dm * dm * f, dl * dm * f],
                        [dl * dm * f, dn - dl * dl * f]])


def kernel_coordinates(N, theta, dl=0, dm=0, T=None):
    """synthesis.py:250-270."""
    l, m = coordinates2(N) * theta
    if T is not None:
        l, m = T[0, 0] * l + T[1, 0] * m, T[0, 1] * l + T[1, 1] * m
    return l + dl, m + dm


def w_kernel_function(l, m, w):
    """exp(2 pi i w (1 - sqrt(1 - l^2 - m^2))).  synthesis.py:273-293."""
    lib = torch if _is_torch(l) else numpy
    r2 = l ** 2 + m ** 2
    assert bool((r2 < 1.0).all()), 'Error in image coordinate system'
    return lib.exp(2j * numpy.pi * w * (1 - lib.sqrt(1.0 - r2)))


def kernel_recentre(cp, theta, w, dl, dm):
    """synthesis.py:296-311."""
    l, m = coordinates2(cp.shape[0]) * theta
    ramp = numpy.exp(-2j * numpy.pi * w * (dl * l + dm * m))
    if _is_torch(cp):
        ramp = torch.as_tensor(ramp, device=cp.device)
    return cp * ramp


def visibility_recentre(uvw, dl, dm):
    """synthesis.py:314-327."""
    out = uvw.clone() if _is_torch(uvw) else numpy.array(uvw, dtype=float)
    out[:, 0] = uvw[:, 0] - uvw[:, 2] * dl
    out[:, 1] = uvw[:, 1] - uvw[:, 2] * dm
    return out


def kernel_oversample(ff, Qpx, s=None):
    """Oversampled convolution kernels [Qpx,Qpx,s,s] (2D far field) or
    [Qpx,s] (1D), computed on the device in FP64 as the zoom DFT equal to
    pad_mid + fft + extract_oversampled.  synthesis.py:330-354."""
    was_torch = _is_torch(ff)
    N = ff.shape[0]
    if s is None:
        s = N
    if ff.ndim == 2:
        assert N == ff.shape[1]
    if ff.ndim not in (1, 2):
        raise AssertionError('Unsupported grid shape for extract_oversampled!')
    assert N * Qpx >= N
    if Qpx * N < s:
        raise ValueError('support larger than the oversampled far field')
    src = _dev(ff, torch.complex128)
    shape = (Qpx, Qpx, s, s) if ff.ndim == 2 else (Qpx, s)
    out = torch.empty(shape, dtype=torch.complex128, device=src.device)
    _lib.check(_lib.load().croc_kernel_oversample(
        _lib.ctx(), _lib.ptr(src), N, ff.ndim, Qpx, s, _lib.ptr(out)))
    return out if was_torch else out.cpu().numpy()


def _w_kernel_args(kwargs):
    dl = float(kwargs.pop('dl', 0))
    dm = float(kwargs.pop('dm', 0))
    T = kwargs.pop('T', None)
    if kwargs:
        raise TypeError('kernel_coordinates() got unexpected keyword argument(s) %s'
                        % ', '.join(sorted(kwargs)))
    Tv = None
    if T is not None:
        Tv = numpy.ascontiguousarray(numpy.asarray(T, dtype=numpy.float64).reshape(4))
    return dl, dm, Tv


def w_kernel_bank(theta, ws, NpixFF, supports, Qpx, invert=True, scale=1.0,
                  double=False, aterm=None, **kwargs):
    """Generate the w-kernels of many planes in one device launch.

    Returns a :class:`KernelBank` (complex64 unless ``double``) whose plane k
    is ``scale * w_kernel(theta, ws[k], NpixFF, supports[k], Qpx, invert,
    **kwargs)``.  Same arithmetic as :func:`w_kernel` (synthesis.py:357-377).
    ``aterm`` (real [NpixFF, NpixFF], optional) multiplies every plane's far
    field before oversampling: an anti-aliasing taper times an A-term gives
    AW kernels, the image-plane form of aw_kernel_fn (synthesis.py:701-734).
    """
    ws = numpy.ascontiguousarray(numpy.asarray(ws, dtype=numpy.float64).reshape(-1))
    S = numpy.ascontiguousarray(numpy.broadcast_to(
        numpy.asarray(supports, dtype=numpy.int32), ws.shape))
    for s in S:
        assert NpixFF > s or (NpixFF == s and Qpx == 1)
    dl, dm, Tv = _w_kernel_args(dict(kwargs))
    sizes = (Qpx * Qpx * S.astype(numpy.int64) ** 2)
    off = numpy.concatenate([[0], numpy.cumsum(sizes)[:-1]]).astype(numpy.int64)
    dtype = torch.complex128 if double else torch.complex64
    out = torch.empty(int(sizes.sum()), dtype=dtype, device=_lib.device())
    at = None
    if aterm is not None:
        at = _dev(aterm, torch.float64)
        if tuple(at.shape) != (int(NpixFF), int(NpixFF)):
            raise ValueError('aterm must have shape [NpixFF, NpixFF]')
    L = _lib.load()
    _lib.check(L.croc_w_kernels(
        _lib.ctx(), float(theta), ws.ctypes.data_as(_lib._P), len(ws), int(NpixFF),
        S.ctypes.data_as(_lib._P), int(Qpx), int(bool(invert)), dl, dm,
        None if Tv is None else Tv.ctypes.data_as(_lib._P),
        off.ctypes.data_as(_lib._P), _lib.ptr(out), int(bool(double)), _lib.ptr(at)))
    if scale != 1.0:
        out.mul_(scale)
    return KernelBank(out, off, S, Qpx, separable=False)


def w_kernel(theta, w, NpixFF, NpixKern, Qpx, invert=True, **kwargs):
    """Oversampled w-kernel [Qpx,Qpx,NpixKern,NpixKern] (complex128,
    unnormalised), generated on the device in FP64.  ``kwargs`` are
    kernel_coordinates' ``dl``, ``dm``, ``T``.  synthesis.py:357-377."""
    bank = w_kernel_bank(theta, [w], NpixFF, [NpixKern], Qpx, invert, double=True,
                         **kwargs)
    return bank.kern.reshape(Qpx, Qpx, NpixKern, NpixKern).cpu().numpy()


def invert_kernel(a):
    """Pseudo-inverse conj(a)/|a|^2 (unused by the reference, kept for the
    API).  synthesis.py:380-386."""
    return (a.conj() if _is_torch(a) else numpy.conj(a)) / (abs(a) ** 2)


def w_conj_kernel_fn(kernel_fn):
    """Fold k(-w) = conj(k(w)).  synthesis.py:682-699."""
    def fn(theta, w, *args, **kw):
        if w < 0:
            k = kernel_fn(theta, -w, *args, **kw)
            return torch.conj_physical(k) if _is_torch(k) else numpy.conj(k)
        return kernel_fn(theta, w, *args, **kw)
    return fn


def aw_kernel_fn(a_kernel_fn, w_kernel_fn=w_kernel, method='auto'):
    """AW kernels: the two antennas' A-kernels convolved (mode 'same') and
    then convolved with every oversampled w sub-kernel, on the device in
    FP64.  synthesis.py:701-734.  Returns a crocodile_amd.kernels.AWKernels
    (A-kernels, w-kernels, antenna-pair kernels and their spectra made once
    and kept across calls) with ``method`` 'fft' (Fourier convolutions with
    shared spectra), 'direct' (direct sums) or 'auto' (direct for small
    kernels, FFT for large); ``method='uncached'`` gives the plain
    per-call direct sums with no state."""
    if method != 'uncached':
        from .kernels import AWKernels
        return AWKernels(a_kernel_fn, w_kernel_fn, method=method)
    from .awkernel import convolve_same, convolve_same_batched

    def fn(theta, w, a1, a2, *src):
        akern = convolve_same(a_kernel_fn(theta, a1, *src), a_kernel_fn(theta, a2, *src))
        return convolve_same_batched(akern, w_kernel_fn(theta, w))
    return fn


# --------------------------------------------------------------------------
# kernel banks (the device form of make_gcf's kernel tables)
# --------------------------------------------------------------------------


class KernelBank:
    """Kernel tables on the device: complex tables concatenated, plane k at
    element ``plane_off[k]`` with support ``plane_s[k]``; layout
    [Q][Q][S][S] (2D) or [Q][S] (separable)."""

    def __init__(self, kern, plane_off, plane_s, Q, separable):
        dev = _lib.device()
        self.kern = kern
        self.plane_off_h = numpy.array(plane_off, dtype=numpy.int64)
        self.plane_s_h = numpy.array(plane_s, dtype=numpy.int32)
        self.plane_off = torch.as_tensor(self.plane_off_h, device=dev)
        self.plane_s = torch.as_tensor(self.plane_s_h, device=dev)
        self.Q = int(Q)
        self.separable = bool(separable)
        self.nplanes = len(self.plane_s_h)
        self.max_support = int(self.plane_s_h.max())

    def as_complex64(self):
        if self.kern.dtype == torch.complex64:
            return self
        return KernelBank(self.kern.to(torch.complex64), self.plane_off_h, self.plane_s_h,
                          self.Q, self.separable)

    @classmethod
    def from_tables(cls, tables, separable=False):
        """Bank from a list of equally-oversampled square tables."""
        dev = _lib.device()
        ts = [_dev(t, torch.complex64).reshape(-1) for t in tables]
        shapes = [tuple(t.shape) for t in tables]
        Q = shapes[0][0]
        S = [s[-1] for s in shapes]
        for s in shapes:
            if s[0] != Q:
                raise ValueError('all kernel tables must share the oversampling')
            if not separable and s[-1] != s[-2]:
                raise NotImplementedError('non-square 2D kernel tables')
        sizes = numpy.array([t.numel() for t in ts], dtype=numpy.int64)
        off = numpy.concatenate([[0], numpy.cumsum(sizes)[:-1]]).astype(numpy.int64)
        kern = torch.cat(ts) if len(ts) > 1 else ts[0]
        return cls(kern.to(dev), off, S, Q, separable)


def make_gcf(src, gcf):
    """Reference accessor (Q, gw, gh, fn) of synthesis.py:451-461, for API
    compatibility; the device path uses :func:`kernel_bank`."""
    src = numpy.array(src)
    k = src.shape[1]
    if gcf.ndim - k == 2:
        Qpx, gw = gcf.shape[-2:]
        return Qpx, gw, gw, lambda s, xf, yf: numpy.outer(gcf[(*s, yf)], gcf[(*s, xf)])
    if gcf.ndim - k == 4:
        Qpx, gh, gw = gcf.shape[-3:]
        return Qpx, gw, gh, lambda s, xf, yf: gcf[(*s, yf, xf)]
    raise AssertionError('unsupported kernel table rank')


def kernel_bank(gcf, src, nv, for_grid):
    """Device bank and per-visibility plane index for ``gcf`` indexed by
    ``src`` rows (make_gcf semantics, synthesis.py:451-461).  Returns
    (bank, plane_mode, vplane)."""
    if isinstance(gcf, KernelBank):
        b = gcf.as_complex64()
        return b, _lib.PLANE_SINGLE, None
    if src is None:
        src_a = numpy.zeros((nv, 0), dtype=numpy.int64)
    elif _is_torch(src):
        src_a = src.cpu().numpy()
    else:
        src_a = numpy.array(src)
    if src_a.ndim != 2:
        src_a = src_a.reshape(len(src_a), -1)
    k = src_a.shape[1]
    d = gcf.ndim - k
    if d == 2:
        separable = True
        Q, S = gcf.shape[-2:]
    elif d == 4:
        separable = False
        Q, gh, gw = gcf.shape[-3:]
        table_hw = (gh, gw)
        S = gw
        if gh != gw:
            if for_grid:
                # convgrid unpacks make_gcf's (Q, gw, gh) as (Q, gh, gw), so
                # the window is transposed against the table and NumPy's
                # broadcasting fails (synthesis.py:458, :477, :480-481).
                raise ValueError('operands could not be broadcast: convgrid '
                                 'with a non-square %dx%d kernel' % (gh, gw))
            # convdegrid's gh x gw window (synthesis.py:497-503) is the
            # S x S window of S = max(gh, gw) with the table zero-padded by
            # S//2 - gh//2 rows above and S//2 - gw//2 columns to the left:
            # row i of the table lands on grid row y - gh//2 + i either way.
            S = max(gh, gw)
            ro, co = S // 2 - gh // 2, S // 2 - gw // 2
            t = _dev(gcf, torch.complex64)
            padded = torch.zeros(tuple(gcf.shape[:-2]) + (S, S), dtype=torch.complex64,
                                 device=t.device)
            padded[..., ro:ro + gh, co:co + gw] = t
            gcf = padded
    else:
        raise AssertionError('unsupported kernel table rank')
    prefix = tuple(gcf.shape[:k])
    nplanes = int(numpy.prod(prefix)) if k else 1
    table = (Q, S) if separable else (Q, Q, S, S)
    tsize = int(numpy.prod(table))
    kern = _dev(gcf, torch.complex64).reshape(-1)
    off = numpy.arange(nplanes, dtype=numpy.int64) * tsize
    bank = KernelBank(kern, off, numpy.full(nplanes, S, dtype=numpy.int32), Q, separable)
    # the reference's window extents (rows, columns) of these tables
    bank.window = (S, S) if separable else tuple(int(x) for x in table_hw)
    if k == 0:
        return bank, _lib.PLANE_SINGLE, None
    idx = src_a.astype(numpy.int64)
    dims = numpy.array(prefix, dtype=numpy.int64)
    idx = numpy.where(idx < 0, idx + dims, idx)
    if (idx < 0).any() or (idx >= dims).any():
        raise IndexError('src index out of bounds for the kernel table')
    vplane = numpy.ravel_multi_index(tuple(idx.T), prefix).astype(numpy.int32)
    return bank, _lib.PLANE_INDEX, torch.as_tensor(vplane, device=_lib.device())


# --------------------------------------------------------------------------
# L2: gridding core                                   (synthesis.py:389-503)
# --------------------------------------------------------------------------


def _check_device(t, what, dtype):
    """Raise ValueError unless ``t`` is a contiguous tensor of ``dtype`` on
    the current device (the C ABI takes raw pointers and trusts sizes)."""
    if not _is_torch(t):
        raise ValueError('%s must be a torch tensor on the device' % what)
    if t.dtype != dtype:
        raise ValueError('%s must be %s, not %s' % (what, dtype, t.dtype))
    if t.device != _lib.device():
        raise ValueError('%s must be on %s, not %s' % (what, _lib.device(), t.device))
    if not t.is_contiguous():
        raise ValueError('%s must be contiguous' % what)


def _check_grid_args(grid_t, uvw_t, data_t, bank, mode, vplane, nv=None):
    """Shapes and types of a gridding / degridding call: grid complex64
    [N, N]; uvw float64 [nv, >=2] (>=3 with w-bins); visibilities complex64
    with nv entries; bank complex64 on the device; vplane int32 [nv]."""
    _check_device(grid_t, 'grid', torch.complex64)
    if grid_t.ndim != 2 or grid_t.shape[0] != grid_t.shape[1]:
        raise ValueError('grid must be square [N, N], not %s' % (tuple(grid_t.shape),))
    if uvw_t is not None:
        _check_device(uvw_t, 'uvw', torch.float64)
        if uvw_t.ndim != 2 or uvw_t.shape[1] < (3 if mode == _lib.PLANE_WBIN else 2):
            raise ValueError('uvw must have shape [nv, >=%d]'
                             % (3 if mode == _lib.PLANE_WBIN else 2))
        nv = uvw_t.shape[0]
    _check_device(data_t, 'visibilities', torch.complex64)
    if data_t.numel() != nv:
        raise ValueError('visibilities hold %d values for %d coordinates'
                         % (data_t.numel(), nv))
    _check_device(bank.kern, 'kernel bank', torch.complex64)
    if mode == _lib.PLANE_INDEX:
        if vplane is None:
            raise ValueError('CROC_PLANE_INDEX needs vplane')
        _check_device(vplane, 'vplane', torch.int32)
        if vplane.numel() != nv:
            raise ValueError('vplane holds %d values for %d coordinates' % (vplane.numel(), nv))


def _grid_call(fn, grid_t, N, uvw_t, lam, data_t, bank, mode, vplane, wstep=1.0, wbin0=0):
    bank = bank.as_complex64()
    _check_grid_args(grid_t, uvw_t, data_t, bank, mode, vplane)
    if int(N) != grid_t.shape[0]:
        raise ValueError('grid size %d differs from N = %d' % (grid_t.shape[0], N))
    _lib.check(fn(_lib.ctx(), _lib.ptr(grid_t), int(N), _lib.ptr(uvw_t), uvw_t.shape[0],
                  uvw_t.shape[1], float(lam), _lib.ptr(data_t), _lib.ptr(bank.kern),
                  bank.Q, int(bank.separable), bank.nplanes, _lib.ptr(bank.plane_off),
                  _lib.ptr(bank.plane_s), bank.max_support, int(mode), _lib.ptr(vplane),
                  float(wstep), int(wbin0)))


def grid_device(bank, guv, uvw, lam, vis, mode=_lib.PLANE_SINGLE, vplane=None,
                wstep=1.0, wbin0=0):
    """Accumulate ``vis`` (complex64 device [nv]) at ``uvw / lam`` into the
    complex64 device grid ``guv`` with ``bank`` (device-resident fast path of
    convgrid / w_cache_imaging)."""
    _grid_call(_lib.load().croc_convgrid, guv, guv.shape[0], uvw, lam, vis, bank,
               mode, vplane, wstep, wbin0)


def degrid_device(bank, guv, uvw, lam, out, mode=_lib.PLANE_SINGLE, vplane=None,
                  wstep=1.0, wbin0=0):
    """Degrid complex64 device grid ``guv`` into ``out`` (complex64 [nv])."""
    _grid_call(_lib.load().croc_convdegrid, guv, guv.shape[0], uvw, lam, out, bank,
               mode, vplane, wstep, wbin0)


def grid_device_pervis(bank, guv, uvw, lam, vis, mode=_lib.PLANE_SINGLE, vplane=None,
                       wstep=1.0, wbin0=0):
    """grid_device through the plain per-visibility kernels
    (croc_convgrid_pervis: FP32 atomics into the grid, no binning), the
    baseline of the w-towers comparison."""
    _grid_call(_lib.load().croc_convgrid_pervis, guv, guv.shape[0], uvw, lam, vis, bank,
               mode, vplane, wstep, wbin0)


def degrid_device_pervis(bank, guv, uvw, lam, out, mode=_lib.PLANE_SINGLE, vplane=None,
                         wstep=1.0, wbin0=0):
    """degrid_device through the plain per-visibility kernels
    (croc_convdegrid_pervis)."""
    _grid_call(_lib.load().croc_convdegrid_pervis, guv, guv.shape[0], uvw, lam, out, bank,
               mode, vplane, wstep, wbin0)


class GridPlan:
    """Binning of one set of coordinates, reused by repeated gridding and
    degridding of the same visibilities with one kernel bank layout (the
    C ABI's croc_plan).  bin() validates every coordinate exactly as
    convgrid does (same exceptions); grid() / degrid() then only enqueue
    device work.  Constructed with ``uvw`` it bins at once.

    ``uvw`` is a float64 device tensor [nv, 2 or 3]; ``bank`` a
    :class:`KernelBank` whose layout (plane offsets, supports) the plan
    keeps; its table values may be replaced by a bank of the same layout.
    The plan's buffers are kept and reused by later bin() calls.
    """

    def __init__(self, bank, N, uvw=None, lam=None, mode=_lib.PLANE_SINGLE, vplane=None,
                 wstep=1.0, wbin0=0):
        self.bank, self.N, self.nv = bank.as_complex64(), int(N), 0
        h = _lib._P()
        _lib.check(_lib.load().croc_plan_create(_lib.ctx(), ctypes.byref(h)))
        self._h = h
        if uvw is not None:
            self.bin(uvw, lam, mode, vplane, wstep, wbin0)

    def bin(self, uvw, lam, mode=_lib.PLANE_SINGLE, vplane=None, wstep=1.0, wbin0=0):
        """Bin coordinates ``uvw / lam`` (plane selection as convgrid)."""
        _check_device(uvw, 'uvw', torch.float64)
        if uvw.ndim != 2 or uvw.shape[1] < (3 if mode == _lib.PLANE_WBIN else 2):
            raise ValueError('uvw must have shape [nv, >=%d]'
                             % (3 if mode == _lib.PLANE_WBIN else 2))
        if mode == _lib.PLANE_INDEX:
            if vplane is None:
                raise ValueError('CROC_PLANE_INDEX needs vplane')
            _check_device(vplane, 'vplane', torch.int32)
            if vplane.numel() != uvw.shape[0]:
                raise ValueError('vplane holds %d values for %d coordinates'
                                 % (vplane.numel(), uvw.shape[0]))
        bank = self.bank
        self.nv = 0
        _lib.check(_lib.load().croc_plan_bin(
            _lib.ctx(), self._h, self.N, _lib.ptr(uvw), uvw.shape[0], uvw.shape[1],
            float(lam), bank.Q, int(bank.separable), bank.nplanes, _lib.ptr(bank.plane_off),
            _lib.ptr(bank.plane_s), bank.max_support, int(mode), _lib.ptr(vplane),
            float(wstep), int(wbin0)))
        self.nv = int(uvw.shape[0])

    def _same_layout(self, bank):
        if bank is not self.bank and not (
                bank.Q == self.bank.Q and bank.separable == self.bank.separable
                and numpy.array_equal(bank.plane_off_h, self.bank.plane_off_h)
                and numpy.array_equal(bank.plane_s_h, self.bank.plane_s_h)):
            raise ValueError('kernel bank layout differs from the planned one')

    def grid(self, guv, vis, bank=None):
        """guv (complex64 device [N, N]) += convolution of vis (complex64 [nv])."""
        bank = (self.bank if bank is None else bank).as_complex64()
        self._same_layout(bank)
        _check_grid_args(guv, None, vis, bank, None, None, self.nv)
        if guv.shape != (self.N, self.N):
            raise ValueError('grid shape differs from the plan')
        _lib.check(_lib.load().croc_convgrid_plan(_lib.ctx(), self._h, _lib.ptr(guv),
                                                  _lib.ptr(vis), _lib.ptr(bank.kern)))

    def degrid(self, guv, out, bank=None):
        """out (complex64 device [nv]) = degridded visibilities of guv."""
        bank = (self.bank if bank is None else bank).as_complex64()
        self._same_layout(bank)
        _check_grid_args(guv, None, out, bank, None, None, self.nv)
        if guv.shape != (self.N, self.N):
            raise ValueError('grid shape differs from the plan')
        _lib.check(_lib.load().croc_convdegrid_plan(_lib.ctx(), self._h, _lib.ptr(guv),
                                                    _lib.ptr(out), _lib.ptr(bank.kern)))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and _lib._lib is not None:
            try:
                _lib.load().croc_plan_destroy(_lib.ctx(), h)
            except Exception:
                pass
            self._h = None


def _square(a):
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise NotImplementedError('only square grids are supported on the device')
    return a.shape[0]


def _nonsquare(a):
    return a.ndim == 2 and a.shape[0] != a.shape[1]


def _embedding(Nh, Nw, S):
    """Even square size holding an [Nh, Nw] grid with S + 2 cells of margin
    on every side, and the grid's row / column offset in it."""
    Nb = max(Nh, Nw) + 2 * (S + 2)
    Nb += Nb % 2
    return Nb, (Nb - Nh) // 2, (Nb - Nw) // 2


def _remap(uv, Nh, Nw, Q, window, nearest, Nb, oy, ox):
    """Coordinates of the [Nh, Nw] grid's cells in the square embedding
    (croc_remap_coords)."""
    out = torch.empty((uv.shape[0], 2), dtype=torch.float64, device=uv.device)
    gh, gw = window
    _lib.check(_lib.load().croc_remap_coords(
        _lib.ctx(), _lib.ptr(uv), uv.shape[0], uv.shape[1], 1.0, int(Nh), int(Nw), int(Q),
        int(gh), int(gw), int(nearest), int(Nb), int(oy), int(ox), _lib.ptr(out)))
    return out


def _window(bank):
    return getattr(bank, 'window', None) or (bank.max_support, bank.max_support)


def _add_into(a, part):
    if _is_torch(a):
        a += part.to(device=a.device, dtype=a.dtype)
    else:
        a += part.cpu().numpy()


def frac_coord(N, Qpx, p):
    """(ix, xf) with ix - xf/Qpx the nearest 1/Qpx approximation of
    N//2 + p N, half-to-even at ties; int64 results.  synthesis.py:419-435."""
    was_torch = _is_torch(p)
    pt = _dev(p, torch.float64).reshape(-1)
    uv = torch.stack([pt, pt], dim=1).contiguous()
    n = uv.shape[0]
    outs = [torch.empty(n, dtype=torch.int32, device=uv.device) for _ in range(4)]
    _lib.check(_lib.load().croc_frac_coords(_lib.ctx(), _lib.ptr(uv), n, 2, 1.0, int(N),
                                            int(N), int(Qpx), *[_lib.ptr(o) for o in outs]))
    ix, xf = outs[0].to(torch.int64), outs[1].to(torch.int64)
    if was_torch:
        return ix, xf
    shape = numpy.shape(p)
    return ix.cpu().numpy().reshape(shape), xf.cpu().numpy().reshape(shape)


def frac_coords(shape, Qpx, p):
    """(x, xf, y, yf) from p[:, 0] (width) and p[:, 1] (height).
    synthesis.py:438-449."""
    was_torch = _is_torch(p)
    h, w = shape
    uv = _uvw(p)
    n = uv.shape[0]
    outs = [torch.empty(n, dtype=torch.int32, device=uv.device) for _ in range(4)]
    _lib.check(_lib.load().croc_frac_coords(_lib.ctx(), _lib.ptr(uv), n, uv.shape[1], 1.0,
                                            int(h), int(w), int(Qpx),
                                            *[_lib.ptr(o) for o in outs]))
    outs = [o.to(torch.int64) for o in outs]
    return tuple(outs) if was_torch else tuple(o.cpu().numpy() for o in outs)


def convgrid(gcf, a, p, src, v):
    """Grid visibilities ``v`` at ``p`` (fraction of the grid) into ``a`` in
    place: a[y-S//2 : y+(S+1)//2, x-S//2 : x+(S+1)//2] += k(s, xf, yf) v.
    No conjugation.  synthesis.py:464-481."""
    uv = _uvw(p)
    bank, mode, vplane = kernel_bank(gcf, src, uv.shape[0], True)
    vt = _dev(v, torch.complex64).reshape(-1)
    if _nonsquare(a):
        Nh, Nw = a.shape
        Nb, oy, ox = _embedding(Nh, Nw, bank.max_support)
        pb = _remap(uv, Nh, Nw, bank.Q, _window(bank), 0, Nb, oy, ox)
        acc = torch.zeros((Nb, Nb), dtype=torch.complex64, device=uv.device)
        grid_device(bank, acc, pb, 1.0, vt, mode, vplane)
        _add_into(a, acc[oy:oy + Nh, ox:ox + Nw])
        return
    N = _square(a)
    if _is_torch(a) and a.is_cuda and a.dtype == torch.complex64 and a.is_contiguous():
        grid_device(bank, a, uv, 1.0, vt, mode, vplane)
        return
    acc = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    grid_device(bank, acc, uv, 1.0, vt, mode, vplane)
    if _is_torch(a):
        a += acc.to(device=a.device, dtype=a.dtype)
    else:
        a += acc.cpu().numpy()


def convdegrid(gcf, a, p, src=None):
    """v_i = sum(a[window_i] * k(s, xf, yf)), no conjugation.
    synthesis.py:484-503."""
    uv = _uvw(p)
    bank, mode, vplane = kernel_bank(gcf, src, uv.shape[0], False)
    gt = _dev(a, torch.complex64)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    if _nonsquare(gt):
        Nh, Nw = gt.shape
        Nb, oy, ox = _embedding(Nh, Nw, bank.max_support)
        pb = _remap(uv, Nh, Nw, bank.Q, _window(bank), 0, Nb, oy, ox)
        big = torch.zeros((Nb, Nb), dtype=torch.complex64, device=uv.device)
        big[oy:oy + Nh, ox:ox + Nw] = gt
        degrid_device(bank, big, pb, 1.0, out, mode, vplane)
        return _out(out, _torch_in(a, p))
    _square(gt)
    degrid_device(bank, gt, uv, 1.0, out, mode, vplane)
    return _out(out, _torch_in(a, p))


def grid(a, p, v):
    """Nearest-cell gridding, cell N//2 + floor(0.5 + N p).
    synthesis.py:389-401."""
    uv = _uvw(p)
    vt = _dev(v, torch.complex64).reshape(-1)
    if _nonsquare(a):
        Nh, Nw = a.shape
        Nb, oy, ox = _embedding(Nh, Nw, 0)
        pb = _remap(uv, Nh, Nw, 1, (1, 1), 1, Nb, oy, ox)
        acc = torch.zeros((Nb, Nb), dtype=torch.complex64, device=uv.device)
        _lib.check(_lib.load().croc_grid_nearest(_lib.ctx(), _lib.ptr(acc), Nb, _lib.ptr(pb),
                                                 pb.shape[0], 2, 1.0, _lib.ptr(vt)))
        _add_into(a, acc[oy:oy + Nh, ox:ox + Nw])
        return
    N = _square(a)
    direct = _is_torch(a) and a.is_cuda and a.dtype == torch.complex64 and a.is_contiguous()
    acc = a if direct else torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    _lib.check(_lib.load().croc_grid_nearest(_lib.ctx(), _lib.ptr(acc), N, _lib.ptr(uv),
                                             uv.shape[0], uv.shape[1], 1.0, _lib.ptr(vt)))
    if not direct:
        if _is_torch(a):
            a += acc.to(device=a.device, dtype=a.dtype)
        else:
            a += acc.cpu().numpy()


def degrid(a, p):
    """Nearest-cell degridding.  synthesis.py:404-416."""
    uv = _uvw(p)
    gt = _dev(a, torch.complex64)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    if _nonsquare(gt):
        Nh, Nw = gt.shape
        Nb, oy, ox = _embedding(Nh, Nw, 0)
        pb = _remap(uv, Nh, Nw, 1, (1, 1), 1, Nb, oy, ox)
        big = torch.zeros((Nb, Nb), dtype=torch.complex64, device=uv.device)
        big[oy:oy + Nh, ox:ox + Nw] = gt
        _lib.check(_lib.load().croc_degrid_nearest(_lib.ctx(), _lib.ptr(big), Nb, _lib.ptr(pb),
                                                   pb.shape[0], 2, 1.0, _lib.ptr(out)))
        return _out(out, _torch_in(a, p))
    N = _square(gt)
    _lib.check(_lib.load().croc_degrid_nearest(_lib.ctx(), _lib.ptr(gt), N, _lib.ptr(uv),
                                               uv.shape[0], uv.shape[1], 1.0, _lib.ptr(out)))
    return _out(out, _torch_in(a, p))


# --------------------------------------------------------------------------
# L3: imaging drivers                                 (synthesis.py:506-927)
# --------------------------------------------------------------------------


def sort_vis_w(p, v=None):
    """Order by w (stable).  synthesis.py:506-515."""
    if _is_torch(p):
        zs = torch.sort(p[:, 2], stable=True).indices
    else:
        zs = numpy.argsort(p[:, 2], kind='stable')
    if v is not None:
        return p[zs], v[zs]
    return p[zs]


def slice_vis(step, uvw, v=None):
    """Chunks of ``step`` visibilities.  synthesis.py:518-532."""
    starts = range(0, len(uvw), step)
    if v is None:
        return [uvw[i:i + step] for i in starts]
    return [(uvw[i:i + step], v[i:i + step]) for i in starts]


def _grid_size(theta, lam):
    N = int(round(theta * lam))
    assert N > 1
    return N


def doweight(theta, lam, p, v):
    """Uniform weights: divide by the visibility count of each cell
    (frac_coords with Qpx=1; the reference indexes its count grid as
    gw[x, y]).  synthesis.py:535-549."""
    N = _grid_size(theta, lam)
    was_torch = _torch_in(p, v)
    uv = _uvw(p)
    n = uv.shape[0]
    outs = [torch.empty(n, dtype=torch.int32, device=uv.device) for _ in range(4)]
    _lib.check(_lib.load().croc_frac_coords(_lib.ctx(), _lib.ptr(uv), n, uv.shape[1],
                                            float(lam), N, N, 1,
                                            *[_lib.ptr(o) for o in outs]))
    x, y = outs[0].to(torch.int64), outs[2].to(torch.int64)
    if ((x < 0) | (x >= N) | (y < 0) | (y >= N)).any():
        raise IndexError('index out of bounds for the weight grid')
    key = x * N + y
    counts = torch.bincount(key, minlength=N * N).to(torch.float64)
    vt = v.to(uv.device) if _is_torch(v) else torch.as_tensor(numpy.asarray(v), device=uv.device)
    res = vt / counts[key]
    return res if was_torch else res.cpu().numpy()


def _vis_dev(vis):
    return _dev(vis, torch.complex64).reshape(-1)


def simple_imaging(theta, lam, uvw, src, vis):
    """Nearest-cell imaging.  synthesis.py:552-562."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    guv = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    vt = _vis_dev(vis)
    _lib.check(_lib.load().croc_grid_nearest(_lib.ctx(), _lib.ptr(guv), N, _lib.ptr(uv),
                                             uv.shape[0], uv.shape[1], float(lam),
                                             _lib.ptr(vt)))
    return _out(guv, _torch_in(uvw, vis))


def simple_predict(theta, lam, uvw, src, guv):
    """Nearest-cell predict.  synthesis.py:565-580."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    gt = _dev(guv, torch.complex64)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    _lib.check(_lib.load().croc_degrid_nearest(_lib.ctx(), _lib.ptr(gt), _square(gt),
                                               _lib.ptr(uv), uv.shape[0], uv.shape[1],
                                               float(lam), _lib.ptr(out)))
    return _out(out, _torch_in(uvw, guv))


def conv_imaging(theta, lam, uvw, src, vis, kv):
    """Convolutional imaging with a user kernel.  synthesis.py:583-598."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    bank, mode, vplane = kernel_bank(kv, src, uv.shape[0], True)
    guv = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    grid_device(bank, guv, uv, lam, _vis_dev(vis), mode, vplane)
    return _out(guv, _torch_in(uvw, vis))


def conv_predict(theta, lam, uvw, src, guv, kv):
    """Convolutional predict with a user kernel.  synthesis.py:600-613."""
    uv = _uvw(uvw)
    bank, mode, vplane = kernel_bank(kv, src, uv.shape[0], False)
    gt = _dev(guv, torch.complex64)
    _square(gt)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    degrid_device(bank, gt, uv, lam, out, mode, vplane)
    return _out(out, _torch_in(uvw, guv))


def _native_w_kernel(kernel_fn):
    return kernel_fn is w_kernel


def _bank_from_fn(kernel_fn, theta, ws, src_rows, kwargs):
    """One plane per key: kernel_fn(theta, w, *src_row, **kwargs).  The
    device generator is used directly when kernel_fn is :func:`w_kernel`
    and no source columns are present."""
    if _native_w_kernel(kernel_fn) and (src_rows is None or src_rows.shape[1] == 0):
        kw = dict(kwargs)
        NpixFF = kw.pop('NpixFF')
        NpixKern = kw.pop('NpixKern')
        Qpx = kw.pop('Qpx')
        invert = kw.pop('invert', True)
        return w_kernel_bank(theta, ws, NpixFF, NpixKern, Qpx, invert, **kw)
    tables = []
    for i, w in enumerate(ws):
        s = () if src_rows is None else tuple(src_rows[i])
        tables.append(kernel_fn(theta, w, *s, **kwargs))
    return KernelBank.from_tables(tables)


def _slice_planes(uv, wstep):
    """Stable w-order, plane index per visibility and mean w per slice."""
    order = torch.sort(uv[:, 2], stable=True).indices
    nv = uv.shape[0]
    nsl = (nv + wstep - 1) // wstep
    sw = uv[order, 2].cpu().numpy()
    means = [float(numpy.mean(sw[i * wstep:(i + 1) * wstep])) for i in range(nsl)]
    vplane = torch.empty(nv, dtype=torch.int32, device=uv.device)
    vplane[order] = (torch.arange(nv, device=uv.device) // wstep).to(torch.int32)
    return means, vplane, order


def _slice_bank(kernel_fn, theta, means, vplane, src, nv, kwargs):
    """Bank and plane index of the w-slices.  Without source columns, one
    plane per slice.  With source columns (k of them) each slice's table
    kernel_fn(theta, w, **kwargs) carries k leading dimensions indexed by a
    visibility's source row (make_gcf, synthesis.py:451-461): the planes
    are every slice's sub-tables, plane = slice * prod(prefix) + the row's
    flat index (negative indices wrap as NumPy's)."""
    src_rows = _src_rows(src, nv)
    if src_rows is None or src_rows.shape[1] == 0:
        return _bank_from_fn(kernel_fn, theta, means, None, kwargs).as_complex64(), vplane
    k = src_rows.shape[1]
    tables = [kernel_fn(theta, w, **kwargs) for w in means]
    shape = tuple(tables[0].shape)
    d = len(shape) - k
    if d not in (2, 4):
        raise AssertionError('unsupported kernel table rank')
    prefix = shape[:k]
    subs = []
    for t in tables:
        if tuple(t.shape) != shape:
            raise ValueError('w-slice kernel tables differ in shape')
        t = t if _is_torch(t) else numpy.asarray(t)
        subs.extend(t.reshape((-1,) + shape[k:])[i] for i in range(int(numpy.prod(prefix))))
    if d == 4 and shape[-1] != shape[-2]:
        raise NotImplementedError('non-square 2D kernel tables')
    bank = KernelBank.from_tables(subs, separable=(d == 2))
    idx = src_rows.astype(numpy.int64)
    dims = numpy.array(prefix, dtype=numpy.int64)
    idx = numpy.where(idx < 0, idx + dims, idx)
    if (idx < 0).any() or (idx >= dims).any():
        raise IndexError('src index out of bounds for the kernel table')
    flat = torch.as_tensor(numpy.ravel_multi_index(tuple(idx.T), prefix).astype(numpy.int32),
                           device=vplane.device)
    return bank, (vplane * int(numpy.prod(prefix)) + flat).to(torch.int32).contiguous()


def w_slice_imaging(theta, lam, uvw, src, vis, wstep=2000, kernel_fn=w_kernel,
                    **kwargs):
    """W-projection by slices of ``wstep`` visibilities sorted by w, kernel at
    each slice's mean w.  The reference refers to undefined ``p, v``
    (NameError, synthesis.py:638); the intended ``uvw, vis`` are used, with
    a stable sort, and ``src`` rows travel with their visibilities (the
    reference passes the unsorted ``src`` to every slice's convgrid).
    synthesis.py:616-644."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    means, vplane, order = _slice_planes(uv, wstep)
    bank, vplane = _slice_bank(kernel_fn, theta, means, vplane, src, uv.shape[0], kwargs)
    guv = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    grid_device(bank, guv, uv, lam, _vis_dev(vis), _lib.PLANE_INDEX, vplane)
    return _out(guv, _torch_in(uvw, vis))


def w_slice_predict(theta, lam, uvw, src, guv, wstep=2000, kernel_fn=w_kernel,
                    **kwargs):
    """Predict counterpart of w_slice_imaging, results in input order, with
    the source rows selecting sub-tables as in w_slice_imaging.  The
    reference uses undefined ``p`` (NameError, synthesis.py:672).
    synthesis.py:647-679."""
    _grid_size(theta, lam)
    uv = _uvw(uvw)
    means, vplane, order = _slice_planes(uv, wstep)
    bank, vplane = _slice_bank(kernel_fn, theta, means, vplane, src, uv.shape[0], kwargs)
    gt = _dev(guv, torch.complex64)
    _square(gt)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    degrid_device(bank, gt, uv, lam, out, _lib.PLANE_INDEX, vplane)
    return _out(out, _torch_in(uvw, guv))


def _src_rows(src, nv):
    if src is None:
        return None
    a = src.cpu().numpy() if _is_torch(src) else numpy.asarray(src)
    if a.ndim == 1 and a.size == 0:
        return None
    return a.reshape(nv, -1) if a.size else None


def w_cache_planes(theta, uv, src, wstep, fn, kwargs):
    """Kernel bank with one plane per distinct (w-bin, src row) key and the
    plane-selection arguments for the gridder.  w-bin = round-half-even of
    w / wstep (numpy.round in synthesis.py:793)."""
    wb = torch.round(uv[:, 2] / wstep).to(torch.int64)
    src_rows = _src_rows(src, uv.shape[0])
    if src_rows is None or src_rows.shape[1] == 0:
        lo, hi = int(wb.min()), int(wb.max())
        present = torch.zeros(hi - lo + 1, dtype=torch.bool, device=uv.device)
        present[wb - lo] = True
        present = present.cpu().numpy()
        if present.all() or (hi - lo + 1) <= 4 * int(present.sum()):
            # dense bin range: the gridder derives the plane from w itself
            bins = numpy.arange(lo, hi + 1)
            bank = _bank_from_fn(fn, theta, [wstep * float(b) for b in bins], None, kwargs)
            return bank.as_complex64(), _lib.PLANE_WBIN, None, lo
        bins, inv = torch.unique(wb, return_inverse=True)
        bins = bins.cpu().numpy()
        bank = _bank_from_fn(fn, theta, [wstep * float(b) for b in bins], None, kwargs)
        return bank.as_complex64(), _lib.PLANE_INDEX, inv.to(torch.int32), 0
    keys = torch.cat([wb[:, None].to(torch.float64),
                      torch.as_tensor(src_rows, dtype=torch.float64, device=uv.device)], 1)
    ukeys, inv = torch.unique(keys, dim=0, return_inverse=True)
    ukeys = ukeys.cpu().numpy()
    ws = [wstep * float(k[0]) for k in ukeys]
    srcs = numpy.array([[_as_index(x) for x in k[1:]] for k in ukeys], dtype=object)
    bank = _bank_from_fn(fn, theta, ws, srcs, kwargs)
    return bank.as_complex64(), _lib.PLANE_INDEX, inv.to(torch.int32), 0


def _as_index(x):
    return int(x) if float(x).is_integer() else float(x)


def w_cache_imaging(theta, lam, uvw, src, vis, wstep=2000, kernel_cache=None,
                    kernel_fn=w_kernel, **kwargs):
    """W-projection with the kernel of each visibility's w-bin
    (wstep * round(w / wstep)); ``kernel_cache`` replaces ``kernel_fn`` when
    given.  Kernels are made once per distinct key and all visibilities are
    gridded in one device pass.  synthesis.py:757-796."""
    fn = kernel_cache if kernel_cache is not None else kernel_fn
    N = int(round(theta * lam))
    uv = _uvw(uvw)
    bank, mode, vplane, wbin0 = w_cache_planes(theta, uv, src, wstep, fn, kwargs)
    guv = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    grid_device(bank, guv, uv, lam, _vis_dev(vis), mode, vplane, wstep, wbin0)
    return _out(guv, _torch_in(uvw, vis))


def w_cache_predict(theta, lam, uvw, src, guv, wstep=2000, kernel_cache=None,
                    kernel_fn=w_kernel, **kwargs):
    """Predict counterpart of w_cache_imaging, results in input order.  The
    reference uses undefined ``nv`` / ``ixs`` (NameError,
    synthesis.py:823/827).  synthesis.py:799-828."""
    fn = kernel_cache if kernel_cache is not None else kernel_fn
    uv = _uvw(uvw)
    bank, mode, vplane, wbin0 = w_cache_planes(theta, uv, src, wstep, fn, kwargs)
    gt = _dev(guv, torch.complex64)
    _square(gt)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    degrid_device(bank, gt, uv, lam, out, mode, vplane, wstep, wbin0)
    return _out(out, _torch_in(uvw, guv))


def mirror_uvw(uvw, vis):
    """Flip baselines with v < 0 and conjugate them.  synthesis.py:831-852."""
    if _is_torch(uvw):
        neg = uvw[:, 1] < 0
        uvw = uvw.clone()
        vis = vis.clone()
        uvw[neg] = -uvw[neg]
        vis[neg] = vis[neg].conj()
        return uvw, vis
    neg = uvw[:, 1] < 0
    uvw = numpy.array(uvw, dtype=float, copy=True)
    vis = numpy.array(vis, copy=True)
    uvw[neg] = -uvw[neg]
    vis[neg] = numpy.conj(vis[neg])
    return uvw, vis


def make_grid_hermitian(guv):
    """In place: add the conjugated point mirror about N//2 (even N: rows and
    columns 1..N-1).  synthesis.py:855-880."""
    N = _square(guv)
    if _is_torch(guv) and guv.is_cuda and guv.is_contiguous() and \
            guv.dtype in (torch.complex64, torch.complex128):
        _lib.check(_lib.load().croc_make_hermitian(
            _lib.ctx(), _lib.ptr(guv), N, int(guv.dtype == torch.complex128)))
        return guv
    t = _dev(guv, torch.complex128)
    _lib.check(_lib.load().croc_make_hermitian(_lib.ctx(), _lib.ptr(t), N, 1))
    if _is_torch(guv):
        guv.copy_(t.to(guv.device, guv.dtype))
    else:
        guv[...] = t.cpu().numpy()
    return guv


def do_imaging(theta, lam, uvw, src, vis, imgfn, **kwargs):
    """Mirror, weight, image, hermitian, ifft; PSF likewise; normalise.
    synthesis.py:883-911."""
    if src is None:
        src = numpy.ndarray((len(vis), 0))
    uvw, vis = mirror_uvw(uvw, vis)
    ones = torch.ones(len(uvw), dtype=torch.float64) if _is_torch(uvw) else numpy.ones(len(uvw))
    wt = doweight(theta, lam, uvw, ones)
    cdrt = imgfn(theta, lam, uvw, src, wt * vis, **kwargs)
    drt = ifft(make_grid_hermitian(cdrt)).real
    c = imgfn(theta, lam, uvw, src, wt, **kwargs)
    psf = ifft(make_grid_hermitian(c)).real
    pmax = psf.max()
    assert pmax > 0.0
    return drt / pmax, psf / pmax, pmax


def do_predict(theta, lam, uvw, src, modelimage, predfn, **kwargs):
    """fft(model image) then predfn.  synthesis.py:914-927."""
    m = modelimage.to(torch.complex128) if _is_torch(modelimage) else modelimage.astype(complex)
    return predfn(theta, lam, uvw, src, fft(m), **kwargs)


def sample_image(image, x, y):
    """Bilinear sample relative to the image middle.  synthesis.py:736-755."""
    x = x + image.shape[0] / 2
    y = y + image.shape[1] / 2
    ix = numpy.floor(x).astype(int)
    dx = x - ix
    iy = numpy.floor(y).astype(int)
    dy = y - iy
    return numpy.sum([image[iy, ix] * ((1 - dx) * (1 - dy)),
                      image[iy, ix + 1] * (dx * (1 - dy)),
                      image[iy + 1, ix] * ((1 - dx) * dy),
                      image[iy + 1, ix + 1] * (dx * dy)], axis=0)


# --------------------------------------------------------------------------
# W-towers drivers (examples/notebooks/wtowers.ipynb, wtowers-predict.ipynb,
# examples/grid/grid.c:482-710) in the imgfn / predfn protocol
# --------------------------------------------------------------------------


class WKernelSet:
    """Kernels of the w-towers sub-grid gridding for the residual w of each
    visibility relative to its w-plane (grid.c's w_kernel_data, grid.h:
    114-122, as used by w_project, grid.c:277-305): ``tables[k]`` serves
    residuals near ``w_min + k * w_step``; a visibility of w-plane wp takes
    table floor((w - wp * wstep - w_min) / w_step + 0.5).  Tables are all
    [Q, Q, S, S] (2D) or all [Q, S] (separable).  The set must cover the
    w-increment [-wstep/2, wstep/2] (grid.c:493-496)."""

    def __init__(self, tables, w_min, w_step):
        t = _dev(tables, torch.complex64)
        if t.ndim not in (3, 5):
            raise AssertionError('w-kernel set tables must be [n, Q, S] or [n, Q, Q, S, S]')
        self.tables = t.contiguous()
        self.w_min, self.w_step = float(w_min), float(w_step)

    @property
    def nplanes(self):
        return self.tables.shape[0]

    @property
    def w_max(self):
        return self.w_min + (self.nplanes - 1) * self.w_step


def w_towers_kernel_set(theta, wincrement, NpixFF, NpixKern, Qpx, step, taper=None,
                        invert=True):
    """Separable w-kernel set for the w-towers drivers: table k is the
    [Qpx, NpixKern] oversampled kernel (kernel_oversample, synthesis.py:
    330-354) of the 1D far field taper(l) * exp(-+ i pi w_k l^2), l =
    theta * coordinates(NpixFF), w_k = w_min + k * step covering
    [-wincrement/2, wincrement/2].  The product over both axes is the
    w-kernel of the residual w_k in the paraxial form w (1 - n) ~ w (l^2 +
    m^2) / 2 (its error, w r^4 / 8, stays below 2e-7 w at r^2 <= 1.25e-3)
    times the separable taper.  ``invert`` gives the imaging kernel
    (exp(-i ...), as w_kernel's 1 / w_kernel_function), else the predict
    kernel."""
    half = 0.5 * float(wincrement)
    n = int(math.ceil(2 * half / step - 1e-9)) + 1
    w_min = -(n - 1) * step / 2
    l = coordinates(NpixFF) * theta
    a = numpy.ones(NpixFF) if taper is None else numpy.asarray(taper, dtype=float)
    sign = -1.0 if invert else 1.0
    ff = numpy.array([a * numpy.exp(sign * 1j * numpy.pi * (w_min + k * step) * l ** 2)
                      for k in range(n)])
    tabs = [kernel_oversample(torch.as_tensor(f, device=_lib.device()), Qpx, NpixKern)
            for f in ff]
    return WKernelSet(torch.stack(tabs), w_min, step)


def _tower_kernel(kv):
    """(device tables or None, Q, S, separable, nwk, wk_min, wk_step) for the
    sub-grid kernel: one table, or a :class:`WKernelSet`."""
    if kv is None:
        return None, 1, 1, 0, 1, 0.0, 0.0
    if isinstance(kv, WKernelSet):
        t, n, w_min, w_step = kv.tables, kv.nplanes, kv.w_min, kv.w_step
        shape = tuple(t.shape[1:])
    else:
        t, n, w_min, w_step = _dev(kv, torch.complex64), 1, 0.0, 0.0
        shape = tuple(t.shape)
    if len(shape) == 2:
        Q, S = shape
        return t.reshape(-1), int(Q), int(S), 1, n, w_min, w_step
    if len(shape) == 4:
        Q, Q2, gh, gw = shape
        if gh != gw or Q != Q2:
            raise NotImplementedError('w-towers need a square [Q,Q,S,S] kernel')
        return t.reshape(-1), int(Q), int(gw), 0, n, w_min, w_step
    raise AssertionError('unsupported kernel table rank')


def w_towers_imaging(theta, lam, uvw, src, vis, wstep=100.0, subgrid_size=128, margin=32,
                     kv=None):
    """W-towers imaging.  Visibilities are binned into chunks of
    subgrid_size - margin cells and w-planes round(w / wstep); every
    (chunk, plane) is gridded on its own sub-grid (nearest cell when ``kv``
    is None, as the notebook's simple_imaging, else convolution with the
    oversampled kernel ``kv``, or with a :class:`WKernelSet` the kernel of
    each visibility's residual w as grid.c's w_project), inverse transformed, multiplied by
    exp(-2 pi i wp wstep (1 - n)), summed per chunk, transformed back and
    added into the grid.  wtowers.ipynb imaging cell; grid.c:482-710."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    if uv.shape[1] < 3:
        raise ValueError('w-towers need uvw with a w column')
    kern, Q, S, sep, nwk, wk0, wks = _tower_kernel(kv)
    guv = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
    vt = _vis_dev(vis)
    _lib.check(_lib.load().croc_wtowers_grid(
        _lib.ctx(), _lib.ptr(guv), N, _lib.ptr(uv), uv.shape[0], uv.shape[1], float(theta),
        _lib.ptr(vt), float(wstep), int(subgrid_size), int(margin), _lib.ptr(kern), Q, S, sep,
        nwk, wk0, wks))
    return _out(guv, _torch_in(uvw, vis))


def w_towers_predict(theta, lam, uvw, src, guv, wstep=100.0, subgrid_size=128, margin=32,
                     kv=None):
    """W-towers predict: per chunk, inverse transform of the grid window;
    per w-plane, multiply by exp(+2 pi i wp wstep (1 - n)), transform and
    degrid (nearest cell or ``kv``).  wtowers-predict.ipynb predict cell."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    if uv.shape[1] < 3:
        raise ValueError('w-towers need uvw with a w column')
    gt = _dev(guv, torch.complex64)
    if _square(gt) != N:
        raise ValueError('grid size does not match round(theta * lam)')
    kern, Q, S, sep, nwk, wk0, wks = _tower_kernel(kv)
    out = torch.empty(uv.shape[0], dtype=torch.complex64, device=uv.device)
    _lib.check(_lib.load().croc_wtowers_degrid(
        _lib.ctx(), _lib.ptr(gt), N, _lib.ptr(uv), uv.shape[0], uv.shape[1], float(theta),
        _lib.ptr(out), float(wstep), int(subgrid_size), int(margin), _lib.ptr(kern), Q, S, sep,
        nwk, wk0, wks))
    return _out(out, _torch_in(uvw, guv))


class WTowersPlan:
    """(chunk, w-plane) binning of one set of coordinates, reused by repeated
    w-towers imaging and predict of the same visibilities (the C ABI's
    croc_tw_plan; :func:`w_towers_imaging` / :func:`w_towers_predict` bin on
    every call).  The arguments are theirs; the kernel ``kv`` may change
    between calls but keeps the shape it had at binning (or stays None).
    bin() validates every coordinate as those calls do."""

    def __init__(self, theta, lam, uvw, wstep=100.0, subgrid_size=128, margin=32, kv=None):
        self.theta, self.wstep = float(theta), float(wstep)
        self.subgrid_size, self.margin = int(subgrid_size), int(margin)
        self.N = _grid_size(theta, lam)
        self.nv = 0
        h = _lib._P()
        _lib.check(_lib.load().croc_wtowers_plan_create(_lib.ctx(), ctypes.byref(h)))
        self._h = h
        self.bin(uvw, kv)

    def bin(self, uvw, kv=None):
        """Bin the coordinates ``uvw`` (wavelengths, with a w column)."""
        uv = _uvw(uvw)
        if uv.shape[1] < 3:
            raise ValueError('w-towers need uvw with a w column')
        kern, Q, S, sep, nwk, wk0, wks = _tower_kernel(kv)
        self.nv = 0
        kind = 0 if kern is None else (1 if sep else 2)
        _lib.check(_lib.load().croc_wtowers_plan_bin(
            _lib.ctx(), self._h, self.N, _lib.ptr(uv), uv.shape[0], uv.shape[1], self.theta,
            self.wstep, self.subgrid_size, self.margin, kind, Q, S, nwk, wk0, wks))
        self._shape = (kind, Q, S, nwk, wk0, wks)
        self.nv, self._device = int(uv.shape[0]), uv.device

    def _kernel(self, kv):
        kern, Q, S, sep, nwk, wk0, wks = _tower_kernel(kv)
        kind = 0 if kern is None else (1 if sep else 2)
        if (kind, Q, S, nwk, wk0, wks) != self._shape:
            raise ValueError('kernel shape or w-kernel set differs from the planned one')
        return kern, sep

    def imaging(self, vis, kv=None):
        """The grid of w_towers_imaging for the binned coordinates."""
        kern, sep = self._kernel(kv)
        vt = _vis_dev(vis)
        if vt.shape[0] != self.nv:
            raise ValueError('visibility count differs from the plan')
        guv = torch.zeros((self.N, self.N), dtype=torch.complex64, device=self._device)
        _lib.check(_lib.load().croc_wtowers_grid_plan(
            _lib.ctx(), self._h, _lib.ptr(guv), _lib.ptr(vt), _lib.ptr(kern), sep))
        return _out(guv, _torch_in(vis))

    def predict(self, guv, kv=None):
        """The visibilities of w_towers_predict for the binned coordinates."""
        kern, sep = self._kernel(kv)
        gt = _dev(guv, torch.complex64)
        if _square(gt) != self.N:
            raise ValueError('grid size does not match round(theta * lam)')
        out = torch.empty(self.nv, dtype=torch.complex64, device=self._device)
        _lib.check(_lib.load().croc_wtowers_degrid_plan(
            _lib.ctx(), self._h, _lib.ptr(gt), _lib.ptr(out), _lib.ptr(kern), sep))
        return _out(out, _torch_in(guv))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and _lib._lib is not None:
            try:
                _lib.load().croc_wtowers_plan_destroy(_lib.ctx(), h)
            except Exception:
                pass
            self._h = None


# --------------------------------------------------------------------------
# W-stacking drivers (examples/notebooks/wstacking.ipynb cells 10 and 12)
# --------------------------------------------------------------------------


def _w_planes(uv, wstep):
    """Per visibility its w-plane round-half-even(w / wstep) (numpy.around,
    wstacking.ipynb cell 10) as int64, and per non-empty plane in ascending
    order: (plane numbers on the host, counts and mean w on the device)."""
    pl = torch.round(uv[:, 2] / float(wstep)).to(torch.int64)
    lo, hi = (int(x) for x in torch.stack([pl.min(), pl.max()]).cpu())
    idx = pl - lo
    counts = torch.bincount(idx, minlength=hi - lo + 1)
    # per-block histograms: index_add_'s FP64 atomics onto a few dozen
    # addresses serialise (half a second for 2e6 visibilities)
    sums = torch.bincount(idx, weights=uv[:, 2], minlength=hi - lo + 1).to(torch.float64)
    nz = torch.nonzero(counts).reshape(-1)
    midw = sums[nz] / counts[nz].to(torch.float64)
    return pl, lo, hi - lo + 1, (nz + lo).cpu().tolist(), counts[nz], midw


def w_stacking_imaging(theta, lam, uvw, src, vis, wstep, max_stack_bytes=4 << 30):
    """Full-grid w-stacking (wstacking.ipynb cell 10, JSON lines 191-240):
    the visibilities of each w-plane round(w / wstep) are imaged nearest-cell
    (simple_imaging) on the full grid, the image divided by the Fresnel
    pattern w_kernel_function(l, m, mean w of the plane) and the images
    summed.  Returns the image sum ([N, N], complex128).

    Planes go through the device in batches of at most ``max_stack_bytes``
    of [N, N] complex64 grids: one gridding launch per batch, a batched
    inverse FFT and one launch that applies every plane's phase and adds the
    batch into the FP64 image.  Like the notebook (cell 8), callers flip
    visibilities to one half-plane and weight them first."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    if uv.shape[1] < 3:
        raise ValueError('w-stacking needs uvw with a w column')
    vt = _vis_dev(vis)
    like_torch = _torch_in(uvw, vis)
    image = torch.zeros((N, N), dtype=torch.complex128, device=uv.device)
    if uv.shape[0] == 0:
        return _out(image, like_torch)
    _, lo, ntab, planes, _, midw = _w_planes(uv, wstep)
    L = _lib.load()
    centred = int(N % 2 == 0)
    batch = max(1, min(len(planes), int(max_stack_bytes) // (N * N * 8)))
    stack = torch.empty((batch, N, N), dtype=torch.complex64, device=uv.device)
    slot_of = torch.empty(ntab, dtype=torch.int32, device=uv.device)
    for b0 in range(0, len(planes), batch):
        sel = planes[b0:b0 + batch]
        P = len(sel)
        slot_of.fill_(-1)
        slot_of[torch.as_tensor(sel, device=uv.device) - lo] = torch.arange(
            P, dtype=torch.int32, device=uv.device)
        st = stack[:P]
        st.zero_()
        _lib.check(L.croc_wstack_grid(_lib.ctx(), _lib.ptr(st), P, N, _lib.ptr(uv),
                                      uv.shape[0], uv.shape[1], float(lam), float(wstep),
                                      lo, ntab, _lib.ptr(slot_of), _lib.ptr(vt), centred))
        if centred:
            img = torch.fft.ifftn(st, dim=(-2, -1))
        else:
            img = torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(st, dim=(-2, -1)),
                                                     dim=(-2, -1)), dim=(-2, -1))
        mw = midw[b0:b0 + P].contiguous()
        _lib.check(L.croc_wstack_add(_lib.ctx(), _lib.ptr(image), _lib.ptr(img), P, N,
                                     float(theta), _lib.ptr(mw), centred))
        del img
    return _out(image, like_torch)


def w_stacking_subgrid_imaging(theta, lam, uvw, src, vis, wstep, npixkern):
    """Per-plane sub-grid w-stacking (wstacking.ipynb cell 12, JSON lines
    249-300): a plane whose mean w lies within wstep / 2 of 0 is gridded
    nearest-cell onto the grid directly; any other plane onto the square
    sub-grid spanning its visibilities' cells plus ``npixkern``, centred at
    the sub-grid middle and the plane's mean w, whose image is divided by the
    Fresnel pattern of that w and transformed back before it is added to the
    grid at the sub-grid's offset.  Returns the grid ([N, N], complex128).
    A sub-grid that would leave the grid raises ValueError (the notebook's
    slices would wrap or clip it silently)."""
    N = _grid_size(theta, lam)
    uv = _uvw(uvw)
    if uv.shape[1] < 3:
        raise ValueError('w-stacking needs uvw with a w column')
    vt = _vis_dev(vis)
    like_torch = _torch_in(uvw, vis)
    grid_sum = torch.zeros((N, N), dtype=torch.complex128, device=uv.device)
    if uv.shape[0] == 0:
        return _out(grid_sum, like_torch)
    pl, lo, ntab, planes, counts, midw = _w_planes(uv, wstep)
    order = torch.sort(pl, stable=True).indices
    us, vs = uv[order].contiguous(), vt[order].contiguous()
    slot = torch.zeros(ntab, dtype=torch.int64, device=uv.device)
    slot[torch.as_tensor(planes, device=uv.device) - lo] = torch.arange(
        len(planes), device=uv.device)
    s_of = slot[pl[order] - lo]
    big = torch.tensor(float('inf'), dtype=torch.float64, device=uv.device)
    mins = big.expand(len(planes), 2).clone().scatter_reduce_(
        0, s_of[:, None].expand(-1, 2), us[:, :2], 'amin')
    maxs = (-big).expand(len(planes), 2).clone().scatter_reduce_(
        0, s_of[:, None].expand(-1, 2), us[:, :2], 'amax')
    host = torch.cat([mins, maxs, midw[:, None], counts[:, None].to(torch.float64)],
                     1).cpu().numpy()
    mid = int(lam * theta) // 2
    L = _lib.load()
    start = 0
    for k in range(len(planes)):
        n = int(host[k, 5])
        pu, pv = us[start:start + n], vs[start:start + n]
        start += n
        w_mid = float(host[k, 4])
        if abs(w_mid) < wstep / 2:
            g = torch.zeros((N, N), dtype=torch.complex64, device=uv.device)
            _lib.check(L.croc_grid_nearest(_lib.ctx(), _lib.ptr(g), N, _lib.ptr(pu), n,
                                           pu.shape[1], float(lam), _lib.ptr(pv)))
            grid_sum += g
            continue
        xy_min = numpy.floor(host[k, 0:2] * theta).astype(int)
        xy_max = numpy.ceil(host[k, 2:4] * theta).astype(int)
        xy_min -= (npixkern + 1) // 2
        xy_max += npixkern // 2
        xy_size = int(numpy.max(xy_max - xy_min))
        xy_maxq = numpy.amax([xy_max, xy_min + xy_size], axis=0)
        uvw_size = xy_size / theta
        uvw_mid = numpy.hstack([(xy_maxq + xy_min) // 2 / theta, w_mid])
        x0, y0 = mid + xy_min
        x1, y1 = mid + xy_max
        if x0 < 0 or y0 < 0 or x1 > N or y1 > N:
            raise ValueError('sub-grid of w-plane %d leaves the grid' % planes[k])
        n_sub = _grid_size(theta, uvw_size)
        rel = (pu[:, :3] - torch.as_tensor(uvw_mid, device=uv.device)).contiguous()
        g = torch.zeros((n_sub, n_sub), dtype=torch.complex64, device=uv.device)
        _lib.check(L.croc_grid_nearest(_lib.ctx(), _lib.ptr(g), n_sub, _lib.ptr(rel), n, 3,
                                       float(uvw_size), _lib.ptr(pv)))
        img = torch.zeros((n_sub, n_sub), dtype=torch.complex128, device=uv.device)
        mw = torch.tensor([w_mid], dtype=torch.float64, device=uv.device)
        gi = _fft2c(g, True)
        _lib.check(L.croc_wstack_add(_lib.ctx(), _lib.ptr(img), _lib.ptr(gi), 1, n_sub,
                                     float(theta), _lib.ptr(mw), 0))
        grid_sum[y0:y1, x0:x1] += _fft2c(img, False)[0:y1 - y0, 0:x1 - x0]
    return _out(grid_sum, like_torch)


# --------------------------------------------------------------------------
# AW-towers imaging (examples/notebooks/awtowers.ipynb cells 17, 19, 22)
# --------------------------------------------------------------------------


def _awt_chunks(theta, lam, subgrid_size, margin, ub, vb):
    """uv_mid [n, 2] and sub-grid corners [n, 2] of chunks (ub, vb), with
    awtowers.ipynb cell 22's float arithmetic (grid_tower's in-place
    `xy_min -=` moves the corner the clipped add uses)."""
    chunk = subgrid_size - margin
    ub = numpy.asarray(ub, dtype=numpy.int64)
    vb = numpy.asarray(vb, dtype=numpy.int64)
    xy_min = chunk * numpy.stack([ub - 0.5, vb - 0.5], 1)
    xy_max = chunk * numpy.stack([ub + 0.5, vb + 0.5], 1)
    uv_mid = (xy_max + xy_min) // 2 / theta
    xy_min -= (margin + 1) // 2
    xy_max += margin // 2
    if not numpy.all(numpy.max(xy_max - xy_min, axis=1) == subgrid_size):
        raise AssertionError('chunk does not fill the sub-grid')
    mid = int(lam * theta) // 2
    corner = mid + xy_min.astype(int)
    return uv_mid, corner


def aw_towers_imaging(theta, lam, uvw, src, vis, wstep, subgrid_size, margin, aa_gcf,
                      aa_subgrid, aa_subgrid2, aimage, wkern, dl=0.0, dm=0.0,
                      max_stack_bytes=4 << 30):
    """AW-towers imaging (awtowers.ipynb cells 17, 19 and 22, the notebook's
    globals as arguments): chunks of subgrid_size - margin cells centred on
    ub = floor(u theta / chunk + 0.5); per (chunk, w-plane round(w / wstep),
    antenna pair src = (a1, a2)) the visibilities relative to (uv_mid,
    wp wstep) are gridded with the separable anti-aliasing table ``aa_gcf``
    [Q, K] on a 2 subgrid_size grid at 2 theta, inverse transformed,
    divided by aa_subgrid2 (x) aa_subgrid2, cut to the middle, multiplied by
    aa_subgrid (x) aa_subgrid and by aimage[a2] / aimage[a1] (``aimage``
    [nant, S, S]); each chunk sums its planes times wkern^-wp (the
    notebook's stepping by wkern^(wp - last_wp) and final division by
    wkern^last_wp), is transformed and added into the grid, clipped at its
    edges.  ``dl``, ``dm``: the visibilities are divided by
    w_kernel_function(dl, dm, w - wp wstep) first (the notebook's shifted
    field).  Returns grid_sum ([N, N], N = ceil(theta lam)); the notebook's
    image is real(ifft(grid_sum)) / (aa (x) aa) / sum(wt).

    The notebook skips the wkern step after a chunk's plane equal to the
    global lowest plane (`if last_wp != wplanes[0]`); as in
    w_towers_imaging that is not reproduced.  Antenna indices outside
    [0, nant) raise IndexError."""
    N = int(math.ceil(theta * lam))
    S = int(subgrid_size)
    if S % 2 or S < 2 or not 0 <= margin < S:
        raise ValueError('AW-towers need an even subgrid_size above the margin')
    M = 2 * S
    uv = _uvw(uvw)
    if uv.shape[1] < 3:
        raise ValueError('AW-towers need uvw with a w column')
    dev = uv.device
    nv = uv.shape[0]
    like_torch = _torch_in(uvw, vis)
    grid = torch.zeros((N, N), dtype=torch.complex64, device=dev)
    if nv == 0:
        return _out(grid, like_torch)
    st = _dev(src, torch.int64).reshape(nv, -1)[:, :2]
    at = _dev(aimage, torch.complex64)
    nant = at.shape[0]
    if tuple(at.shape[1:]) != (S, S):
        raise ValueError('aimage must be [nant, subgrid_size, subgrid_size]')
    if bool(((st < 0) | (st >= nant)).any()):
        raise IndexError('antenna index outside the A-screens')
    gcf = _dev(aa_gcf, torch.complex64)
    if gcf.ndim != 2:
        raise ValueError('aa_gcf must be a separable [Q, K] table')
    Q, K = int(gcf.shape[0]), int(gcf.shape[1])
    chunk = S - margin
    wp = torch.round(uv[:, 2] / float(wstep)).to(torch.int64)
    ub = torch.floor(uv[:, 0] * theta / chunk + 0.5).to(torch.int64)
    vb = torch.floor(uv[:, 1] * theta / chunk + 0.5).to(torch.int64)
    lo = torch.stack([wp.min(), ub.min(), vb.min()]).cpu().tolist()
    hi = torch.stack([wp.max(), ub.max(), vb.max()]).cpu().tolist()
    P, nub, nvb = (hi[k] - lo[k] + 1 for k in range(3))
    cid = (ub - lo[1]) * nvb + (vb - lo[2])
    key = ((cid * P + (wp - lo[0])) * nant + st[:, 0]) * nant + st[:, 1]
    key, order = torch.sort(key, stable=True)
    gkey, gcount = torch.unique_consecutive(key, return_counts=True)
    G = gkey.shape[0]
    ga2 = (gkey % nant).to(torch.int32)
    ga1 = ((gkey // nant) % nant).to(torch.int32)
    gwp_rel = (gkey // (nant * nant)) % P
    gcid = gkey // (nant * nant * P)
    gwp = gwp_rel + lo[0]
    # chunks in key order, and each group's chunk slot
    ccid, ccount = torch.unique_consecutive(gcid, return_counts=True)
    C = ccid.shape[0]
    gchunk = torch.repeat_interleave(torch.arange(C, device=dev), ccount)
    ccid_h = ccid.cpu().numpy()
    uv_mid, corner = _awt_chunks(theta, lam, S, margin, ccid_h // nvb + lo[1],
                                 ccid_h % nvb + lo[2])
    mid_g = torch.cat([torch.as_tensor(uv_mid, device=dev)[gchunk],
                       (gwp.to(torch.float64) * float(wstep))[:, None]], 1)
    gid = torch.repeat_interleave(torch.arange(G, device=dev, dtype=torch.int32), gcount)
    rel = (uv[order, :3] - mid_g[gid.long()]).contiguous()
    vc = _dev(vis, torch.complex128).reshape(-1)[order]
    if dl != 0 or dm != 0:
        vc = vc / w_kernel_function(torch.tensor(float(dl), dtype=torch.float64, device=dev),
                                    torch.tensor(float(dm), dtype=torch.float64, device=dev),
                                    rel[:, 2])
    vt = vc.to(torch.complex64).contiguous()
    lam_sub = S / theta
    L = _lib.load()
    _lib.check(L.croc_awt_pregrid(_lib.ctx(), None, M, _lib.ptr(rel), 0, nv, float(lam_sub),
                                  Q, K, _lib.ptr(gcf), _lib.ptr(vt), _lib.ptr(gid), 0, 1))
    # image-space tables of the sub-grid
    a2 = numpy.asarray(aa_subgrid2, dtype=float)[S // 2:S // 2 + S]
    a1 = numpy.asarray(aa_subgrid, dtype=float)
    win = torch.as_tensor(numpy.outer(a1, a1) / numpy.outer(a2, a2), dtype=torch.float32,
                          device=dev).contiguous()
    wk = numpy.asarray(wkern.cpu().numpy() if _is_torch(wkern) else wkern, dtype=complex)
    turns = torch.as_tensor(numpy.angle(wk) / (2 * numpy.pi), device=dev).contiguous()
    lmag = torch.as_tensor(numpy.log2(numpy.abs(wk)), dtype=torch.float32,
                           device=dev).contiguous()
    # group ranges: visibility offsets and chunk offsets
    vstart = torch.zeros(G + 1, dtype=torch.int64, device=dev)
    vstart[1:] = torch.cumsum(gcount, 0)
    cstart = torch.zeros(C + 1, dtype=torch.int64, device=dev)
    cstart[1:] = torch.cumsum(ccount, 0)
    vstart_h, cstart_h = vstart.cpu().numpy(), cstart.cpu().numpy()
    bins = torch.zeros((C, S, S), dtype=torch.complex64, device=dev)
    gwp32 = gwp.to(torch.int32).contiguous()
    gmax = max(1, int(max_stack_bytes) // (M * M * 8))
    stack = torch.empty((min(G, gmax), M, M), dtype=torch.complex64, device=dev)
    for g0 in range(0, G, gmax):
        g1 = min(G, g0 + gmax)
        sb = stack[:g1 - g0]
        sb.zero_()
        _lib.check(L.croc_awt_pregrid(_lib.ctx(), _lib.ptr(sb), M, _lib.ptr(rel),
                                      int(vstart_h[g0]), int(vstart_h[g1]), float(lam_sub),
                                      Q, K, _lib.ptr(gcf), _lib.ptr(vt), _lib.ptr(gid), g0, 0))
        img = torch.fft.ifftn(sb, dim=(-2, -1))
        c_lo = int(numpy.searchsorted(cstart_h, g0, side='right')) - 1
        c_hi = int(numpy.searchsorted(cstart_h, g1, side='left'))
        for c0 in range(c_lo, c_hi, 65535):
            c1 = min(c_hi, c0 + 65535)
            # group range of each chunk within the batch
            starts = numpy.clip(cstart_h[c0:c1 + 1], g0, g1) - g0
            st_t = torch.as_tensor(starts.astype(numpy.int32), device=dev)
            _lib.check(L.croc_awt_apply(
                _lib.ctx(), _lib.ptr(bins), c0, c1 - c0, _lib.ptr(img), S, _lib.ptr(st_t),
                _lib.ptr(gwp32[g0:g1]), _lib.ptr(ga1[g0:g1]), _lib.ptr(ga2[g0:g1]),
                _lib.ptr(at), _lib.ptr(win), _lib.ptr(turns), _lib.ptr(lmag)))
        del img
    fb = torch.fft.fftn(bins, dim=(-2, -1))
    corner_t = torch.as_tensor(corner.astype(numpy.int32), device=dev).contiguous()
    for c0 in range(0, C, 65535):
        c1 = min(C, c0 + 65535)
        _lib.check(L.croc_awt_place(_lib.ctx(), _lib.ptr(grid), N, _lib.ptr(fb[c0:c1]),
                                    c1 - c0, S, _lib.ptr(corner_t[c0:c1])))
    return _out(grid, like_torch)
