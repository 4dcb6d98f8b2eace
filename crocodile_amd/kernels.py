"""Kernel generation at scale: AW kernels by FFT convolution with shared
spectra, and a persistent kernel cache.

``AWKernels`` is ``aw_kernel_fn`` (crocodile/synthesis.py:701-734) computed
with Fourier transforms instead of direct 2D convolutions, as
KERNEL_WORK.md:485-520 suggests, with the exact ``mode='same'`` border
behaviour of scipy.signal.convolve2d kept (zero padding to the full linear
convolution, then the centred extract).  The transforms are shared: the
spectrum of each A-kernel (per antenna, time, frequency) and of every
oversampled w sub-kernel (per w) is computed once and cached; an antenna
pair's A x A kernel (and its spectrum) once per pair; each AW kernel is then
one product and one inverse transform per sub-kernel.  All in FP64 on the
device (rocFFT batched transforms); results equal the direct convolutions to
rounding (~1e-15 relative).

``KernelCache`` is the kernel cache of w_cache_imaging / w_cache_predict
(``kernel_cache``, synthesis.py:757-828; image_dataset.py:117-171 builds it
from pylru.FunctionCacheManager): a callable with the kernel function's
signature that keeps the kernels it made on the device, least recently used
first out beyond a byte budget, and optionally persists them on disk in the
layout scripts/generate_wkernels.py writes into its HDF5 file
(``wkern/<theta>/<w>/kern``, :55-65), one ``.npy`` file per kernel, so that
later processes reuse them.
"""

import collections
import os

import numpy
import torch

from . import _lib


def _c128(x):
    if isinstance(x, torch.Tensor):
        return x.to(device=_lib.device(), dtype=torch.complex128).resolve_conj().contiguous()
    return torch.as_tensor(numpy.ascontiguousarray(numpy.asarray(x, dtype=numpy.complex128)),
                           device=_lib.device())


def _spectrum(k, P):
    """FFT of k ([..., h, w]) zero-padded at the far end to [..., P, P]."""
    h, w = k.shape[-2:]
    pad = torch.zeros(k.shape[:-2] + (P, P), dtype=torch.complex128, device=k.device)
    pad[..., :h, :w] = k
    return torch.fft.fft2(pad)


def _same(full, ah, aw, bh, bw):
    """scipy's mode='same' window of a full linear convolution of an
    [ah, aw] array with a [bh, bw] one: [(bh-1)//2 : +ah, (bw-1)//2 : +aw]."""
    y0, x0 = (bh - 1) // 2, (bw - 1) // 2
    return full[..., y0:y0 + ah, x0:x0 + aw]


class AWKernels:
    """aw_kernel_fn(a_kernel_fn, w_kernel_fn) with FFT convolutions and
    shared spectra: ``AWKernels(a_fn, w_fn)(theta, w, a1, a2, *src)`` is
    convolve2d(convolve2d(A1, A2, 'same'), wk[yf, xf], 'same') for every
    sub-kernel of ``w_fn(theta, w)`` ([Q, Q, S, S]): a NumPy complex128
    array when the A-kernel function returns NumPy (as the reference does),
    else a complex128 device tensor (``numpy_out`` forces either).

    ``method``: 'fft'; 'direct' (the device's direct FP64 sums of
    croc_convolve_same on the cached A and w kernels); 'auto' (the
    default): direct sums while the A x A kernel and the w sub-kernels are
    small (S_A * S_w < 576), where they are faster on the GPU (one AW set of
    1024 sub-kernels: S = 16 direct 0.12 ms, FFT 0.22 ms; S = 32 direct
    0.74 ms, FFT 0.20 ms), FFTs beyond.
    ``max_entries`` bounds each kernel / spectrum cache (least recently used
    out)."""

    FFT_FROM = 576

    def __init__(self, a_kernel_fn, w_kernel_fn=None, max_entries=4096, numpy_out=None,
                 method='auto'):
        if w_kernel_fn is None:
            from .synthesis import w_kernel as w_kernel_fn
        if method not in ('auto', 'fft', 'direct'):
            raise ValueError("method must be 'auto', 'fft' or 'direct'")
        self.method = method
        self.a_fn, self.w_fn = a_kernel_fn, w_kernel_fn
        self.max_entries = int(max_entries)
        self.numpy_out = numpy_out
        self._a = collections.OrderedDict()   # (theta, a, *src) -> A kernel
        self._a_numpy = False
        self._aa = collections.OrderedDict()  # (theta, a1, a2, *src) -> AA spectrum
        self._w = collections.OrderedDict()   # (theta, w) -> [w kernel, {P: spectra}]
        self.transforms = 0                   # forward + inverse 2D FFTs done

    def _get(self, cache, key, make):
        v = cache.get(key)
        if v is None:
            v = make()
            cache[key] = v
            if len(cache) > self.max_entries:
                cache.popitem(last=False)
        else:
            cache.move_to_end(key)
        return v

    def _akern(self, theta, a, src):
        def make():
            k = self.a_fn(theta, a, *src)
            self._a_numpy = not isinstance(k, torch.Tensor)
            return _c128(k)
        return self._get(self._a, (theta, a) + src, make)

    def _aa_spectrum(self, theta, a1, a2, src, wshape):
        """Spectrum, at the size of the second convolution, of
        convolve2d(A1, A2, 'same')."""
        def make():
            k1, k2 = self._akern(theta, a1, src), self._akern(theta, a2, src)
            (h1, w1), (h2, w2) = k1.shape, k2.shape
            P1 = max(h1 + h2 - 1, w1 + w2 - 1)
            full = torch.fft.ifft2(_spectrum(k1, P1) * _spectrum(k2, P1))
            aa = _same(full, h1, w1, h2, w2)
            self.transforms += 3
            P2 = max(h1 + wshape[0] - 1, w1 + wshape[1] - 1)
            self.transforms += 1
            return aa.shape, P2, _spectrum(aa, P2)
        return self._get(self._aa, (theta, a1, a2, wshape) + src, make)

    def _w_kernel(self, theta, w):
        """(sub-kernel shape [Q, Q, S, S], per-size spectra) of w: made once."""
        def make():
            wk = _c128(self.w_fn(theta, w))
            if wk.ndim != 4:
                raise AssertionError('w kernels must be [Q, Q, S, S]')
            return [wk, {}]
        return self._get(self._w, (theta, float(w)), make)

    def _aa_direct(self, theta, a1, a2, src):
        from .awkernel import convolve_same

        def make():
            return convolve_same(self._akern(theta, a1, src), self._akern(theta, a2, src))
        return self._get(self._aa, ('direct', theta, a1, a2) + src, make)

    def __call__(self, theta, w, a1, a2, *src):
        src = tuple(src)
        entry = self._w_kernel(theta, w)
        wk, spectra = entry
        bh, bw = wk.shape[-2:]
        as_numpy = self._a_numpy if self.numpy_out is None else self.numpy_out
        method = self.method
        if method == 'auto':
            k1 = self._akern(theta, a1, src)
            as_numpy = self._a_numpy if self.numpy_out is None else self.numpy_out
            method = 'fft' if max(k1.shape) * max(bh, bw) >= self.FFT_FROM else 'direct'
        if method == 'direct':
            from .awkernel import convolve_same_batched
            out = convolve_same_batched(self._aa_direct(theta, a1, a2, src), wk)
            as_numpy = self._a_numpy if self.numpy_out is None else self.numpy_out
            return out.cpu().numpy() if as_numpy else out
        (ah, aw), P, aspec = self._aa_spectrum(theta, a1, a2, src, (bh, bw))
        wspec = spectra.get(P)
        if wspec is None:
            wspec = spectra[P] = _spectrum(wk, P)
            self.transforms += wk.shape[0] * wk.shape[1]
        full = torch.fft.ifft2(aspec * wspec)
        self.transforms += wk.shape[0] * wk.shape[1]
        out = _same(full, ah, aw, bh, bw).contiguous()
        as_numpy = self._a_numpy if self.numpy_out is None else self.numpy_out
        return out.cpu().numpy() if as_numpy else out


def _key_part(x):
    if isinstance(x, torch.Tensor):
        x = x.item()
    if isinstance(x, (float, numpy.floating)) and float(x).is_integer():
        return float(x)
    return x


class KernelCache:
    """Persistent kernel cache: ``KernelCache(kernel_fn, **kwargs)(theta, w,
    *src)`` returns ``kernel_fn(theta, w, *src, **kwargs)``, made once per
    key and kept on the device (complex128), least recently used first out
    once ``max_bytes`` is exceeded.  With ``directory`` every kernel made is
    also saved there as ``wkern/<theta>/<w>[/<src>...]/kern.npy`` (the
    naming of scripts/generate_wkernels.py:55-65) and kernels found there are
    loaded instead of made, across processes.  Counters: hits, misses,
    loads (from disk), made."""

    def __init__(self, kernel_fn, max_bytes=8 << 30, directory=None, **kwargs):
        self.fn, self.kwargs = kernel_fn, kwargs
        self.max_bytes = int(max_bytes)
        self.directory = directory
        self._c = collections.OrderedDict()
        self.bytes = 0
        self.hits = self.misses = self.loads = self.made = 0

    def path(self, theta, w, *src):
        parts = ['wkern', repr(float(theta)), repr(float(w))] + [repr(_key_part(s)) for s in src]
        return os.path.join(self.directory, *parts, 'kern.npy')

    def _make(self, theta, w, src):
        if self.directory is not None:
            p = self.path(theta, w, *src)
            if os.path.exists(p):
                self.loads += 1
                return _c128(numpy.load(p, allow_pickle=False))
        k = _c128(self.fn(theta, w, *src, **self.kwargs))
        self.made += 1
        if self.directory is not None:
            p = self.path(theta, w, *src)
            os.makedirs(os.path.dirname(p), exist_ok=True)
            tmp = p + '.tmp.npy'
            numpy.save(tmp, k.cpu().numpy(), allow_pickle=False)
            os.replace(tmp, p)
        return k

    def __call__(self, theta, w, *src):
        key = (float(theta), float(w)) + tuple(_key_part(s) for s in src)
        k = self._c.get(key)
        if k is not None:
            self.hits += 1
            self._c.move_to_end(key)
            return k
        self.misses += 1
        k = self._make(theta, w, key[2:])
        self._c[key] = k
        self.bytes += k.numel() * k.element_size()
        while self.bytes > self.max_bytes and len(self._c) > 1:
            _, old = self._c.popitem(last=False)
            self.bytes -= old.numel() * old.element_size()
        return k

    def __len__(self):
        return len(self._c)

    def clear(self):
        self._c.clear()
        self.bytes = 0


def generate_w_kernels(directory, theta, wstep, wcount, NpixFF, NpixKern, Qpx=8,
                       overwrite=False):
    """scripts/generate_wkernels.py:50-65 on the device: the w-kernels of
    w = iw * wstep for iw in [-wcount, wcount], all made in one launch
    (w_kernel_bank), written as ``wkern/<theta>/<w>/kern.npy`` under
    ``directory`` (existing files kept unless ``overwrite``).  Returns the
    list of w written."""
    from .synthesis import w_kernel_bank
    ws = [iw * wstep for iw in range(-wcount, wcount + 1)]
    cache = KernelCache(None, directory=directory)
    todo = [w for w in ws if overwrite or not os.path.exists(cache.path(theta, w))]
    if not todo:
        return []
    bank = w_kernel_bank(theta, todo, NpixFF, NpixKern, Qpx, True, double=True)
    n = Qpx * Qpx * NpixKern * NpixKern
    for k, w in enumerate(todo):
        o = int(bank.plane_off_h[k])
        t = bank.kern[o:o + n].reshape(Qpx, Qpx, NpixKern, NpixKern).cpu().numpy()
        p = cache.path(theta, w)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        numpy.save(p + '.tmp.npy', t, allow_pickle=False)
        os.replace(p + '.tmp.npy', p)
    return todo


def generate_w_kernels_h5(filename, theta, wstep, wcount, NpixFF, NpixKern, Qpx=8,
                          overwrite=False):
    """scripts/generate_wkernels.py itself: the w-kernels of w = iw * wstep,
    iw in [-wcount, wcount], made in one launch and stored in the HDF5 file
    ``filename`` as ``wkern/<theta>/<w>/kern`` [Qpx, Qpx, NpixKern, NpixKern]
    complex128 (the names and layout that hdf5.c:594-700 and
    image_dataset.py:114-129 read).  An existing file is extended, its
    kernels kept unless ``overwrite`` (the script opens it in append mode);
    the whole file is rewritten.  Returns the list of w written."""
    from . import formats, h5
    from .synthesis import w_kernel_bank
    ws = [iw * wstep for iw in range(-wcount, wcount + 1)]
    tree = {}
    if os.path.exists(filename):
        f = h5.File(filename)

        def collect(g, path):
            for k in g.keys():
                o = g[k]
                if isinstance(o, h5.Group):
                    collect(o, path + k + '/')
                else:
                    tree[path + k] = o[()]
        collect(f, '')
        del f
    name = lambda w: 'wkern/%s/%s/kern' % (formats._name(theta), formats._name(w))  # noqa: E731
    todo = [w for w in ws if overwrite or name(w) not in tree]
    if not todo:
        return []
    bank = w_kernel_bank(theta, todo, NpixFF, NpixKern, Qpx, True, double=True)
    n = Qpx * Qpx * NpixKern * NpixKern
    for k, w in enumerate(todo):
        o = int(bank.plane_off_h[k])
        tree[name(w)] = bank.kern[o:o + n].reshape(Qpx, Qpx, NpixKern, NpixKern).cpu().numpy()
    h5.write(filename + '.tmp', tree)
    os.replace(filename + '.tmp', filename)
    return todo
