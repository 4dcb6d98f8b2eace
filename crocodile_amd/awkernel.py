"""AW-kernel building blocks on the device (crocodile/synthesis.py:701-734).

``convolve_same(a, b)`` is scipy.signal.convolve2d(a, b, mode='same');
``convolve_same_batched(a, wkern)`` applies it to every oversampled
sub-kernel of ``wkern`` ([Q, Q, S, S]) in one launch, the inner loop of
aw_kernel_fn (:729-731).  FP64 throughout; NumPy in, NumPy out (torch in,
torch out).
"""

import numpy
import torch

from . import _lib


def _c128(x):
    if isinstance(x, torch.Tensor):
        return x.to(device=_lib.device(), dtype=torch.complex128).contiguous()
    return torch.as_tensor(numpy.ascontiguousarray(numpy.asarray(x, dtype=numpy.complex128)),
                           device=_lib.device())


def _conv(a, b, batch_shape):
    at, bt = _c128(a), _c128(b)
    if at.ndim != 2:
        raise ValueError('convolve2d inputs must both be 2-D arrays')
    bh, bw = bt.shape[-2:]
    nb = int(numpy.prod(batch_shape)) if batch_shape else 1
    out = torch.empty(tuple(batch_shape) + tuple(at.shape), dtype=torch.complex128,
                      device=at.device)
    _lib.check(_lib.load().croc_convolve_same(_lib.ctx(), _lib.ptr(at), at.shape[0],
                                              at.shape[1], _lib.ptr(bt), bh, bw, nb,
                                              _lib.ptr(out)))
    return out


def convolve_same(a, b):
    """scipy.signal.convolve2d(a, b, mode='same') on the device."""
    out = _conv(a, b, ())
    return out if isinstance(a, torch.Tensor) else out.cpu().numpy()


def convolve_same_batched(a, wkern):
    """[convolve2d(a, wk, 'same') for every sub-kernel wk of wkern], shaped
    wkern.shape[:-2] + a.shape."""
    batch = tuple(wkern.shape[:-2])
    out = _conv(a, wkern, batch)
    return out if isinstance(a, torch.Tensor) else out.cpu().numpy()
