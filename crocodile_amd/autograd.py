"""Differentiable degridding: image -> visibilities as a torch.autograd.Function.

The forward model is crocodile's do_predict with w_cache_predict
(synthesis.py:914-927, :799-828): V = degrid_K(fft(I)) with the
per-visibility w-kernel K of each w-bin.  It is linear in the image, so the
gradient of a real loss is the adjoint applied to the incoming gradient:

    dL/dI = Re( fft^H degrid_K^H g ),  fft^H = N^2 ifft (centred),
    degrid_K^H g = grid_{conj K}(g) = conj( grid_K(conj g) ).

The backward pass is therefore the adjoint gridding kernel of libcrocgpu run
on conj(g) with the forward's own kernel bank: taps are recomputed from the
bank, nothing per visibility is stored between the passes.  The coordinates
are binned once, when the operator is built (a plan), since they stay fixed
while the image is fitted.  Both passes use the FFT of torch (rocFFT) for
the full-image transform.
"""

import numpy
import torch

from . import _lib
from . import synthesis as G


class WProjectionOperator:
    """Linear operator image [N, N] (real) -> visibilities [nv] (complex64).

    ``kernel_fn`` is not taken: the w-kernels are generated on the device
    with :func:`crocodile_amd.synthesis.w_kernel_bank` (invert=False, the
    predict kernels) for every w-bin present, normalised by NpixFF^2.
    ``support`` is an int or a callable w -> support.
    """

    def __init__(self, theta, lam, uvw, wstep, npix_ff, support, Qpx):
        self.theta, self.lam, self.wstep = float(theta), float(lam), float(wstep)
        self.N = int(round(theta * lam))
        dev = _lib.device()
        self.uvw = (uvw.to(dev, torch.float64) if isinstance(uvw, torch.Tensor)
                    else torch.as_tensor(numpy.asarray(uvw, dtype=numpy.float64), device=dev)
                    ).contiguous()
        wb = torch.round(self.uvw[:, 2] / self.wstep).to(torch.int64)
        self.wbin0, hi = int(wb.min()), int(wb.max())
        ws = [self.wstep * b for b in range(self.wbin0, hi + 1)]
        S = [support(w) if callable(support) else int(support) for w in ws]
        self.bank = G.w_kernel_bank(theta, ws, npix_ff, S, Qpx, invert=False,
                                    scale=1.0 / npix_ff ** 2)
        self.plan = G.GridPlan(self.bank, self.N, self.uvw, lam, _lib.PLANE_WBIN, None,
                               self.wstep, self.wbin0)

    @property
    def nv(self):
        return self.uvw.shape[0]

    def apply(self, image):
        """V = degrid(fft(image))."""
        grid = G.fft(image.to(torch.complex64)).contiguous()
        out = torch.empty(self.nv, dtype=torch.complex64, device=grid.device)
        self.plan.degrid(grid, out)
        return out

    def adjoint(self, g):
        """fft^H degrid^H g (complex image)."""
        acc = torch.zeros((self.N, self.N), dtype=torch.complex64, device=g.device)
        self.plan.grid(acc, torch.conj_physical(g).to(torch.complex64).contiguous())
        return G.ifft(torch.conj_physical(acc)) * (self.N * self.N)


class DegridFunction(torch.autograd.Function):
    """forward: image -> visibilities; backward: adjoint gridding."""

    @staticmethod
    def forward(ctx, image, op):
        ctx.op = op
        ctx.real_input = not image.is_complex()
        return op.apply(image)

    @staticmethod
    def backward(ctx, grad_v):
        g = ctx.op.adjoint(grad_v.to(torch.complex64))
        return (g.real if ctx.real_input else g), None


def degrid_image(image, op):
    """Differentiable V = op(image)."""
    return DegridFunction.apply(image, op)
