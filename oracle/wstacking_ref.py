"""FP64 NumPy restatement of the w-stacking imaging variants of
examples/notebooks/wstacking.ipynb (oracle, TEST INFRASTRUCTURE ONLY).

* ``w_stacking_imaging`` -- the full-grid form (cell 10, wstacking.ipynb
  JSON lines 191-240): visibilities grouped by w-plane round-half-even(w /
  wstep); each plane imaged with simple_imaging on the full grid, the image
  divided by the Fresnel pattern w_kernel_function(l, m, mean w of the
  plane) and summed.  Returns the image sum.
* ``w_stacking_subgrid_imaging`` -- the per-plane sub-grid form (cell 12,
  lines 249-300): a plane whose mean w is within wstep / 2 of 0 is gridded
  directly; any other plane is gridded on the square sub-grid spanning its
  visibilities' cells plus the kernel size, centred at the sub-grid middle
  and at the plane's mean w, the sub-grid's image divided by the Fresnel
  pattern of that w, transformed back and added to the grid at the
  sub-grid's offset.  Returns the grid.

The notebook flips visibilities to one half-plane and weights them before
these loops (cell 8); callers do that here.  The notebook slices the grid
with the window's start wherever it lies; a window that leaves the grid is
an error here (NumPy would wrap or clip silently).
"""

import numpy

from . import synthesis_ref as R


def _planes(uvw, wstep):
    wplane = numpy.around(uvw[:, 2] / wstep).astype(int)
    return wplane, sorted(set(wplane.tolist()))


def w_stacking_imaging(theta, lam, uvw, src, vis, wstep):
    N = int(round(theta * lam))
    wplane, planes = _planes(uvw, wstep)
    l, m = theta * R.coordinates2(N)
    image_sum = numpy.zeros((N, N), dtype=complex)
    for wp in planes:
        sel = wplane == wp
        puvw, pvis = uvw[sel], vis[sel]
        midw = numpy.mean(puvw[:, 2])
        cdrt = R.simple_imaging(theta, lam, puvw, None, pvis)
        image_sum += R.ifft(cdrt) / R.w_kernel_function(l, m, midw)
    return image_sum


def w_stacking_subgrid_imaging(theta, lam, uvw, src, vis, wstep, npixkern):
    N = int(round(theta * lam))
    wplane, planes = _planes(uvw, wstep)
    grid_sum = numpy.zeros((N, N), dtype=complex)
    mid = int(lam * theta) // 2
    for wp in planes:
        sel = wplane == wp
        puvw, pvis = uvw[sel], vis[sel]
        midw = numpy.mean(puvw[:, 2])
        if abs(midw) < wstep / 2:
            grid_sum += R.simple_imaging(theta, lam, puvw, None, pvis)
            continue
        xy_min = numpy.floor(numpy.amin(puvw[:, :2], axis=0) * theta).astype(int)
        xy_max = numpy.ceil(numpy.amax(puvw[:, :2], axis=0) * theta).astype(int)
        xy_min -= (npixkern + 1) // 2
        xy_max += npixkern // 2
        xy_size = int(numpy.max(xy_max - xy_min))
        xy_maxq = numpy.amax([xy_max, xy_min + xy_size], axis=0)
        uvw_size = xy_size / theta
        uvw_mid = numpy.hstack([(xy_maxq + xy_min) // 2 / theta, midw])
        pgrid = R.simple_imaging(theta, uvw_size, puvw - uvw_mid, None, pvis)
        l, m = theta * R.coordinates2(xy_size)
        pgrid_w = R.fft(R.ifft(pgrid) / R.w_kernel_function(l, m, midw))
        x0, y0 = mid + xy_min
        x1, y1 = mid + xy_max
        assert x0 >= 0 and y0 >= 0 and x1 <= N and y1 <= N, 'sub-grid leaves the grid'
        grid_sum[y0:y1, x0:x1] += pgrid_w[0:y1 - y0, 0:x1 - x0]
    return grid_sum
