"""FP64 NumPy restatement of AW-towers imaging (oracle, TEST INFRASTRUCTURE
ONLY): examples/notebooks/awtowers.ipynb cells 17 (pregrid_tower_plane,
JSON lines 374-399), 19 (grid_tower_plane, :426-440) and 22 (grid_tower,
grid_towers, :486-577), with the notebook's globals made arguments.

* Chunks: ub = floor(u theta / (subgrid_size - margin) + 0.5), vb likewise;
  chunk (ub, vb) spans xy_min = (subgrid_size - margin) (ub - 0.5, vb - 0.5)
  to xy_max = (.. + 0.5), centred at uv_mid = (xy_max + xy_min) // 2 / theta;
  the sub-grid starts at mid + int(xy_min - (margin + 1) // 2) of the grid
  (the notebook's in-place `xy_min -=` changes the caller's array, so the
  clipped add uses the extended corner).
* Per (chunk, w-plane wp = around(w / wstep), antenna pair (a1, a2)): the
  visibilities relative to (uv_mid, wp wstep) (divided by
  w_kernel_function(dl, dm, w - wp wstep) when dl or dm is non-zero) are
  gridded with the anti-aliasing kernel aa_gcf on a 2 subgrid_size grid at
  2 theta, transformed (centred ifft), divided by aa_subgrid2 (x) aa_subgrid2,
  cut to the middle subgrid_size, multiplied by aa_subgrid (x) aa_subgrid and
  by the A-screen ratio aimage[a2] / aimage[a1], and summed over pairs.
* Per chunk, the plane images are stacked with wkern powers as the notebook
  does (multiply the sum by wkern ** (wp - last_wp) before adding plane wp,
  divide by wkern ** last_wp at the end), transformed (centred fft) and
  added into the grid, clipped at its edges.

The notebook skips the wkern step after a chunk's plane equal to the global
lowest plane (`if last_wp != wplanes[0]`); as for w-towers that is a bug
(grid.c steps always) and is not reproduced.  Returns grid_sum; the
notebook's image is real(ifft(grid_sum)) / (aa (x) aa) / sum(wt).
"""

import numpy

from . import synthesis_ref as R


def chunk_geometry(theta, lam, subgrid_size, margin, ub, vb):
    """(uv_mid [2], sub-grid corner (x0, y0)) of chunk (ub, vb), with the
    notebook's float arithmetic (awtowers.ipynb cell 22)."""
    xy_min = (subgrid_size - margin) * numpy.array([ub - 0.5, vb - 0.5])
    xy_max = (subgrid_size - margin) * numpy.array([ub + 0.5, vb + 0.5])
    uv_mid = (xy_max + xy_min) // 2 / theta
    xy_min -= (margin + 1) // 2
    xy_max += margin // 2
    assert numpy.all(numpy.max(xy_max - xy_min) == subgrid_size)
    mid = int(lam * theta) // 2
    x0, y0 = mid + xy_min.astype(int)
    return uv_mid, int(x0), int(y0)


def pregrid_tower_plane(theta, lam, uvw_mid, puvw, pvis, aa_gcf, aa_subgrid, aa_subgrid2,
                        subgrid_size, dl=0.0, dm=0.0):
    puvw_new = puvw - uvw_mid
    if dl != 0 or dm != 0:
        pvis = pvis / R.w_kernel_function(dl, dm, puvw_new[:, 2])
    psrc = numpy.empty((len(puvw), 0))
    pgrid = R.conv_imaging(2 * theta, lam, puvw_new, psrc, pvis, aa_gcf)
    pimage = R.ifft(pgrid)
    pimage /= numpy.outer(aa_subgrid2, aa_subgrid2)
    pimage = R.extract_mid(pimage, subgrid_size)
    return pimage * numpy.outer(aa_subgrid, aa_subgrid)


def aw_towers_imaging(theta, lam, uvw, src, vis, wstep, subgrid_size, margin, aa_gcf,
                      aa_subgrid, aa_subgrid2, aimage, wkern, dl=0.0, dm=0.0):
    N = int(numpy.ceil(theta * lam))
    chunk = subgrid_size - margin
    wplane = numpy.around(uvw[:, 2] / wstep).astype(int)
    wplanes = numpy.arange(numpy.min(wplane), numpy.max(wplane) + 1)
    ubin = numpy.floor(uvw[:, 0] * theta / chunk + 0.5).astype(int)
    vbin = numpy.floor(uvw[:, 1] * theta / chunk + 0.5).astype(int)
    uvw_size = subgrid_size / theta
    grid_sum = numpy.zeros((N, N), dtype=complex)
    for ub in range(numpy.min(ubin), numpy.max(ubin) + 1):
        for vb in range(numpy.min(vbin), numpy.max(vbin) + 1):
            sel = (ubin == ub) & (vbin == vb)
            if not numpy.any(sel):
                continue
            uv_mid, x0, y0 = chunk_geometry(theta, lam, subgrid_size, margin, ub, vb)
            tuvw, tsrc, tvis, twp = uvw[sel], src[sel], vis[sel], wplane[sel]
            bin_sum = numpy.zeros((subgrid_size, subgrid_size), dtype=complex)
            last_wp = None
            for wp in wplanes:
                slc = twp == wp
                if not numpy.any(slc):
                    continue
                if last_wp is not None:
                    bin_sum *= wkern ** (wp - last_wp)
                last_wp = wp
                uvw_mid = numpy.hstack([uv_mid, [wp * wstep]])
                puvw, psrc, pvis = tuvw[slc], tsrc[slc], tvis[slc]
                for a1, a2 in numpy.unique(psrc, axis=0):
                    s = (psrc[:, 0] == a1) & (psrc[:, 1] == a2)
                    plane = pregrid_tower_plane(theta, uvw_size, uvw_mid, puvw[s], pvis[s],
                                                aa_gcf, aa_subgrid, aa_subgrid2, subgrid_size,
                                                dl, dm)
                    bin_sum += plane * aimage[a2] / aimage[a1]
            bin_sum /= wkern ** last_wp
            bin_grid = R.fft(bin_sum)
            x1, y1 = x0 + subgrid_size, y0 + subgrid_size
            x0b, y0b = max(x0, 0), max(y0, 0)
            x1b, y1b = min(x1, N), min(y1, N)
            if x1b > x0b and y1b > y0b:
                grid_sum[y0b:y1b, x0b:x1b] += bin_grid[y0b - y0:y1b - y0, x0b - x0:x1b - x0]
    return grid_sum


def dft_image(theta, N, uvw, src, vis, ascreen_fn, pixels):
    """Direct Fourier dirty image with the A-correction, at the given
    (x, y) pixels: sum_k vis_k A[a2](l, m) / A[a1](l, m)
    exp(2 pi i (u l + v m + w (n - 1))), l = theta (x - N//2) / N, m likewise;
    ``ascreen_fn(a, l, m)`` gives antenna a's screen.  The check the
    AW-towers image is held to."""
    out = numpy.empty(len(pixels), dtype=complex)
    for i, (x, y) in enumerate(pixels):
        l = theta * (x - N // 2) / N
        m = theta * (y - N // 2) / N
        n1 = numpy.sqrt(1 - l * l - m * m) - 1
        ph = numpy.exp(2j * numpy.pi * (uvw[:, 0] * l + uvw[:, 1] * m + uvw[:, 2] * n1))
        a = numpy.array([ascreen_fn(s[1], l, m) / ascreen_fn(s[0], l, m) for s in src])
        out[i] = numpy.sum(vis * a * ph)
    return out
