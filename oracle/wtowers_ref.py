"""FP64 NumPy restatement of the w-towers imaging / predict drivers (oracle,
TEST INFRASTRUCTURE ONLY).

Follows examples/notebooks/wtowers.ipynb (imaging cell, JSON lines 297-372),
examples/notebooks/wtowers-predict.ipynb (predict cell, lines 349-433) and
examples/grid/grid.c:482-710 (grid_wtowers, make_wtransfer :418-480), with
one chunk geometry for both directions (the two notebooks bin differently;
grid.c's grid-index chunking is used):

* chunk = subgrid_size - margin; a visibility's chunk is
  (floor(u*theta + 0.5) + N//2) // chunk along each axis (grid.c:528-531);
* chunk (cx, cy) owns the subgrid of absolute cells starting at
  x0 = cx*chunk - margin//2 (grid.c:664), centred on
  u_mid = (x0 + sg//2 - N//2) / theta;
* w-plane wp = round-half-even(w / wstep) (numpy.around in the notebooks);
* per (chunk, plane): grid the plane's visibilities at (u - u_mid, v - v_mid)
  on the subgrid (nearest cell as the imaging notebook's simple_imaging,
  or with a convolution kernel as the predict notebook's conv_predict),
  centred ifft, multiply by exp(-2 pi i wp wstep (1 - n)) and accumulate;
  centred fft of the sum and add the subgrid into the grid, clipped at the
  grid edge.  Multiplying by the Fresnel pattern directly for each plane
  equals grid.c's incremental cipow stepping (:629-641) in exact arithmetic
  and avoids its rounding drift;
* predict reverses it: cut the subgrid out of the grid (zeros outside),
  centred ifft, per plane multiply by exp(+2 pi i wp wstep (1 - n)),
  centred fft, degrid.

The imaging notebook skips the Fresnel step for the second plane when a
bin's first plane is the global first one (`if last_wp != wplanes[0]`); that
is a bug (grid.c always steps), not reproduced here.

The kernel ``kv`` is one table (the notebooks' form: every visibility of a
job uses it) or a w-kernel set ``(tables, w_min, w_step)`` as grid.c's
w_kernel_data: a visibility of plane wp uses table
floor((w - wp wstep - w_min) / w_step + 0.5) (w_project, grid.c:277-305,
with the residual w relative to the plane's w, grid.c:609-617).
"""

import numpy

from . import synthesis_ref as R


def plan(theta, lam, uvw, wstep, subgrid_size, margin):
    """Chunk indices (cx, cy), w-plane index and grid size."""
    N = int(round(theta * lam))
    chunk = subgrid_size - margin
    assert chunk > 0 and margin >= 0
    cellx = numpy.floor(uvw[:, 0] * theta + 0.5).astype(int) + N // 2
    celly = numpy.floor(uvw[:, 1] * theta + 0.5).astype(int) + N // 2
    cx = cellx // chunk
    cy = celly // chunk
    wp = numpy.around(uvw[:, 2] / wstep).astype(int)
    return N, chunk, cx, cy, wp


def subgrid_origin(c, chunk, margin):
    return c * chunk - margin // 2


def fresnel_phase(subgrid_size, theta):
    """1 - sqrt(1 - l^2 - m^2) on the subgrid's image pixels."""
    l, m = R.kernel_coordinates(subgrid_size, theta)
    return 1 - numpy.sqrt(1 - l ** 2 - m ** 2)


def _kernel_planes(kv, uvw, wp, wstep):
    """(tables, plane index per visibility) for a single table or a
    w-kernel set (grid.c:286-289)."""
    if isinstance(kv, tuple):
        tables, w_min, w_step = kv
        dw = uvw[:, 2] - wp * wstep
        k = numpy.floor((dw - w_min) / w_step + 0.5).astype(int)
        assert (k >= 0).all() and (k < len(tables)).all(), 'w-kernel set too small'
        return tables, k
    return [kv], numpy.zeros(len(uvw), dtype=int)


def _sub_coords(uvw, theta, N, sg, x0, y0):
    u_mid = (x0 + sg // 2 - N // 2) / theta
    v_mid = (y0 + sg // 2 - N // 2) / theta
    p = numpy.empty((len(uvw), 2))
    p[:, 0] = (uvw[:, 0] - u_mid) / (sg / theta)
    p[:, 1] = (uvw[:, 1] - v_mid) / (sg / theta)
    return p


def w_towers_imaging(theta, lam, uvw, src, vis, wstep, subgrid_size, margin, kv=None):
    N, chunk, cx, cy, wp = plan(theta, lam, uvw, wstep, subgrid_size, margin)
    sg = subgrid_size
    ph = fresnel_phase(sg, theta)
    guv = numpy.zeros((N, N), dtype=complex)
    if kv is not None:
        tables, kidx = _kernel_planes(kv, uvw, wp, wstep)
    keys = sorted(set(zip(cy.tolist(), cx.tolist())))
    for (ky, kx) in keys:
        in_chunk = (cx == kx) & (cy == ky)
        x0 = subgrid_origin(kx, chunk, margin)
        y0 = subgrid_origin(ky, chunk, margin)
        img = numpy.zeros((sg, sg), dtype=complex)
        for w in sorted(set(wp[in_chunk].tolist())):
            sel = in_chunk & (wp == w)
            p = _sub_coords(uvw[sel], theta, N, sg, x0, y0)
            sub = numpy.zeros((sg, sg), dtype=complex)
            if kv is None:
                R.grid(sub, p, vis[sel])
            else:
                for k in sorted(set(kidx[sel].tolist())):
                    sk = kidx[sel] == k
                    R.convgrid(tables[k], sub, p[sk], None, vis[sel][sk])
            img += R.ifft(sub) * numpy.exp(-2j * numpy.pi * (w * wstep) * ph)
        sub = R.fft(img)
        xa, xb = max(x0, 0), min(x0 + sg, N)
        ya, yb = max(y0, 0), min(y0 + sg, N)
        guv[ya:yb, xa:xb] += sub[ya - y0:yb - y0, xa - x0:xb - x0]
    return guv


def w_towers_predict(theta, lam, uvw, src, guv, wstep, subgrid_size, margin, kv=None):
    N, chunk, cx, cy, wp = plan(theta, lam, uvw, wstep, subgrid_size, margin)
    sg = subgrid_size
    ph = fresnel_phase(sg, theta)
    out = numpy.zeros(len(uvw), dtype=complex)
    if kv is not None:
        tables, kidx = _kernel_planes(kv, uvw, wp, wstep)
    keys = sorted(set(zip(cy.tolist(), cx.tolist())))
    for (ky, kx) in keys:
        in_chunk = (cx == kx) & (cy == ky)
        x0 = subgrid_origin(kx, chunk, margin)
        y0 = subgrid_origin(ky, chunk, margin)
        sub = numpy.zeros((sg, sg), dtype=complex)
        xa, xb = max(x0, 0), min(x0 + sg, N)
        ya, yb = max(y0, 0), min(y0 + sg, N)
        sub[ya - y0:yb - y0, xa - x0:xb - x0] = guv[ya:yb, xa:xb]
        img = R.ifft(sub)
        for w in sorted(set(wp[in_chunk].tolist())):
            sel = in_chunk & (wp == w)
            p = _sub_coords(uvw[sel], theta, N, sg, x0, y0)
            psub = R.fft(img * numpy.exp(2j * numpy.pi * (w * wstep) * ph))
            if kv is None:
                out[sel] = R.degrid(psub, p)
            else:
                idx = numpy.flatnonzero(sel)
                for k in sorted(set(kidx[sel].tolist())):
                    sk = kidx[sel] == k
                    out[idx[sk]] = R.convdegrid(tables[k], psub, p[sk])
    return out
