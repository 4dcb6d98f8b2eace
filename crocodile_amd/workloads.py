"""Seeded synthetic inputs for the five benchmark workloads.

Visibilities come from point sources through the phase-tracked direct
Fourier transform of crocodile/simulate.py:210-230 (simulate_point), with
(u, v, w) either drawn from analytic densities or from antenna layouts rotated
by xyz_to_uvw (simulate.py:74-100).  Everything here is input generation, off
the gridding path; it runs on the host in NumPy, or on the device through
torch when ``device`` is given for the large layouts.

Workload names:

* ``plain512``   N=512, theta=0.1, lam=5120, 1e5 vis in a disk, w=0,
                 separable PSWF kernel S=8 Q=8.
* ``wproj2048``  N=2048, theta=0.1, lam=20480, 2e6 vis with surface density
                 r^-1.5, w uniform +-1500, w-kernels every 20 wavelengths with
                 support 16..32 growing with |w|, Q=8.
* ``towers4096`` N=4096, theta=0.08, lam=51200, 1e7 vis from 128 antennas
                 (Gaussian, sigma 1.5 km) over 6 h at 1.4 GHz, dec -45 deg,
                 AW-style kernel S=16 Q=16.
* ``diff1024``   1024^2 image of Gaussian blobs, 1e6 vis, w +-500.
* ``towers8192`` N=8192, theta=0.05, lam=163840, 5e7 vis from 256 antennas
                 (Gaussian core plus log-spiral arms to 40 km) over 8 h at
                 150 MHz, dec -45 deg, AW-style kernel S=16 Q=32, 256
                 w-planes.
* ``stream32k``  subgrid-streaming degrid (examples/grid/streamer.c): the
                 32k-8k-1k recombination geometry (image 32768, subgrids
                 xM=1024 spaced xA=832, bench_recombine.c:154-165) with the
                 LOWBD2-style visibility spec (bench_recombine.c:20-36: 512
                 stations, 128 dumps in chunks of 64, channels from 260 MHz
                 in steps of 40 MHz / 2048, chunks of 64, grid theta = field
                 of view = image / 210000); one GPU's share is one 64-channel
                 frequency chunk of every baseline, 1.07e9 visibilities;
                 separable PSWF kernel S=8, Q=1024.
* ``wstack2048`` full-grid w-stacking (wstacking.ipynb cell 10): N=2048,
                 theta=0.1, 2e6 vis in a disk of radius 0.4 lam, w uniform
                 +-4000, 64 w-planes of 125 wavelengths, 20 point sources.
* ``awtowers1800`` AW-towers at the geometry of awtowers.ipynb (cells 3-22):
                 theta=0.1, lam=18000 (N=1800), sub-grids 128 with margin 30,
                 wstep 50, 27 antennas (core and three arms to 18 km) over
                 36 dumps at 5 m, dec 45 deg; per-antenna ionospheric phase
                 screens as the notebook's ion_sample; anti-aliasing PSWF
                 support 6 oversampled 4096 times.
"""

import math

import numpy

SPEC = {
    'plain512': dict(N=512, theta=0.1, lam=5120.0, nv=100_000, S=8, Q=8,
                     nsrc=10, seed=1001),
    'wproj2048': dict(N=2048, theta=0.1, lam=20480.0, nv=2_000_000, Q=8,
                      wmax=1500.0, wstep=20.0, S_min=16, S_max=32, NpixFF=64,
                      nsrc=20, seed=1002),
    'towers4096': dict(N=4096, theta=0.08, lam=51200.0, nv=10_000_000, Q=16,
                       S=16, nant=128, hours=6.0, freq=1.4e9, dec=-45.0,
                       subgrid=128, nsrc=20, seed=1003),
    'diff1024': dict(N=1024, theta=0.1, lam=10240.0, nv=1_000_000, Q=8,
                     wmax=500.0, nblob=50, seed=1004),
    'towers8192': dict(N=8192, theta=0.05, lam=163840.0, nv=50_000_000, Q=32,
                       S=16, nant=256, hours=8.0, freq=150e6, dec=-45.0,
                       subgrid=128, nsrc=20, seed=1005),
    'wstack2048': dict(N=2048, theta=0.1, lam=20480.0, nv=2_000_000, wmax=4000.0, wstep=125.0,
                       nsrc=20, seed=1007),
    'awtowers1800': dict(theta=0.1, lam=18000.0, subgrid=128, margin=30, wstep=50.0, nant=27,
                         ndump=36, wvl=5.0, dec=45.0, hours=6.0, aa_support=6, aa_over=4096,
                         nsrc=25, seed=1008),
    'stream32k': dict(image_size=32768, theta=32768 / 210000, xA=832, xM=1024, nant=512,
                      ndump=128, time_chunk=64, nchan=64, freq_chunk=64, freq0=260e6,
                      dfreq=40e6 / 2048, hours=4.0, dec=-45.0, S=8, Q=1024, seed=1006),
}


def wproj_support(w, wmax=1500.0, S_min=16, S_max=32):
    """Even kernel support growing linearly from S_min at w=0 to S_max at
    |w| = wmax (the wproj2048 w-kernel size rule)."""
    steps = (S_max - S_min) // 2
    k = int(math.ceil(steps * min(abs(float(w)) / wmax, 1.0) - 1e-9))
    return S_min + 2 * k


def point_sources(rng, n, theta):
    """n sources with l, m uniform in +-theta/2 and lognormal(0,1) flux."""
    l = rng.uniform(-theta / 2, theta / 2, n)
    m = rng.uniform(-theta / 2, theta / 2, n)
    flux = rng.lognormal(0.0, 1.0, n)
    return l, m, flux


def dft_vis(uvw, l, m, flux, chunk=1 << 20):
    """sum_k F_k exp(-2 pi i (u l_k + v m_k + w (n_k - 1))) in FP64, the
    phase-tracked point-source DFT of simulate.py:228-230."""
    nv = uvw.shape[0]
    out = numpy.zeros(nv, dtype=numpy.complex128)
    n1 = numpy.sqrt(1.0 - l ** 2 - m ** 2) - 1.0
    for s in range(0, nv, chunk):
        u = uvw[s:s + chunk]
        acc = numpy.zeros(u.shape[0], dtype=numpy.complex128)
        for k in range(len(l)):
            ph = u[:, 0] * l[k] + u[:, 1] * m[k] + u[:, 2] * n1[k]
            acc += flux[k] * numpy.exp(-2j * numpy.pi * ph)
        out[s:s + chunk] = acc
    return out


def add_noise(rng, vis, sigma=0.01):
    return vis + sigma * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))


def plain512(nv=None, seed=None):
    """uvw [nv,3] (wavelengths, w=0), vis [nv] complex128, sources."""
    c = SPEC['plain512']
    rng = numpy.random.default_rng(c['seed'] if seed is None else seed)
    nv = c['nv'] if nv is None else nv
    R = 0.45 * c['lam']
    r = R * numpy.sqrt(rng.uniform(0, 1, nv))
    phi = rng.uniform(0, 2 * numpy.pi, nv)
    uvw = numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi), numpy.zeros(nv)], axis=1)
    l, m, f = point_sources(rng, c['nsrc'], c['theta'])
    vis = add_noise(rng, dft_vis(uvw, l, m, f))
    return uvw, vis


def wproj2048(nv=None, seed=None):
    """uvw [nv,3] with surface density ~ r^-1.5 out to 0.45 lam, w uniform
    in +-wmax; vis from 20 point sources plus noise."""
    c = SPEC['wproj2048']
    rng = numpy.random.default_rng(c['seed'] if seed is None else seed)
    nv = c['nv'] if nv is None else nv
    R = 0.45 * c['lam']
    r = R * rng.uniform(0, 1, nv) ** 2
    phi = rng.uniform(0, 2 * numpy.pi, nv)
    w = rng.uniform(-c['wmax'], c['wmax'], nv)
    uvw = numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi), w], axis=1)
    l, m, f = point_sources(rng, c['nsrc'], c['theta'])
    vis = add_noise(rng, dft_vis(uvw, l, m, f))
    return uvw, vis


def xyz_at_latitude(local_xyz, lat):
    """Local (east, north, up) antenna coordinates rotated to celestial
    (x east, z towards the pole) at latitude ``lat`` (simulate.py:48-67)."""
    x, y, z = local_xyz[:, 0], local_xyz[:, 1], local_xyz[:, 2]
    c = numpy.pi / 2 - lat
    return numpy.stack([x, -z * numpy.sin(c) + y * numpy.cos(c),
                        z * numpy.cos(c) + y * numpy.sin(c)], axis=1)


def xyz_to_uvw(xyz, ha, dec):
    """Celestial (x, y, z) rotated to (u, v, w) for hour angle ``ha`` and
    declination ``dec`` (simulate.py:74-100); xyz [n, 3], ha scalar or [n]."""
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    u = x * numpy.cos(ha) - y * numpy.sin(ha)
    v0 = x * numpy.sin(ha) + y * numpy.cos(ha)
    w = z * numpy.sin(dec) - v0 * numpy.cos(dec)
    v = z * numpy.cos(dec) + v0 * numpy.sin(dec)
    return numpy.stack([u, v, w], axis=1)


def antenna_layout(rng, nant, sigma_m=1500.0, arms_to_m=None):
    """Antenna (x, y, z) in metres: a 2D Gaussian core (z small), optionally
    plus three log-spiral arms reaching ``arms_to_m``."""
    if arms_to_m is None:
        xy = rng.normal(0.0, sigma_m, (nant, 2))
    else:
        ncore = nant // 2
        core = rng.normal(0.0, sigma_m, (ncore, 2))
        narm = nant - ncore
        t = numpy.linspace(0.0, 1.0, narm)
        rad = sigma_m * (arms_to_m / sigma_m) ** t
        ang = 2.0 * numpy.pi * (numpy.arange(narm) % 3) / 3 + 1.5 * numpy.log(rad / sigma_m)
        arms = numpy.stack([rad * numpy.cos(ang), rad * numpy.sin(ang)], axis=1)
        xy = numpy.concatenate([core, arms])
    z = rng.normal(0.0, 5.0, (nant, 1))
    return numpy.concatenate([xy, z], axis=1)


LATITUDE = -26.82 * numpy.pi / 180   # SKA-Low site
C_LIGHT = 299792458.0


def baseline_tracks(xyz, hours, freq, dec_deg, ndump):
    """uvw in wavelengths [ndump * nbl, 3] of every antenna pair over
    ``hours`` of hour angle centred on transit, time-major."""
    nant = xyz.shape[0]
    a1, a2 = numpy.triu_indices(nant, 1)
    bl = xyz_at_latitude(xyz, LATITUDE)
    bl = bl[a2] - bl[a1]
    dec = dec_deg * numpy.pi / 180
    has = (numpy.arange(ndump) + 0.5) / ndump - 0.5
    has = has * hours / 24 * 2 * numpy.pi
    lam_m = C_LIGHT / freq
    out = numpy.empty((ndump, bl.shape[0], 3))
    for t, ha in enumerate(has):
        out[t] = xyz_to_uvw(bl, ha, dec) / lam_m
    return out.reshape(-1, 3)


def towers(name, nv=None, seed=None, device=None, margin_cells=48):
    """(uvw [nv, 3] float64, vis [nv] complex) of the ``towers4096`` /
    ``towers8192`` workloads: antenna tracks, samples whose (u, v) would lie
    within ``margin_cells`` cells of the grid edge dropped, the first ``nv``
    kept; visibilities of 20 point sources plus noise.  With ``device`` the
    arrays are torch tensors there (the direct Fourier sum then runs on
    the device, FP64)."""
    c = SPEC[name]
    rng = numpy.random.default_rng(c['seed'] if seed is None else seed)
    nv = c['nv'] if nv is None else nv
    if name == 'towers8192':
        xyz = antenna_layout(rng, c['nant'], sigma_m=1500.0, arms_to_m=40000.0)
    else:
        xyz = antenna_layout(rng, c['nant'], sigma_m=1500.0)
    nbl = c['nant'] * (c['nant'] - 1) // 2
    lim = (c['N'] // 2 - margin_cells) / c['theta']
    ndump = max(1, -(-nv // nbl))
    while True:
        uvw = baseline_tracks(xyz, c['hours'], c['freq'], c['dec'], ndump)
        keep = (numpy.abs(uvw[:, 0]) < lim) & (numpy.abs(uvw[:, 1]) < lim)
        if keep.sum() >= nv:
            break
        ndump = int(ndump * 1.25 * nv / max(1, keep.sum())) + 1
    uvw = uvw[keep][:nv]
    l, m, f = point_sources(rng, c['nsrc'], c['theta'])
    if device is None:
        return uvw, add_noise(rng, dft_vis(uvw, l, m, f))
    import torch
    ut = torch.as_tensor(uvw, device=device)
    n1 = numpy.sqrt(1.0 - l ** 2 - m ** 2) - 1.0
    vis = torch.zeros(nv, dtype=torch.complex128, device=device)
    for k in range(len(l)):
        ph = ut[:, 0] * l[k] + ut[:, 1] * m[k] + ut[:, 2] * n1[k]
        vis += f[k] * torch.exp(-2j * numpy.pi * ph)
    g = torch.Generator(device=device).manual_seed(int(c['seed']))
    vis += 0.01 * torch.complex(torch.randn(nv, generator=g, device=device, dtype=torch.float64),
                                torch.randn(nv, generator=g, device=device, dtype=torch.float64))
    return ut, vis


def _towers_antennas(rng, name):
    """Antenna layout of the ``towers4096`` / ``towers8192`` workloads (the
    first draws of towers()' generator)."""
    c = SPEC[name]
    if name == 'towers8192':
        return antenna_layout(rng, c['nant'], sigma_m=1500.0, arms_to_m=40000.0)
    return antenna_layout(rng, c['nant'], sigma_m=1500.0)


def towers_sample(name, n, seed, margin_cells=48):
    """A seeded uniform subsample of ``n`` visibilities of the full
    ``towers4096`` / ``towers8192`` workload: random (dump, baseline) pairs of
    the full track set (the same antennas, sources, hour-angle sampling and
    edge filter as towers() at its full size), each computed directly, with
    noise from ``seed``.  Returns (uvw [n, 3], vis [n] complex128) on the
    host: the full set's distribution at a cost proportional to n, for
    oracle checks and CPU timings."""
    c = SPEC[name]
    grng = numpy.random.default_rng(c['seed'])
    xyz = _towers_antennas(grng, name)
    ls, ms, fs = point_sources(grng, c['nsrc'], c['theta'])
    nant = xyz.shape[0]
    nbl = nant * (nant - 1) // 2
    a1, a2 = numpy.triu_indices(nant, 1)
    bl = xyz_at_latitude(xyz, LATITUDE)
    bl = bl[a2] - bl[a1]
    lim = (c['N'] // 2 - margin_cells) / c['theta']
    lam_m = C_LIGHT / c['freq']
    dec = c['dec'] * numpy.pi / 180

    def hour_angle(t, ndump):
        return ((t + 0.5) / ndump - 0.5) * c['hours'] / 24 * 2 * numpy.pi

    # the dump count of the full set: towers() takes enough dumps for nv
    # samples to survive the edge filter (estimated from 32 hour angles)
    keep = numpy.mean([((numpy.abs(u[:, 0]) < lim) & (numpy.abs(u[:, 1]) < lim)).mean()
                       for u in (xyz_to_uvw(bl, hour_angle(t, 32), dec) / lam_m
                                 for t in range(32))])
    ndump = max(1, int(numpy.ceil(c['nv'] / (nbl * max(keep, 1e-9)))))
    rng = numpy.random.default_rng(seed)
    parts, have = [], 0
    while have < n:
        k = 2 * (n - have) + 64
        t = rng.integers(0, ndump, k)
        b = rng.integers(0, nbl, k)
        u = xyz_to_uvw(bl[b], hour_angle(t, ndump), dec) / lam_m
        u = u[(numpy.abs(u[:, 0]) < lim) & (numpy.abs(u[:, 1]) < lim)]
        parts.append(u)
        have += u.shape[0]
    uvw = numpy.concatenate(parts)[:n]
    return uvw, add_noise(rng, dft_vis(uvw, ls, ms, fs))


def aw_far_field(n, theta, c_pswf=None):
    """Far field [n, n] of an AW-style kernel without its w part: the PSWF
    anti-aliasing function times a Gaussian primary-beam A-term (FWHM =
    the field of view).  Input to kernel_oversample."""
    import scipy.special
    x = numpy.arange(-n // 2, n - n // 2) / n
    c = n // 2 if c_pswf is None else c_pswf
    aa = scipy.special.pro_ang1(0, 0, c, numpy.clip(2 * numpy.abs(x), 0, 0.999999))[0]
    ll = x * theta
    beam = numpy.exp(-4 * numpy.log(2) * (ll / theta) ** 2)
    a1 = aa * beam
    return numpy.outer(a1, a1)


def aw_taper(n, theta, support):
    """1D image-plane taper of the AW kernels of the towers workloads,
    sampled on n pixels of the field theta: crocodile's PSWF anti-aliasing
    function (synthesis.py:199-218, c = 1.3 * support as in
    wtowers-predict.ipynb) times a Gaussian primary beam whose FWHM is the
    field of view.  The 2D taper is its outer product."""
    import scipy.special
    x = (numpy.arange(n) - n // 2) / n
    aa = scipy.special.pro_ang1(0, 0, 1.3 * support, (2 - 1 / n / 4) * x)[0]
    return aa * beam_taper(x * theta, theta)


def beam_taper(l, theta):
    """Gaussian primary beam exp(-4 ln 2 (l / theta)^2): half power at the
    field edge l = +-theta/2."""
    return numpy.exp(-4 * numpy.log(2) * (numpy.asarray(l) / theta) ** 2)


def gaussian_blobs(rng, N, nblob=50, sigma=(1.0, 4.0)):
    """float32 [N, N] image of Gaussian blobs with lognormal amplitudes,
    centres in the inner 80% of the field."""
    img = numpy.zeros((N, N), dtype=numpy.float64)
    yy, xx = numpy.mgrid[0:N, 0:N]
    for _ in range(nblob):
        cx, cy = rng.uniform(0.1 * N, 0.9 * N, 2)
        s = rng.uniform(*sigma)
        a = rng.lognormal(0.0, 1.0)
        r = int(5 * s) + 1
        x0, x1 = max(int(cx) - r, 0), min(int(cx) + r + 1, N)
        y0, y1 = max(int(cy) - r, 0), min(int(cy) + r + 1, N)
        img[y0:y1, x0:x1] += a * numpy.exp(-((xx[y0:y1, x0:x1] - cx) ** 2 +
                                            (yy[y0:y1, x0:x1] - cy) ** 2) / (2 * s * s))
    return img.astype(numpy.float32)


def diff_uvw(rng, nv, lam, wmax):
    """uvw with surface density ~ r^-1.5 out to 0.45 lam, w uniform."""
    R = 0.45 * lam
    r = R * rng.uniform(0, 1, nv) ** 2
    phi = rng.uniform(0, 2 * numpy.pi, nv)
    return numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi),
                        rng.uniform(-wmax, wmax, nv)], axis=1)


def diff1024(nv=None, N=None, seed=None):
    """(image, uvw) of the differentiable-degridding workload."""
    c = SPEC['diff1024']
    rng = numpy.random.default_rng(c['seed'] if seed is None else seed)
    N = c['N'] if N is None else N
    lam = N / c['theta']
    image = gaussian_blobs(rng, N, c['nblob'])
    uvw = diff_uvw(rng, c['nv'] if nv is None else nv, lam, c['wmax'])
    return image, uvw


# --------------------------------------------------------------------------
# Kernels and the DFT accuracy check of the towers workloads
# --------------------------------------------------------------------------

RES_STEP = 0.5      # residual w-kernel step of the w-towers sets (wavelengths)
NPLANES = 256       # w-planes of the towers workloads
MARGIN = 32         # w-towers sub-grid margin (cells); sub-grids are 128
PLAIN_FF = 128      # far-field pixels of the plain path's 2D AW w-kernels


def tower_margin(wmax, theta):
    """W-towers sub-grid margin for |w| up to wmax: the w-term moves a
    visibility's contribution by up to |w| theta^2 / 2 cells (the phase
    gradient w l at the field edge l = theta / 2), which must stay within
    the margin of its 128-cell sub-grid: MARGIN (32) while that spread fits,
    else 64, the most a 128 sub-grid takes with a chunk left (towers8192:
    31.6 cells -> 32; towers4096: 63.8 -> 64, predict error against the DFT
    0.126 -> 0.030)."""
    return MARGIN if abs(wmax) * theta * theta / 2 <= MARGIN else 64


def plain_support(w, theta, S, cap=64):
    """Support of the plain path's w-kernel at w: the AW kernel's S plus
    the w-term's spread, 2 ceil(|w| theta^2 / 2) cells (the phase gradient
    w l at the field edge l = theta/2, in cells of 1/theta), at most the
    library's 64."""
    return int(min(cap, S + 2 * math.ceil(abs(w) * theta * theta / 2 - 1e-9)))


def towers_kernels(c, wstep, ws_plain):
    """Kernels of the towers workloads, all normalised to unit DC gain.

    * w-towers: separable residual w-kernel sets (imaging, predict) of the
      AW taper sampled on the full grid's N pixels, residuals in steps of
      RES_STEP over [-wstep/2, wstep/2] (G.w_towers_kernel_set);
    * plain path: 2D AW w-kernels (taper on PLAIN_FF pixels times the
      w-term of each w-plane's centre) with supports growing with |w|
      (plain_support), and the same at the fixed support S."""
    import time
    import torch
    from . import synthesis as G
    theta, S, Q, N = c['theta'], c['S'], c['Q'], c['N']
    t0 = time.perf_counter()
    taper = aw_taper(N, theta, S)
    norm1 = float(G.kernel_oversample(taper, Q, S)[0].sum().real)
    kset_img = G.w_towers_kernel_set(theta, wstep, N, S, Q, RES_STEP, taper / norm1, True)
    kset_pred = G.w_towers_kernel_set(theta, wstep, N, S, Q, RES_STEP, taper / norm1, False)
    tp = aw_taper(PLAIN_FF, theta, S)
    norm2 = float(G.kernel_oversample(tp, Q, S)[0].sum().real) ** 2
    aterm = numpy.outer(tp, tp)
    sup = [plain_support(w, theta, S) for w in ws_plain]
    banks = {}
    for name, supports in (('rule', sup), ('fixed', [S] * len(ws_plain))):
        banks[name] = tuple(G.w_kernel_bank(theta, ws_plain, PLAIN_FF, supports, Q, inv,
                                            1.0 / norm2, aterm=aterm) for inv in (True, False))
    torch.cuda.synchronize()
    return dict(set_img=kset_img, set_pred=kset_pred, banks=banks, supports=sup,
                gen_ms=1e3 * (time.perf_counter() - t0), taper_norm=norm1)


def dft_predict_error(c, uvw_d, predict_fns, seed):
    """Accuracy of predict against the direct Fourier transform, the metric
    of wtowers-predict.ipynb (error = sqrt(mean |V - V_dft|^2) / mean
    |V_dft|): a model image of 20 pixel-centred point sources within 3/8 of
    the field of its centre, divided by the PSWF anti-aliasing function and
    FFT'd to a grid; each predict function degrids that grid; the reference
    is the phase-tracked DFT of the sources seen through the primary beam
    (which the AW kernels apply)."""
    import torch
    from . import synthesis as Gs
    dev = uvw_d.device
    N, theta, S = c['N'], c['theta'], c['S']
    lam = N / theta
    rng = numpy.random.default_rng(seed)
    lim = 3 * N // 8
    px = rng.integers(-lim, lim + 1, 20)
    py = rng.integers(-lim, lim + 1, 20)
    flux = rng.lognormal(0.0, 1.0, 20)
    aa = Gs.anti_aliasing_function(N, 0, 1.3 * S)
    image = torch.zeros((N, N), dtype=torch.float64, device=dev)
    for x, y, f in zip(px, py, flux):
        image[N // 2 + int(y), N // 2 + int(x)] += float(f / (aa[N // 2 + y] * aa[N // 2 + x]))
    grid = Gs.fft(image.to(torch.complex128)).to(torch.complex64).contiguous()
    del image
    ref = torch.zeros(uvw_d.shape[0], dtype=torch.complex128, device=dev)
    for x, y, f in zip(px, py, flux):
        l, m = x / lam, y / lam
        n1 = math.sqrt(1.0 - l * l - m * m) - 1.0
        a = f * float(beam_taper(l, theta) * beam_taper(m, theta))
        ref += a * torch.exp(-2j * math.pi * (uvw_d[:, 0] * l + uvw_d[:, 1] * m + uvw_d[:, 2] * n1))
    mref = float(ref.abs().mean())
    out = {}
    for name, fn in predict_fns.items():
        v = fn(grid).to(torch.complex128)
        out[name] = float((v - ref).abs().pow(2).mean().sqrt()) / mref
        del v
    del grid, ref
    return out


# --------------------------------------------------------------------------
# Subgrid-streaming degrid workload
# --------------------------------------------------------------------------

def stream(name='stream32k', nant=None, ndump=None, nchan=None, device=None):
    """Visibility spec of the streaming degrid workload: (uvw_m [nbl, ndump,
    3] metres of every station pair, time-major per baseline, freq [nchan]
    Hz), as torch device tensors when ``device`` is given.  Stations: a
    Gaussian core plus log-spiral arms to 40 km; dumps spread over
    ``hours`` of hour angle centred on transit (ha_to_uvw, grid.h:64-89)."""
    c = SPEC[name]
    rng = numpy.random.default_rng(c['seed'])
    nant = c['nant'] if nant is None else nant
    ndump = c['ndump'] if ndump is None else ndump
    nchan = c['nchan'] if nchan is None else nchan
    xyz = xyz_at_latitude(antenna_layout(rng, nant, sigma_m=1500.0, arms_to_m=40000.0),
                          LATITUDE)
    a1, a2 = numpy.triu_indices(nant, 1)
    bl = xyz[a2] - xyz[a1]
    dec = c['dec'] * numpy.pi / 180
    has = ((numpy.arange(ndump) + 0.5) / ndump - 0.5) * c['hours'] / 24 * 2 * numpy.pi
    uvw = numpy.empty((bl.shape[0], ndump, 3))
    for t, ha in enumerate(has):
        uvw[:, t] = xyz_to_uvw(bl, ha, dec)
    freq = c['freq0'] + c['dfreq'] * numpy.arange(nchan)
    if device is None:
        return uvw, freq
    import torch
    return (torch.as_tensor(uvw, device=device), torch.as_tensor(freq, device=device))


def stream_kernel(name='stream32k'):
    """Separable PSWF degridding kernel [Q, S] (FP64, rows normalised to
    unit sum) and its 1D anti-aliasing function on xM pixels."""
    import scipy.special
    c = SPEC[name]
    S, Q, n = c['S'], c['Q'], 64
    x = (numpy.arange(n) - n // 2) / n
    aa = scipy.special.pro_ang1(0, 0, 0.5 * numpy.pi * S, (2 - 1 / n / 4) * x)[0]
    # kernel_oversample (synthesis.py:330-354): pad to n Q, centred FFT,
    # every Q-th sample from n Q / 2 - Q (S // 2) + f
    big = numpy.zeros(n * Q, dtype=complex)
    big[n * Q // 2 - n // 2:n * Q // 2 - n // 2 + n] = aa
    ft = numpy.fft.fftshift(numpy.fft.fft(numpy.fft.ifftshift(big))).real
    start = n * Q // 2 - Q * (S // 2)
    k = numpy.array([ft[start + f:start + f + Q * S:Q] for f in range(Q)])
    return k / k.sum(axis=1, keepdims=True)


def stream_subgrids(name='stream32k', umax=None):
    """Subgrid offsets (iu * xA, iv * xA) cells of the streamer's work
    assignment (config.c:505-516): iu >= 0 only (visibilities at u < 0 are
    degridded as conjugates of their mirror image), every subgrid inside
    the image and reaching |u|, |v| <= umax wavelengths."""
    c = SPEC[name]
    xA, xM, N, th = c['xA'], c['xM'], c['image_size'], c['theta']
    kmax = (N // 2 - xM // 2) // xA
    if umax is not None:
        kmax = min(kmax, int(math.ceil((umax * th - xA / 2) / xA)))
    return [(iu * xA, iv * xA) for iv in range(-kmax, kmax + 1) for iu in range(0, kmax + 1)]


def wstack2048(nv=None, seed=None):
    """uvw [nv, 3] (wavelengths), visibilities of 20 point sources, and the
    spec of the w-stacking workload."""
    c = dict(SPEC['wstack2048'])
    nv = nv or c['nv']
    rng = numpy.random.default_rng(c['seed'] if seed is None else seed)
    r = 0.4 * c['lam'] * numpy.sqrt(rng.uniform(0, 1, nv))
    ph = rng.uniform(0, 2 * numpy.pi, nv)
    uvw = numpy.stack([r * numpy.cos(ph), r * numpy.sin(ph),
                       rng.uniform(-c['wmax'], c['wmax'], nv)], 1)
    l, m, flux = point_sources(rng, c['nsrc'], 0.8 * c['theta'])
    return c, uvw, dft_vis(uvw, l, m, flux)


def awtowers1800(nsample=None, seed=None):
    """The AW-towers workload: flipped uvw (v >= 0), antenna pairs, the
    visibilities of point sources seen through per-antenna phase screens,
    and the notebook's kernels and image-space tables (awtowers.ipynb cells
    3, 11-22).  ``nsample``: a seeded subset of the visibilities."""
    from . import synthesis as G
    c = dict(SPEC['awtowers1800'])
    rng = numpy.random.default_rng(c['seed'] if seed is None else seed)
    xyz = antenna_layout(rng, c['nant'], sigma_m=2000.0, arms_to_m=18000.0)
    uvw = baseline_tracks(xyz, c['hours'], C_LIGHT / c['wvl'], c['dec'], c['ndump'])
    a1, a2 = numpy.triu_indices(c['nant'], 1)
    src = numpy.tile(numpy.stack([a1, a2], 1), (c['ndump'], 1))
    coef = rng.uniform(-0.5, 0.5, (c['nant'], 4))

    def screen(a, l, m):
        return numpy.exp(2j * numpy.pi * (coef[a, 0] + 3 * coef[a, 1] * l + 3 * coef[a, 2] * m
                                          + 60 * coef[a, 3] * (l * l + m * m)))

    theta = c['theta']
    ls, ms, flux = point_sources(rng, c['nsrc'], 0.6 * theta)
    vis = numpy.zeros(len(uvw), dtype=complex)
    for l, m, f in zip(ls, ms, flux):
        n1 = numpy.sqrt(1 - l * l - m * m) - 1
        ratio = screen(src[:, 0], l, m) / screen(src[:, 1], l, m)
        vis += f * ratio * numpy.exp(-2j * numpy.pi * (uvw[:, 0] * l + uvw[:, 1] * m
                                                       + uvw[:, 2] * n1))
    flip = uvw[:, 1] < 0
    uvw[flip] = -uvw[flip]
    vis[flip] = numpy.conj(vis[flip])
    src[flip] = src[flip, ::-1]
    if nsample:
        pick = numpy.sort(rng.choice(len(uvw), nsample, replace=False))
        uvw, src, vis = uvw[pick], src[pick], vis[pick]
    N = int(math.ceil(theta * c['lam']))
    S, sup = c['subgrid'], c['aa_support']
    aa = G.anti_aliasing_function(N, 0, sup)
    l, m = theta * G.coordinates2(S)
    k = dict(aa=aa, aa_gcf=G.kernel_oversample(aa, c['aa_over'], sup),
             aa_sub=G.anti_aliasing_function(S, 0, sup),
             aa_sub2=G.anti_aliasing_function(2 * S, 0, sup),
             aimage=numpy.array([screen(a, l, m) for a in range(c['nant'])]),
             wkern=G.w_kernel_function(l, m, c['wstep']))
    return c, N, uvw, src, vis, k
