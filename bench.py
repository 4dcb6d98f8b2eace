#!/usr/bin/env python3
"""Grid + degrid throughput of crocodile's gridding path on one MI355X.

Workload ``towers8192`` (the default, the largest configuration and the one
the w-towers pipeline targets): N=8192 grid, theta=0.05, 5e7 visibilities
of 256 antennas over 8 h at 150 MHz, 256 w-planes, AW kernels of support 16
and oversampling 32.  One step = one w-towers round trip (bin, image,
predict) through the C ABI with inputs resident in HBM; see run_towers.
``towers4096`` is the same on the 4096^2 configuration.

Workload ``wproj2048``: N=2048 grid, theta=0.1, lam=20480,
2e6 visibilities with surface density ~ r^-1.5 out to 0.45 lam, w uniform in
+-1500 wavelengths, one w-kernel per 20-wavelength w-bin (151 bins, support
16..32 growing with |w|, oversampling 8, far field 64 px), visibilities from
20 point sources plus noise.  One step = w_cache_imaging's gridding of all
visibilities into a cleared grid followed by w_cache_predict's degridding of
that grid (per-visibility kernel of its w-bin, crocodile/synthesis.py:757-828)
through the C ABI with inputs resident in HBM: the coordinates are binned
once per step (croc_plan_bin) and used by both croc_convgrid_plan and
croc_convdegrid_plan.  The one-shot croc_convgrid + croc_convdegrid pair
(binning on each call) is timed too and reported under "one_shot".  Kernel
generation is done once before timing (it is reported separately).

Prints one JSON line (metric visibilities/s, higher is better) with the
roofline of the dominant kernel measured by HIP events over the timed steps
and the CPU baseline (the FP64 NumPy restatement of the same per-visibility
loops, one core, on a bounded sample of the same workload).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# w-towers imaging: jobs of at most this many visibilities are gridded by the
# summing kernel (k_tw_image_runs), the rest by the job kernel (wtowers.hip)
TW_SMALL_VIS = 128
FP16_MFMA_PEAK_TFLOPS = 2500.0  # dense
LDS_PEAK_TBS = 78.6  # 256 CUs x 128 B/clk x 2.4 GHz


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='towers8192',
                    choices=['wproj2048', 'plain512', 'towers4096', 'towers8192', 'diff1024',
                             'dft128', 'stream32k', 'wstack2048', 'awtowers1800'])
    ap.add_argument('--nv', type=int, default=0,
                    help='towers workloads: visibilities (0 = the workload\'s own count)')
    ap.add_argument('--cpu-sample', type=int, default=0,
                    help='visibilities in the CPU-baseline sample (0: the workload default, '
                         '20000 for the towers workloads, 300000 for wproj2048, about 10-15 s '
                         'of one core; for stream32k a count of chunk slots, default 32000; '
                         'negative disables it)')
    ap.add_argument('--copy-probe-mib', type=int, default=0,
                    help='before warm-up, copy a buffer of this many MiB once (a plain copy '
                         'of known size, to scale HBM counters of a profiled run)')
    ap.add_argument('--autotune', action='store_true',
                    help='measure a few tile / chunk settings during warm-up and keep the '
                         'fastest (the default is the library default: automatic tile, '
                         'chunk 512)')
    ap.add_argument('--headline-only', action='store_true',
                    help='towers workloads: time the headline w-towers steps only (no '
                         'comparison paths, DFT accuracy or CPU baseline), for profiled runs')
    ap.add_argument('--tile', type=int, default=0)
    ap.add_argument('--chunk', type=int, default=512)
    ap.add_argument('--group-kib', type=int, default=0)
    return ap.parse_args()


def build_wproj(G, W, torch, dev, nv=None):
    c = W.SPEC['wproj2048']
    uvw, vis = W.wproj2048(nv=nv)
    wb = numpy_round(uvw[:, 2] / c['wstep'])
    lo, hi = int(wb.min()), int(wb.max())
    ws = [c['wstep'] * b for b in range(lo, hi + 1)]
    S = [W.wproj_support(w, c['wmax'], c['S_min'], c['S_max']) for w in ws]
    scale = 1.0 / c['NpixFF'] ** 2
    t0 = time.perf_counter()
    bank_img = G.w_kernel_bank(c['theta'], ws, c['NpixFF'], S, c['Q'], True, scale)
    bank_pred = G.w_kernel_bank(c['theta'], ws, c['NpixFF'], S, c['Q'], False, scale)
    torch.cuda.synchronize()
    kgen_ms = 1e3 * (time.perf_counter() - t0)
    return dict(c=c, uvw=uvw, vis=vis, ws=ws, S=S, lo=lo, scale=scale,
                bank_img=bank_img, bank_pred=bank_pred, kgen_ms=kgen_ms,
                kernel_bytes=8 * int(bank_img.kern.numel()))


def tap_count(config, uvw):
    """Kernel taps per launch: sum over visibilities of S(w-bin)^2."""
    import numpy
    if config['workload'] == 'plain512':
        return len(uvw) * config['support'] ** 2
    from crocodile_amd import workloads as W
    c = W.SPEC['wproj2048']
    wb = numpy.round(uvw[:, 2] / c['wstep']).astype(int)
    bins, counts = numpy.unique(wb, return_counts=True)
    S = numpy.array([W.wproj_support(c['wstep'] * b, c['wmax'], c['S_min'], c['S_max'])
                     for b in bins])
    return int((counts * S.astype(numpy.int64) ** 2).sum())


def numpy_round(x):
    import numpy
    return numpy.round(x).astype(int)


def cpu_baseline(G, W, torch, dev, n):
    """FP64 NumPy restatement (oracle) of w_cache_imaging + w_cache_predict,
    single core, on an ``n``-visibility sample of the workload; also the
    device result on the same sample for the error figures."""
    import numpy
    from oracle import synthesis_ref as R

    c = W.SPEC['wproj2048']
    uvw, vis = W.wproj2048(nv=n, seed=c['seed'] + 77)
    wb = numpy.round(uvw[:, 2] / c['wstep']).astype(int)
    tables_img, tables_pred = {}, {}
    for b in numpy.unique(wb):
        w = c['wstep'] * b
        S = W.wproj_support(w, c['wmax'], c['S_min'], c['S_max'])
        tables_img[w] = R.w_kernel(c['theta'], w, c['NpixFF'], S, c['Q']) / c['NpixFF'] ** 2
        tables_pred[w] = R.w_kernel(c['theta'], w, c['NpixFF'], S, c['Q'],
                                    invert=False) / c['NpixFF'] ** 2
    t0 = time.perf_counter()
    g64 = R.w_cache_imaging(c['theta'], c['lam'], uvw, None, vis, wstep=c['wstep'],
                            kernel_cache=lambda th, w: tables_img[w])
    t1 = time.perf_counter()
    v64 = R.w_cache_predict(c['theta'], c['lam'], uvw, None, g64, wstep=c['wstep'],
                            kernel_cache=lambda th, w: tables_pred[w])
    t2 = time.perf_counter()
    # device result on the same sample (kernels generated on the device)
    kw = dict(NpixFF=c['NpixFF'], Qpx=c['Q'])
    g32 = G.w_cache_imaging(c['theta'], c['lam'], uvw, None, vis, wstep=c['wstep'],
                            kernel_fn=lambda th, w: G.w_kernel(
                                th, w, kw['NpixFF'],
                                W.wproj_support(w, c['wmax'], c['S_min'], c['S_max']),
                                kw['Qpx']) / c['NpixFF'] ** 2)
    v32 = G.w_cache_predict(c['theta'], c['lam'], uvw, None, g64, wstep=c['wstep'],
                            kernel_fn=lambda th, w: G.w_kernel(
                                th, w, kw['NpixFF'],
                                W.wproj_support(w, c['wmax'], c['S_min'], c['S_max']),
                                kw['Qpx'], invert=False) / c['NpixFF'] ** 2)

    def err(x, ref):
        d = x - ref
        return (float(numpy.linalg.norm(d) / numpy.linalg.norm(ref)),
                float(numpy.abs(d).max() / numpy.abs(ref).max()))

    ge, ve = err(g32, g64), err(v32, v64)
    base = {
        'value': n / (t2 - t0), 'unit': 'visibilities/s', 'cores': 1, 'kind': 'port',
        'sample': '%d visibilities of wproj2048 (seed %d), FP64 NumPy per-visibility '
                  'w_cache_imaging + w_cache_predict loops (oracle/synthesis_ref.py), '
                  'kernels precomputed; grid %.2fs + degrid %.2fs'
                  % (n, c['seed'] + 77, t1 - t0, t2 - t1),
        'host_cpus': os.cpu_count(),
    }
    parity = {'sample_nv': n, 'grid_rms_rel': ge[0], 'grid_max_rel': ge[1],
              'vis_rms_rel': ve[0], 'vis_max_rel': ve[1]}
    return base, parity


def run_towers(args, torch, G, W, _lib, dev):
    """The north-star workload: the w-towers grid + degrid round trip on one
    GPU (towers8192: N = 8192, 5e7 antenna-track visibilities, 256 w-planes,
    AW kernels S = 16, Q = 32), next to the plain paths on the same data.

    One step = bin the coordinates into (chunk, w-plane) jobs
    (croc_wtowers_plan_bin), w-towers imaging into a cleared grid
    (croc_wtowers_grid_plan) and w-towers predict from it
    (croc_wtowers_degrid_plan).  The sub-grid gridding uses grid.c's
    w_project rule: each visibility takes the kernel of its residual w
    relative to its plane from a separable residual w-kernel set.

    Comparisons on the same visibilities: the binned tile path with one 2D
    AW w-kernel per w-plane (PLANE_WBIN) at supports growing with |w|
    ('plain') and at the fixed S ('plain_s16'), and the plain per-visibility
    kernels ('per_visibility').  Each path's predict error against the DFT
    of a point-source model is measured at full size; the FP64 oracle
    (oracle/wtowers_ref.py) is timed on a seeded sample of the same
    distribution and the device result compared with it there."""
    import numpy
    name = args.workload
    c = W.SPEC[name]
    N, theta, S, Q = c['N'], c['theta'], c['S'], c['Q']
    lam = N / theta
    nv = args.nv or c['nv']
    print('[bench] generating %s (%d visibilities)' % (name, nv), file=sys.stderr, flush=True)
    uvw_d, vis64 = W.towers(name, nv=nv, device=dev)
    vis_d = vis64.to(torch.complex64).contiguous()
    del vis64
    w = uvw_d[:, 2]
    wstep = (float(w.max()) - float(w.min())) / (W.NPLANES - 1)
    wb = torch.round(w / wstep)
    lo, hi = int(wb.min()), int(wb.max())
    margin = W.tower_margin(float(w.abs().max()), theta)
    ws_plain = [wstep * b for b in range(lo, hi + 1)]
    print('[bench] kernels: %d w-planes, wstep %.2f' % (len(ws_plain), wstep), file=sys.stderr,
          flush=True)
    K = W.towers_kernels(c, wstep, ws_plain)
    kimg, kpred = K['set_img'], K['set_pred']
    guv = torch.zeros((N, N), dtype=torch.complex64, device=dev)
    out = torch.empty(nv, dtype=torch.complex64, device=dev)
    L = _lib.load()

    def kargs(ks):
        return (_lib.ptr(ks.tables), int(ks.nplanes), float(ks.w_min), float(ks.w_step))

    copy_probe(torch, dev, args.copy_probe_mib)
    twp = _lib._P()
    _lib.check(L.croc_wtowers_plan_create(_lib.ctx(), ctypes.byref(twp)))

    def tw_bin():
        _lib.check(L.croc_wtowers_plan_bin(_lib.ctx(), twp, N, _lib.ptr(uvw_d), nv, 3, theta,
                                           wstep, 128, margin, 1, Q, S, kimg.nplanes,
                                           kimg.w_min, kimg.w_step))

    def tw_grid(ks):
        _lib.check(L.croc_wtowers_grid_plan(_lib.ctx(), twp, _lib.ptr(guv), _lib.ptr(vis_d),
                                            _lib.ptr(ks.tables), 1))
        g = _lib.last_tap_kernel_ms()
        return g[0], g[1]

    def tw_degrid(ks, dst):
        _lib.check(L.croc_wtowers_degrid_plan(_lib.ctx(), twp, _lib.ptr(guv), _lib.ptr(dst),
                                              _lib.ptr(ks.tables), 1))
        d = _lib.last_tap_kernel_ms()
        return d[2], d[3]

    def towers_step():
        tw_bin()
        guv.zero_()
        gms, gn = tw_grid(kimg)
        dms, dn = tw_degrid(kpred, out)
        return gms, gn, dms, dn

    def towers_one_shot():
        guv.zero_()
        _lib.check(L.croc_wtowers_grid(_lib.ctx(), _lib.ptr(guv), N, _lib.ptr(uvw_d), nv, 3,
                                       theta, _lib.ptr(vis_d), wstep, 128, margin,
                                       _lib.ptr(kimg.tables), Q, S, 1, *kargs(kimg)[1:]))
        _lib.check(L.croc_wtowers_degrid(_lib.ctx(), _lib.ptr(guv), N, _lib.ptr(uvw_d), nv, 3,
                                         theta, _lib.ptr(out), wstep, 128, margin,
                                         _lib.ptr(kpred.tables), Q, S, 1, *kargs(kpred)[1:]))
        t = _lib.last_tap_kernel_ms()
        return t[0], t[1], t[2], t[3]

    plans = {}

    def plain_step(which):
        bi, bp = K['banks'][which]
        if which not in plans:
            plans[which] = G.GridPlan(bi, N)
        plan = plans[which]
        plan.bin(uvw_d, lam, _lib.PLANE_WBIN, None, wstep, lo)
        guv.zero_()
        plan.grid(guv, vis_d)
        plan.degrid(guv, out, bp)
        t = _lib.last_tap_kernel_ms()
        return t[0], t[1], t[2], t[3]

    def pervis_step():
        bi, bp = K['banks']['rule']
        guv.zero_()
        G.grid_device_pervis(bi, guv, uvw_d, lam, vis_d, _lib.PLANE_WBIN, None, wstep, lo)
        G.degrid_device_pervis(bp, guv, uvw_d, lam, out, _lib.PLANE_WBIN, None, wstep, lo)
        t = _lib.last_tap_kernel_ms()
        return t[0], t[1], t[2], t[3]

    def timed(step, label, steps, warmup):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        print('[bench] timing %d %s steps' % (steps, label), file=sys.stderr, flush=True)
        rec = []
        t0 = time.perf_counter()
        for _ in range(steps):
            rec.append(step())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        a = numpy.array(rec, dtype=float)
        return dict(value=nv * steps / dt, ms=1e3 * dt / steps, grid_tap_ms=a[:, 0].mean(),
                    grid_launches=int(a[:, 1].mean()), degrid_tap_ms=a[:, 2].mean(),
                    degrid_launches=int(a[:, 3].mean()))

    res = timed(towers_step, 'w-towers (planned)', args.steps, args.warmup)
    if args.headline_only:
        _lib.check(L.croc_wtowers_plan_destroy(_lib.ctx(), twp))
        print(json.dumps({'metric': 'visibilities/s grid+degrid', 'value': round(res['value'], 1),
                          'unit': 'visibilities/s', 'ms_per_step': round(res['ms'], 3),
                          'kernel_ms_per_step': {'k_tw_grid_jobs': round(res['grid_tap_ms'], 3),
                                                 'k_tw_degrid_jobs': round(res['degrid_tap_ms'], 3)},
                          'launches_per_step': [res['grid_launches'], res['degrid_launches']],
                          'config': {'workload': name, 'headline_only': True}}), flush=True)
        return
    few = max(2, args.steps // 4)
    one = timed(towers_one_shot, 'w-towers (one-shot)', few, 1)
    plain = timed(lambda: plain_step('rule'), 'plain tile path (support rule)', few, 1)
    plain16 = timed(lambda: plain_step('fixed'), 'plain tile path (S fixed)', few, 1)
    plans.clear()
    pv = timed(pervis_step, 'per-visibility', 2, 1)

    # job structure of the headline
    key = (torch.floor(uvw_d[:, 0] * theta + 0.5).to(torch.int64) + N // 2) // (128 - margin)
    key = key * 100000 + (torch.floor(uvw_d[:, 1] * theta + 0.5).to(torch.int64) + N // 2) \
        // (128 - margin)
    key = key * 4096 + (wb.to(torch.int64) - lo)
    _, jcount = torch.unique(key, return_counts=True)
    njobs = int(jcount.numel())
    # jobs of at most TW_SMALL_VIS visibilities are gridded by the summing
    # kernel (k_tw_image_runs), the rest by the job kernel
    small = jcount <= TW_SMALL_VIS
    nv_small, njobs_small = int(jcount[small].sum()), int(small.sum())
    nchunks = int(torch.unique(key // 4096).numel())
    del key, jcount, small

    # predict accuracy against the DFT, full size
    print('[bench] predict error against the DFT', file=sys.stderr, flush=True)
    single = G.WKernelSet(kpred.tables[kpred.nplanes // 2:kpred.nplanes // 2 + 1], 0.0, 1.0)
    nores = G.WTowersPlan(theta, lam, uvw_d, wstep, 128, margin, single.tables[0])

    def towers_pred(grid):
        guv.copy_(grid)
        v = torch.empty(nv, dtype=torch.complex64, device=dev)
        tw_degrid(kpred, v)
        return v

    def plain_pred(which):
        def fn(grid):
            v = torch.empty(nv, dtype=torch.complex64, device=dev)
            G.degrid_device(K['banks'][which][1], grid, uvw_d, lam, v, _lib.PLANE_WBIN, None,
                            wstep, lo)
            return v
        return fn

    tw_bin()
    dft_err = W.dft_predict_error(c, uvw_d, {
        'towers': towers_pred,
        'towers_no_residual_kernels': lambda g: nores.predict(g, single.tables[0]),
        'plain': plain_pred('rule'), 'plain_s16': plain_pred('fixed')}, c['seed'] + 99)
    del nores
    _lib.check(L.croc_wtowers_plan_destroy(_lib.ctx(), twp))

    # roofline of the dominant tap kernel, from the live HIP events of the
    # timed steps: 8 FP32 flop per kernel tap (one complex multiply-add),
    # S^2 taps per visibility and direction
    # separable sets of support <= 16: gridding of the jobs above
    # TW_SMALL_VIS visibilities on the matrix cores (k_tw_grid_jobs_mfma);
    # predict transforms each plane sub-grid and degrids from it in LDS
    # (k_tw_predict_sep: FFT and taps in one launch)
    split = S <= 16
    gname = 'k_tw_grid_jobs_mfma' if split else 'k_tw_grid_jobs'
    dname = 'k_tw_predict_sep' if split else 'k_tw_degrid_jobs'
    nv_g = nv - nv_small if split else nv
    nj_g = njobs - njobs_small if split else njobs
    if res['grid_tap_ms'] >= res['degrid_tap_ms']:
        kname, kms, kn = gname, res['grid_tap_ms'], res['grid_launches']
        taps = float(nv_g) * S * S
        # each gridded visibility's 16-byte record and 8-byte value read
        # once, each job's 128^2 complex64 sub-grid written once
        alg_bytes = (24.0 * nv_g + 8.0 * 128 * 128 * nj_g) / max(kn, 1)
    else:
        kname, kms, kn = dname, res['degrid_tap_ms'], res['degrid_launches']
        taps = float(nv) * S * S
        # each visibility's record read and value written once, each chunk
        # image read once (the plane sub-grids stay in LDS)
        alg_bytes = (24.0 * nv + 8.0 * 128 * 128 * nchunks) / max(kn, 1)
    achieved = 8.0 * taps / (kms * 1e-3) / 1e12
    roofline = {
        'bound': 'valu', 'achieved': round(achieved, 3), 'peak': FP32_PEAK_TFLOPS,
        'unit': 'TFLOP/s', 'frac': round(achieved / FP32_PEAK_TFLOPS, 5),
        'bound_note': 'complex multiply-add per kernel tap, 8 FP32 flop/tap (the reference '
                      'convention), over the taps the kernel forms, against the FP32 vector '
                      'peak: the gridding job kernel forms the products of the jobs above %d '
                      'visibilities (%d of %d jobs, %.1f %% of the visibilities) on the matrix '
                      'cores (fp16 hi/lo split, FP32 accumulation) and is bound by the vector '
                      'work and memory latency around them; the predict kernel transforms each '
                      'plane sub-grid in LDS and degrids from it one visibility per thread on '
                      'the vector ALUs (its time includes the transforms)'
                      % (TW_SMALL_VIS, nj_g, njobs, 100.0 * nv_g / nv),
        'kernel': kname, 'launches_per_step': kn,
        'kernel_ms_per_launch': round(kms / max(kn, 1), 4),
        'kernel_ms_per_step': {gname: round(res['grid_tap_ms'], 3),
                               dname: round(res['degrid_tap_ms'], 3)},
        'algorithmic_flops_per_launch': 8.0 * taps / max(kn, 1),
        'algorithmic_bytes_per_launch': round(alg_bytes),
        'traffic': None,
    }
    pmc = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                       '%s_%s_pmc.json' % (kname, name))
    if os.path.exists(pmc):
        prof = json.load(open(pmc))
        roofline['traffic'] = prof.get('hbm_bytes_per_launch')
        roofline['traffic_over_algorithmic'] = round(roofline['traffic'] / alg_bytes, 3)
        roofline['traffic_profile'] = {'file': os.path.relpath(pmc, os.path.dirname(
            os.path.abspath(__file__))), 'kernel_us': prof.get('kernel_us'),
            'note': 'calibrated FETCH_SIZE + WRITE_SIZE per launch from rocprofv3 --pmc of the '
                    'same command; compare kernel_us with kernel_ms_per_launch for staleness'}

    cpu, parity = None, None
    n_cpu = args.cpu_sample or 20000
    if n_cpu > 0:
        cpu, parity = towers_cpu_and_parity(G, W, c, name, wstep, kimg, kpred, n_cpu, margin)

    line = {
        'metric': 'visibilities/s grid+degrid', 'value': round(res['value'], 1),
        'unit': 'visibilities/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(res['ms'], 3), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': name, 'N': N, 'theta': theta, 'nv': nv, 'w_planes': W.NPLANES,
                   'wstep': round(wstep, 3), 'support': S, 'oversampling': Q,
                   'subgrid': 128, 'margin': margin, 'jobs': njobs, 'chunks': nchunks,
                   'residual_w_kernels': kimg.nplanes, 'residual_w_step': W.RES_STEP,
                   'kernels': 'separable AW (PSWF c=1.3S x Gaussian beam) x residual w-term',
                   'step': 'croc_wtowers_plan_bin + croc_wtowers_grid_plan + '
                           'croc_wtowers_degrid_plan',
                   'kernel_gen_ms': round(K['gen_ms'], 1)},
        'roofline': roofline, 'cpu_baseline': cpu, 'parity_vs_fp64_oracle': parity,
        'predict_error_vs_dft': {k: float('%.4g' % v) for k, v in dft_err.items()},
        'towers_one_shot': {'value': round(one['value'], 1), 'ms_per_step': round(one['ms'], 3),
                            'note': 'croc_wtowers_grid + croc_wtowers_degrid, each binning'},
        'plain': {'value': round(plain['value'], 1), 'ms_per_step': round(plain['ms'], 3),
                  'grid_tap_ms': round(plain['grid_tap_ms'], 3),
                  'degrid_tap_ms': round(plain['degrid_tap_ms'], 3),
                  'supports': [min(K['supports']), max(K['supports'])],
                  'kernel_bank_mib': round(8 * K['banks']['rule'][0].kern.numel() / 2 ** 20, 1),
                  'note': 'binned tile path, croc_plan_bin + croc_convgrid_plan + '
                          'croc_convdegrid_plan, 2D AW w-kernel per w-plane, support '
                          'S + 2 ceil(|w| theta^2 / 2) capped at 64'},
        'plain_s16': {'value': round(plain16['value'], 1), 'ms_per_step': round(plain16['ms'], 3),
                      'note': 'the same at support 16 for every w-plane'},
        'per_visibility': {'value': round(pv['value'], 1), 'ms_per_step': round(pv['ms'], 3),
                           'note': 'croc_convgrid_pervis + croc_convdegrid_pervis (one '
                                   'wavefront per visibility, FP32 atomics), support rule bank'},
        'speedup_vs_plain': round(res['value'] / plain['value'], 3),
        'speedup_vs_plain_s16': round(res['value'] / plain16['value'], 3),
        'speedup_vs_per_visibility': round(res['value'] / pv['value'], 3),
    }
    print(json.dumps(line), flush=True)


def towers_cpu_and_parity(G, W, c, name, wstep, kimg, kpred, n, margin):
    """The FP64 NumPy restatement of the w-towers drivers
    (oracle/wtowers_ref.py) timed on one core over a seeded n-visibility
    sample of the same distribution (W.towers_sample), with the device
    kernel-set tables; and the device result on the same sample against
    it."""
    import numpy
    from oracle import wtowers_ref as WT
    theta, N = c['theta'], c['N']
    lam = N / theta
    uvw, vis = W.towers_sample(name, n, c['seed'] + 77)
    ki = (kimg.tables.cpu().numpy().astype(numpy.complex128), kimg.w_min, kimg.w_step)
    kp = (kpred.tables.cpu().numpy().astype(numpy.complex128), kpred.w_min, kpred.w_step)
    print('[bench] cpu baseline: FP64 oracle w-towers on %d visibilities' % n, file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    g64 = WT.w_towers_imaging(theta, lam, uvw, None, vis, wstep, 128, margin, ki)
    t1 = time.perf_counter()
    v64 = WT.w_towers_predict(theta, lam, uvw, None, g64, wstep, 128, margin, kp)
    t2 = time.perf_counter()
    g32 = G.w_towers_imaging(theta, lam, uvw, None, vis, wstep, 128, margin, kimg)
    v32 = G.w_towers_predict(theta, lam, uvw, None, g64, wstep, 128, margin, kpred)

    def err(x, ref):
        d = x - ref
        return (float(numpy.linalg.norm(d) / numpy.linalg.norm(ref)),
                float(numpy.abs(d).max() / numpy.abs(ref).max()))

    ge, ve = err(g32, g64), err(v32, v64)
    del g32, g64
    base = {'value': n / (t2 - t0), 'unit': 'visibilities/s', 'cores': 1, 'kind': 'port',
            'sample': '%d visibilities of %s (towers_sample seed %d): FP64 NumPy w-towers '
                      'imaging + predict (oracle/wtowers_ref.py) with the same residual w-kernel '
                      'sets; grid %.2fs + degrid %.2fs; the rate is per visibility of this '
                      'sample (its per-job sub-grid FFTs grow more slowly than nv, so it does '
                      'not extrapolate linearly to the full set)'
                      % (n, name, c['seed'] + 77, t1 - t0, t2 - t1),
            'host_cpus': os.cpu_count()}
    parity = {'sample_nv': n, 'grid_rms_rel': ge[0], 'grid_max_rel': ge[1],
              'vis_rms_rel': ve[0], 'vis_max_rel': ve[1]}
    return base, parity


def run_dft(args, torch, _lib, dev):
    """The w-towers sub-grid transform stage on its own: a batch of 2048
    centred 2D transforms of 128 x 128 complex64 sub-grids (crocodile's fft,
    synthesis.py:83-93, per sub-grid), as the dense DFT Y = D X D on the
    matrix cores (croc_dft2d_batched, fp16x3 split products with FP32
    accumulation) and as the radix 16 x 8 FFT on the vector ALUs
    (croc_fft2d_batched).  The roofline is the matrix-core one: algorithmic
    16 n^3 flop per transform (two complex GEMMs) against the fp16 dense
    peak; the fp16x3 split executes three times that on the matrix cores."""
    n, batch = 128, 2048
    gen = torch.Generator(device=dev)
    gen.manual_seed(1005)
    x = torch.randn((batch, n, n), dtype=torch.complex64, device=dev, generator=gen)
    y = torch.empty_like(x)
    L = _lib.load()

    def run(entry):
        fn = getattr(L, entry)
        for _ in range(args.warmup):
            _lib.check(fn(_lib.ctx(), _lib.ptr(x), _lib.ptr(y), n, batch, 0))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            _lib.check(fn(_lib.ctx(), _lib.ptr(x), _lib.ptr(y), n, batch, 0))
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    print('[bench] timing %d dft128 steps' % args.steps, file=sys.stderr, flush=True)
    dft_ms = run('croc_dft2d_batched')
    fft_ms = run('croc_fft2d_batched')
    flops = 16.0 * n ** 3 * batch
    achieved = flops / (dft_ms * 1e-3) / 1e12
    line = {
        'metric': 'sub-grid 2D transforms/s (128x128 complex64, centred DFT)',
        'value': round(batch / (dft_ms * 1e-3), 1), 'unit': 'transforms/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(dft_ms, 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'fp16x3 (fp32 accumulate)', 'data': 'synthetic',
        'config': {'workload': 'dft128', 'n': n, 'batch': batch},
        'roofline': {'bound': 'mfma', 'achieved': round(achieved, 2),
                     'peak': FP16_MFMA_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': round(achieved / FP16_MFMA_PEAK_TFLOPS, 5),
                     'executed_tflops': round(3 * achieved, 2), 'kernel': 'k_dft2d',
                     'traffic': None},
        'fft': {'entry': 'croc_fft2d_batched', 'ms_per_step': round(fft_ms, 4),
                'transforms_per_s': round(batch / (fft_ms * 1e-3), 1),
                'note': 'radix 16 x 8 FFT on the vector ALUs, the w-towers default'},
    }
    print(json.dumps(line), flush=True)


def run_diff(args, torch, G, W, _lib, dev):
    """Differentiable degridding on the 1024^2 Gaussian-blob image with 1e6
    visibilities (w +-500, w-kernels per 20-wavelength bin, S 16..32, Q 8):
    one step = forward V = degrid(fft(I)) (croc_convdegrid_plan), the
    least-squares loss sum |V - V_obs|^2 and its gradient by the adjoint
    gridding (croc_convgrid_plan) and ifft; the coordinates are binned once
    when the operator is built."""
    from crocodile_amd.autograd import WProjectionOperator, degrid_image
    c = W.SPEC['diff1024']
    image, uvw = W.diff1024()
    N = c['N']
    lam = N / c['theta']
    sup = lambda w: W.wproj_support(w, c['wmax'], 16, 32)
    t0 = time.perf_counter()
    op = WProjectionOperator(c['theta'], lam, uvw, 20.0, 64, sup, c['Q'])
    torch.cuda.synchronize()
    setup_ms = 1e3 * (time.perf_counter() - t0)
    img_true = torch.as_tensor(image, device=dev)
    v_obs = op.apply(img_true)
    x = torch.zeros((N, N), dtype=torch.float32, device=dev, requires_grad=True)

    def step():
        x.grad = None
        v = degrid_image(x, op)
        loss = (v - v_obs).abs().pow(2).sum()
        loss.backward()
        return _lib.last_kernel_ms()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    gms, dms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g, d = step()
        gms.append(g)
        dms.append(d)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nv = op.nv
    line = {
        'metric': 'visibilities/s grid+degrid', 'value': round(nv * args.steps / dt, 1),
        'unit': 'visibilities/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(1e3 * dt / args.steps, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'diff1024', 'N': N, 'nv': nv, 'w_planes': op.bank.nplanes,
                   'support': [16, 32], 'oversampling': c['Q'],
                   'step': 'loss and gradient: degrid forward, adjoint grid backward',
                   'operator_setup_ms': round(setup_ms, 1)},
        'kernel_ms': {'k_grid_tiles': round(sum(gms) / len(gms), 4),
                      'k_degrid_tiles': round(sum(dms) / len(dms), 4)},
    }
    print(json.dumps(line), flush=True)


def run_stream(args, torch, W, _lib, dev):
    """Subgrid-streaming degrid (examples/grid/streamer.c:366-600) on the
    stream32k workload: 130816 baselines x 128 dumps x one 64-channel
    frequency chunk (1.07e9 visibilities, uvw in metres, frequencies in Hz:
    the streamer's bl_data layout) against the subgrids of the 32k-8k-1k
    recombination (xM = 1024 spaced xA = 832, iu >= 0 half of the uv plane).
    One step = every subgrid of the work assignment degridded into queue
    slots (separable PSWF kernel S = 8, Q = 1024), so each visibility is
    degridded exactly once (directly or as the conjugate of its mirror
    image): all subgrids in one croc_streamer_degrid_batch launch (the
    headline), and the reference-shaped one croc_streamer_degrid call per
    subgrid (reported beside it).  The subgrids are random complex64 arrays
    resident in HBM (the facet recombination that would produce them is not
    part of this path); chunk lists are selected before timing by
    streamer_degrid_chunk's bounding-box test.  The queue writer (HDF5
    output) is not timed."""
    import numpy
    from crocodile_amd import bl as BL
    c = W.SPEC['stream32k']
    S, Q, tc, fc = c['S'], c['Q'], c['time_chunk'], c['freq_chunk']
    theta, xA, xM = c['theta'], c['xA'], c['xM']
    print('[bench] building workload stream32k', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    uvw_h, freq_h = W.stream()
    vis = BL.VisData.uniform(uvw_h, freq_h, device=dev)
    kern_h = W.stream_kernel()
    kern = BL.SepKernelData(kern_h)
    st = BL.Streamer(vis, theta, xA, xM, tc, fc, kern)
    umax = float(numpy.abs(uvw_h[..., :2]).max() * freq_h.max() / W.C_LIGHT)
    offs = [o for o in W.stream_subgrids(umax=umax) if len(st.chunks_for(*o))]
    wins, ch6 = st.batch(offs)
    nch = ch6.shape[0]
    per_sub = numpy.bincount(ch6[:, 0].cpu().numpy(), minlength=len(offs))
    first = numpy.concatenate([[0], numpy.cumsum(per_sub)])
    ch5 = ch6[:, 1:].contiguous()
    gen = torch.Generator(device=dev)
    gen.manual_seed(c['seed'])
    subs = torch.randn((len(offs), xM, xM), dtype=torch.complex64, device=dev, generator=gen)
    queue = torch.empty((nch, tc * fc), dtype=torch.complex64, device=dev)
    nd = torch.empty(nch, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    copy_probe(torch, dev, args.copy_probe_mib)
    per_vis = 4 * (1 + S) * S      # grid.c:222: flops per degridded visibility
    nvis_total = int(vis.vis.shape[0])

    def step_batch():
        BL.streamer_degrid_batch(vis, subs, theta, wins, ch6, kern, tc, fc, vis_queue=queue,
                                 ndegrid=nd, want_flops=False)
        return _lib.last_tap_kernel_ms()[2]

    def step_single():
        for i in range(len(offs)):
            a, b = int(first[i]), int(first[i + 1])
            w = wins[i]
            BL.streamer_degrid_worker(vis, subs[i], xM, theta, w[0], w[1], ch5[a:b], w[2], w[3],
                                      w[4], w[5], kern, tc, fc, vis_queue=queue[a:b],
                                      ndegrid=nd[a:b], want_flops=False)

    print('[bench] stream32k: %d subgrids, %d chunk slots, warm-up' % (len(offs), nch),
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step_batch()
        step_single()
    torch.cuda.synchronize()
    kms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        kms += step_batch()
        print('[bench] step %d done' % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n_step = int(nd.sum())
    t1 = time.perf_counter()
    for i in range(args.steps):
        step_single()
    torch.cuda.synchronize()
    dt_single = time.perf_counter() - t1
    assert int(nd.sum()) == n_step
    kernel_ms = kms / args.steps
    achieved = n_step * per_vis / (kernel_ms * 1e-3) / 1e12
    line = {
        'metric': 'visibilities/s degridded from streamed subgrids',
        'value': round(n_step * args.steps / dt, 1), 'unit': 'visibilities/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(1e3 * dt / args.steps, 3), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 (f64 coordinates)',
        'data': 'synthetic (random subgrids, station tracks)',
        'config': {'workload': 'stream32k', 'image_size': c['image_size'],
                   'theta': round(theta, 6), 'xA': xA, 'xM': xM, 'baselines': vis.bl_count,
                   'dumps': c['ndump'], 'channels': c['nchan'], 'time_chunk': tc,
                   'freq_chunk': fc, 'support': S, 'oversampling': Q,
                   'subgrids': len(offs), 'chunk_slots': nch, 'visibilities': nvis_total,
                   'degridded_per_step': n_step,
                   # chunks are selected by the bounding box of their corner
                   # dumps and channels (streamer.c:520-535): a visibility of
                   # a curved track outside that box is missed, as there
                   'degridded_fraction': round(n_step / nvis_total, 6),
                   'step': 'all subgrids in one croc_streamer_degrid_batch launch',
                   'setup_s': round(setup_s, 1)},
        'per_subgrid_calls': {
            'entry': 'croc_streamer_degrid, one call per subgrid (streamer_degrid_worker)',
            'ms_per_step': round(1e3 * dt_single / args.steps, 3),
            'visibilities_per_s': round(n_step * args.steps / dt_single, 1)},
        'roofline': {'bound': 'valu', 'achieved': round(achieved, 3),
                     'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': round(achieved / FP32_PEAK_TFLOPS, 5), 'kernel': 'k_bl_degrid',
                     'kernel_ms_per_step': round(kernel_ms, 3),
                     'flop_per_visibility': per_vis,
                     # queue slots written once (zeros included), subgrids read
                     # once, each chunk's uvw rows and channel frequencies
                     'algorithmic_bytes_per_launch': int(8 * nch * tc * fc + 8 * len(offs) * xM * xM
                                                         + nch * (24 * tc + 8 * fc)),
                     'traffic': None,
                     'note': 'algorithmic flops of degrid_conv_uv (grid.c:222, 4 (1 + S) S '
                             'per visibility) / k_bl_degrid launch time (HIP events)'},
    }
    pmc = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                       'k_bl_degrid_stream32k_pmc.json')
    if os.path.exists(pmc):
        prof = json.load(open(pmc))
        r = line['roofline']
        r['traffic'] = prof.get('hbm_bytes_per_launch')
        r['traffic_over_algorithmic'] = round(r['traffic'] / r['algorithmic_bytes_per_launch'], 3)
        r['traffic_profile'] = {'file': 'profiles/k_bl_degrid_stream32k_pmc.json',
                                'kernel_us': prof.get('kernel_us'),
                                'note': 'calibrated FETCH_SIZE + WRITE_SIZE of the batched launch '
                                        'from rocprofv3 --pmc of the same command'}
    if args.cpu_sample >= 0:
        from oracle import blgrid as OB
        # the busiest subgrid's chunks, a prefix of them: about 5-10 s of one core
        i = int(numpy.argmax(per_sub))
        a = int(first[i])
        ncs = min(int(per_sub[i]), args.cpu_sample if args.cpu_sample > 0 else 32000)
        ch_h = ch5[a:a + ncs].cpu().numpy()
        bls = numpy.unique(ch_h[:, 0])
        hs = OB.BlSet.uniform(uvw_h[bls], freq_h, with_vis=False)
        chs = ch_h.copy()
        chs[:, 0] = numpy.searchsorted(bls, ch_h[:, 0])
        sub_h = subs[i].cpu().numpy().astype(numpy.complex128)
        t1 = time.perf_counter()
        ref, nd_ref = OB.streamer_degrid(sub_h, theta, *wins[i], hs, chs, tc, fc, kern_h)
        cpu_s = time.perf_counter() - t1
        q = queue[a:a + ncs].cpu().numpy()
        zero = ref == 0
        d = q - ref
        line['cpu_baseline'] = {
            'value': round(float(nd_ref.sum()) / cpu_s, 1), 'unit': 'visibilities/s',
            'cores': 1, 'kind': 'port',
            'sample': '%d chunk slots (%d degridded visibilities) of the busiest subgrid, '
                      'oracle/blgrid.c streamer_degrid (FP64 C restatement of '
                      'streamer.c:366-486), %.1f s' % (ncs, int(nd_ref.sum()), cpu_s)}
        line['parity'] = {
            'vs': 'oracle/blgrid.c on the CPU-baseline sample (slots of the timed run)',
            'selection_exact': bool((nd[a:a + ncs].cpu().numpy() == nd_ref).all() and
                                    ((q == 0) == zero).all()),
            'rms_rel': float(numpy.linalg.norm(d) / numpy.linalg.norm(ref)),
            'max_rel': float(numpy.abs(d).max() / numpy.abs(ref).max())}
    print(json.dumps(line), flush=True)


def run_imaging_extra(args, torch, G, W, _lib, dev):
    """The notebook imaging drivers outside the w-towers headline, timed on
    inputs resident in HBM: full-grid w-stacking (wstack2048,
    wstacking.ipynb cell 10) and AW-towers with per-antenna A-screens
    (awtowers1800, awtowers.ipynb cells 17-22).  One step is one imaging
    call.  The CPU baseline is the FP64 oracle on a seeded sample, and the
    device result on that sample is compared with it."""
    import numpy
    name = args.workload
    if name == 'wstack2048':
        c, uvw, vis = W.wstack2048()
        theta, lam = c['theta'], c['lam']
        uvw_d = torch.as_tensor(uvw, device=dev)
        vis_d = torch.as_tensor(vis, device=dev)

        def step():
            return G.w_stacking_imaging(theta, lam, uvw_d, None, vis_d, c['wstep'])
        nv = uvw.shape[0]
        planes = len(numpy.unique(numpy.round(uvw[:, 2] / c['wstep'])))
        config = {'workload': name, 'N': c['N'], 'theta': theta, 'nv': nv, 'w_planes': planes,
                  'wstep': c['wstep'], 'step': 'w_stacking_imaging (croc_wstack_grid, batched '
                  'rocFFT, croc_wstack_add)'}
    else:
        c, N, uvw, src, vis, k = W.awtowers1800()
        theta, lam = c['theta'], c['lam']
        uvw_d = torch.as_tensor(uvw, device=dev)
        src_d = torch.as_tensor(src, device=dev)
        vis_d = torch.as_tensor(vis, device=dev)
        kd = {n: torch.as_tensor(numpy.asarray(k[n]), device=dev) for n in ('aa_gcf', 'aimage',
                                                                           'wkern')}

        def step():
            return G.aw_towers_imaging(theta, lam, uvw_d, src_d, vis_d, c['wstep'],
                                       c['subgrid'], c['margin'], kd['aa_gcf'], k['aa_sub'],
                                       k['aa_sub2'], kd['aimage'], kd['wkern'])
        nv = uvw.shape[0]
        config = {'workload': name, 'N': N, 'theta': theta, 'nv': nv, 'antennas': c['nant'],
                  'dumps': c['ndump'], 'subgrid': c['subgrid'], 'margin': c['margin'],
                  'wstep': c['wstep'], 'aa_support': c['aa_support'], 'aa_over': c['aa_over'],
                  'step': 'aw_towers_imaging (croc_awt_pregrid, batched rocFFT, '
                          'croc_awt_apply, rocFFT, croc_awt_place)'}
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print('[bench] timing %d %s steps' % (args.steps, name), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps

    cpu, parity = None, None
    n_cpu = args.cpu_sample or (20000 if name == 'wstack2048' else 1500)
    if n_cpu > 0:
        from oracle import wstacking_ref as WS, awtowers_ref as AWT
        print('[bench] cpu baseline: FP64 oracle on %d visibilities' % n_cpu, file=sys.stderr,
              flush=True)
        if name == 'wstack2048':
            c_s, uvw_s, vis_s = W.wstack2048(nv=n_cpu, seed=c['seed'] + 1)
            t0 = time.perf_counter()
            ref = WS.w_stacking_imaging(theta, lam, uvw_s, None, vis_s, c['wstep'])
            t_cpu = time.perf_counter() - t0
            got = G.w_stacking_imaging(theta, lam, uvw_s, None, vis_s, c['wstep'])
            what = ('%d visibilities of wstack2048 (seed %d) on the full 2048^2 grid: the '
                    'planes\' FFTs dominate, so the rate does not extrapolate linearly'
                    % (n_cpu, c['seed'] + 1))
        else:
            _, _, uvw_s, src_s, vis_s, ks = W.awtowers1800(nsample=n_cpu)
            args_s = (theta, lam, uvw_s, src_s, vis_s, c['wstep'], c['subgrid'], c['margin'],
                      numpy.asarray(ks['aa_gcf']), ks['aa_sub'], ks['aa_sub2'], ks['aimage'],
                      ks['wkern'])
            t0 = time.perf_counter()
            ref = AWT.aw_towers_imaging(*args_s)
            t_cpu = time.perf_counter() - t0
            got = G.aw_towers_imaging(*args_s)
            what = ('%d visibilities of awtowers1800 (a seeded subset): one (2S)^2 transform '
                    'per (chunk, plane, antenna pair) group' % n_cpu)
        cpu = {'value': n_cpu / t_cpu, 'unit': 'visibilities/s', 'cores': 1, 'kind': 'port',
               'sample': what + '; %.2f s' % t_cpu}
        d = numpy.abs(got - ref)
        parity = {'sample_nv': n_cpu,
                  'max_rel': float(d.max() / numpy.abs(ref).max()),
                  'rms_rel': float(numpy.sqrt((d ** 2).mean() / (numpy.abs(ref) ** 2).mean()))}

    line = {
        'metric': 'visibilities/s imaging', 'value': round(nv / dt, 1),
        'unit': 'visibilities/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(1e3 * dt, 3), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic', 'config': config,
        'roofline': None,
        'roofline_note': 'a notebook driver beside the towers8192 headline, reported for '
                         'throughput and parity; its transforms are rocFFT\'s',
        'cpu_baseline': cpu, 'parity_vs_fp64_oracle': parity,
    }
    print(json.dumps(line), flush=True)


def copy_probe(torch, dev, mib):
    """One device copy of ``mib`` MiB (a blit kernel of known traffic, to
    scale the HBM counters of a profiled run); nothing when mib <= 0."""
    if mib <= 0:
        return
    src = torch.full((mib << 18,), 1.0, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def measured_copy_gbs(torch, dev, mib=1024, reps=5):
    """HBM bandwidth re-measured on this device: a 1 GiB device-to-device
    copy (torch copy_), read + write bytes over the mean of ``reps`` copies
    timed with events; reported beside the 8 TB/s spec peak."""
    src = torch.ones(mib << 18, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return round(2.0 * (mib << 20) / (ms * 1e-3) / 1e9, 1)


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path '
              '(--gpus %d requested); run with --gpus 1' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from crocodile_amd import _lib, workloads as W, synthesis as G

    dev = _lib.device()
    _lib.set_tuning(args.tile, args.chunk, args.group_kib)
    if args.workload.startswith('towers'):
        return run_towers(args, torch, G, W, _lib, dev)
    if args.workload == 'dft128':
        return run_dft(args, torch, _lib, dev)
    if args.workload == 'stream32k':
        return run_stream(args, torch, W, _lib, dev)
    if args.workload == 'diff1024':
        return run_diff(args, torch, G, W, _lib, dev)
    if args.workload in ('wstack2048', 'awtowers1800'):
        return run_imaging_extra(args, torch, G, W, _lib, dev)
    print('[bench] building workload %s' % args.workload, file=sys.stderr, flush=True)
    if args.workload == 'wproj2048':
        wl = build_wproj(G, W, torch, dev)
        c = wl['c']
        N, lam = c['N'], c['lam']
        bank_img, bank_pred = wl['bank_img'], wl['bank_pred']
        mode, wstep, wbin0 = _lib.PLANE_WBIN, c['wstep'], wl['lo']
        config = {'workload': 'wproj2048', 'N': N, 'theta': c['theta'], 'lam': lam,
                  'nv': c['nv'], 'w_planes': len(wl['ws']), 'wstep': wstep,
                  'support': [min(wl['S']), max(wl['S'])], 'oversampling': c['Q'],
                  'npix_ff': c['NpixFF'], 'kernel_gen_ms': round(wl['kgen_ms'], 1)}
        uvw, vis = wl['uvw'], wl['vis']
        kernel_bytes = wl['kernel_bytes']
    else:
        c = W.SPEC['plain512']
        uvw, vis = W.plain512()
        N, lam = c['N'], c['lam']
        aaf = G.anti_aliasing_function(N, 0, c['S'])
        kv = G.kernel_oversample(aaf, c['Q'], c['S'])
        kv = kv / kv[0].sum()
        bank_img = bank_pred = G.KernelBank.from_tables([kv], separable=True)
        mode, wstep, wbin0 = _lib.PLANE_SINGLE, 1.0, 0
        config = {'workload': 'plain512', 'N': N, 'nv': c['nv'], 'support': c['S'],
                  'oversampling': c['Q'], 'kernel': 'separable PSWF'}
        kernel_bytes = 8 * int(bank_img.kern.numel())
    nv = uvw.shape[0]
    uvw_d = torch.as_tensor(uvw, device=dev).contiguous()
    vis_d = torch.as_tensor(vis, device=dev).to(torch.complex64).contiguous()
    guv = torch.zeros((N, N), dtype=torch.complex64, device=dev)
    out = torch.empty(nv, dtype=torch.complex64, device=dev)

    copy_gbs = measured_copy_gbs(torch, dev)
    copy_probe(torch, dev, args.copy_probe_mib)

    # With --autotune the warm-up also measures a few tile / chunk settings on
    # this workload and keeps the fastest (results do not depend on them).
    if args.autotune:
        from crocodile_amd import tuning
        choice, _ = tuning.autotune(bank_img, N, uvw_d, lam, vis_d, mode, None, wstep, wbin0,
                                    repeats=3, tiles=(0, 48), chunks=(512, 1024, 2048))
        config['tuning'] = {'tile': choice[0], 'chunk': choice[1], 'group_kib': choice[2]}
    else:
        config['tuning'] = {'tile': args.tile, 'chunk': args.chunk, 'group_kib': args.group_kib}

    # One step: bin the coordinates once (croc_plan_bin, which validates
    # them), grid into a cleared grid, degrid it (croc_convgrid_plan /
    # croc_convdegrid_plan).  The one-shot calls, which bin on every call,
    # are timed the same way for comparison.
    plan = G.GridPlan(bank_img, N)

    def step():
        plan.bin(uvw_d, lam, mode, None, wstep, wbin0)
        guv.zero_()
        plan.grid(guv, vis_d)
        plan.degrid(guv, out, bank_pred)
        return _lib.last_kernel_ms()

    def step_one_shot():
        guv.zero_()
        G.grid_device(bank_img, guv, uvw_d, lam, vis_d, mode, None, wstep, wbin0)
        G.degrid_device(bank_pred, guv, uvw_d, lam, out, mode, None, wstep, wbin0)
        return _lib.last_kernel_ms()

    def timed(fn, label):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        print('[bench] timing %d %s steps' % (args.steps, label), file=sys.stderr, flush=True)
        gms, dms = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g, d = fn()
            gms.append(g)
            dms.append(d)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return t1 - t0, gms, dms

    dt1, _, _ = timed(step_one_shot, 'one-shot')
    dt, gms, dms = timed(step, 'planned')
    ms_per_step = 1e3 * dt / args.steps
    value = nv * args.steps / dt

    g_avg = sum(gms) / len(gms)
    d_avg = sum(dms) / len(dms)
    # Algorithmic work per launch (implementation independent): 8 FP32 flop
    # per kernel tap (one complex multiply-add), and the compulsory bytes --
    # uvw (24 B) + visibility (8 B) per visibility, the grid written or read
    # once (8 B per cell) and the kernel bank read once.
    taps = tap_count(config, uvw)
    flops = 8.0 * taps
    alg_bytes = nv * (24 + 8) + N * N * 8 + kernel_bytes
    name, t_ms = ('k_grid_tiles', g_avg) if g_avg >= d_avg else ('k_degrid_tiles', d_avg)
    t_flop, t_byte = flops / FP32_PEAK_TFLOPS / 1e12, alg_bytes / HBM_PEAK_GBS / 1e9
    if t_flop >= t_byte:
        achieved = flops / (t_ms * 1e-3) / 1e12
        roofline = {'bound': 'valu', 'achieved': round(achieved, 3), 'peak': FP32_PEAK_TFLOPS,
                    'unit': 'TFLOP/s', 'frac': round(achieved / FP32_PEAK_TFLOPS, 5),
                    'bound_note': 'complex multiply-add per kernel tap on the vector ALUs, 8 '
                                  'FP32 flop/tap (the reference convention); no matrix cores '
                                  'on this path'}
    else:
        achieved = alg_bytes / (t_ms * 1e-3) / 1e9
        roofline = {'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS,
                    'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 5)}
    traffic = None
    pmc = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                       '%s_%s_pmc.json' % (name, config['workload']))
    if os.path.exists(pmc):
        traffic = json.load(open(pmc)).get('hbm_bytes_per_launch')
    # LDS side (supplementary): gridding adds 16 B per tap (two 64-bit
    # integer atomics), degridding reads 8 B per tap; peak 128 B/clk/CU.
    lds_bytes = (16.0 if name == 'k_grid_tiles' else 8.0) * taps
    # traffic_frac: the profiled bytes beyond L2 (kernel-table taps missing an
    # XCD's L2) per launch at this launch duration, against the HBM peak --
    # the bound the kernel actually runs into.
    roofline.update({'kernel': name, 'traffic': traffic, 'algorithmic_bytes': alg_bytes,
                     'algorithmic_flops': flops,
                     'traffic_frac': (round(traffic / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                      if traffic else None),
                     'hbm_frac': round(alg_bytes / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                     'lds_bytes': lds_bytes,
                     'lds_frac': round(lds_bytes / (t_ms * 1e-3) / (LDS_PEAK_TBS * 1e12), 5),
                     'kernel_ms': {'k_grid_tiles': round(g_avg, 4),
                                   'k_degrid_tiles': round(d_avg, 4)},
                     'hbm_copy_measured_gbs': copy_gbs})

    cpu, parity = None, None
    n_cpu = args.cpu_sample or 300_000
    if n_cpu > 0 and config['workload'] == 'wproj2048':
        print('[bench] cpu baseline on %d visibilities' % n_cpu, file=sys.stderr,
              flush=True)
        cpu, parity = cpu_baseline(G, W, torch, dev, n_cpu)

    line = {
        'metric': 'visibilities/s grid+degrid', 'value': round(value, 1),
        'unit': 'visibilities/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic', 'config': config, 'roofline': roofline,
        'cpu_baseline': cpu, 'parity_vs_fp64_oracle': parity,
        'one_shot': {'value': round(nv * args.steps / dt1, 1),
                     'ms_per_step': round(1e3 * dt1 / args.steps, 4),
                     'note': 'croc_convgrid + croc_convdegrid, each binning the coordinates'},
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
