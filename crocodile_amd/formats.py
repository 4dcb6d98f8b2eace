"""crocodile's on-disk formats, read and written without h5py (through
crocodile_amd.h5, the classic-HDF5 subset those files use):

* visibility files (examples/grid/hdf5.c:199-392 load_vis, :395-478
  create_vis_group): group ``vis`` holding ``<a1>/<a2>`` baseline groups of
  ``frequency`` [nf] (Hz), ``time`` [nt], ``uvw`` [nt, 3] (metres) and
  ``vis`` [nt, nf, 1] complex, or the legacy flat form (a ``vis`` group with
  ``type`` = 'Visibility' and one visibility per row: ``antenna1``,
  ``antenna2``, ``time``, ``uvw``, ``vis`` [n, nf, 1]);
* w-kernel files (hdf5.c:594-700 load_wkern, scripts/generate_wkernels.py):
  ``wkern/<theta>/<w>/kern`` [Q, Q, S, S] complex, planes indexed by w;
* separable kernels ``sepkern/kern`` [Q, S] real (hdf5.c:540-592);
* A-kernels ``akern/<theta>/<antenna>/<time>/<freq>/kern`` [S, S] complex
  (hdf5.c:712-875 load_akern), indexed by antenna, time and frequency.

The loaders return this package's host/device structures
(:class:`crocodile_amd.bl.VisData`, ``WKernelData``, ``SepKernelData``), so
a file goes straight into the native gridding / degridding entry points.
"""

import math

import numpy

from . import h5


def _name(x):
    """A float as the group names use it: Python's str (generate_wkernels.py:
    55) and C's %g (hdf5.c:604) agree on these values."""
    return str(float(x))


def _theta_group(f, kind, theta):
    for cand in (_name(theta), '%g' % theta, str(theta)):
        path = '%s/%s' % (kind, cand)
        if path in f:
            return f[path], path
    raise KeyError("no '%s/%g' group in the file" % (kind, theta))


# ---------------------------------------------------------------------------
# visibilities
# ---------------------------------------------------------------------------

def _baseline(g, a1, a2, min_len, max_len):
    """BlData of one baseline group, or None when hdf5.c:91-139 skips it
    (wrong shapes, or its first uvw's length outside [min_len, max_len))."""
    from .bl import BlData
    try:
        freq, time = g['frequency'][()], g['time'][()]
        uvw, vis = g['uvw'][()], g['vis'][()]
    except KeyError:
        return None
    if not (freq.ndim == 1 and time.ndim == 1 and uvw.shape == (time.shape[0], 3) and
            vis.ndim == 3 and vis.shape == (time.shape[0], freq.shape[0], 1) and
            numpy.iscomplexobj(vis)):
        return None
    length = math.hypot(uvw[0, 0], uvw[0, 1]) if uvw.shape[0] else 0.0
    if length < min_len or length >= max_len:
        return None
    return BlData(a1, a2, time, freq, uvw, vis.reshape(-1))


def read_baselines(filename, min_len=-math.inf, max_len=math.inf):
    """(baselines, antenna count) of a visibility file (hdf5.c:199-392)."""
    from .bl import BlData
    f = h5.File(filename)
    if 'vis' not in f:
        raise KeyError("no 'vis' group in visibility file %s" % filename)
    g = f['vis']
    if isinstance(g, h5.Group) and g.attrs.get('type') == 'Visibility':
        # legacy flat form: one baseline per visibility row (hdf5.c:224-309)
        whole = _baseline(g, 0, 0, -math.inf, math.inf)
        if whole is None:
            raise ValueError('flat visibility group of %s has the wrong shapes' % filename)
        a1, a2 = g['antenna1'][()], g['antenna2'][()]
        if a1.shape != (whole.time_count,) or a2.shape != (whole.time_count,):
            raise ValueError('antenna datasets of %s do not match the rows' % filename)
        nf = whole.freq_count
        bls, nant = [], 0
        for i in range(whole.time_count):
            u = whole.uvw_m[i]
            length = math.hypot(u[0], u[1])
            if length < min_len or length >= max_len:
                continue
            bls.append(BlData(a1[i], a2[i], whole.time[i:i + 1], whole.freq, u[None, :],
                              whole.vis[i * nf:(i + 1) * nf]))
            nant = max(nant, int(a1[i]), int(a2[i]))
        return bls, nant
    # baseline groups vis/<a1>/<a2>, antennas 0 .. (number of a1 groups)
    nant = len(g) + 1
    bls = []
    for a1 in range(nant - 1):
        if str(a1) not in g:
            continue
        g1 = g[str(a1)]
        for a2 in range(a1 + 1, nant):
            if str(a2) not in g1:
                continue
            bl = _baseline(g1[str(a2)], a1, a2, min_len, max_len)
            if bl is not None:
                bls.append(bl)
    return bls, nant


def load_vis(filename, min_len=-math.inf, max_len=math.inf, device=None):
    """Visibility file -> :class:`crocodile_amd.bl.VisData` on the device
    (hdf5.c:199 load_vis: baselines whose first uvw is shorter than
    ``min_len`` or at least ``max_len`` metres are skipped)."""
    from .bl import VisData
    bls, nant = read_baselines(filename, min_len, max_len)
    return VisData(bls, nant, device=device)


def write_vis(filename, baselines, extra=None, chunks=None):
    """Write baselines (BlData) as a visibility file of baseline groups
    ``vis/<a1>/<a2>`` (the form hdf5.c:395-478 creates and :330-362 reads).
    ``chunks`` (time_chunk, freq_chunk): store each ``vis`` dataset in chunks
    of [time_chunk, freq_chunk, 1], as hdf5.c:416-466 does.  ``extra``: more
    {path: array} entries for the same file."""
    tree, chunked = {}, {}
    for b in baselines:
        base = 'vis/%d/%d/' % (b.antenna1, b.antenna2)
        tree[base + 'frequency'] = numpy.asarray(b.freq, dtype=numpy.float64)
        tree[base + 'time'] = numpy.asarray(b.time, dtype=numpy.float64)
        tree[base + 'uvw'] = numpy.asarray(b.uvw_m, dtype=numpy.float64)
        tree[base + 'vis'] = numpy.asarray(b.vis, dtype=numpy.complex128).reshape(
            b.time_count, b.freq_count, 1)
        if chunks is not None:
            chunked['/' + base + 'vis'] = (int(chunks[0]), int(chunks[1]), 1)
    tree.update(extra or {})
    h5.write(filename, tree, chunks=chunked)


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------

def read_wkernels(filename, theta):
    """{w: kernel [Q, Q, S, S]} of ``wkern/<theta>`` (hdf5.c:594-660:
    kernels of another shape than the first are ignored)."""
    f = h5.File(filename)
    g, _ = _theta_group(f, 'wkern', theta)
    out, shape = {}, None
    for name in g.keys():
        if name + '/kern' not in g:
            continue
        k = g[name + '/kern'][()]
        if k.ndim != 4 or not numpy.iscomplexobj(k):
            continue
        shape = shape or k.shape
        if k.shape != shape or shape[0] != shape[1]:
            continue
        out[float(name)] = k
    if not out:
        raise ValueError("no w-kernels in %s" % filename)
    return out


def load_wkern(filename, theta, device=None):
    """w-kernel file -> :class:`crocodile_amd.bl.WKernelData` (hdf5.c:594-700):
    planes at w_min + k w_step, w_step = (w_max - w_min) / (planes - 1), each
    taking the file's kernel closest to its w."""
    from .bl import WKernelData
    ks = read_wkernels(filename, theta)
    ws = numpy.array(sorted(ks))
    n = len(ws)
    w_min, w_max = ws[0], ws[-1]
    w_step = (w_max - w_min) / (n - 1) if n > 1 else 0.0
    planes = []
    for i in range(n):
        w = w_min + i * w_step
        planes.append(ks[ws[numpy.argmin(numpy.abs(ws - w))]])
    return WKernelData(numpy.stack(planes), w_min, w_step if n > 1 else 1.0, device=device)


def load_sepkern(filename, device=None):
    """``sepkern/kern`` [Q, S] real -> :class:`crocodile_amd.bl.SepKernelData`
    (hdf5.c:540-592)."""
    from .bl import SepKernelData
    k = h5.File(filename)['sepkern/kern'][()]
    if k.ndim != 2 or numpy.iscomplexobj(k) or k.dtype.itemsize != 8:
        raise ValueError("'sepkern/kern' has the wrong format in %s" % filename)
    return SepKernelData(k, device=device)


class AKernelData:
    """A-kernels indexed by antenna, time and frequency (grid.h:125-140,
    hdf5.c:712-875): ``data`` [antennas, times, freqs, S, S] complex128,
    slot (a, t, f) holding the kernel closest to t_min + t t_step,
    f_min + f f_step."""

    def __init__(self, data, antennas, t_min, t_step, f_min, f_step):
        self.data = data
        self.antennas = antennas
        self.antenna_count, self.time_count, self.freq_count = data.shape[:3]
        self.size_y, self.size_x = data.shape[3:]
        self.t_min, self.t_step, self.f_min, self.f_step = t_min, t_step, f_min, f_step

    def kernel(self, antenna, t, f):
        return self.data[antenna, t, f]


def load_akern(filename, theta):
    """A-kernel file ``akern/<theta>/<antenna>/<time>/<freq>/kern`` ->
    :class:`AKernelData` (hdf5.c:712-875)."""
    f = h5.File(filename)
    g, _ = _theta_group(f, 'akern', theta)
    ants = g.keys()
    if not ants:
        raise ValueError('no A-kernels in %s' % filename)
    first_t = g[ants[0]].keys()
    nt, nf = len(first_t), len(g[ants[0]][first_t[0]].keys())
    ks = {}
    for a in ants:
        ga = g[a]
        if len(ga) != nt:
            continue
        for t in ga.keys():
            gt = ga[t]
            if len(gt) != nf:
                continue
            for fr in gt.keys():
                k = gt[fr + '/kern'][()]
                if k.ndim == 2 and numpy.iscomplexobj(k):
                    ks[(int(a), float(t), float(fr))] = k
    if not ks:
        raise ValueError('no A-kernels in %s' % filename)
    ts = sorted({k[1] for k in ks})
    fs = sorted({k[2] for k in ks})
    t_min, f_min = ts[0], fs[0]
    t_step = (ts[-1] - ts[0]) / (nt - 1) if nt > 1 else 0.0
    f_step = (fs[-1] - fs[0]) / (nf - 1) if nf > 1 else 0.0
    antennas = sorted({k[0] for k in ks})
    shape = next(iter(ks.values())).shape
    data = numpy.zeros((len(antennas), nt, nf) + shape, dtype=numpy.complex128)
    for ia, a in enumerate(antennas):
        mine = [k for k in ks if k[0] == a]
        for it in range(nt):
            for jf in range(nf):
                t, fr = t_min + it * t_step, f_min + jf * f_step
                best = min(mine, key=lambda k: abs(k[1] - t) + abs(k[2] - fr))
                data[ia, it, jf] = ks[best]
    return AKernelData(data, antennas, t_min, t_step, f_min, f_step)


def write_kernels(filename, wkern=None, sepkern=None, akern=None):
    """Write a kernel file: ``wkern`` {(theta, w): [Q, Q, S, S]},
    ``sepkern`` [Q, S] real, ``akern`` {(theta, antenna, time, freq): [S, S]}
    (the layouts hdf5.c reads and generate_wkernels.py writes)."""
    tree = {}
    for (theta, w), k in (wkern or {}).items():
        tree['wkern/%s/%s/kern' % (_name(theta), _name(w))] = numpy.asarray(k, numpy.complex128)
    if sepkern is not None:
        tree['sepkern/kern'] = numpy.asarray(sepkern, numpy.float64)
    for (theta, a, t, fr), k in (akern or {}).items():
        tree['akern/%s/%d/%s/%s/kern' % (_name(theta), a, _name(t), _name(fr))] = \
            numpy.asarray(k, numpy.complex128)
    h5.write(filename, tree)


class H5KernelFile:
    """A ``kernel_cache`` / ``kernel_fn`` over a kernel file, as
    scripts/image_dataset.py:117-129 uses an h5py file behind pylru:
    ``fn(theta, w)`` returns the file's ``wkern/<theta>/<w'>/kern`` for the
    w' closest to w (read once and kept); ``wsteps(theta)`` lists the
    file's w values and ``wstep(theta)`` their spacing (image_dataset.py:
    117-119).  Pass it as w_cache_imaging's ``kernel_cache``."""

    def __init__(self, filename):
        self.f = h5.File(filename)
        self._kept, self._ws = {}, {}

    def wsteps(self, theta):
        theta = float(theta)
        if theta not in self._ws:
            g, path = _theta_group(self.f, 'wkern', theta)
            names = {float(n): n for n in g.keys()}
            self._ws[theta] = (numpy.array(sorted(names)), names, path)
        return self._ws[theta][0]

    def wstep(self, theta):
        ws = self.wsteps(theta)
        return float(ws[1] - ws[0])

    def __call__(self, theta, w, *src):
        ws = self.wsteps(theta)
        _, names, path = self._ws[float(theta)]
        kw = float(ws[numpy.argmin(numpy.abs(ws - float(w)))])
        key = (float(theta), kw)
        if key not in self._kept:
            self._kept[key] = self.f['%s/%s/kern' % (path, names[kw])][()]
        return self._kept[key]
