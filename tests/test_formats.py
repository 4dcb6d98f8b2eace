"""The classic-HDF5 reader / writer (crocodile_amd.h5) and the file formats
of crocodile (crocodile_amd.formats: examples/grid/hdf5.c,
scripts/generate_wkernels.py, scripts/image_dataset.py:117-129), on the
host.

The reader is checked against an HDF5 file that neither this package nor
its writer produced: scipy's test file testhdf5_7.4_GLNX86.mat (MATLAB 7.3
format, written by the HDF5 library in 2008: a 512-byte user block, a
version-0 superblock, a symbol-table root group, a version-2 layout
message, a fixed-length string attribute) holds ``testdouble``, whose
values scipy itself reads from the non-HDF5 MAT-file of the same test set.
"""
import glob
import math
import os

import numpy
import pytest
from numpy.testing import assert_array_equal

from crocodile_amd import formats, h5

SCIPY_DATA = None
try:
    import scipy.io
    SCIPY_DATA = os.path.join(os.path.dirname(scipy.io.__file__), 'matlab', 'tests', 'data')
except ImportError:       # pragma: no cover
    pass


@pytest.mark.skipif(SCIPY_DATA is None or not glob.glob(os.path.join(
    SCIPY_DATA or '', 'testhdf5_7.4_GLNX86.mat')), reason='scipy MAT-file test data absent')
def test_reads_an_hdf5_file_written_by_the_hdf5_library():
    f = h5.File(os.path.join(SCIPY_DATA, 'testhdf5_7.4_GLNX86.mat'))
    assert f.keys() == ['testdouble']
    ds = f['testdouble']
    assert ds.attrs['MATLAB_class'] == 'double'
    got = ds[()]
    ref = scipy.io.loadmat(os.path.join(SCIPY_DATA, 'testdouble_7.1_GLNX86.mat'))['testdouble']
    # MATLAB stores column-major: the HDF5 dataset is the transpose
    assert got.shape == ref.T.shape and got.dtype == numpy.float64
    assert_array_equal(got, ref.T)
    assert_array_equal(got[:, 0], numpy.arange(9) * numpy.pi / 4)


def test_round_trip_dtypes_shapes_and_attributes(tmp_path):
    rng = numpy.random.default_rng(0)
    tree = {
        'c128': rng.standard_normal((4, 4, 8, 8)) + 1j * rng.standard_normal((4, 4, 8, 8)),
        'c64': (rng.standard_normal(7) + 1j).astype(numpy.complex64),
        'f64': rng.standard_normal((5, 3)), 'f32': rng.standard_normal(6).astype(numpy.float32),
        'i64': numpy.arange(-5, 5, dtype=numpy.int64), 'u8': numpy.arange(7, dtype=numpy.uint8),
        'i32': numpy.arange(4, dtype=numpy.int32).reshape(2, 2), 'scalar': numpy.array(2.5),
        'empty': numpy.zeros((0, 3)), 'bytes': numpy.array([b'abc', b'de'], dtype='S3'),
        'deep/er/still': numpy.ones(2),
    }
    p = str(tmp_path / 'a.h5')
    h5.write(p, tree, attrs={'/': {'title': 'root'}, '/deep': {'n': numpy.int64(3)},
                             '/f64': {'unit': 'm', 'scale': numpy.array([1.0, 2.0])}})
    f = h5.File(p)
    for k, v in tree.items():
        got = f[k][()]
        assert got.dtype == v.dtype and got.shape == v.shape, k
        assert_array_equal(got, v)
    assert f.attrs['title'] == 'root' and f['deep'].attrs['n'] == 3
    assert f['f64'].attrs['unit'] == 'm'
    assert_array_equal(f['f64'].attrs['scale'], [1.0, 2.0])
    assert 'deep/er' in f and 'deep/nope' not in f and f['deep/er'].keys() == ['still']
    assert f['c128'].dtype == numpy.complex128 and f['c128'].shape == (4, 4, 8, 8)


@pytest.mark.parametrize('n', [0, 1, 8, 9, 300])
def test_groups_of_any_size(tmp_path, n):
    """Symbol table nodes hold 8 entries and B-tree nodes 32 children: 300
    members need a two-level tree; names come back in HDF5's (byte) order."""
    p = str(tmp_path / 'g.h5')
    tree = {'g/%d' % i: numpy.full(1, i) for i in range(n)}
    h5.write(p, tree if n else {'g': {}})
    g = h5.File(p)['g']
    assert g.keys() == sorted(str(i) for i in range(n))
    for i in range(n):
        assert g[str(i)][()][0] == i


def test_not_hdf5_and_missing(tmp_path):
    p = tmp_path / 'x.h5'
    p.write_bytes(b'\0' * 1024)
    with pytest.raises(ValueError):
        h5.File(str(p))
    q = str(tmp_path / 'y.h5')
    h5.write(q, {'a': numpy.ones(2)})
    with pytest.raises(KeyError):
        h5.File(q)['b']


def baselines(rng, nant=5, nt=4, nf=3):
    from crocodile_amd.bl import BlData
    out = []
    for a1 in range(nant):
        for a2 in range(a1 + 1, nant):
            uvw = rng.normal(0, 500.0, (nt, 3))
            vis = rng.standard_normal(nt * nf) + 1j * rng.standard_normal(nt * nf)
            out.append(BlData(a1, a2, numpy.arange(nt) * 10.0, 1e8 + numpy.arange(nf) * 1e6, uvw,
                              vis))
    return out


@pytest.mark.parametrize('shape,chunks', [((37, 11, 1), (8, 4, 1)), ((70, 5), (1, 5)),
                                          ((5,), (5,)), ((0, 3), (2, 3))])
def test_chunked_datasets(tmp_path, shape, chunks):
    """Chunked layout under version-1 chunk B-trees (70 chunks: two levels),
    edge chunks cut at the dataset's extent."""
    rng = numpy.random.default_rng(5)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p = str(tmp_path / 'c.h5')
    h5.write(p, {'d': a}, chunks={'/d': chunks})
    ds = h5.File(p)['d']
    assert ds.shape == shape
    assert_array_equal(ds[()], a)


def test_variable_length_string_attribute(tmp_path):
    p = str(tmp_path / 's.h5')
    h5.write(p, {'g/x': numpy.ones(1)}, attrs={'/g': {'type': h5.VlenStr('Visibility'),
                                                      'name': h5.VlenStr('LOW-\u00e9')}})
    a = h5.File(p)['g'].attrs
    assert a['type'] == 'Visibility' and a['name'] == 'LOW-\u00e9'


def test_visibility_file_round_trip_and_length_cut(tmp_path):
    """vis/<a1>/<a2> groups (hdf5.c:330-362): every baseline back, in
    (a1, a2) order; baselines whose first uvw is shorter than min_len or at
    least max_len are skipped (hdf5.c:124-127)."""
    rng = numpy.random.default_rng(1)
    bls = baselines(rng)
    p = str(tmp_path / 'vis.h5')
    formats.write_vis(p, bls)
    got, nant = formats.read_baselines(p)
    assert nant == 5 and len(got) == len(bls)
    q = str(tmp_path / 'vis_chunked.h5')          # vis in [2, 2, 1] chunks (hdf5.c:416-466)
    formats.write_vis(q, bls, chunks=(2, 2))
    for a, b in zip(formats.read_baselines(q)[0], bls):
        assert_array_equal(a.vis, b.vis)
    for a, b in zip(got, bls):
        assert (a.antenna1, a.antenna2) == (b.antenna1, b.antenna2)
        assert_array_equal(a.uvw_m, b.uvw_m)
        assert_array_equal(a.freq, b.freq)
        assert_array_equal(a.time, b.time)
        assert_array_equal(a.vis, b.vis)
    lens = numpy.array([math.hypot(*b.uvw_m[0, :2]) for b in bls])
    lo, hi = numpy.quantile(lens, [0.25, 0.75])
    cut, _ = formats.read_baselines(p, min_len=lo, max_len=hi)
    keep = [(b.antenna1, b.antenna2) for b, l in zip(bls, lens) if lo <= l < hi]
    assert [(b.antenna1, b.antenna2) for b in cut] == keep


def test_legacy_flat_visibility_group(tmp_path):
    """A 'vis' group of type Visibility: one single-dump baseline per row
    (hdf5.c:224-309); the antenna count is the largest antenna index."""
    rng = numpy.random.default_rng(2)
    n, nf = 6, 2
    uvw = rng.normal(0, 300.0, (n, 3))
    vis = rng.standard_normal((n, nf, 1)) + 1j * rng.standard_normal((n, nf, 1))
    a1 = numpy.array([0, 0, 1, 1, 2, 3], dtype=numpy.int64)
    a2 = numpy.array([1, 2, 2, 3, 3, 4], dtype=numpy.int64)
    p = str(tmp_path / 'flat.h5')
    h5.write(p, {'vis/frequency': numpy.array([1e8, 2e8]), 'vis/time': numpy.arange(n, dtype=float),
                 'vis/uvw': uvw, 'vis/vis': vis, 'vis/antenna1': a1, 'vis/antenna2': a2},
             attrs={'/vis': {'type': h5.VlenStr('Visibility')}})   # as h5py writes a str
    bls, nant = formats.read_baselines(p)
    assert nant == 4 and len(bls) == n
    for i, b in enumerate(bls):
        assert (b.antenna1, b.antenna2, b.time_count) == (a1[i], a2[i], 1)
        assert_array_equal(b.uvw_m[0], uvw[i])
        assert_array_equal(b.vis, vis[i, :, 0])


def test_kernel_files(tmp_path):
    """wkern/<theta>/<w>/kern as generate_wkernels.py names them, read by
    w (hdf5.c:594-660) and through the image_dataset.py-style cache (the
    file's closest w); sepkern and akern (hdf5.c:540-592, :712-875)."""
    rng = numpy.random.default_rng(3)
    ws = [-100.0, -50.0, 0.0, 50.0, 100.0]
    wk = {(0.1, w): rng.standard_normal((4, 4, 6, 6)) + 1j * rng.standard_normal((4, 4, 6, 6))
          for w in ws}
    sep = rng.standard_normal((8, 6))
    ak = {(0.1, a, t, f): rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))
          for a in range(3) for t in (10.0, 20.0) for f in (1e8, 1.5e8, 2e8)}
    p = str(tmp_path / 'k.h5')
    formats.write_kernels(p, wkern=wk, sepkern=sep, akern=ak)
    got = formats.read_wkernels(p, 0.1)
    assert sorted(got) == ws
    for w in ws:
        assert_array_equal(got[w], wk[(0.1, w)])
    assert h5.File(p)['wkern/0.1/-100.0/kern'].shape == (4, 4, 6, 6)
    cache = formats.H5KernelFile(p)
    assert cache.wstep(0.1) == 50.0
    assert_array_equal(cache(0.1, 61.0), wk[(0.1, 50.0)])
    assert_array_equal(cache(0.1, -1000.0), wk[(0.1, -100.0)])
    assert cache(0.1, 61.0) is cache(0.1, 40.0)
    assert_array_equal(h5.File(p)['sepkern/kern'][()], sep)
    a = formats.load_akern(p, 0.1)
    assert (a.antenna_count, a.time_count, a.freq_count) == (3, 2, 3)
    assert a.t_min == 10.0 and a.t_step == 10.0 and a.f_min == 1e8 and a.f_step == 5e7
    for (th, ant, t, f), k in ak.items():
        it, jf = int(round((t - a.t_min) / a.t_step)), int(round((f - a.f_min) / a.f_step))
        assert_array_equal(a.kernel(ant, it, jf), k)
