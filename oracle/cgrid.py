"""ctypes binding of the C oracle (cgrid.c) -- TEST INFRASTRUCTURE.

Kernel-bank helpers mirror the reference's kernel selection: one table for
convgrid/convdegrid (synthesis.py:464-503), or one table per w-bin for
w_cache_imaging / w_cache_predict (synthesis.py:757-828).
"""

import ctypes
import os
import subprocess

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_EDOMAIN, _EBOUNDS = -1, -2


def build():
    subprocess.check_call(['make', '-s', '-C', _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'libcgrid.so')
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.cg_convgrid.argtypes = [P, ctypes.c_int, P, i64, i64, P, P,
                                  ctypes.c_int, ctypes.c_int, P, P, P]
        L.cg_convdegrid.argtypes = L.cg_convgrid.argtypes
        L.cg_frac_coord.argtypes = [ctypes.c_int, ctypes.c_int, P, i64, i64, P, P]
        L.cg_wbin.argtypes = [P, i64, i64, ctypes.c_double, P]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc == _EDOMAIN:
        raise AssertionError("p outside [-0.5, 0.5)")
    if rc == _EBOUNDS:
        raise ValueError("kernel window outside the grid")
    if rc != 0:
        raise RuntimeError("cgrid error %d" % rc)


class Bank:
    """Flat complex128 kernel bank with per-plane offsets and supports."""

    def __init__(self, tables, separable=False):
        tables = [numpy.ascontiguousarray(t, dtype=numpy.complex128) for t in tables]
        self.Q = tables[0].shape[0]
        self.separable = separable
        self.plane_s = numpy.array([t.shape[-1] for t in tables], dtype=numpy.int32)
        sizes = numpy.array([t.size for t in tables], dtype=numpy.int64)
        self.plane_off = numpy.concatenate([[0], numpy.cumsum(sizes)[:-1]]).astype(numpy.int64)
        self.kern = numpy.concatenate([t.ravel() for t in tables])


def frac_coord(N, Q, p):
    p = numpy.ascontiguousarray(p, dtype=numpy.float64)
    ix = numpy.empty(p.shape[0], dtype=numpy.int64)
    xf = numpy.empty(p.shape[0], dtype=numpy.int64)
    _check(lib().cg_frac_coord(N, Q, _ptr(p), p.shape[0], 1, _ptr(ix), _ptr(xf)))
    return ix, xf


def convgrid(bank, a, p, v, vplane=None):
    """In-place FP64 gridding into complex128 ``a``."""
    assert a.dtype == numpy.complex128 and a.flags.c_contiguous
    p = numpy.ascontiguousarray(p, dtype=numpy.float64)
    v = numpy.ascontiguousarray(v, dtype=numpy.complex128)
    vp = None if vplane is None else numpy.ascontiguousarray(vplane, dtype=numpy.int32)
    _check(lib().cg_convgrid(_ptr(a), a.shape[0], _ptr(p), p.shape[0], p.shape[1],
                             _ptr(v), _ptr(bank.kern), bank.Q, int(bank.separable),
                             _ptr(bank.plane_off), _ptr(bank.plane_s), _ptr(vp)))


def convdegrid(bank, a, p, vplane=None):
    a = numpy.ascontiguousarray(a, dtype=numpy.complex128)
    p = numpy.ascontiguousarray(p, dtype=numpy.float64)
    out = numpy.empty(p.shape[0], dtype=numpy.complex128)
    vp = None if vplane is None else numpy.ascontiguousarray(vplane, dtype=numpy.int32)
    _check(lib().cg_convdegrid(_ptr(a), a.shape[0], _ptr(p), p.shape[0], p.shape[1],
                               _ptr(out), _ptr(bank.kern), bank.Q, int(bank.separable),
                               _ptr(bank.plane_off), _ptr(bank.plane_s), _ptr(vp)))
    return out


def wbin(w, wstep):
    w = numpy.ascontiguousarray(w, dtype=numpy.float64)
    out = numpy.empty(w.shape[0], dtype=numpy.int64)
    _check(lib().cg_wbin(_ptr(w), w.shape[0], 1, float(wstep), _ptr(out)))
    return out
