"""The C ABI library loads and exports every entry point include/crocgpu.h
declares; host-side argument checks and error reporting work without a GPU."""
import ctypes
import os
import re

import pytest

from crocodile_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'crocgpu.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char \*)\s*(croc_\w+)\s*\(', text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert 'croc_convgrid' in names and 'croc_convdegrid' in names
    assert len(names) >= 14


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared()


def test_version_and_error_path_without_device():
    import torch
    L = _lib.load()
    assert L.croc_version() == 1
    if torch.cuda.is_available():
        pytest.skip('a GPU is present; the no-device error path is not reachable')
    h = ctypes.c_void_p()
    rc = L.croc_ctx_create(0, ctypes.byref(h))
    assert rc == _lib.EHIP
    assert b'hipGetDeviceCount' in L.croc_last_error()


def test_null_context_rejected():
    L = _lib.load()
    assert L.croc_ctx_set_stream(None, None) == _lib.EINVAL
    assert L.croc_ctx_set_tuning(None, 0, 4096, 2048) == _lib.EINVAL
    with pytest.raises(ValueError):
        _lib.check(L.croc_convgrid(None, None, 16, None, 0, 3, 1.0, None, None, 1, 0, 1,
                                   None, None, 1, 0, None, 1.0, 0))


def test_check_maps_reference_exceptions():
    with pytest.raises(AssertionError):
        _lib.check(_lib.EDOMAIN)
    with pytest.raises(ValueError):
        _lib.check(_lib.EBOUNDS)
    with pytest.raises(_lib.CrocError):
        _lib.check(_lib.EHIP)


def test_lazy_conjugate_views_are_refused():
    import torch
    t = torch.zeros(4, dtype=torch.complex64).conj()
    with pytest.raises(ValueError):
        _lib.ptr(t)
    assert _lib.ptr(torch.conj_physical(torch.zeros(4, dtype=torch.complex64))) is not None
