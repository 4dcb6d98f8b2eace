"""ctypes binding of libcrocgpu.so (the C ABI declared in include/crocgpu.h).

The library is required: importing the GPU entry points without a built
library, or calling them without a ROCm device, raises immediately.  There is
no CPU fallback.
"""

import ctypes
import os
import threading

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libcrocgpu.so')

OK = 0
EINVAL, EDOMAIN, EBOUNDS, EHIP, ENOMEM, EPLANE = -1, -2, -3, -4, -5, -6
PLANE_SINGLE, PLANE_INDEX, PLANE_WBIN = 0, 1, 2

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double


class BlSetC(ctypes.Structure):
    """croc_bl_set: a packed set of baselines (device arrays)."""
    _fields_ = [('nbl', _I64), ('uvw_rows', _I64), ('freq_len', _I64), ('nvis', _I64),
                ('time_count', _P), ('freq_count', _P), ('uvw_first', _P),
                ('freq_first', _P), ('vis_first', _P), ('uvw_m', _P), ('freq', _P),
                ('vis', _P)]


_BL = ctypes.POINTER(BlSetC)
_U64P = ctypes.POINTER(ctypes.c_uint64)

# name -> argtypes, in the order of include/crocgpu.h
SIGNATURES = {
    'croc_version': [],
    'croc_last_error': [],
    'croc_ctx_create': [_I, ctypes.POINTER(_P)],
    'croc_ctx_destroy': [_P],
    'croc_ctx_set_stream': [_P, _P],
    'croc_ctx_set_tuning': [_P, _I, _I, _I],
    'croc_last_kernel_ms': [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    'croc_last_tap_kernel_ms': [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_I),
                                ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_I)],
    'croc_frac_coords': [_P, _P, _I64, _I64, _D, _I, _I, _I, _P, _P, _P, _P],
    'croc_convgrid': [_P, _P, _I, _P, _I64, _I64, _D, _P, _P, _I, _I, _I, _P, _P,
                      _I, _I, _P, _D, _I],
    'croc_convdegrid': [_P, _P, _I, _P, _I64, _I64, _D, _P, _P, _I, _I, _I, _P, _P,
                        _I, _I, _P, _D, _I],
    'croc_convgrid_pervis': [_P, _P, _I, _P, _I64, _I64, _D, _P, _P, _I, _I, _I, _P, _P,
                        _I, _I, _P, _D, _I],
    'croc_convdegrid_pervis': [_P, _P, _I, _P, _I64, _I64, _D, _P, _P, _I, _I, _I, _P, _P,
                        _I, _I, _P, _D, _I],
    'croc_plan_create': [_P, ctypes.POINTER(_P)],
    'croc_plan_destroy': [_P, _P],
    'croc_plan_bin': [_P, _P, _I, _P, _I64, _I64, _D, _I, _I, _I, _P, _P, _I, _I, _P, _D, _I],
    'croc_convgrid_plan': [_P, _P, _P, _P, _P],
    'croc_convdegrid_plan': [_P, _P, _P, _P, _P],
    'croc_convdegrid_noalloc': [_P, _P, _I, _P, _I64, _I64, _D, _P, _P, _I, _I, _I, _P, _P,
                                _I, _I, _P, _D, _I],
    'croc_remap_coords': [_P, _P, _I64, _I64, _D, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P],
    'croc_grid_nearest': [_P, _P, _I, _P, _I64, _I64, _D, _P],
    'croc_wstack_grid': [_P, _P, _I, _I, _P, _I64, _I64, _D, _D, _I64, _I, _P, _P, _I],
    'croc_wstack_add': [_P, _P, _P, _I, _I, _D, _P, _I],
    'croc_awt_pregrid': [_P, _P, _I, _P, _I64, _I64, _D, _I, _I, _P, _P, _P, _I, _I],
    'croc_awt_apply': [_P, _P, _I, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    'croc_awt_place': [_P, _P, _I, _P, _I, _I, _P],
    'croc_degrid_nearest': [_P, _P, _I, _P, _I64, _I64, _D, _P],
    'croc_w_kernels': [_P, _D, _P, _I, _I, _P, _I, _I, _D, _D, _P, _P, _P, _I, _P],
    'croc_kernel_oversample': [_P, _P, _I, _I, _I, _I, _P],
    'croc_make_hermitian': [_P, _P, _I, _I],
    'croc_dft2d_batched': [_P, _P, _P, _I, _I64, _I],
    'croc_fft2d_batched': [_P, _P, _P, _I, _I64, _I],
    'croc_convolve_same': [_P, _P, _I, _I, _P, _I, _I, _I64, _P],
    'croc_wtowers_grid': [_P, _P, _I, _P, _I64, _I64, _D, _P, _D, _I, _I, _P, _I, _I, _I,
                          _I, _D, _D],
    'croc_wtowers_degrid': [_P, _P, _I, _P, _I64, _I64, _D, _P, _D, _I, _I, _P, _I, _I, _I,
                            _I, _D, _D],
    'croc_wtowers_plan_create': [_P, ctypes.POINTER(_P)],
    'croc_wtowers_plan_destroy': [_P, _P],
    'croc_wtowers_plan_bin': [_P, _P, _I, _P, _I64, _I64, _D, _D, _I, _I, _I, _I, _I, _I, _D,
                              _D],
    'croc_wtowers_grid_plan': [_P, _P, _P, _P, _P, _I],
    'croc_wtowers_degrid_plan': [_P, _P, _P, _P, _P, _I],
    'croc_bl_weight': [_P, _P, _I, _D, _BL],
    'croc_bl_grid_simple': [_P, _P, _I, _D, _BL, _U64P],
    'croc_bl_degrid_simple': [_P, _P, _I, _D, _BL, _U64P],
    'croc_degrid_conv_uv': [_P, _P, _I, _D, _P, _I64, _P, _I, _I, _P, _U64P],
    'croc_degrid_conv_bl': [_P, _P, _I, _D, _D, _D, _D, _D, _D, _D, _BL, _I, _I, _I, _I, _P,
                            _I, _I, _U64P],
    'croc_degrid_conv': [_P, _P, _I, _D, _D, _D, _BL, _P, _I, _I, _U64P],
    'croc_grid_wprojection': [_P, _P, _I, _D, _BL, _P, _I, _D, _D, _I, _I, _U64P],
    'croc_streamer_degrid': [_P, _P, _I, _D, _D, _D, _D, _D, _D, _D, _BL, _P, _I64, _I, _I,
                             _P, _I, _I, _P, _P, _U64P],
    'croc_streamer_degrid_batch': [_P, _P, _I, _I, _D, _P, _BL, _P, _I64, _I, _I, _P, _I, _I,
                                   _P, _P, _U64P],
}

_lib = None
_lock = threading.Lock()
_ctx = {}
_tuning = (0, 512, 0)  # tile, chunk, group_kib of croc_ctx_set_tuning


class CrocError(RuntimeError):
    """Failure reported by libcrocgpu that has no reference exception type."""


def load():
    """Load libcrocgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    'libcrocgpu.so is not built (%s); run python -m crocodile_amd.build'
                    % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = ctypes.c_char_p if name == 'croc_last_error' else ctypes.c_int
            _lib = L
    return _lib


def check(rc):
    """Map a libcrocgpu status to the exception the reference raises."""
    if rc == OK:
        return
    msg = load().croc_last_error().decode(errors='replace')
    if rc == EDOMAIN:
        raise AssertionError(msg)
    if rc in (EBOUNDS, EPLANE):
        raise ValueError(msg)
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise CrocError('libcrocgpu error %d: %s' % (rc, msg))


def device():
    if not torch.cuda.is_available():
        raise RuntimeError('crocodile_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


def ctx():
    """Library context of torch's current stream on the current device.

    A context owns the scratch buffers of its calls (fixed-point
    accumulator, scale state, work lists, w-towers sub-grids), and calls
    return without waiting for the device.  One context per (device,
    stream) keeps calls queued on different streams from sharing those
    buffers; calls on one stream are ordered by the stream itself."""
    dev = device()
    L = load()
    stream = torch.cuda.current_stream(dev).cuda_stream
    key = (dev.index, stream)
    c = _ctx.get(key)
    if c is None:
        with _lock:
            c = _ctx.get(key)
            if c is None:
                h = _P()
                check(L.croc_ctx_create(dev.index, ctypes.byref(h)))
                check(L.croc_ctx_set_stream(h, _P(stream)))
                check(L.croc_ctx_set_tuning(h, *_tuning))
                c = h
                _ctx[key] = c
    return c


def last_kernel_ms():
    a, b = ctypes.c_float(), ctypes.c_float()
    check(load().croc_last_kernel_ms(ctx(), ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def last_tap_kernel_ms():
    """(grid ms, grid launches, degrid ms, degrid launches) of the tap
    kernels of the last gridding / degridding call on this stream."""
    a, b = ctypes.c_float(), ctypes.c_float()
    na, nb = _I(), _I()
    check(load().croc_last_tap_kernel_ms(ctx(), ctypes.byref(a), ctypes.byref(na),
                                         ctypes.byref(b), ctypes.byref(nb)))
    return a.value, na.value, b.value, nb.value


def set_tuning(tile=0, chunk=512, group_kib=0):
    """Tuning knobs (croc_ctx_set_tuning) for every context, present and
    future: results do not depend on them."""
    global _tuning
    _tuning = (int(tile), int(chunk), int(group_kib))
    ctx()
    for c in list(_ctx.values()):
        check(load().croc_ctx_set_tuning(c, *_tuning))


def ptr(t):
    """Device address of a tensor (None stays NULL).  Tensors carrying torch's
    lazy conjugate / negative bits are refused: their memory does not hold
    the values they represent."""
    if t is None:
        return None
    if t.is_conj() or t.is_neg():
        raise ValueError('pass a materialised tensor (resolve_conj / resolve_neg), '
                         'not a lazy conjugate or negative view')
    return _P(t.data_ptr())
