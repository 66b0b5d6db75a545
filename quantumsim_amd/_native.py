"""ctypes binding of libqsv.so (the C ABI declared in include/qsv.h).

The library is loaded after ``torch`` so that it binds to the HIP runtime
torch already mapped (both carry the soname libamdhip64.so.7); device
pointers and streams therefore come straight from torch tensors and
``torch.cuda.current_stream()``.  There is no fallback: if the library or a
GPU is missing, every operation raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module doc)

__all__ = ['lib', 'check', 'QSV_F32', 'QSV_F64', 'NativeError', 'library_path', 'Tile',
           'dtype_code']

QSV_F32 = 0
QSV_F64 = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, 'lib', 'libqsv.so')


class NativeError(RuntimeError):
    """A HIP runtime failure reported by libqsv."""


_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_f64p = ctypes.POINTER(ctypes.c_double)


class Tile(ctypes.Structure):
    """qsv_tile (include/qsv.h): launch shape of the batched apply and the
    PTM gradient chosen by the autotuner; zero fields = library's choice."""
    _fields_ = [('kernel', ctypes.c_int32), ('tile_orbits', ctypes.c_int32),
                ('grad_blocks', ctypes.c_int32), ('reserved', ctypes.c_int32)]

    def key(self):
        return (self.kernel, self.tile_orbits, self.grad_blocks)


KERNEL_AUTO, KERNEL_DIRECT, KERNEL_LDS = 0, 1, 2
_vp = ctypes.c_void_p

_SIGNATURES = {
    'qsv_last_error': (ctypes.c_char_p, []),
    'qsv_abi_version': (ctypes.c_int, []),
    'qsv_apply_ptm_workspace_bytes': (ctypes.c_int64, [ctypes.c_int, ctypes.c_int,
                                                       _c_i32p, _c_i32p]),
    'qsv_apply_ptm': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, _c_i64p,
                                     ctypes.c_int, _c_i32p, _c_i32p, _c_f64p, _vp,
                                     ctypes.c_int64, _vp]),
    'qsv_apply_ptm_strided': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, _c_i64p,
                                             _c_i64p, _c_i64p, ctypes.c_int, _c_i32p, _c_i32p,
                                             _c_f64p, _vp, ctypes.c_int64, _vp]),
    'qsv_apply_ptm_dev': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, _c_i64p,
                                         ctypes.c_int, _c_i32p, _c_i32p, _vp, _vp,
                                         ctypes.c_int64, _vp]),
    'qsv_apply_ptm_dev_strided': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, _c_i64p,
                                                 _c_i64p, _c_i64p, ctypes.c_int, _c_i32p,
                                                 _c_i32p, _vp, _vp, ctypes.c_int64, _vp]),
    'qsv_apply_ptm_batched': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                             _c_i64p, ctypes.c_int, _c_i32p, _c_i32p, _vp,
                                             ctypes.c_int64, _vp]),
    'qsv_ptm_grad_workspace_bytes': (ctypes.c_int64, [ctypes.c_int, ctypes.c_int, _c_i64p,
                                                      ctypes.c_int, _c_i32p, _c_i32p]),
    'qsv_ptm_grad': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _c_i64p,
                                    ctypes.c_int, _c_i32p, _c_i32p, _vp, ctypes.c_int, _vp,
                                    ctypes.c_int64, _vp]),
    'qsv_ptm_grad_noalloc': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                            _c_i64p, ctypes.c_int, _c_i32p, _c_i32p, _vp,
                                            ctypes.c_int, _vp]),
    'qsv_ptm_grad_noalloc_scratch_bytes': (ctypes.c_int64, []),
    'qsv_emulate_ptm_grad': (ctypes.c_int, [ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int, _c_i64p,
                                            ctypes.c_int, _c_i32p, _c_i32p, _c_f64p,
                                            ctypes.c_int32, _c_i32p]),
    'qsv_apply_ptm_batched_backward': (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                                      ctypes.c_int, _c_i64p, ctypes.c_int,
                                                      _c_i32p, _vp, ctypes.c_int64, _vp,
                                                      ctypes.c_int, _vp, ctypes.c_int64, _vp]),
    'qsv_emulate_apply_ptm_batched_backward': (ctypes.c_int, [ctypes.c_int, _c_f64p, _c_f64p,
                                                              _c_f64p, ctypes.c_int, _c_i64p,
                                                              ctypes.c_int, _c_i32p, _c_f64p,
                                                              _c_f64p, ctypes.c_int32]),
    'qsv_emulate_apply_ptm_direct': (ctypes.c_int, [ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int,
                                                    _c_i64p, ctypes.c_int, _c_i32p, _c_f64p,
                                                    ctypes.c_int32]),
    'qsv_window_blocks': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                         _vp, _vp, ctypes.c_int64, _vp]),
    'qsv_window_blocks_backward': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                                  ctypes.c_int64, _vp, _vp, _vp]),
    'qsv_emulate_window_blocks': (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, _c_f64p, _c_f64p,
                                                 _vp, _c_f64p, _c_f64p, ctypes.c_int64, _c_f64p,
                                                 _c_f64p, _c_f64p]),
    'qsv_circuits_max_qubits': (ctypes.c_int, [ctypes.c_int]),
    'qsv_circuits_lds_bytes': (ctypes.c_int64, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    'qsv_circuits_forward': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                            ctypes.c_int, _vp]),
    'qsv_circuits_adjoint': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                            _vp, ctypes.c_int, _vp]),
    'qsv_emulate_circuits': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int64, _c_f64p,
                                            ctypes.c_int64, _c_f64p, ctypes.c_int64, _c_f64p,
                                            ctypes.c_int64, _c_f64p, _c_f64p, ctypes.c_int]),
    'qsv_ptm_grad_tiled': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                          _c_i64p, ctypes.c_int, _c_i32p, _c_i32p, _vp,
                                          ctypes.c_int, _vp, ctypes.c_int64, _vp, _vp]),
    'qsv_apply_ptm_batched_tiled': (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.c_int,
                                                   ctypes.c_int, _c_i64p, ctypes.c_int, _c_i32p,
                                                   _c_i32p, _vp, ctypes.c_int64, _vp, _vp]),
    'qsv_emulate_apply_ptm_tiled': (ctypes.c_int, [ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int,
                                                   _c_i64p, ctypes.c_int, _c_i32p, _c_i32p,
                                                   _c_f64p, _vp, _c_i32p]),
    'qsv_multitake': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                     _c_i64p, _c_i64p, _vp, _c_i32p, _vp]),
    'qsv_multitake_strided': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                             _c_i64p, _c_i64p, _c_i64p, _vp, _c_i32p, _vp]),
    'qsv_axis_sum_workspace_bytes': (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int,
                                                      ctypes.c_int64]),
    'qsv_axis_sum': (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_int64, _vp, _vp, ctypes.c_int64, _vp]),
    'qsv_scale': (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_double, _vp]),
    'qsv_selftest_fastdiv': (ctypes.c_int64, [ctypes.c_int32, ctypes.c_uint64]),
    'qsv_emulate_apply_ptm': (ctypes.c_int, [ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int,
                                             _c_i64p, _c_i64p, _c_i64p, ctypes.c_int, _c_i32p,
                                             _c_i32p, _c_f64p, _c_i32p]),
}

_lib = None


def _load():
    if not os.path.exists(library_path):
        raise ImportError(
            'libqsv.so not found at {}; build it with '
            '`python -m quantumsim_amd.build`'.format(library_path))
    h = ctypes.CDLL(library_path, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    return h


def lib():
    """The loaded library (loaded on first use)."""
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(rc, what):
    """Map a libqsv return code onto the reference's exception types:
    ValueError for bad arguments (quantumsim/pauli_vectors/numpy.py:24-54),
    NotImplementedError for unsupported requests (cuda.py:125-127)."""
    if rc >= 0:
        return rc
    msg = '{}: {}'.format(what, lib().qsv_last_error().decode(errors='replace'))
    if rc == -1:
        raise ValueError(msg)
    if rc == -3:
        raise NotImplementedError(msg)
    raise NativeError(msg)


def dtype_code(dtype):
    if dtype == torch.float64:
        return QSV_F64
    if dtype == torch.float32:
        return QSV_F32
    raise ValueError('unsupported Pauli-vector dtype {}'.format(dtype))


def i64(seq):
    seq = list(seq)
    return (ctypes.c_int64 * max(1, len(seq)))(*seq)


def i32(seq):
    seq = list(seq)
    return (ctypes.c_int32 * max(1, len(seq)))(*seq)
