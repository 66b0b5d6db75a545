"""Physical layout of a Pauli vector in HBM.

The logical tensor is C-order over the qudits (qudit 0 slowest).  Axes
whose contiguous stride reaches PAD_FROM bytes (8 MB) get PAD bytes (4 KB)
of padding per step, so the strides of the outermost axes are not powers
of two: a gate on two outer qudits reads 16 operand streams whose addresses
would otherwise coincide in the HBM channel/bank mapping (measured on
MI355X: 4.4-4.9 TB/s unpadded, 5.6-5.8 TB/s with 4 KB padding for bonds
(0,1), (1,2), (2,3) of a 14-qubit float64 vector).  QSV_PAD_BYTES
overrides the padding (0 disables it).
"""
import math
import os

__all__ = ['physical_strides', 'contiguous_strides']

PAD_FROM = 8 << 20


def _pad_bytes():
    v = os.environ.get('QSV_PAD_BYTES')
    return 4096 if v is None else int(v)


def contiguous_strides(shape):
    strides, acc = [0] * len(shape), 1
    for q in reversed(range(len(shape))):
        strides[q] = acc
        acc *= shape[q]
    return tuple(strides), acc


def physical_strides(shape, elem_bytes):
    """(strides, extent) in elements for a vector of ``shape``."""
    pad = _pad_bytes() // elem_bytes
    strides, acc = [0] * len(shape), 1
    for q in reversed(range(len(shape))):
        s = acc
        if pad > 0 and s * elem_bytes >= PAD_FROM:
            s += pad
        strides[q] = s
        acc = s * shape[q]
    return tuple(strides), max(acc, 1)
