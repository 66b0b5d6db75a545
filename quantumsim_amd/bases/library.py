"""Basis factories.  The element order of each factory fixes the memory
layout of every Pauli-vector axis that uses it, so it is identical to
quantumsim/bases/library.py:10-105 (an interface that has to match).
"""
from functools import lru_cache

import numpy as np

from .pauli_basis import PauliBasis

__all__ = ['general', 'gell_mann']

_R = np.sqrt(0.5)


@lru_cache(maxsize=64)
def general(dim_hilbert):
    """|i><i| for i < d first, then for each i, each j < i the pair
    X_ij = (|i><j| + |j><i|)/sqrt2, Y_ij = i(|i><j| - |j><i|)/sqrt2
    (quantumsim/bases/library.py:10-53).  general(2) = [0, 1, X10, Y10]."""
    d = int(dim_hilbert)
    mats, labels = [], []
    for i in range(d):
        m = np.zeros((d, d), complex)
        m[i, i] = 1
        mats.append(m)
        labels.append(str(i))
    for i in range(d):
        for j in range(i):
            x = np.zeros((d, d), complex)
            x[i, j] = x[j, i] = _R
            y = np.zeros((d, d), complex)
            y[i, j] = 1j * _R
            y[j, i] = -1j * _R
            mats += [x, y]
            labels += ['X{}{}'.format(i, j), 'Y{}{}'.format(i, j)]
    return PauliBasis(np.array(mats), np.array(labels, dtype=object))


@lru_cache(maxsize=64)
def gell_mann(dim_hilbert):
    """Generalised Gell-Mann matrices, (i, j) in row-major order: i == j
    generalised diagonals (I/sqrt(d) for i = 0), i < j symmetric, i > j
    antisymmetric (quantumsim/bases/library.py:56-105)."""
    d = int(dim_hilbert)
    mats, labels = [], []
    for i in range(d):
        for j in range(d):
            m = np.zeros((d, d), complex)
            if i == j:
                if i == 0:
                    np.fill_diagonal(m, 1 / np.sqrt(d))
                else:
                    diag = np.zeros(d)
                    diag[:i] = 1
                    diag[i] = -i
                    np.fill_diagonal(m, diag / np.sqrt(i * (i + 1)))
            elif i < j:
                m[i, j] = m[j, i] = _R
            else:
                m[i, j] = 1j * _R
                m[j, i] = -1j * _R
            mats.append(m)
            labels.append('γ{}{}'.format(i, j))
    return PauliBasis(np.array(mats), np.array(labels, dtype=object))
