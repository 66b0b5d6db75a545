"""Oracle restatement of the basis factories (test infrastructure only).

Follows quantumsim/bases/library.py:10-105 and pauli_basis.py:22-106.
Written independently of quantumsim_amd.bases (explicit index formulas)
so that each pins the other.
"""
import numpy as np


class Basis:
    def __init__(self, vectors, labels):
        self.vectors = np.asarray(vectors, dtype=complex)
        self.labels = list(labels)

    @property
    def dim_hilbert(self):
        return self.vectors.shape[1]

    @property
    def dim_pauli(self):
        return self.vectors.shape[0]

    @property
    def computational_basis_vectors(self):
        # [j, x] = P_x[j, j]   (pauli_basis.py:33-34, einsum "xii -> ix")
        d = self.dim_hilbert
        return np.array([[self.vectors[x, j, j] for x in range(self.dim_pauli)]
                         for j in range(d)])

    def subbasis(self, indices):
        indices = list(indices)
        return Basis(self.vectors[indices], [self.labels[i] for i in indices])


def general(d):
    """library.py:10-53: d projectors, then for i in range(d), j in range(i):
    (X_ij, Y_ij)."""
    n = d * d
    v = np.zeros((n, d, d), complex)
    labels = []
    for i in range(d):
        v[i, i, i] = 1.0
        labels.append(str(i))
    pos = d
    r = np.sqrt(0.5)   # library.py:7 uses sqrt(0.5), 1 ULP from 1/sqrt(2)
    for i in range(d):
        for j in range(i):
            v[pos, i, j] = r
            v[pos, j, i] = r
            v[pos + 1, i, j] = 1j * r
            v[pos + 1, j, i] = -1j * r
            labels += ['X%d%d' % (i, j), 'Y%d%d' % (i, j)]
            pos += 2
    return Basis(v, labels)


def gell_mann(d):
    """library.py:56-105: element number i*d + j."""
    v = np.zeros((d * d, d, d), complex)
    r = np.sqrt(0.5)   # library.py:7 uses sqrt(0.5), 1 ULP from 1/sqrt(2)
    for i in range(d):
        for j in range(d):
            x = i * d + j
            if i == j and i == 0:
                for a in range(d):
                    v[x, a, a] = 1 / np.sqrt(d)
            elif i == j:
                norm = np.sqrt(i * (i + 1))
                for a in range(i):
                    v[x, a, a] = 1 / norm
                v[x, i, i] = -i / norm
            elif i < j:
                v[x, i, j] = v[x, j, i] = r
            else:
                v[x, i, j] = 1j * r
                v[x, j, i] = -1j * r
    return Basis(v, ['g%d%d' % (i, j) for i in range(d) for j in range(d)])
