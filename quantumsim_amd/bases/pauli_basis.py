"""Pauli bases: orthonormal bases of Hermitian d x d matrices.

Behaviour follows quantumsim/bases/pauli_basis.py:9-128.  A basis holds
``vectors`` of shape (dim_pauli, dim_hilbert, dim_hilbert); the order of its
elements is the memory order of the corresponding Pauli-vector axis.
"""
import numpy as np

__all__ = ['PauliBasis']


class PauliBasis:
    """An orthonormal (w.r.t. tr(A B^dagger)) basis of Hermitian matrices.

    Parameters
    ----------
    vectors : array, shape (dim_pauli, dim_hilbert, dim_hilbert)
    labels : sequence of str
    superbasis : PauliBasis or None
        The basis this one was cut from by :meth:`subbasis`.
    """

    def __init__(self, vectors, labels, superbasis=None):
        vectors = np.asarray(vectors)
        if vectors.ndim != 3 or vectors.shape[1] != vectors.shape[2]:
            # quantumsim/bases/pauli_basis.py:23-26
            raise ValueError(
                "Pauli basis vectors must be square matrices, got shape {}"
                .format(vectors.shape[1:]))
        self.vectors = vectors
        self.labels = labels
        self._superbasis = superbasis
        # C[j, x] = <j| P_x |j>: row j extracts population j from a Pauli
        # vector (quantumsim/bases/pauli_basis.py:33-34)
        self.computational_basis_vectors = np.ascontiguousarray(
            np.diagonal(vectors, axis1=1, axis2=2).T)
        # level j -> index x of the element |j><j|, or None
        # (quantumsim/bases/pauli_basis.py:38-40)
        self.computational_basis_indices = {
            j: self._unit_index(row)
            for j, row in enumerate(self.computational_basis_vectors)}
        # index of the element proportional to the identity, or None
        # (quantumsim/bases/pauli_basis.py:43-46)
        traces = np.trace(vectors, axis1=1, axis2=2) / np.sqrt(self.dim_hilbert)
        self.trace_index = self._unit_index(traces)

    def __eq__(self, other):
        # allclose equality (quantumsim/bases/pauli_basis.py:48-53)
        if not isinstance(other, PauliBasis) and not hasattr(other, 'vectors'):
            return False
        ov = np.asarray(other.vectors)
        return self.vectors.shape == ov.shape and np.allclose(self.vectors, ov)

    def __hash__(self):
        return hash(np.ascontiguousarray(self.vectors).tobytes())

    @property
    def dim_hilbert(self):
        return self.vectors.shape[1]

    @property
    def dim_pauli(self):
        return self.vectors.shape[0]

    @property
    def superbasis(self):
        return self._superbasis or self

    def subbasis(self, indices):
        """Basis of the listed elements, in the listed order
        (quantumsim/bases/pauli_basis.py:70-83)."""
        indices = list(indices)
        return PauliBasis(self.vectors[indices],
                          [self.labels[i] for i in indices], self)

    def computational_subbasis(self):
        """Subbasis of the |j><j| elements present
        (quantumsim/bases/pauli_basis.py:85-89)."""
        return self.subbasis([x for x in self.computational_basis_indices.values()
                              if x is not None])

    def hilbert_to_pauli_vector(self, rho):
        """x -> tr(P_x rho)."""
        return np.einsum('xab,ba->x', self.vectors, rho)

    def is_orthonormal(self):
        gram = np.einsum('xab,yba->xy', self.vectors, self.vectors)
        assert np.allclose(gram, np.eye(self.dim_pauli))

    @staticmethod
    def _unit_index(v):
        # a vector summing to 1 with exactly one nonzero entry after rounding
        # to 8 decimals names that entry (quantumsim/bases/pauli_basis.py:99-106)
        if np.allclose(np.sum(v), 1):
            nz = np.flatnonzero(np.round(v, 8))
            if nz.size == 1:
                return int(nz[0])
        return None

    def __repr__(self):
        names = ' '.join(self.labels) if self.labels is not None else 'unnamed basis'
        return '<{} d_hilbert={}, d_pauli={}, {}>'.format(
            type(self).__name__, self.dim_hilbert, self.dim_pauli, names)
