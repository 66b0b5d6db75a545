"""Interface every Pauli-vector backend implements.

Mirrors quantumsim/pauli_vectors/pauli_vector.py:6-116 (same method set,
size guard and error types) so Circuit/State/Operation code written against
the reference drives this package unchanged.
"""
import abc
import math

from ..algebra.algebra import dm_to_pv, pv_to_dm

__all__ = ['PauliVectorBase']


class PauliVectorBase(metaclass=abc.ABCMeta):
    """A density matrix of several qudits, expanded in per-qudit Pauli bases.

    Parameters
    ----------
    bases : list of PauliBasis
        Basis of every qudit; qudit 0 is the most significant axis.
    pv : array or None
        Expansion coefficients of shape ``dim_pauli``; ``None`` puts 1 at
        index (0, ..., 0).
    force : bool
        Density matrices with more than 2**22 elements ((prod d_hilbert)**2)
        are refused unless ``force=True`` (pauli_vector.py:27,31-37).
    """
    _size_max = 2 ** 22

    @abc.abstractmethod
    def __init__(self, bases, pv=None, *, force=False):
        self.bases = list(bases)
        if self.size > self._size_max and not force:
            raise ValueError(
                'Density matrix of the system is going to have {} items. It '
                'is probably too much. If you know what you are doing, '
                'pass `force=True` argument to the constructor.'.format(self.size))

    @classmethod
    def from_pv(cls, pv, bases, *, force=False):
        return cls(bases, pv, force=force)

    @abc.abstractmethod
    def to_pv(self):
        """Expansion coefficients as a numpy array."""

    @classmethod
    def from_dm(cls, dm, bases, *, force=False):
        if not hasattr(bases, '__iter__'):
            n_qubits = len(dm) // bases.dim_hilbert
            bases = [bases] * n_qubits
        return cls(bases, dm_to_pv(dm, bases), force=force)

    def to_dm(self):
        return pv_to_dm(self.to_pv(), self.bases)

    @property
    def n_qubits(self):
        return len(self.bases)

    @property
    def dim_hilbert(self):
        return tuple(b.dim_hilbert for b in self.bases)

    @property
    def size(self):
        return math.prod(self.dim_hilbert) ** 2

    @property
    def dim_pauli(self):
        return tuple(b.dim_pauli for b in self.bases)

    @abc.abstractmethod
    def apply_ptm(self, ptm, *qubits):
        """Apply a Pauli transfer matrix to the listed qudits."""

    @abc.abstractmethod
    def diagonal(self, *, get_data=True):
        """Diagonal of the density matrix (length prod d_hilbert)."""

    @abc.abstractmethod
    def trace(self):
        """Trace of the density matrix."""

    @abc.abstractmethod
    def partial_trace(self, *qubits):
        """New Pauli vector of the listed qudits, all others traced out."""

    @abc.abstractmethod
    def meas_prob(self, qubit):
        """Populations of the levels of one qudit."""

    @abc.abstractmethod
    def renormalize(self):
        """Scale to unit trace; return the trace before scaling."""

    @abc.abstractmethod
    def copy(self):
        """Deep copy."""

    def _validate_qubit(self, number, name):
        if number < 0 or number >= self.n_qubits:
            raise ValueError(
                "`{name}` number {n} does not exist in the system, "
                "it contains {n_qubits} qubits in total."
                .format(name=name, n=number, n_qubits=self.n_qubits))

    def _validate_ptm_shape(self, ptm, target_shape, name):
        if ptm.shape != target_shape:
            raise ValueError(
                "`{name}` shape must be {target_shape}, got {real_shape}"
                .format(name=name, target_shape=target_shape, real_shape=ptm.shape))
