"""The caller side of ``apply_ptm``.

``PTMOp`` reproduces what quantumsim's ``PTMOperation.__call__`` does
around the backend call (quantumsim/operations/operation.py:529-551): if the
Pauli vector's bases differ from the operation's input bases, the PTM is
re-expressed in the vector's bases (ptm_convert_basis), applied, and then the
*caller* writes the output bases into ``pauli_vector.bases``.
"""
import numpy as np

from ..algebra.algebra import kraus_to_ptm, ptm_convert_basis

__all__ = ['PTMOp', 'from_kraus', 'from_ptm']


class PTMOp:
    def __init__(self, ptm, bases_in, bases_out=None):
        self.bases_in = tuple(bases_in)
        self.bases_out = tuple(bases_out) if bases_out is not None else self.bases_in
        self.ptm = np.asarray(ptm, dtype=np.float64)
        shape = tuple(b.dim_pauli for b in self.bases_out + self.bases_in)
        if self.ptm.shape != shape:
            raise ValueError('Shape of `ptm` {} is not compatible with the bases {}'
                             .format(self.ptm.shape, shape))

    @property
    def num_qubits(self):
        return len(self.bases_in)

    def set_bases(self, bases_in=None, bases_out=None):
        b_in = tuple(bases_in) if bases_in is not None else self.bases_in
        b_out = tuple(bases_out) if bases_out is not None else self.bases_out
        if b_in == self.bases_in and b_out == self.bases_out:
            return self
        return PTMOp(ptm_convert_basis(self.ptm, self.bases_in, self.bases_out, b_in, b_out),
                     b_in, b_out)

    def __call__(self, pauli_vector, *qubits):
        if len(qubits) != self.num_qubits:
            raise ValueError('This is a {}-qubit operation, but number of qubits provided '
                             'is {}'.format(self.num_qubits, len(qubits)))
        op = self
        if any(pauli_vector.bases[q] != b for q, b in zip(qubits, self.bases_in)):
            op = self.set_bases(bases_in=tuple(pauli_vector.bases[q] for q in qubits))
        pauli_vector.apply_ptm(op.ptm, *qubits)
        for q, b in zip(qubits, op.bases_out):
            pauli_vector.bases[q] = b


def from_kraus(kraus, bases_in, bases_out=None):
    bases_out = bases_out or bases_in
    return PTMOp(kraus_to_ptm(np.asarray(kraus), bases_in, bases_out), bases_in, bases_out)


def from_ptm(ptm, bases_in, bases_out=None):
    return PTMOp(ptm, bases_in, bases_out)
