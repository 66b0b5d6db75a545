"""Oracle restatement of PauliVectorNumpy (test infrastructure only).

Every operation is one numpy einsum over the float64 expansion, as in
quantumsim/pauli_vectors/numpy.py:8-120, including its quirks:
* default state = 1 at index (0, ..., 0)                 (numpy.py:39-41)
* apply_ptm never touches ``bases``                      (numpy.py:50-64)
* diagonal has length prod(d_hilbert)                    (numpy.py:66-80)
* partial_trace: data axes ascending, bases in caller order (numpy.py:86-95)
* renormalize returns the trace before scaling           (numpy.py:109-117)
"""
import math
import warnings

import numpy as np

SIZE_MAX = 2 ** 22


class PauliVectorNumpy:
    def __init__(self, bases, pv=None, *, force=False):
        self.bases = list(bases)
        if self.size > SIZE_MAX and not force:
            raise ValueError('too large')          # pauli_vector.py:33-37
        if pv is not None:
            if self.dim_pauli != pv.shape:
                raise ValueError('shape')           # numpy.py:25-30
            if pv.dtype not in (np.float16, np.float32, np.float64):
                raise ValueError('dtype')           # numpy.py:31-35
            self._data = pv
        else:
            self._data = np.zeros(self.dim_pauli)
            self._data[tuple([0] * self.n_qubits)] = 1

    @property
    def n_qubits(self):
        return len(self.bases)

    @property
    def dim_hilbert(self):
        return tuple(b.dim_hilbert for b in self.bases)

    @property
    def dim_pauli(self):
        return tuple(b.dim_pauli for b in self.bases)

    @property
    def size(self):
        return math.prod(self.dim_hilbert) ** 2

    def to_pv(self):
        return self._data

    def apply_ptm(self, ptm, *qubits):
        """numpy.py:50-64: einsum(data, [0..n), ptm, out+in, data idx with the
        targets renamed to the PTM output labels)."""
        n, k = self.n_qubits, len(qubits)
        if ptm.ndim != 2 * k:
            raise ValueError('ndim')
        data_idx = list(range(n))
        out_labels = list(range(n, n + k))
        res_idx = list(data_idx)
        for q, lab in zip(qubits, out_labels):
            res_idx[q] = lab
        self._data = np.einsum(self._data, data_idx, ptm, out_labels + list(qubits),
                               res_idx, optimize=True)

    def diagonal(self):
        """numpy.py:66-80."""
        n = self.n_qubits
        args = [self._data, list(range(n))]
        for i, b in enumerate(self.bases):
            args += [b.computational_basis_vectors, [n + i, i]]
        res = np.einsum(*args, list(range(n, 2 * n)), optimize=True)
        return res.real.reshape(math.prod(self.dim_hilbert))

    def trace(self):
        return np.sum(self.diagonal())             # numpy.py:82-84

    def partial_trace(self, *qubits):
        """numpy.py:86-95 (implicit-output einsum: ascending axes)."""
        n = self.n_qubits
        args = [self._data, list(range(n))]
        for i, b in enumerate(self.bases):
            if i not in qubits:
                args += [b.vectors, [i, n + i, n + i]]
        traced = np.einsum(*args, optimize=True).real
        return PauliVectorNumpy([self.bases[q] for q in qubits], traced)

    def meas_prob(self, qubit):
        """numpy.py:97-107."""
        n = self.n_qubits
        args = [self._data, list(range(n))]
        for i, b in enumerate(self.bases):
            args += [b.vectors, [i, n + i, n + i]]
        args.append([n + qubit])
        return np.einsum(*args, optimize=True).real

    def renormalize(self):
        tr = self.trace()
        if tr > 1e-8:
            self._data *= self.trace() ** -1
        else:
            warnings.warn('Density matrix trace is 0')
        return tr

    def copy(self):
        return PauliVectorNumpy(self.bases, self._data.copy(), force=True)


def multitake(x, idx):
    """x[np.ix_(*idx)] with index -1 giving 0 (primitives.cu:92-130 extended
    by the missing-level rule of numpy.py:66-80)."""
    idx = [np.asarray(i, dtype=np.int64) for i in idx]
    safe = [np.where(i < 0, 0, i) for i in idx]
    out = x[np.ix_(*safe)] if idx else x.copy()
    for ax, i in enumerate(idx):
        bad = i < 0
        if bad.any():
            sl = [slice(None)] * out.ndim
            sl[ax] = bad
            out[tuple(sl)] = 0
    return out
