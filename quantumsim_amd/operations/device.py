"""The operation -> PTM pipeline on the device.

quantumsim builds every gate on the host: ``Operation.from_kraus`` turns
Kraus operators into a PTM (quantumsim/operations/operation.py:132-175 ->
quantumsim/algebra/algebra.py:11-30), ``PTMOperation.set_bases`` re-expresses
it in the bases the state is in (operation.py:508-519 ->
algebra.py:33-41), and ``PTMOperation.__call__`` hands the numpy PTM to the
backend and then writes the output bases into ``pauli_vector.bases``
(operation.py:529-551).

Here the same pipeline runs in torch on the state's device:

* ``kraus_to_ptm``: PTM[x, y] = sum_k Re tr(P^out_x K_k P^in_y K_k^dagger)
  for Kraus tensors of any leading batch shape, per-qudit Hilbert
  dimensions allowed (the reference assumes one, algebra.py:12,17);
* ``ptm_convert_basis``: P' = T_out P T_in with
  T_out[x, y] = tr(P^new_x P^old_y) and T_in[z, w] = tr(P^old_z P^new_w)
  (real: both factors are Hermitian);
* ``PTMOperation``: the reference's class with its PTM as a torch tensor.
  Calling it on a ``PauliVectorHIP`` converts the bases on the device and
  applies the PTM through ``qsv_apply_ptm_dev_strided`` (dense 3- and
  4-qubit PTMs on the MFMA kernels): no host copy, no synchronisation.

Every step is differentiable: Kraus operators or PTMs built from trainable
parameters carry gradients through ``kraus_to_ptm``, ``set_bases`` and
``quantumsim_amd.autograd.apply``.
"""
from functools import lru_cache, reduce
from itertools import chain

import numpy as np
import torch

__all__ = ['basis_tensor', 'kraus_to_ptm', 'ptm_convert_basis', 'PTMOperation']

_DEVICE_CACHE = {}


def _kron_vectors(bases):
    return reduce(np.kron, [np.asarray(b.vectors) for b in bases])


@lru_cache(maxsize=256)
def _kron_host(bases):
    return np.ascontiguousarray(_kron_vectors(bases), dtype=np.complex128)


def _on(key, host, device, dtype):
    k = (key, str(torch.device(device)), dtype)
    t = _DEVICE_CACHE.get(k)
    if t is None:
        t = torch.as_tensor(host, dtype=dtype, device=device)
        _DEVICE_CACHE[k] = t
    return t


def basis_tensor(bases, device='cpu', dtype=torch.complex128):
    """(prod dim_pauli, D, D) Kronecker products of the basis matrices of
    ``bases`` (qudit 0 most significant; quantumsim/algebra/algebra.py:6-8),
    cached per device."""
    bases = tuple(bases)
    return _on(('kron', bases), _kron_host(bases), device, dtype)


def _check_pairing(bases_in, bases_out):
    if len(bases_in) != len(bases_out):
        raise ValueError("Input and output bases must contain the same number"
                         " of elements")
    for q, (bi, bo) in enumerate(zip(bases_in, bases_out)):
        if bi.dim_hilbert != bo.dim_hilbert:
            raise ValueError('qudit {}: input basis of Hilbert dimension {} but output basis '
                             'of dimension {}'.format(q, bi.dim_hilbert, bo.dim_hilbert))


def kraus_to_ptm(kraus, bases_in, bases_out=None):
    """PTM of rho -> sum_k K_k rho K_k^dagger (quantumsim/algebra/algebra.py:11-30).

    ``kraus``: torch tensor (complex or real) of shape (..., K, D, D) or
    (D, D), D = prod of the qudits' Hilbert dimensions (qudit 0 most
    significant).  Returns a real float64 tensor of shape
    (..., dims_out..., dims_in...) on ``kraus``'s device, differentiable
    with respect to ``kraus``."""
    bases_in = tuple(bases_in)
    bases_out = tuple(bases_out) if bases_out is not None else bases_in
    _check_pairing(bases_in, bases_out)
    if not torch.is_tensor(kraus):
        kraus = torch.as_tensor(np.asarray(kraus))
    kraus = kraus.to(torch.complex128)
    if kraus.dim() == 2:
        kraus = kraus.unsqueeze(0)
    D = int(np.prod([b.dim_hilbert for b in bases_in]))
    if kraus.dim() < 3 or kraus.shape[-1] != D or kraus.shape[-2] != D:
        raise ValueError('Shape of the Kraus operator for bases provided must be '
                         '{0}x{0}, got {1}x{2} instead'
                         .format(D, kraus.shape[-2] if kraus.dim() > 1 else '?',
                                 kraus.shape[-1]))
    b_in = basis_tensor(bases_in, kraus.device)
    b_out = basis_tensor(bases_out, kraus.device)
    # K P_y K^dagger for every y, summed over k; then <P_x, .> = tr(P_x .)
    kpk = torch.einsum('...kab,ybc,...kdc->...yad', kraus, b_in, kraus.conj())
    ptm = torch.einsum('xda,...yad->...xy', b_out, kpk).real
    shape = tuple(b.dim_pauli for b in bases_out) + tuple(b.dim_pauli for b in bases_in)
    return ptm.reshape(tuple(kraus.shape[:-3]) + shape)


@lru_cache(maxsize=256)
def _transfer_host(bases_new, bases_old, new_first):
    """tr(P^new_x P^old_y) (new_first) or tr(P^old_z P^new_w), real."""
    bn = _kron_host(bases_new)
    bo = _kron_host(bases_old)
    if new_first:
        t = np.einsum('xij,yji->xy', bn, bo, optimize=True)
    else:
        t = np.einsum('zkl,wlk->zw', bo, bn, optimize=True)
    return np.ascontiguousarray(t.real)


def ptm_convert_basis(ptm, bi_old, bo_old, bi_new, bo_new):
    """Re-express ``ptm`` (torch, (..., dims_out_old..., dims_in_old...)) in
    new input/output bases (quantumsim/algebra/algebra.py:33-41).  Same
    device and dtype as ``ptm``; differentiable with respect to it."""
    bi_old, bo_old, bi_new, bo_new = map(tuple, (bi_old, bo_old, bi_new, bo_new))
    k = len(bi_old)
    if not (len(bo_old) == len(bi_new) == len(bo_new) == k):
        raise ValueError('all four basis tuples must have one basis per qudit')
    d_in = int(np.prod([b.dim_pauli for b in bi_old]))
    d_out = int(np.prod([b.dim_pauli for b in bo_old]))
    lead = tuple(ptm.shape[:ptm.dim() - 2 * k])
    if tuple(ptm.shape[len(lead):]) != tuple(b.dim_pauli for b in chain(bo_old, bi_old)):
        raise ValueError('PTM of shape {} does not match the old bases {}'
                         .format(tuple(ptm.shape), tuple(b.dim_pauli
                                                         for b in chain(bo_old, bi_old))))
    dt = ptm.dtype if ptm.dtype in (torch.float32, torch.float64) else torch.float64
    t_out = _on(('tout', bo_new, bo_old), _transfer_host(bo_new, bo_old, True), ptm.device, dt)
    t_in = _on(('tin', bi_new, bi_old), _transfer_host(bi_new, bi_old, False), ptm.device, dt)
    res = t_out @ ptm.to(dt).reshape(lead + (d_out, d_in)) @ t_in
    return res.reshape(lead + tuple(b.dim_pauli for b in chain(bo_new, bi_new)))


class PTMOperation:
    """quantumsim's ``PTMOperation`` (quantumsim/operations/operation.py:443-551)
    with the PTM as a torch tensor (float64, any device, may require grad).

    ``from_kraus`` / ``from_ptm`` build it; ``set_bases`` / ``ptm`` convert
    it (on its device); calling it on a Pauli vector applies it and then sets
    ``pauli_vector.bases[q]`` to the output bases, as the reference's caller
    does.  A PTM on the vector's device is applied without leaving it."""

    def __init__(self, ptm, bases_in, bases_out=None):
        bases_in = tuple(bases_in)
        bases_out = tuple(bases_out) if bases_out is not None else bases_in
        if not torch.is_tensor(ptm):
            ptm = torch.as_tensor(np.asarray(ptm, dtype=np.float64))
        if ptm.is_complex():
            raise ValueError('PTM must be real, got dtype {}'.format(ptm.dtype))
        self._ptm = ptm if ptm.dtype == torch.float64 else ptm.to(torch.float64)
        self.bases_in = bases_in
        self.bases_out = bases_out
        self._num_qubits = len(bases_in)
        self._validate_bases(bases_out=bases_out)
        shape = tuple(b.dim_pauli for b in chain(bases_out, bases_in))
        if tuple(self._ptm.shape) != shape:
            raise ValueError(
                'Shape of `ptm` is not compatible with the `bases` '
                'dimensionality: \n'
                '- expected shape from provided `bases`: {}\n'
                '- `ptm` shape: {}'.format(shape, tuple(self._ptm.shape)))

    # ---------------------------------------------------------- builders
    @staticmethod
    def from_ptm(ptm, bases_in, bases_out=None):
        return PTMOperation(ptm, bases_in, bases_out)

    @staticmethod
    def from_kraus(kraus, bases_in, bases_out=None):
        """Kraus operators (torch or array-like, (K, D, D) or (D, D))."""
        bases_in = tuple(bases_in)
        bases_out = tuple(bases_out) if bases_out else bases_in
        if not torch.is_tensor(kraus):
            kraus = torch.as_tensor(np.asarray(kraus))
        if kraus.dim() not in (2, 3):
            raise ValueError('`kraus` should be a 2D or 3D array, got shape {}'
                             .format(tuple(kraus.shape)))
        return PTMOperation(kraus_to_ptm(kraus, bases_in, bases_out), bases_in, bases_out)

    # -------------------------------------------------------- properties
    @property
    def num_qubits(self):
        return self._num_qubits

    @property
    def dim_hilbert(self):
        """Hilbert dimension per qudit (a tuple: qudits may differ)."""
        return tuple(b.dim_hilbert for b in self.bases_in)

    @property
    def shape(self):
        return tuple(self._ptm.shape)

    @property
    def device(self):
        return self._ptm.device

    def to(self, device):
        return PTMOperation(self._ptm.to(device), self.bases_in, self.bases_out)

    def _validate_bases(self, **kwargs):
        # quantumsim/operations/operation.py:361-377, per qudit
        for name, bases in kwargs.items():
            if not hasattr(bases, '__iter__'):
                raise ValueError("`{n}` must be list-like, got {t}."
                                 .format(n=name, t=type(bases)))
            bases = tuple(bases)
            if self.num_qubits != len(bases):
                raise ValueError("Number of basis elements in `{}` ({}) does "
                                 "not match number of qubits in the "
                                 "operation ({})."
                                 .format(name, len(bases), self.num_qubits))
            for b, ref in zip(bases, self.bases_in):
                if b.dim_hilbert != ref.dim_hilbert:
                    raise ValueError(
                        "Expected bases with Hilbert dimensionality {}, "
                        "but {} has elements with Hilbert dimensionality {}."
                        .format(ref.dim_hilbert, name, b.dim_hilbert))

    # -------------------------------------------------------- conversion
    def set_bases(self, bases_in=None, bases_out=None):
        """The same operation in other bases (operation.py:60-83,508-519)."""
        if bases_in is None and bases_out is None:
            raise ValueError('Either bases_in or bases_out must be specified.')
        if bases_in is not None:
            self._validate_bases(bases_in=bases_in)
        if bases_out is not None:
            self._validate_bases(bases_out=bases_out)
        b_in = tuple(bases_in) if bases_in else self.bases_in
        b_out = tuple(bases_out) if bases_out else self.bases_out
        if b_in == self.bases_in and b_out == self.bases_out:
            return self
        return PTMOperation(ptm_convert_basis(self._ptm, self.bases_in, self.bases_out,
                                              b_in, b_out), b_in, b_out)

    def ptm(self, bases_in, bases_out=None):
        bases_out = bases_out or bases_in
        if tuple(bases_in) == self.bases_in and tuple(bases_out) == self.bases_out:
            return self._ptm
        return self.set_bases(bases_in, bases_out)._ptm

    # -------------------------------------------------------- application
    def __call__(self, pauli_vector, *qubit_indices):
        """Applies the operation (operation.py:529-551): converts to the
        vector's input bases when they differ, applies, updates
        ``pauli_vector.bases`` for the targets."""
        if len(qubit_indices) != self.num_qubits:
            raise ValueError('This is a {}-qubit operation, but number of '
                             'qubits provided is {}'
                             .format(self.num_qubits, len(qubit_indices)))
        op = self
        for q, b in zip(qubit_indices, self.bases_in):
            if pauli_vector.bases[q] != b:
                op = self.set_bases(bases_in=tuple(pauli_vector.bases[q]
                                                   for q in qubit_indices))
                break
        ptm = op._ptm.detach()
        dev = getattr(pauli_vector, 'device', None)
        if dev is not None and torch.device(dev).type == 'cuda':
            ptm = ptm.to(dev)               # stays on (or moves to) the vector's GPU
        else:
            ptm = ptm.cpu().numpy()         # host backends get what the reference passes
        pauli_vector.apply_ptm(ptm, *qubit_indices)
        for q, b in zip(qubit_indices, op.bases_out):
            pauli_vector.bases[q] = b

    def apply_to(self, x, bases, qubit_indices):
        """Differentiable form of ``__call__`` for batched states: ``x`` is a
        [B, N] tensor of states in ``bases`` (one basis per qudit); the PTM
        is converted to the targets' bases and applied with
        ``quantumsim_amd.autograd.apply`` (PTMs up to 16 x 16), so gradients
        reach ``x`` and whatever the PTM was built from.  Returns
        (y, new_bases)."""
        from .. import autograd
        qubit_indices = tuple(int(q) for q in qubit_indices)
        if len(qubit_indices) != self.num_qubits:
            raise ValueError('This is a {}-qubit operation, but number of '
                             'qubits provided is {}'
                             .format(self.num_qubits, len(qubit_indices)))
        bases = list(bases)
        op = self
        if any(bases[q] != b for q, b in zip(qubit_indices, self.bases_in)):
            op = self.set_bases(bases_in=tuple(bases[q] for q in qubit_indices))
        shape = tuple(b.dim_pauli for b in bases)
        dims_out = tuple(b.dim_pauli for b in op.bases_out)
        d_in = int(np.prod([shape[q] for q in qubit_indices]))
        ptm = op._ptm.to(x.device).reshape(1, int(np.prod(dims_out)), d_in)
        y = autograd.apply(x, ptm, shape, qubit_indices, dims_out)
        for q, b in zip(qubit_indices, op.bases_out):
            bases[q] = b
        return y, bases

    def compose(self, other):
        """PTM of applying ``self`` then ``other`` on the same qudits (the
        reference's ``_Chain.ptm`` for one shared qudit set,
        operation.py:610-623): other's input bases are matched to self's
        output bases first."""
        if other.num_qubits != self.num_qubits:
            raise ValueError('operations act on {} and {} qudits'
                             .format(self.num_qubits, other.num_qubits))
        o2 = other.set_bases(bases_in=self.bases_out)
        d_mid = int(np.prod([b.dim_pauli for b in self.bases_out]))
        d_in = int(np.prod([b.dim_pauli for b in self.bases_in]))
        d_out = int(np.prod([b.dim_pauli for b in o2.bases_out]))
        p2 = o2._ptm.to(self._ptm.device).reshape(d_out, d_mid)
        p = p2 @ self._ptm.reshape(d_mid, d_in)
        return PTMOperation(p.reshape(tuple(b.dim_pauli for b in chain(o2.bases_out,
                                                                       self.bases_in))),
                            self.bases_in, o2.bases_out)
