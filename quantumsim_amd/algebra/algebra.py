"""Host-side PTM algebra for the Pauli-vector backend.

Semantics follow quantumsim/algebra/algebra.py:6-62:
* ``kraus_to_ptm``: PTM[x, y] = sum_k Re tr(P_x K_k P_y K_k^dagger);
* ``ptm_convert_basis``: re-express a PTM in other input/output bases;
* ``dm_to_pv`` / ``pv_to_dm``: pv_x = Re tr(P_x rho), rho = sum_x pv_x P_x.
Multi-qudit bases are Kronecker products with qudit 0 most significant.
"""
from functools import lru_cache, reduce

import numpy as np

__all__ = ['kraus_to_ptm', 'ptm_convert_basis', 'dm_to_pv', 'pv_to_dm', 'bases_kron']


@lru_cache(maxsize=128)
def bases_kron(bases):
    """(prod dim_pauli, D, D) stack of Kronecker-product basis matrices for a
    tuple of bases (quantumsim/algebra/algebra.py:6-8)."""
    return reduce(np.kron, [b.vectors for b in bases])


def kraus_to_ptm(kraus, bases_in, bases_out):
    """PTM of the channel rho -> sum_k K_k rho K_k^dagger, shape
    (out_0..out_{n-1}, in_0..in_{n-1}) (quantumsim/algebra/algebra.py:11-30).

    ``kraus``: (K, D, D) or (D, D) with D = d^n (qudit 0 most significant)."""
    bases_in, bases_out = tuple(bases_in), tuple(bases_out)
    if len(bases_in) != len(bases_out):
        raise ValueError("Input and output bases must contain the same number"
                         " of elements")
    kraus = np.asarray(kraus)
    if kraus.ndim == 2:
        kraus = kraus[None]
    b_out = bases_kron(bases_out)           # (X, D, D)
    b_in = bases_kron(bases_in)             # (Y, D, D)
    # K P_y K^dagger for all k, y, then <P_x, .> = tr(P_x .)
    kpk = np.einsum('kab,ybc,kdc->yad', kraus, b_in, kraus.conj(), optimize=True)
    ptm = np.einsum('xda,yad->xy', b_out, kpk, optimize=True).real
    shape = tuple(b.dim_pauli for b in bases_out) + tuple(b.dim_pauli for b in bases_in)
    return ptm.reshape(shape)


def ptm_convert_basis(ptm, bi_old, bo_old, bi_new, bo_new):
    """P' = T_out . P . T_in with T_out[x, y] = tr(P^new_x P^old_y) and
    T_in[z, w] = tr(P^old_z P^new_w) (quantumsim/algebra/algebra.py:33-41)."""
    bi_old, bo_old, bi_new, bo_new = map(tuple, (bi_old, bo_old, bi_new, bo_new))
    d_in = int(np.prod([b.dim_pauli for b in bi_old]))
    d_out = int(np.prod([b.dim_pauli for b in bo_old]))
    t_out = np.einsum('xij,yji->xy', bases_kron(bo_new), bases_kron(bo_old), optimize=True)
    t_in = np.einsum('zkl,wlk->zw', bases_kron(bi_old), bases_kron(bi_new), optimize=True)
    res = (t_out @ np.asarray(ptm).reshape(d_out, d_in) @ t_in).real
    shape = tuple(b.dim_pauli for b in bo_new) + tuple(b.dim_pauli for b in bi_new)
    return res.reshape(shape)


def dm_to_pv(dm, bases):
    """Pauli vector of a density matrix of uniform qudit dimension
    (quantumsim/algebra/algebra.py:44-52)."""
    bases = tuple(bases)
    n = len(bases)
    if n == 0:
        return np.asarray(dm, dtype=complex).reshape(()).real.copy()
    d = bases[0].dim_hilbert
    rho = np.asarray(dm).reshape((d ** n, d ** n))
    b = bases_kron(bases)                            # (X, D, D)
    pv = np.einsum('xji,ij->x', b, rho, optimize=True).real
    return pv.reshape(tuple(bb.dim_pauli for bb in bases))


def pv_to_dm(pv, bases):
    """Density matrix sum_x pv_x P_x (quantumsim/algebra/algebra.py:55-62)."""
    bases = tuple(bases)
    n = len(bases)
    if n == 0:
        return np.asarray(pv, dtype=complex).reshape((1, 1))
    d = bases[0].dim_hilbert
    b = bases_kron(bases)
    return np.einsum('x,xij->ij', np.asarray(pv).reshape(-1), b,
                     optimize=True).reshape((d ** n,) * 2)
