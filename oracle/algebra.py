"""Oracle restatement of the PTM algebra (test infrastructure only).

Follows quantumsim/algebra/algebra.py:6-62 in meaning; computed by a
different route than quantumsim_amd.algebra (superoperators on row-major
vectorised matrices) so that each pins the other.
"""
from functools import reduce

import numpy as np


def kron_bases(bases):
    return reduce(np.kron, [b.vectors for b in bases])


def kraus_to_ptm(kraus, bases_in, bases_out):
    """PTM[x, y] = Re sum_k tr(P_x K P_y K^dag)  (algebra.py:11-30), via the
    superoperator S = sum_k K (x) conj(K) acting on row-major vec(rho) and
    tr(A B) = vec(A^T) . vec(B)."""
    kraus = np.asarray(kraus, dtype=complex)
    if kraus.ndim == 2:
        kraus = kraus[None]
    D = kraus.shape[1]
    S = sum(np.kron(k, k.conj()) for k in kraus)
    bo = kron_bases(bases_out).reshape(-1, D, D)
    bi = kron_bases(bases_in).reshape(-1, D, D)
    left = np.array([p.T.reshape(-1) for p in bo])       # vec(P_x^T)
    right = np.array([p.reshape(-1) for p in bi]).T      # vec(P_y) columns
    ptm = (left @ S @ right).real
    shape = [b.dim_pauli for b in bases_out] + [b.dim_pauli for b in bases_in]
    return ptm.reshape(shape)


def dm_to_pv(dm, bases):
    """pv_x = Re tr(P_x rho)  (algebra.py:44-52)."""
    bk = kron_bases(bases)
    D = bk.shape[1]
    rho = np.asarray(dm).reshape(D, D)
    pv = np.array([np.trace(p @ rho) for p in bk]).real
    return pv.reshape([b.dim_pauli for b in bases])


def pv_to_dm(pv, bases):
    """rho = sum_x pv_x P_x  (algebra.py:55-62)."""
    bk = kron_bases(bases)
    return np.tensordot(np.asarray(pv).reshape(-1), bk, axes=(0, 0))


def ptm_convert_basis(ptm, bi_old, bo_old, bi_new, bo_new):
    """algebra.py:33-41."""
    t_out = np.array([[np.trace(a @ b) for b in kron_bases(bo_old)]
                      for a in kron_bases(bo_new)])
    t_in = np.array([[np.trace(a @ b) for b in kron_bases(bi_new)]
                     for a in kron_bases(bi_old)])
    d_out = t_out.shape[1]
    d_in = t_in.shape[0]
    res = (t_out @ np.asarray(ptm).reshape(d_out, d_in) @ t_in).real
    return res.reshape([b.dim_pauli for b in bo_new] + [b.dim_pauli for b in bi_new])
