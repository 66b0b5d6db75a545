"""Density matrix <-> Pauli vector and expectation values on the device.

The reference does these on the host, one einsum each
(quantumsim/algebra/algebra.py:44-62 ``dm_to_pv`` / ``pv_to_dm``,
quantumsim/states/state.py:43-83 ``State.exp_value``), and assumes one
Hilbert dimension for all qudits (``algebra.py:46,57``).  Here every qudit
keeps its own basis and dimension, and the per-qudit contractions run on the
apply kernels of libqsv.

Complex arithmetic on the real kernels: a complex tensor is stored as a real
one with a trailing axis of size 2 (real part, imaginary part), and a
complex matrix M acting on qudit q is the real 2-target PTM on (q, that
axis)

    [[Re M, -Im M],
     [Im M,  Re M]]     (layout (out_q, out_ri, in_q, in_ri)).

* pv[x] = Re tr(rho P_x): rho's row and column index of every qudit are
  paired into one axis y = i d + j, then qudit by qudit M[x, y] = P_x[j, i].
* rho = sum_x pv[x] P_x: per qudit N[y, x] = P_x[i, j], then the pairs are
  split back into rows and columns.
* <O> = tr(O rho): for a Pauli string, per qudit the covector
  c[x] = tr(sigma P_x) (a 1 x d^2 complex map); for a dense operator,
  tr(O rho) with rho from pv_to_dm.
"""
import math

import numpy as np
import torch

from .. import _native as nat

__all__ = ['SIGMA', 'dm_to_pv', 'pv_to_dm', 'exp_value']

SIGMA = {
    'I': np.array([[1, 0], [0, 1]], dtype=complex),
    'X': np.array([[0, 1], [1, 0]], dtype=complex),
    'Y': np.array([[0, -1j], [1j, 0]], dtype=complex),
    'Z': np.array([[1, 0], [0, -1]], dtype=complex),
}


def _complex_ptm(M):
    M = np.asarray(M, dtype=complex)
    mo, mi = M.shape
    R = np.zeros((mo, 2, mi, 2))
    R[:, 0, :, 0] = M.real
    R[:, 0, :, 1] = -M.imag
    R[:, 1, :, 0] = M.imag
    R[:, 1, :, 1] = M.real
    return R


def _stream(device):
    s = torch.cuda.current_stream(device).cuda_stream
    return s if s else None


def _apply(x, dims, target, M):
    """Complex M (m_out x dims[target]) applied to axis `target` of the
    complex tensor x (real, shape dims + [2]); returns (y, new dims)."""
    from ..pauli_vectors.hip import _Scratch
    n = len(dims)
    full = list(dims) + [2]
    mo = M.shape[0]
    out_dims = list(dims)
    out_dims[target] = mo
    y = torch.empty(math.prod(out_dims) * 2, dtype=torch.float64, device=x.device)
    P = np.ascontiguousarray(_complex_ptm(M))
    L = nat.lib()
    tg, do = nat.i32([target, n]), nat.i32([mo, 2])
    ws = nat.check(L.qsv_apply_ptm_workspace_bytes(nat.QSV_F64, 2, nat.i32([dims[target], 2]), do),
                   'apply_ptm')
    buf = _Scratch.get(x.device, 'ptm', ws).data_ptr() if ws else None
    rc = L.qsv_apply_ptm(nat.QSV_F64, x.data_ptr(), y.data_ptr(), n + 1, nat.i64(full), 2, tg, do,
                         P.ctypes.data_as(nat._c_f64p), buf, ws, _stream(x.device))
    nat.check(rc, 'apply_ptm')
    return y, out_dims


def _device(device):
    if device is None:
        return torch.device('cuda', torch.cuda.current_device())
    return torch.device(device)


def dm_to_pv(dm, bases, device=None):
    """Pauli vector (float64 device tensor of shape dim_pauli) of the density
    matrix ``dm`` (numpy or torch, (prod d) x (prod d)) in ``bases`` (one
    basis per qudit, any Hilbert dimensions)."""
    dev = _device(device)
    dh = [b.dim_hilbert for b in bases]
    n = len(dh)
    D = math.prod(dh)
    rho = torch.as_tensor(np.asarray(dm) if not torch.is_tensor(dm) else dm,
                          dtype=torch.complex128, device=dev)
    if tuple(rho.shape) != (D, D):
        raise ValueError('density matrix of shape {} does not match the bases ({} x {})'
                         .format(tuple(rho.shape), D, D))
    perm = [v for q in range(n) for v in (q, n + q)]
    x = rho.reshape(dh + dh).permute(perm).contiguous().reshape(-1)
    x = torch.view_as_real(x).reshape(-1).contiguous()
    dims = [d * d for d in dh]
    for q, b in enumerate(bases):
        d = b.dim_hilbert
        V = np.asarray(b.vectors)                      # [x, a, b] = P_x[a, b]
        M = V.transpose(0, 2, 1).reshape(V.shape[0], d * d)   # M[x, i d + j] = P_x[j, i]
        x, dims = _apply(x, dims, q, M)
    return x.reshape(dims + [2])[..., 0].contiguous()


def pv_to_dm(pv, bases, device=None):
    """Complex128 device density matrix of the Pauli vector ``pv`` (device
    or host array of shape dim_pauli) in ``bases``."""
    dev = _device(device)
    dp = [b.dim_pauli for b in bases]
    dh = [b.dim_hilbert for b in bases]
    n = len(dp)
    v = torch.as_tensor(pv if torch.is_tensor(pv) else np.asarray(pv), dtype=torch.float64,
                        device=dev).reshape(-1)
    if v.numel() != math.prod(dp):
        raise ValueError('Pauli vector of {} entries does not match the bases {}'
                         .format(v.numel(), dp))
    x = torch.stack([v, torch.zeros_like(v)], dim=-1).reshape(-1).contiguous()
    dims = list(dp)
    for q, b in enumerate(bases):
        d = b.dim_hilbert
        V = np.asarray(b.vectors)                      # [x, i, j] = P_x[i, j]
        N = V.transpose(1, 2, 0).reshape(d * d, V.shape[0])   # N[i d + j, x] = P_x[i, j]
        x, dims = _apply(x, dims, q, N)
    rho = torch.view_as_complex(x.reshape(dims + [2]))
    pairs = []
    for d in dh:
        pairs += [d, d]
    perm = list(range(0, 2 * n, 2)) + list(range(1, 2 * n, 2))
    D = math.prod(dh)
    return rho.reshape(pairs).permute(perm).reshape(D, D)


def exp_value(pv, bases, operator, device=None):
    """tr(O rho) for the Pauli vector ``pv`` in ``bases``
    (quantumsim/states/state.py:43-83): ``operator`` is a string of I, X, Y,
    Z (one per qubit) or an array with 2 n indices (rows then columns) or of
    shape (D, D).  Returns a complex number, as the reference's einsum."""
    dev = _device(device)
    n = len(bases)
    if isinstance(operator, str):
        if len(operator) != n:
            raise ValueError('operator string must have the same length as number of qubits '
                             'in the state')
        try:
            sig = [SIGMA[ch.upper()] for ch in operator]
        except KeyError as ex:
            raise ValueError('operator string must contain only I, X, Y or Z') from ex
        dp = [b.dim_pauli for b in bases]
        v = torch.as_tensor(pv if torch.is_tensor(pv) else np.asarray(pv), dtype=torch.float64,
                            device=dev).reshape(-1)
        x = torch.stack([v, torch.zeros_like(v)], dim=-1).reshape(-1).contiguous()
        dims = list(dp)
        for q, (b, s) in enumerate(zip(bases, sig)):
            V = np.asarray(b.vectors)
            if V.shape[1] != s.shape[0]:
                raise ValueError('Pauli operators act on qubits; qudit {} has dimension {}'
                                 .format(q, V.shape[1]))
            c = np.einsum('ab,xba->x', s, V)[None, :]  # c[x] = tr(sigma P_x)
            x, dims = _apply(x, dims, q, c)
        r = x.reshape(2).cpu().numpy()
        return complex(r[0], r[1])
    O = torch.as_tensor(np.asarray(operator) if not torch.is_tensor(operator) else operator,
                        dtype=torch.complex128, device=dev)
    D = math.prod(b.dim_hilbert for b in bases)
    O = O.reshape(D, D)
    rho = pv_to_dm(pv, bases, device=dev)
    return complex((O * rho.transpose(0, 1)).sum().item())
