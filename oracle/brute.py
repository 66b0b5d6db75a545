"""Hilbert-space brute force for small systems (test infrastructure only).

Independent of every einsum in the oracle and the product: density
matrices as dense complex matrices, channels as Kraus sums, partial traces
by explicit index loops.
"""
import numpy as np


def embed_kraus(kraus, targets, dims):
    """Kraus operators on ``targets`` (in that order, qudit 0 most
    significant within the operator) embedded into the full Hilbert space of
    qudits with dimensions ``dims``."""
    kraus = np.asarray(kraus, dtype=complex)
    if kraus.ndim == 2:
        kraus = kraus[None]
    n = len(dims)
    D = int(np.prod(dims))
    tdims = [dims[t] for t in targets]
    out = []
    for K in kraus:
        K = K.reshape(tdims + tdims)
        full = np.zeros((D, D), complex)
        for row in range(D):
            rd = np.unravel_index(row, dims)
            for col in range(D):
                cd = np.unravel_index(col, dims)
                if any(rd[i] != cd[i] for i in range(n) if i not in targets):
                    continue
                idx = tuple(rd[t] for t in targets) + tuple(cd[t] for t in targets)
                full[row, col] = K[idx]
        out.append(full)
    return out


def apply_channel(rho, kraus_full):
    return sum(K @ rho @ K.conj().T for K in kraus_full)


def partial_trace(rho, dims, keep):
    """Reduced density matrix of the qudits in ``keep`` (ascending order)."""
    keep = sorted(keep)
    n = len(dims)
    t = np.asarray(rho).reshape(list(dims) + list(dims))
    for ax in reversed([i for i in range(n) if i not in keep]):
        m = t.ndim // 2
        t = np.trace(t, axis1=ax, axis2=ax + m)
    dk = int(np.prod([dims[i] for i in keep]))
    return t.reshape(dk, dk)


def random_density_matrix(D, rng):
    a = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


def random_unitary(D, rng):
    """Haar unitary: QR of a complex Ginibre matrix with the phase fix of
    Mezzadri (2007)."""
    z = (rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))
