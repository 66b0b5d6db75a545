"""Qubit gates and channels as PTMs in the ``general(2)`` basis.

Kraus operators follow quantumsim/operations/qubits.py:11-220; each
function returns a ``PTMOp`` whose ``ptm`` has shape (4, 4) or (4, 4, 4, 4).
``t1t2_rates`` (no counterpart in the reference, whose T1/T2 tutorial is a
placeholder) turns relaxation times into the damping rates used by
``amp_phase_damping``.
"""
import numpy as np

from .. import bases
from .operation import from_kraus, PTMOp

__all__ = ['rotate_euler', 'rotate_x', 'rotate_y', 'rotate_z', 'hadamard', 'cphase',
           'cnot', 'amp_damping', 'phase_damping', 'amp_phase_damping', 't1t2_rates']

_B1 = (bases.general(2),)
_B2 = _B1 * 2


def rotate_euler(phi, theta, lamda):
    """R_Z(phi) R_X(theta) R_Z(lamda) (qubits.py:11-35)."""
    ep, el = np.exp(1j * phi), np.exp(1j * lamda)
    s, c = np.sin(theta / 2), np.cos(theta / 2)
    return from_kraus(np.array([[c, -1j * el * s], [-1j * ep * s, ep * el * c]]), _B1)


def rotate_x(angle=np.pi):
    s, c = np.sin(angle / 2), np.cos(angle / 2)
    return from_kraus(np.array([[c, -1j * s], [-1j * s, c]]), _B1)


def rotate_y(angle=np.pi):
    s, c = np.sin(angle / 2), np.cos(angle / 2)
    return from_kraus(np.array([[c, -s], [s, c]]), _B1)


def rotate_z(angle=np.pi):
    e = np.exp(-1j * angle / 2)
    return from_kraus(np.diag([e, e.conjugate()]), _B1)


def hadamard():
    return from_kraus(np.sqrt(0.5) * np.array([[1, 1], [1, -1]]), _B1)


def cphase(angle=np.pi):
    return from_kraus(np.diag([1, 1, 1, np.exp(1j * angle)]), _B2)


def cnot():
    return from_kraus(np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]), _B2)


def amp_damping(total_rate=None, *, exc_rate=None, damp_rate=None):
    """Amplitude damping (qubits.py:182-198).  ``total_rate``: decay to |0>
    with probability ``total_rate`` (Kraus form, ``general(2)`` basis).
    Otherwise generalised amplitude damping with excitation rate
    ``exc_rate`` and decay rate ``damp_rate``, given as its PTM in the
    normalised Pauli basis ``gell_mann(2)`` = (I, X, Y, Z) / sqrt(2): X and
    Y shrink by sqrt(1 - r), Z by 1 - r (r = exc_rate + damp_rate) and
    moves by damp_rate - exc_rate towards |0>."""
    if total_rate is not None:
        g = total_rate
        return from_kraus(np.array([[[1, 0], [0, np.sqrt(1 - g)]],
                                    [[0, np.sqrt(g)], [0, 0]]]), _B1)
    if exc_rate is None or damp_rate is None:
        raise ValueError('Either the total_rate or both the exc_rate and damp_rate '
                         'must be provided')
    r = exc_rate + damp_rate
    ptm = np.zeros((4, 4))
    ptm[0, 0] = 1.0
    ptm[1, 1] = ptm[2, 2] = np.sqrt(1 - r)
    ptm[3, 3] = 1 - r
    ptm[3, 0] = damp_rate - exc_rate
    return PTMOp(ptm, (bases.gell_mann(2),))


def phase_damping(total_rate=None, *, x_deph_rate=None, y_deph_rate=None, z_deph_rate=None):
    """Phase damping (qubits.py:201-214).  ``total_rate``: the Kraus form
    that shrinks the coherences by sqrt(1 - total_rate) (``general(2)``
    basis).  Otherwise Pauli-axis dephasing: X, Y, Z of the normalised
    Pauli basis ``gell_mann(2)`` shrink by 1 - x_deph_rate, 1 - y_deph_rate,
    1 - z_deph_rate."""
    if total_rate is not None:
        lam = total_rate
        return from_kraus(np.array([[[1, 0], [0, np.sqrt(1 - lam)]],
                                    [[0, 0], [0, np.sqrt(lam)]]]), _B1)
    rates = (x_deph_rate, y_deph_rate, z_deph_rate)
    if any(r is None for r in rates):
        raise ValueError('Either the total_rate or the dephasing rates along each of '
                         'the three axis must be provided')
    return PTMOp(np.diag([1.0] + [1.0 - r for r in rates]), (bases.gell_mann(2),))


def amp_phase_damping(damp_rate, deph_rate):
    """Amplitude damping followed by phase damping, as one PTM
    (qubits.py:217-220 builds the same sequence)."""
    a = amp_damping(damp_rate).ptm
    p = phase_damping(deph_rate).ptm
    return PTMOp(p @ a, _B1)


def t1t2_rates(t, t1, t2):
    """(gamma, lambda) for a wait of ``t`` under relaxation T1 and coherence
    T2 <= 2 T1: gamma = 1 - exp(-t/T1); 1/T_phi = 1/T2 - 1/(2 T1);
    lambda = 1 - exp(-2 t / T_phi), so that sqrt(1-gamma) sqrt(1-lambda) =
    exp(-t/T2)."""
    gamma = 1.0 - np.exp(-t / t1)
    inv_tphi = 1.0 / t2 - 0.5 / t1
    lam = 1.0 - np.exp(-2.0 * t * inv_tphi)
    return gamma, lam
