"""A portable circuit file: the gate list of a Pauli-vector simulation in
one ``.npz`` (plain arrays, no pickled objects), for exchanging circuits
with the reference and for replaying benchmark circuits.

Contents (``numpy.savez``; every entry is a plain array):

* ``format``            : the string ``quantumsim-amd-circuit/1``
* ``n_bases``           : number of distinct bases in the table
* ``basis{j}_vectors``  : complex (dim_pauli, d, d), element order of the basis
* ``basis{j}_labels``   : its labels (unicode array)
* ``initial_bases``     : int32 per qudit, index into the table
* ``initial_pv``        : (optional) float64 initial Pauli vector; absent =
                          the default state (index (0, ..., 0) = 1,
                          quantumsim/pauli_vectors/numpy.py:39-41)
* ``n_gates``
* ``gate{g}_ptm``       : float64, layout (out..., in...) in qudit order
* ``gate{g}_qudits``    : int32 target qudits (PTM axis order)
* ``gate{g}_bases_out`` : int32 per target, index of the basis the target
                          is in after the gate (the caller's update of
                          ``bases``, quantumsim/operations/operation.py:549-551)

A PTM acts in the bases the qudits are in at that point, as
``PauliVector.apply_ptm`` requires.  ``load`` returns plain numpy data and
``Basis`` objects of this package; ``run`` replays the file on a
``PauliVectorHIP``.
"""
import numpy as np

from .bases.pauli_basis import PauliBasis

__all__ = ['FORMAT', 'save', 'load', 'run']

FORMAT = 'quantumsim-amd-circuit/1'


def _basis_key(b):
    v = np.asarray(b.vectors)
    return (v.shape, v.tobytes(), tuple(b.labels))


def save(path, bases, gates, initial_pv=None):
    """``bases``: the initial basis of every qudit; ``gates``: list of
    (ptm, qudits) or (ptm, qudits, bases_out) with bases_out one basis per
    target (omitted: the targets keep their bases)."""
    table, index = [], {}

    def idx(b):
        k = _basis_key(b)
        if k not in index:
            index[k] = len(table)
            table.append(b)
        return index[k]

    cur = list(bases)
    arrays = {'format': np.array(FORMAT), 'initial_bases': np.array([idx(b) for b in cur],
                                                                     dtype=np.int32)}
    for g, gate in enumerate(gates):
        ptm, qudits = gate[0], tuple(int(q) for q in gate[1])
        if any(not 0 <= q < len(cur) for q in qudits) or len(set(qudits)) != len(qudits):
            raise ValueError('gate %d: bad qudits %s for %d qudits' % (g, qudits, len(cur)))
        outs = list(gate[2]) if len(gate) > 2 and gate[2] is not None else [cur[q] for q in qudits]
        if len(outs) != len(qudits):
            raise ValueError('gate %d: %d output bases for %d qudits' % (g, len(outs), len(qudits)))
        ptm = np.asarray(ptm, dtype=np.float64)
        if ptm.ndim != 2 * len(qudits):
            raise ValueError('gate %d: PTM of %d dimensions for %d qudits'
                             % (g, ptm.ndim, len(qudits)))
        arrays['gate%d_ptm' % g] = ptm
        arrays['gate%d_qudits' % g] = np.array(qudits, dtype=np.int32)
        arrays['gate%d_bases_out' % g] = np.array([idx(b) for b in outs], dtype=np.int32)
        for q, b in zip(qudits, outs):
            cur[q] = b
    arrays['n_gates'] = np.array(len(gates))
    arrays['n_bases'] = np.array(len(table))
    for j, b in enumerate(table):
        arrays['basis%d_vectors' % j] = np.asarray(b.vectors, dtype=complex)
        arrays['basis%d_labels' % j] = np.array([str(x) for x in b.labels])
    if initial_pv is not None:
        arrays['initial_pv'] = np.asarray(initial_pv, dtype=np.float64)
    np.savez(path, **arrays)


def load(path):
    """Returns (bases, gates, initial_pv): initial bases per qudit, a list
    of (ptm, qudits, bases_out), and the initial Pauli vector or None."""
    with np.load(path, allow_pickle=False) as f:
        if str(f['format']) != FORMAT:
            raise ValueError('not a %s file: %r' % (FORMAT, str(f['format'])))
        table = [PauliBasis(f['basis%d_vectors' % j], [str(x) for x in f['basis%d_labels' % j]])
                 for j in range(int(f['n_bases']))]
        bases = [table[j] for j in f['initial_bases']]
        gates = []
        for g in range(int(f['n_gates'])):
            gates.append((np.array(f['gate%d_ptm' % g]),
                          tuple(int(q) for q in f['gate%d_qudits' % g]),
                          [table[j] for j in f['gate%d_bases_out' % g]]))
        pv = np.array(f['initial_pv']) if 'initial_pv' in f.files else None
    return bases, gates, pv


def run(path, dtype=None, device=None, force=True):
    """Replays a circuit file on a PauliVectorHIP (the caller's bases
    update after every gate included); returns the vector."""
    from .pauli_vectors.hip import PauliVectorHIP
    bases, gates, pv = load(path)
    v = PauliVectorHIP(bases, pv, force=force, dtype=dtype, device=device)
    for ptm, qudits, outs in gates:
        v.apply_ptm(ptm, *qudits)
        for q, b in zip(qudits, outs):
            v.bases[q] = b
    return v
