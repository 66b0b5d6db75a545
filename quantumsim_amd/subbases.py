"""Subbasis optimisation of a gate list: keep, per qudit and per point of the
circuit, only the basis elements that can be nonzero there and that can
still influence the readout, so the Pauli vector shrinks (a qubit known to
be classical needs 2 elements instead of 4; one in |0><0|, 1).

This is the behaviour of the reference compiler's ``optimal_bases``
(quantumsim/operations/compiler.py:303-362, run by
``stage_align_bases`` over its circuit graph, compiler.py:42-98), restated
for a plain gate list:

* forward: the support of the state before gate t (per qudit, as indices
  into the full basis) starts at ``initial``; a gate maps the support of
  its inputs to the output elements in the column space of its PTM
  restricted to them (SVD, singular values above ``sv_cutoff``, entries of
  the left singular vectors above 1e-13, taken per qudit);
* backward: the elements the readout needs start at ``final``; a gate needs
  the input elements in the row space of its PTM restricted to the needed
  outputs (right singular vectors, same thresholds);
* the state before gate t lives on the intersection of the two, and every
  gate becomes the rectangular restriction of its PTM to (output support,
  input support) with the subbases attached, ready for
  ``PauliVector.apply_ptm`` plus the caller's bases update.

Dropping elements outside the backward set is exact (they cannot reach the
readout); dropping those outside the forward set drops entries that are
zero up to the cut-offs (as in the reference).
"""
import numpy as np

__all__ = ['optimal_subbases']


def _axis_unique(mask, dims):
    """Per-axis indices that appear in a boolean mask over prod(dims)."""
    idx = np.nonzero(mask.reshape(dims))
    return [sorted(set(int(i) for i in ax)) for ax in idx]


def _row_space_mask(M, sv_cutoff):
    """Columns of M touched by its row space (right singular vectors of
    singular values above the cut-off)."""
    if M.size == 0:
        return np.zeros(M.shape[1], dtype=bool)
    u, s, vh = np.linalg.svd(M, full_matrices=False)
    r = int(np.sum(s > sv_cutoff))
    return np.any(np.abs(vh[:r]) > 1e-13, axis=0)


def optimal_subbases(full_bases, gates, initial=None, final=None, sv_cutoff=1e-5):
    """``full_bases``: the basis of every qudit in which the gates' PTMs are
    written (square PTMs, layout (out..., in...)); ``gates``: [(ptm,
    qudits), ...]; ``initial``: per qudit the indices of the initial state's
    support (None = all); ``final``: per qudit the indices the readout uses
    (None = all; e.g. the |j><j| elements for populations).

    Returns (initial_bases, gates_out, final_bases): the subbases the state
    starts in, [(ptm, qudits, bases_out), ...] with rectangular PTMs, and
    the subbases after the last gate."""
    n = len(full_bases)
    dims = [b.dim_pauli for b in full_bases]
    full_sets = [list(range(d)) for d in dims]
    R = [list(full_sets[q]) if initial is None or initial[q] is None else sorted(set(initial[q]))
         for q in range(n)]
    gates = [(np.asarray(p, dtype=np.float64), tuple(int(x) for x in qs)) for p, qs in gates]
    for g, (P, qs) in enumerate(gates):
        k = len(qs)
        if P.shape != tuple(dims[q] for q in qs) * 2:
            raise ValueError('gate %d: PTM of shape %s, expected %s in the full bases'
                             % (g, P.shape, tuple(dims[q] for q in qs) * 2))
    # forward supports R_t (before gate t)
    fwd = [list(map(list, R))]
    for P, qs in gates:
        k = len(qs)
        dq = [dims[q] for q in qs]
        sub = P[(slice(None),) * k + np.ix_(*[R[q] for q in qs])]
        M = sub.reshape(int(np.prod(dq)), -1)
        mask = _row_space_mask(M.T, sv_cutoff)          # column space of M = row space of M^T
        outs = _axis_unique(mask, dq)
        for q, o in zip(qs, outs):
            R[q] = o
        fwd.append(list(map(list, R)))
    # backward needs N_t (before gate t)
    N = [list(full_sets[q]) if final is None or final[q] is None else sorted(set(final[q]))
         for q in range(n)]
    bwd = [None] * (len(gates) + 1)
    bwd[-1] = list(map(list, N))
    for g in range(len(gates) - 1, -1, -1):
        P, qs = gates[g]
        k = len(qs)
        dq = [dims[q] for q in qs]
        sub = P[np.ix_(*[N[q] for q in qs]) + (slice(None),) * k]
        M = sub.reshape(-1, int(np.prod(dq)))
        mask = _row_space_mask(M, sv_cutoff)
        ins = _axis_unique(mask, dq)
        for q, i in zip(qs, ins):
            N[q] = i
        bwd[g] = list(map(list, N))
    # supports = intersections; restricted PTMs
    S = [[sorted(set(fwd[t][q]) & set(bwd[t][q])) for q in range(n)]
         for t in range(len(gates) + 1)]
    for t in range(len(gates) + 1):
        for q in range(n):
            if not S[t][q]:
                # an empty support: the readout does not depend on the state
                # here; keep one element so shapes stay valid
                S[t][q] = [fwd[t][q][0] if fwd[t][q] else 0]
    out = []
    for g, (P, qs) in enumerate(gates):
        k = len(qs)
        sel = np.ix_(*([S[g + 1][q] for q in qs] + [S[g][q] for q in qs]))
        out.append((P[sel], qs, [full_bases[q].subbasis(S[g + 1][q]) for q in qs]))
    initial_bases = [full_bases[q].subbasis(S[0][q]) for q in range(n)]
    final_bases = [full_bases[q].subbasis(S[-1][q]) for q in range(n)]
    return initial_bases, out, final_bases
