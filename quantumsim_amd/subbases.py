"""Subbasis optimisation of a gate list: keep, per qudit and per point of the
circuit, only the basis elements that can be nonzero there and that can
still influence the readout, so the Pauli vector shrinks (a qubit known to
be classical needs 2 elements instead of 4; one in |0><0|, 1).

This is the behaviour of the reference compiler's ``optimal_bases``
(quantumsim/operations/compiler.py:303-362, run by
``stage_align_bases`` over its circuit graph until no node changes,
compiler.py:38-98), restated for a plain gate list:

* every wire segment (a qudit between two of its gates, before its first
  gate or after its last) carries a support, a set of indices into the
  qudit's full basis: the initial supports and the readout's needs
  (``final``) fix the outer segments, the inner ones start full;
* a gate is refined with its PTM restricted to (output supports, input
  supports): the SVD of that matrix (singular values above ``sv_cutoff``)
  keeps the outputs touched by its left singular vectors and the inputs
  touched by its right singular vectors (entries above 1e-13, taken per
  qudit), which shrinks the segments on both sides of the gate;
* sweeps over the gates, forwards then backwards, repeat until no segment
  changes (the reference's queue reaches the same fixed point), and every
  gate becomes the rectangular restriction of its PTM to its segments,
  with the subbases attached, ready for ``PauliVector.apply_ptm`` plus the
  caller's bases update.

Restricting both sides of a gate at once keeps pairings: an element of one
qudit that only ever appears together with elements of another qudit that
nothing downstream needs is dropped too.  Dropped outputs cannot reach the
readout; dropped inputs are zero up to the cut-offs (as in the reference).
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


def _refine(P, k, rows, cols, sv_cutoff):
    """optimal_bases of one gate (compiler.py:330-354): P restricted to the
    output supports ``rows`` and input supports ``cols`` (per target, index
    lists), SVD, and the per-target indices (into rows / cols) that the
    kept singular vectors touch."""
    sub = P[np.ix_(*(rows + cols))]
    dout = [len(r) for r in rows]
    din = [len(c) for c in cols]
    M = sub.reshape(int(np.prod(dout)), int(np.prod(din)))
    if M.size == 0:
        return [[] for _ in rows], [[] for _ in cols]
    u, s, vh = np.linalg.svd(M, full_matrices=False)
    r = int(np.sum(s > sv_cutoff))
    mask_out = np.any(np.abs(u[:, :r]) > 1e-13, axis=1)
    mask_in = np.any(np.abs(vh[:r]) > 1e-13, axis=0)
    return _axis_unique(mask_out, dout), _axis_unique(mask_in, din)


def optimal_subbases(full_bases, gates, initial=None, final=None, sv_cutoff=1e-5,
                     max_sweeps=100):
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
    gates = [(np.asarray(p, dtype=np.float64), tuple(int(x) for x in qs)) for p, qs in gates]
    for g, (P, qs) in enumerate(gates):
        if P.shape != tuple(dims[q] for q in qs) * 2:
            raise ValueError('gate %d: PTM of shape %s, expected %s in the full bases'
                             % (g, P.shape, tuple(dims[q] for q in qs) * 2))
    # wire segments: seg_in[g][j] / seg_out[g][j] for target j of gate g
    cur = list(range(n))                 # current segment of every qudit
    sup = []
    for q in range(n):
        sup.append(list(range(dims[q])) if initial is None or initial[q] is None
                   else sorted(set(int(i) for i in initial[q])))
    seg_in, seg_out = [], []
    for P, qs in gates:
        seg_in.append([cur[q] for q in qs])
        outs = []
        for q in qs:
            sup.append(list(range(dims[q])))
            cur[q] = len(sup) - 1
            outs.append(cur[q])
        seg_out.append(outs)
    for q in range(n):
        if final is not None and final[q] is not None:
            sup[cur[q]] = sorted(set(sup[cur[q]]) & set(int(i) for i in final[q]))
    last = list(cur)
    # sweeps until no segment changes (the compiler queue's fixed point)
    order = list(range(len(gates)))
    for _ in range(max_sweeps):
        changed = False
        for g in order + order[::-1]:
            P, qs = gates[g]
            rows = [sup[s] for s in seg_out[g]]
            cols = [sup[s] for s in seg_in[g]]
            outs, ins = _refine(P, len(qs), rows, cols, sv_cutoff)
            for s, sel, base in zip(seg_out[g] + seg_in[g], outs + ins, rows + cols):
                new = [base[i] for i in sel]
                if new != sup[s]:
                    sup[s] = new
                    changed = True
        if not changed:
            break
    for s in range(len(sup)):
        if not sup[s]:
            # an empty support: the readout does not depend on the state
            # here; keep one element so shapes stay valid
            sup[s] = [0]
    out = []
    for g, (P, qs) in enumerate(gates):
        sel = np.ix_(*([sup[s] for s in seg_out[g]] + [sup[s] for s in seg_in[g]]))
        out.append((P[sel], qs, [full_bases[q].subbasis(sup[s])
                                 for q, s in zip(qs, seg_out[g])]))
    initial_bases = [full_bases[q].subbasis(sup[q]) for q in range(n)]
    final_bases = [full_bases[q].subbasis(sup[last[q]]) for q in range(n)]
    return initial_bases, out, final_bases
