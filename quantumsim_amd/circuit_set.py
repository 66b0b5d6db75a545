"""Many circuits of any qudit shapes -- qubits, qutrits, mixed -- simulated
together, differentiable in every gate's PTM.

A circuit is its per-qudit bases (as PauliVector takes them: ``dim_pauli``
4 for a qubit in general(2), 9 for a qutrit in general(3) or gell_mann(3),
...) and a gate list ``[(ptm, targets), ...]``: ``ptm`` a square tensor
[D, D] (D = prod dim_pauli of the targets; row (out...), column (in...),
the flattened form of the PTM PauliVector.apply_ptm takes), on the host or
the device, possibly requiring grad.

``CircuitSet`` groups the circuits by qudit shape and, inside a shape, by
gate structure: a circuit joins the batch of a longer circuit whose target
sequence starts with its own.  Each batch is ordered by decreasing depth,
so the circuits still running at gate position p are a prefix of it, and
runs through autograd.run_schedule: one batched launch per gate position
(qsv_apply_ptm_batched -- the inline kernels up to 16 x 16, the MFMA kernel
for dense 64 x 64, the offset-table kernel for any other size up to 1024
entries a side: a qutrit pair is 81 x 81, a qutrit + qubit pair 36 x 36)
and, backward, the adjoint apply plus qsv_ptm_grad (MFMA up to 64 x 64, the
LDS-staged float64 kernel above that, up to 8192 PTM entries; a larger
gate that requires grad is an error when the backward reaches it).  Shorter
circuits retire by narrowing the prefix; a circuit that shares no prefix
with any other is a batch of one.

Initial states are the reference's default (entry 0 of every qudit, the
ground state in the general bases: quantumsim/pauli_vectors/numpy.py:39-41)
unless the caller gives them.

The reference simulates one circuit per PauliVector object and has no
batching or differentiation; the per-gate arithmetic is its apply_ptm
(quantumsim/pauli_vectors/numpy.py:50-64).
"""
import math

import numpy as np
import torch

from .autograd import run_schedule, diagonal

__all__ = ['CircuitSet', 'plan_batches', 'MAX_PTM_SIDE', 'MAX_GRAD_ENTRIES']

# largest PTM side of the batched apply (qsv_apply_ptm_batched), and the
# most PTM entries whose gradient qsv_ptm_grad forms (a 90 x 90 PTM; three
# qutrits, 729 x 729, run forward only)
MAX_PTM_SIDE = 1024
MAX_GRAD_ENTRIES = 8192


def _targets_of(gates):
    return tuple(tuple(int(q) for q in t) for _, t in gates)


def plan_batches(shapes, structures):
    """Group circuits for batched runs.

    shapes[i]: circuit i's tuple of Pauli dimensions; structures[i]: its
    tuple of gate target tuples.  Returns a list of (shape, structure,
    members): members (circuit indices) by decreasing depth, each member's
    structure a prefix of ``structure`` (the longest member's)."""
    by_shape = {}
    for i, s in enumerate(shapes):
        by_shape.setdefault(tuple(s), []).append(i)
    out = []
    for shape in sorted(by_shape, key=lambda s: (len(s), s)):
        idx = sorted(by_shape[shape], key=lambda i: -len(structures[i]))
        batches = []
        for i in idx:
            st = structures[i]
            for b in batches:
                if b[1][:len(st)] == st:
                    b[2].append(i)
                    break
            else:
                batches.append((shape, st, [i]))
        out += batches
    return out


def segments_of(depths):
    """run_schedule segments of a batch with member depths ``depths``
    (decreasing): gate positions are applied while the number of circuits
    that reach them stays the same; the circuits that end there are final."""
    B = len(depths)
    segs, pos, active = [], 0, B
    while active > 0:
        end = depths[active - 1]
        cont = sum(1 for d in depths if d > end)
        segs.append((active, cont, [(p, 0) for p in range(pos, end)]))
        pos, active = end, cont
    return segs


class CircuitSet:
    """Circuits of any qudit shapes, run together (see the module text).

    circuits: sequence of (bases, gates).  dtype: the states' element type
    (float64 or float32; PTMs are cast to it, gradients flow back in the
    PTMs' own type).  checkpoint_every: the recompute interval of the
    adjoint (1: keep every intermediate state)."""

    def __init__(self, circuits, dtype=torch.float64, device=None, checkpoint_every=1):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError('dtype must be torch.float32 or torch.float64')
        self.device = torch.device(device) if device is not None else torch.device(
            'cuda', torch.cuda.current_device())
        self.dtype = dtype
        self.k = max(1, int(checkpoint_every))
        self.bases, self.gates = [], []
        for c, (bases, gates) in enumerate(circuits):
            bases = list(bases)
            dims = [b.dim_pauli for b in bases]
            gates = list(gates)
            for g, (ptm, targets) in enumerate(gates):
                targets = tuple(int(q) for q in targets)
                if not targets or len(set(targets)) != len(targets) or \
                        not all(0 <= q < len(dims) for q in targets):
                    raise ValueError('circuit %d gate %d: bad targets %s' % (c, g, targets))
                D = math.prod(dims[q] for q in targets)
                if D > MAX_PTM_SIDE:
                    raise ValueError('circuit %d gate %d: a %d x %d PTM is above the batched '
                                     'kernels\' %d entries a side' % (c, g, D, D, MAX_PTM_SIDE))
                if tuple(ptm.shape) not in ((D, D), tuple(dims[q] for q in targets) * 2):
                    raise ValueError('circuit %d gate %d: PTM of shape %s on qudits %s '
                                     '(Pauli dimension %d)' % (c, g, tuple(ptm.shape),
                                                               targets, D))
            self.bases.append(bases)
            self.gates.append(gates)
        shapes = [tuple(b.dim_pauli for b in bs) for bs in self.bases]
        self.batches = plan_batches(shapes, [_targets_of(g) for g in self.gates])

    def __len__(self):
        return len(self.gates)

    def _initial(self, members, shape, initial):
        N = math.prod(shape)
        x = torch.zeros((len(members), N), dtype=self.dtype, device=self.device)
        for b, i in enumerate(members):
            if initial is None or initial[i] is None:
                x[b, 0] = 1.0
            else:
                v = initial[i]
                v = v if torch.is_tensor(v) else torch.as_tensor(np.asarray(v))
                if v.numel() != N:
                    raise ValueError('initial state %d has %d entries, not %d' % (i, v.numel(), N))
                x[b] = v.reshape(-1).to(device=self.device, dtype=self.dtype)
        return x

    def run(self, initial=None):
        """Final Pauli vectors, one [prod dim_pauli] tensor per circuit in
        input order (views into the batches' outputs), differentiable in
        every gate PTM and in ``initial`` (a sequence of per-circuit states or
        None entries, each flat or shaped as dim_pauli)."""
        if initial is not None and len(initial) != len(self.gates):
            raise ValueError('one initial state (or None) per circuit')
        out = [None] * len(self.gates)
        for shape, structure, members in self.batches:
            B = len(members)
            depths = [len(self.gates[i]) for i in members]
            x0 = self._initial(members, shape, initial)
            steps = []
            for p, tg in enumerate(structure):
                D = math.prod(shape[q] for q in tg)
                rows = [self.gates[i][p][0].to(device=self.device, dtype=self.dtype)
                        .reshape(D, D) for i, d in zip(members, depths) if d > p]
                if len(rows) < B:
                    rows.append(torch.zeros((B - len(rows), D, D), dtype=self.dtype,
                                            device=self.device))
                    steps.append(torch.cat([torch.stack(rows[:-1]), rows[-1]])[None])
                else:
                    steps.append(torch.stack(rows)[None])
            fin = run_schedule(x0, [s.contiguous() for s in steps], list(structure),
                               segments_of(depths), shape, self.k)
            for b, i in enumerate(members):
                out[i] = fin[b]
        return out

    def populations(self, states):
        """Diagonals of the density matrices of ``states`` (run's output):
        float64 [prod dim_hilbert] per circuit, differentiable (gather
        bases: general(d) and its subbases)."""
        return [diagonal(s.reshape(1, -1), bs)[0] for s, bs in zip(states, self.bases)]
