"""Many noisy circuits of different widths and depths, simulated together.

Narrow circuits (up to ``persistent.max_qubits()`` qubits, 7 in float32 on
MI355X) all run in ONE persistent kernel launch (forward) and one more
(adjoint): a workgroup per circuit, its Pauli vector in LDS, every width and
depth in the same launch (quantumsim_amd/persistent.py).  The PTMs of all
their gates are computed by torch from the circuit parameters in a handful
of batched ops (all layers of a parity at once) into one pool, so autograd
carries dL/dP from the adjoint kernel to the angles and T1/T2.

Wider circuits (whose states, 256 KB to 16 MB each, do not fit in LDS and
are streamed at HBM bandwidth anyway) are grouped by qubit count; inside a
group the states form one [B, 4^n] tensor ordered by decreasing depth, so
the circuits still running at any layer are a prefix of the batch.  Layers
are taken two at a time and their gates composed (torch, differentiable)
into n // 2 dense blocks of up to three qubits
(workloads.brickwork_window_plan), each one qsv_apply_ptm_batched launch
over that prefix (64 x 64 PTMs on the MFMA kernel): about half the passes
over the states of one launch per gate.  A circuit that ends inside a
two-layer window gets identity PTMs for its missing layer; circuits that
have finished are retired by narrowing the prefix, without copies.  A
group's schedule is one autograd node (autograd.run_schedule), whose adjoint
keeps the prefix's gradient in place and writes every block's dL/dP into
one gradient tensor per block position.
Forward, the float64 population readout and the adjoint
(checkpointed recompute + qsv_ptm_grad) are all device work with no host
synchronisation, so the whole step can be captured once in a HIP graph and
replayed (``CircuitBatch.capture``).  The checkpoint interval is chosen
from a memory budget: every intermediate state is kept when they fit (no
recompute: on the 512-circuit batch the states of the wide circuits take
~175 GB of the 288 GB HBM, and the step runs 0.31 s instead of 0.38 s with
a checkpoint every 16 blocks), else the shortest interval that fits.

The reference simulates one circuit per PauliVector object; this module has
no counterpart there.  Its per-gate arithmetic is the reference's
apply_ptm (quantumsim/pauli_vectors/numpy.py:50-64).
"""
import math

import numpy as np
import torch

from . import persistent
from .autograd import Gate, run_schedule, diagonal, shared_constant, cz_ptm
from .autotune import tuning
from .bases import general
from .fusion import fuse_batched, fuse_plan
from .workloads import (rotation_ptms, damping_ptms, fused_layer_gates, layer_targets,
                        brickwork_window_plan)
from .windows import WindowPlan, window_blocks

__all__ = ['CircuitBatch', 'plain_forward_backward', 'choose_checkpoint_interval']

CHECKPOINT_INTERVALS = (1, 2, 4, 8, 16)
# share of the free device memory the kept states may take by default
DEFAULT_BUDGET_FRACTION = 0.4


def choose_checkpoint_interval(kept_bytes, budget, intervals=CHECKPOINT_INTERVALS):
    """The shortest interval k of ``intervals`` whose kept states,
    ``kept_bytes(k)``, fit ``budget`` bytes (the longest if none does)."""
    for k in intervals:
        if kept_bytes(k) <= budget:
            return k
    return intervals[-1]


class _Group:
    def __init__(self, n, members, circuits, device, param_dtype):
        self.n = n
        self.members = members                  # indices into the circuit list
        L = [circuits[i]['layers'] for i in members]
        self.layers = L                         # decreasing
        Lmax = L[0]
        ang = np.zeros((len(members), Lmax, n, 2))
        t1 = np.zeros((len(members), n))
        t2 = np.zeros((len(members), n))
        for b, i in enumerate(members):
            c = circuits[i]
            ang[b, :c['layers']] = c['angles']
            t1[b], t2[b] = c['t1'], c['t2']
        mk = lambda a: torch.tensor(a, dtype=param_dtype, device=device)
        self.angles, self.t1, self.t2 = mk(ang), mk(t1), mk(t2)


class CircuitBatch:
    """Batched simulation of ``circuits`` (see workloads.random_circuit_batch).

    Parameters are the leaf tensors ``group.angles`` [B, L, n, 2] and
    ``group.t1``/``group.t2`` [B, n] (microseconds) of each group; the step
    returns float64 populations per circuit (divided by the circuit's
    float64 trace unless ``renormalize=False``) and, with a loss, their
    gradients with respect to every parameter."""

    def __init__(self, circuits, dtype=torch.float32, device=None, checkpoint_every=None,
                 requires_grad=True, persistent_kernel=True, fuse_qudits=3,
                 checkpoint_budget=None, renormalize=True, overlap_narrow=True,
                 block_kernel=True):
        self.device = device or torch.device('cuda', torch.cuda.current_device())
        self.dtype = dtype
        self.renormalize = renormalize
        self.traces = None
        # circuits in batched launches: layers taken two at a time and fused
        # into dense blocks of up to fuse_qudits qubits (3: 64 x 64 PTMs on
        # the MFMA kernels; 2 or less: the layer's own gates, one at a time)
        self.fuse_qudits = int(fuse_qudits or 0)
        self.window = 2 if self.fuse_qudits > 2 else 1
        # dense 3-qubit window blocks built by one kernel launch per group
        # (windows.window_blocks) instead of torch products
        self.block_kernel = bool(block_kernel)
        self._consts = {}
        self.circuits = circuits
        # one stream (default priority) for parameters, eager steps and the
        # captured graph
        self.stream = torch.cuda.Stream(device=self.device)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        # the narrow circuits' persistent kernels (one workgroup per CU,
        # latency-bound) and part of the batched groups run on a second
        # stream, also of default priority, beside the other groups'
        # launches (filling each other's launch gaps and kernel tails);
        # autograd runs each group's adjoint on the stream of its forward,
        # and the step joins the two
        self.side = torch.cuda.Stream(device=self.device) if overlap_narrow else None
        by_n = {}
        for i, c in enumerate(circuits):
            by_n.setdefault(c['n'], []).append(i)
        self.groups = []
        with torch.cuda.stream(self.stream):
            for n in sorted(by_n):
                members = sorted(by_n[n], key=lambda i: -circuits[i]['layers'])
                g = _Group(n, members, circuits, self.device, torch.float64)
                for t in (g.angles, g.t1, g.t2):
                    t.requires_grad_(requires_grad)
                self.groups.append(g)
        self.graph = None
        # narrow circuits: one persistent launch (tables built once here)
        self.lds_qubits = persistent.max_qubits(dtype) if persistent_kernel else 0
        self.small = [gi for gi, g in enumerate(self.groups) if g.n <= self.lds_qubits]
        # batched groups split over the two streams by their state traffic
        # (largest first, each to the stream with less so far; the first
        # stays on the main stream): on the benchmark batch the 11-qubit
        # group runs beside the 8-10-qubit ones
        self.side_groups = self._balance_streams() if overlap_narrow else []
        self.ptable = self._persistent_table() if self.small else None
        # checkpoint interval of the batched launches: None = the shortest
        # of CHECKPOINT_INTERVALS whose kept states fit checkpoint_budget
        # bytes (default: DEFAULT_BUDGET_FRACTION of the device memory free
        # now, leaving room for other users of the device; a caller that
        # owns the device passes a larger budget)
        if checkpoint_every is None:
            if checkpoint_budget is None:
                checkpoint_budget = DEFAULT_BUDGET_FRACTION * torch.cuda.mem_get_info(
                    self.device)[0]
            checkpoint_every = choose_checkpoint_interval(self.kept_bytes, checkpoint_budget)
        self.k = max(1, int(checkpoint_every))

    def _pool_layout(self):
        """(group index, parity, gate index within the layer, D, layers of
        that parity) of every piece of the PTM pool, in pool order."""
        out = []
        for gi in self.small:
            g = self.groups[gi]
            for p in (0, 1):
                Lp = len(range(p, g.layers[0], 2))
                if Lp == 0:
                    continue
                for j, t in enumerate(layer_targets(g.n, p)):
                    out.append((gi, p, j, 4 ** len(t), Lp))
        return out

    def _persistent_table(self):
        base, off = {}, 0
        for gi, p, j, D, Lp in self._pool_layout():
            base[(gi, p, j)] = (off, Lp)
            off += len(self.groups[gi].members) * Lp * D * D
        circuits = []
        for gi in self.small:
            g = self.groups[gi]
            targets = [layer_targets(g.n, p) for p in (0, 1)]
            for b, L in enumerate(g.layers):
                gates = []
                for l in range(L):
                    p, lp = l % 2, l // 2
                    for j, t in enumerate(targets[p]):
                        b0, Lp = base[(gi, p, j)]
                        D = 4 ** len(t)
                        gates.append((t, b0 + (b * Lp + lp) * D * D))
                circuits.append((g.n, gates))
        return persistent.CircuitTable(circuits, off, device=self.device)

    def _persistent_populations(self):
        """Populations of every narrow circuit: one forward launch over all
        of them; per group [B, 2^n] in member order."""
        pieces = []
        for gi in self.small:
            g = self.groups[gi]
            B, n = len(g.members), g.n
            rot = rotation_ptms(g.angles)            # [B, L, n, 4, 4]
            damp = damping_ptms(g.t1, g.t2)          # [B, n, 4, 4]
            for p in (0, 1):
                Lp = len(range(p, g.layers[0], 2))
                if Lp == 0:
                    continue
                rot_p = rot[:, p::2].reshape(B * Lp, n, 4, 4)
                damp_p = damp[:, None].expand(B, Lp, n, 4, 4).reshape(B * Lp, n, 4, 4)
                for gate in fused_layer_gates(n, p, rot_p, damp_p):
                    pieces.append(gate.ptm.reshape(-1))
        pool = torch.cat(pieces)
        pops = persistent.run(pool, self.ptable, self.dtype)
        out, pos = [], 0
        for gi in self.small:
            g = self.groups[gi]
            size = len(g.members) * 2 ** g.n
            out.append(pops[pos:pos + size].reshape(len(g.members), 2 ** g.n))
            pos += size
        return out

    def parameters(self):
        return [t for g in self.groups for t in (g.angles, g.t1, g.t2)]

    def _layer_ptms(self, g):
        """PTMs of every layer of group g, computed for all layers of a
        parity in one batch: per parity p, a list over the layer's gates of
        ([layers of that parity, B, D, D] in the state's element type,
        targets); layer l is step l // 2 of parity l % 2."""
        n, B = g.n, len(g.members)
        rot = rotation_ptms(g.angles)            # [B, L, n, 4, 4]
        damp = damping_ptms(g.t1, g.t2)          # [B, n, 4, 4]
        out = []
        for p in (0, 1):
            Lp = len(range(p, g.layers[0], 2))
            if Lp == 0:
                out.append([])
                continue
            rot_p = rot[:, p::2].reshape(B * Lp, n, 4, 4)
            damp_p = damp[:, None].expand(B, Lp, n, 4, 4).reshape(B * Lp, n, 4, 4)
            out.append([(gt.ptm.reshape(B, Lp, *gt.ptm.shape[1:]).transpose(0, 1)
                         .to(self.dtype).contiguous(), gt.targets)
                        for gt in fused_layer_gates(n, p, rot_p, damp_p)])
        return out

    def _cached(self, key, make):
        """Constant device tensors (masks, identities) made once, in an eager
        step, so a captured graph never copies from the host; readable from
        both streams once made."""
        t = self._consts.get(key)
        if t is None:
            t = shared_constant(make())
            self._consts[key] = t
        return t

    def _window_blocks(self, g):
        """Dense blocks of every two-layer window of group g, composed for
        all windows and circuits at once (the windows share one fusion plan):
        a list over block positions of ([n_windows, B, D, D] in the state's
        element type, block qubits), so window w's PTMs for the running
        prefix are a contiguous slice.  Layer 2w + 1 of a
        circuit with at most 2w + 1 layers (and the padding layer of an odd
        maximum depth) is the identity: its state is final after the
        window."""
        n, B = g.n, len(g.members)
        Lmax = g.layers[0]
        nw = (Lmax + 1) // 2
        ang = g.angles
        if 2 * nw > Lmax:                 # pad to whole windows (masked below)
            ang = torch.cat([ang, torch.zeros_like(ang[:, :1])], 1)
        rot = rotation_ptms(ang)                                  # [B, 2 nw, n, 4, 4]
        damp = damping_ptms(g.t1, g.t2)                           # [B, n, 4, 4]
        if self.block_kernel and self.fuse_qudits == 3:
            # every block of every window and circuit in one launch (and one
            # for the backward): qsv_window_blocks
            plan = self._consts.get(('wplan', id(g)))
            if plan is None:
                plan = WindowPlan(n, nw, B, self.device)
                shared_constant(plan.table_dev)
                self._consts[('wplan', id(g))] = plan
            keep8 = self._cached(('keep8', id(g)), lambda: torch.tensor(
                [[g.layers[b] > 2 * w + 1 for w in range(nw)] for b in range(B)],
                dtype=torch.uint8, device=self.device))
            cz = self._cached(('cz16',), lambda: cz_ptm(self.device).reshape(16, 16)
                              .contiguous())
            blocks = window_blocks(rot, damp, keep8, plan, self.dtype, cz)
            return [(m, qs) for m, (qs, _, _) in zip(blocks, plan.positions)]
        damp_w = damp[:, None].expand(B, nw, n, 4, 4).reshape(B * nw, n, 4, 4)
        keep = self._cached(('keep', id(g)), lambda: torch.tensor(
            [g.layers[b] > 2 * w + 1 for b in range(B) for w in range(nw)],
            dtype=torch.bool, device=self.device)[:, None, None])
        gates = []
        for par in (0, 1):
            rot_p = rot[:, par::2].reshape(B * nw, n, 4, 4)
            for gt in fused_layer_gates(n, par, rot_p, damp_w):
                ptm = gt.ptm
                if par == 1:
                    D = ptm.shape[-1]
                    eye = self._cached(('eye', D, ptm.dtype), lambda: torch.eye(
                        D, dtype=ptm.dtype, device=self.device))
                    ptm = torch.where(keep, ptm, eye)
                gates.append(Gate(ptm, gt.targets))
        out = []
        # the right-to-left brickwork tiling (n // 2 blocks of up to three
        # qubits) when blocks of three are allowed, else the greedy pass
        plan = brickwork_window_plan(n) if self.fuse_qudits == 3 else None
        for m, q in fuse_batched(gates, [4] * n, self.fuse_qudits, plan=plan):
            D = m.shape[-1]
            out.append((m.reshape(B, nw, D, D).transpose(0, 1).to(self.dtype).contiguous(), q))
        return out

    def _segments(self, g):
        """Runs of layers over which the running circuits of group g (a
        prefix of its depth-sorted batch) do not change, one segment of
        run_schedule each: [(first layer, end layer, running, still running
        after)]."""
        deeper = lambda layer: sum(1 for d in g.layers if d > layer)
        out = []
        layer, active = 0, len(g.members)
        while active > 0:
            first = layer
            while layer < g.layers[0] and deeper(layer) == active:
                layer += self.window
            nxt = deeper(layer)
            out.append((first, layer, active, nxt))
            active = nxt
        return out

    def _gates_in(self, g, first, end):
        """Applications the batched launches make for layers [first, end)."""
        n = g.n
        if self.window == 2:
            if self.fuse_qudits == 3:
                per = len(brickwork_window_plan(n))
            else:
                per = len(fuse_plan(layer_targets(n, 0) + layer_targets(n, 1), [4] * n,
                                    self.fuse_qudits))
            return (end - first) // 2 * per
        return sum(len(layer_targets(n, l % 2)) for l in range(first, end))

    def kept_bytes(self, k):
        """Device bytes the batched launches hold for the adjoint with a
        checkpoint every k applications (an estimate of the step's peak):
        each segment's input (the storage of the previous segment's output)
        and every k-th state, the final states, and for the group whose
        adjoint runs, its incoming gradient, two adjoint buffers and the
        states one checkpoint interval recomputes."""
        es = torch.empty((), dtype=self.dtype).element_size()
        total = extra = 0
        for gi, g in enumerate(self.groups):
            if gi in self.small:
                continue
            S, B = 4 ** g.n * es, len(g.members)
            total += B * S
            prev = B
            for first, end, active, _ in self._segments(g):
                G = self._gates_in(g, first, end)
                total += (prev + (G - 1) // k * active) * S
                prev = active
                extra = max(extra, ((min(k, G) - 1) * active + 3 * B) * S)
        return total + extra

    def _group_populations(self, g):
        n, B = g.n, len(g.members)
        if self.window == 2:
            blocks = self._window_blocks(g)
            steps = [m for m, _ in blocks]
            targets = [q for _, q in blocks]
            step_of = lambda layer: [(p, layer // 2) for p in range(len(blocks))]
        else:
            ptms = self._layer_ptms(g)
            steps = [P for par in ptms for P, _ in par]
            targets = [t for par in ptms for _, t in par]
            first = [0, len(ptms[0])]
            step_of = lambda layer: [(first[layer % 2] + i, layer // 2)
                                     for i in range(len(ptms[layer % 2]))]
        # one segment per run of layers over which the running circuits (a
        # prefix of the depth-sorted batch) do not change
        segments = []
        for lo, hi, active, nxt in self._segments(g):
            gates = [pw for layer in range(lo, hi, self.window) for pw in step_of(layer)]
            assert len(gates) == self._gates_in(g, lo, hi)
            segments.append((active, nxt, gates))
        x = torch.zeros((B, 4 ** n), dtype=self.dtype, device=self.device)
        x[:, 0] = 1
        states = run_schedule(x, steps, targets, segments, (4,) * n, self.k)
        return diagonal(states, [general(2)] * n)

    def populations(self):
        """float64 populations per group ([B_n, 2^n] tensors, depth-sorted).

        With ``renormalize`` (the default) each circuit's populations are
        divided by their float64 sum, the state's trace, when it exceeds
        1e-8 (quantumsim/pauli_vectors/numpy.py:109-117, without the
        warning, which would need a host synchronisation): the float32
        states' rounding drift of the trace is removed in float64.  The
        traces before normalisation ([B_n] float64 per group, what the
        reference's renormalize returns) are kept in ``self.traces``."""
        out = [None] * len(self.groups)
        cur = torch.cuda.current_stream(self.device)
        side = self.side if self._side_groups() else None
        if side is not None:
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                small = self._persistent_populations() if self.small else []
                for gi in self.side_groups:
                    out[gi] = self._group_populations(self.groups[gi])
        else:
            small = self._persistent_populations() if self.small else []
        for gi, g in enumerate(self.groups):
            if gi not in self.small and out[gi] is None:
                out[gi] = self._group_populations(g)
        for gi, p in zip(self.small, small):
            out[gi] = p
        if side is not None:
            cur.wait_stream(side)
            for gi in self._side_groups():
                out[gi].record_stream(cur)
        self.traces = [p.detach().sum(-1) for p in out]
        if self.renormalize:
            out = [torch.where(t[:, None] > 1e-8, p / p.sum(-1, keepdim=True), p)
                   for p, t in zip(out, self.traces)]
        return out

    def _step(self, weights):
        pops = self.populations()
        loss = sum((p * w).sum() for p, w in zip(pops, weights))
        grads = torch.autograd.grad(loss, self.parameters())
        if self.side is not None and self._side_groups():
            # the side stream's groups ran their adjoint there: join it
            cur = torch.cuda.current_stream(self.device)
            cur.wait_stream(self.side)
            for gi in self._side_groups():
                for t in grads[3 * gi:3 * gi + 3]:
                    t.record_stream(cur)
        return loss, pops, grads

    def _balance_streams(self):
        """Batched groups for the side stream: longest-processing-time
        assignment of the groups' state traffic (sum over circuits of
        windows x 4^n) to the two streams, the largest on the main one."""
        cost = {gi: sum((d + 1) // 2 for d in g.layers) * 4 ** g.n
                for gi, g in enumerate(self.groups) if gi not in self.small}
        load, side = [0, 0], []
        for gi in sorted(cost, key=lambda gi: -cost[gi]):
            s = 0 if load[0] <= load[1] else 1
            load[s] += cost[gi]
            if s:
                side.append(gi)
        return sorted(side)

    def _side_groups(self):
        """Groups whose work runs on the side stream: the persistent ones
        and the batched ones of small states (launch-bound)."""
        return (list(self.small) + list(self.side_groups)) if self.side is not None else []

    def step(self, weights):
        """Forward + adjoint of loss = sum over circuits of <w_c, p_c>;
        returns (loss, populations, gradients of every parameter)."""
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            out = self._step(weights)
        cur.wait_stream(self.stream)
        return out

    def random_weights(self, seed):
        g = np.random.default_rng(seed)
        return [torch.tensor(g.standard_normal((len(gr.members), 2 ** gr.n)),
                             dtype=torch.float64, device=self.device) for gr in self.groups]

    def capture(self, weights, autotune=True, ledger=None):
        """Record one forward + adjoint step in a HIP graph (after an eager
        warm-up step that sizes every buffer and, with ``autotune``, times
        the launch shapes of every batched apply and PTM gradient of the
        step: see quantumsim_amd.autotune); replay() re-runs it.  A
        ``ledger.Ledger`` given here counts (and times) the warm-up step's
        launches: the step the graph records, with the tuned shapes."""
        if autotune:
            with tuning():
                self.step(weights)               # tunes each new geometry once
            torch.cuda.synchronize(self.device)
        if ledger is not None:
            with ledger:
                self.step(weights)               # warm-up, counted
        else:
            self.step(weights)                   # warm-up: allocations, tables
        torch.cuda.synchronize(self.device)
        # the graph allocates from a pool of its own: hand the blocks the
        # eager steps cached back first (as large as the graph's own needs
        # when every state is kept)
        torch.cuda.empty_cache()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.g_loss, self.g_pops, self.g_grads = self._step(weights)
            self.g_traces = self.traces
        torch.cuda.synchronize(self.device)
        return self.graph

    def replay(self):
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            self.graph.replay()
        cur.wait_stream(self.stream)
        return self.g_loss, self.g_pops, self.g_grads


def plain_forward_backward(circuit, weight, dtype=torch.float32, device=None,
                           checkpoint_every=None, fuse_qudits=3):
    """One circuit on its own (batch of 1, eager launches, the same gate
    fusion and kernels): the per-circuit baseline the batched engine is
    compared with."""
    device = device or torch.device('cuda', torch.cuda.current_device())
    b = CircuitBatch([circuit], dtype=dtype, device=device, checkpoint_every=checkpoint_every,
                     persistent_kernel=False, fuse_qudits=fuse_qudits)
    return b.step([weight])
