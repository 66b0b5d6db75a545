"""Gate fusion: turn a gate list into fewer, denser PTM applications.

A circuit applied gate by gate reads and writes the whole Pauli vector once
per gate.  Runs of gates that together act on at most ``max_qudits``
qudits compose into one dense PTM (the reference builds such PTMs with
``_Chain.ptm``, quantumsim/operations/operation.py:610-623, and applies a
3-qubit one in test_operations.py:174-195), applied in a single pass: the
16x16 streaming kernel for 2 qudits, the MFMA kernel for 3-4 qubits.

The pass keeps several open blocks on disjoint qudit sets (gates on
disjoint qudits commute, so their blocks can stay open side by side).  A
gate touching open blocks joins the largest subset of them whose union with
its own qudits fits ``max_qudits`` (ties: the subset holding more gates);
the touched blocks left out are emitted first (they act on a qudit of the
gate, so they must precede it).  At the end every open block is emitted.
Square PTMs (same input and output dimension per qudit) only.
"""
import itertools
import numpy as np

__all__ = ['fuse', 'fuse_plan', 'FusedGate', 'embed_ptm', 'embed_ptm_batched', 'fuse_batched',
           'apply_circuit']


class FusedGate:
    """A dense PTM over ``qudits`` (ascending), layout (out..., in...);
    ``sources``: indices of the input gates it is the product of, in
    application order."""
    __slots__ = ('ptm', 'qudits', 'n_source', 'sources')

    def __init__(self, ptm, qudits, n_source, sources=()):
        self.ptm = ptm
        self.qudits = tuple(qudits)
        self.n_source = n_source
        self.sources = tuple(sources)

    def __iter__(self):          # (ptm, qudits) unpacking, like the inputs
        yield self.ptm
        yield self.qudits


def embed_ptm(ptm, qudits, block, dims):
    """PTM of ``ptm`` (acting on ``qudits`` in that order) as a square
    matrix over ``block`` (sorted), identity on the other block qudits."""
    k = len(qudits)
    pos = [block.index(q) for q in qudits]
    m = len(block)
    din = [dims[q] for q in block]
    p = np.asarray(ptm, dtype=np.float64).reshape([dims[q] for q in qudits] * 2)
    # einsum labels: output axes 0..m-1, input axes m..2m-1
    args = [p, [pos[j] for j in range(k)] + [m + pos[j] for j in range(k)]]
    for b in range(m):
        if b not in pos:
            args += [np.eye(din[b]), [b, m + b]]
    out = np.einsum(*args, list(range(2 * m)), optimize=True)
    D = int(np.prod(din))
    return out.reshape(D, D)


class _Block:
    __slots__ = ('qudits', 'mat', 'n', 'src')

    def __init__(self, qudits, mat, n, src):
        self.qudits = qudits      # sorted tuple
        self.mat = mat            # D x D in the block's qudit order
        self.n = n
        self.src = src            # source gate indices, application order


def _merge(blocks, dims):
    """Blocks on disjoint qudits -> one block (their PTMs commute)."""
    qs = tuple(sorted(q for b in blocks for q in b.qudits))
    mat = np.eye(int(np.prod([dims[q] for q in qs])))
    for b in blocks:
        mat = embed_ptm(b.mat.reshape([dims[q] for q in b.qudits] * 2), b.qudits, qs,
                        dims) @ mat
    return _Block(qs, mat, sum(b.n for b in blocks), [i for b in blocks for i in b.src])


def fuse(gates, dims, max_qudits=2):
    """Fuse ``gates`` = [(ptm, qudits), ...] on a vector with per-qudit Pauli
    dimensions ``dims``.  Returns a list of FusedGate whose sequential
    application equals that of ``gates``."""
    open_blocks = {}              # qudit -> block
    out = []

    def emit(b):
        shape = [dims[q] for q in b.qudits] * 2
        out.append(FusedGate(b.mat.reshape(shape), b.qudits, b.n, b.src))
        for q in b.qudits:
            del open_blocks[q]

    for gi, (ptm, qudits) in enumerate(gates):
        qudits = tuple(int(q) for q in qudits)
        p = np.asarray(ptm)
        k = len(qudits)
        if p.shape[:k] != p.shape[k:] or any(p.shape[j] != dims[q] for j, q in enumerate(qudits)):
            raise ValueError('fuse() needs square PTMs matching the qudit dimensions')
        touched = []
        for q in qudits:
            b = open_blocks.get(q)
            if b is not None and b not in touched:
                touched.append(b)
        # largest subset of the touched blocks that fits together with the gate
        best, best_key = (), None
        for r in range(len(touched), -1, -1):
            for sub in itertools.combinations(touched, r):
                u = set(qudits).union(*(b.qudits for b in sub))
                if len(u) > max_qudits and sub:
                    continue
                key = (len(u) if sub else 0, sum(b.n for b in sub))
                if best_key is None or key > best_key:
                    best, best_key = sub, key
        for b in touched:
            if b not in best:
                emit(b)
        union = set(qudits).union(*(b.qudits for b in best))
        qs = tuple(sorted(union))
        if best:
            base = _merge(list(best), dims) if len(best) > 1 else best[0]
            for b in best:
                for q in b.qudits:
                    open_blocks.pop(q, None)
            prev = base.mat
            if base.qudits != qs:
                prev = embed_ptm(prev.reshape([dims[q] for q in base.qudits] * 2),
                                 base.qudits, qs, dims)
            nb = _Block(qs, embed_ptm(p, qudits, qs, dims) @ prev, base.n + 1, base.src + [gi])
        else:
            nb = _Block(qs, embed_ptm(p, qudits, qs, dims), 1, [gi])
        for q in nb.qudits:
            open_blocks[q] = nb
    seen = []
    for b in open_blocks.values():
        if b not in seen:
            seen.append(b)
    for b in seen:
        emit(b)
    return out


_PLANS = {}
_EYES = {}


def fuse_plan(qudit_sets, dims, max_qudits=3):
    """The blocks ``fuse`` forms for gates on ``qudit_sets`` (the grouping
    depends only on which qudits the gates touch): [(block qudits, source
    gate indices in application order), ...].  Cached."""
    key = (tuple(tuple(int(q) for q in t) for t in qudit_sets), tuple(dims), max_qudits)
    plan = _PLANS.get(key)
    if plan is None:
        probe = [(np.eye(int(np.prod([dims[q] for q in t]))).reshape([dims[q] for q in t] * 2), t)
                 for t in key[0]]
        plan = [(f.qudits, f.sources) for f in fuse(probe, dims, max_qudits)]
        _PLANS[key] = plan
    return plan


def embed_ptm_batched(ptm, qudits, block, dims):
    """Torch form of ``embed_ptm`` for a batch: ptm [B, D, D] acting on
    ``qudits`` -> [B, Db, Db] over ``block`` (sorted), identity elsewhere."""
    import torch
    qudits, block = tuple(qudits), tuple(block)
    if qudits == block:
        return ptm
    k, m = len(qudits), len(block)
    pos = [block.index(q) for q in qudits]
    B = ptm.shape[0]
    p = ptm.reshape([B] + [dims[q] for q in qudits] * 2)
    # labels: batch 0, outputs 1..m, inputs m+1..2m
    ops = [p, [0] + [1 + pos[j] for j in range(k)] + [1 + m + pos[j] for j in range(k)]]
    for b in range(m):
        if b not in pos:
            key = (dims[block[b]], ptm.dtype, ptm.device)
            eye = _EYES.get(key)
            if eye is None:
                from .autograd import shared_constant
                eye = _EYES[key] = shared_constant(
                    torch.eye(key[0], dtype=ptm.dtype, device=ptm.device))
            ops += [eye, [1 + b, 1 + m + b]]
    out = torch.einsum(*ops, list(range(2 * m + 1)))
    D = int(np.prod([dims[q] for q in block]))
    return out.reshape(B, D, D)


def fuse_batched(gates, dims, max_qudits=3, plan=None):
    """Fuse batched gates (objects with ``.ptm`` [B or 1, D, D] torch
    tensors and ``.targets``) into dense blocks of up to ``max_qudits``
    qudits with torch ops, so autograd reaches the source PTMs.  ``plan``:
    [(block qudits, source indices in order), ...] to use instead of
    ``fuse_plan``'s.  Returns [(ptm [B, Db, Db], block qudits), ...] in
    application order."""
    if plan is None:
        plan = fuse_plan([g.targets for g in gates], dims, max_qudits)
    out = []
    for block, src in plan:
        mat = None
        for i in src:
            e = embed_ptm_batched(gates[i].ptm, gates[i].targets, block, dims)
            mat = e if mat is None else e @ mat
        out.append((mat, block))
    return out


def apply_circuit(pv, gates, max_qudits=3):
    """Apply ``gates`` = [(ptm, qudits), ...] (PTMs in ``pv``'s bases) to the
    Pauli vector ``pv``, fused into blocks of up to ``max_qudits`` qudits
    (0 or 1: gate by gate).  Returns the number of PTM applications made."""
    if max_qudits is None or max_qudits <= 1:
        for ptm, qudits in gates:
            pv.apply_ptm(np.asarray(ptm), *qudits)
        return len(gates)
    fused = fuse(gates, list(pv.dim_pauli), max_qudits)
    for f in fused:
        pv.apply_ptm(f.ptm, *f.qudits)
    return len(fused)
