"""Seeded synthetic circuits of the benchmark configurations.

A circuit is a list of ``(ptm, qubits)`` with float64 PTMs in the
``general(2)`` basis (layout (out..., in...)), ready for ``apply_ptm``.

* ``noisy_layers``: per layer a Haar-random 1-qubit unitary on every qubit,
  CZ on a set of disjoint pairs, then amplitude + phase damping of every
  qubit for a 20 ns gate with T1 ~ U(20, 40) us, T2 ~ U(10, T1) us drawn once
  per qubit.
  - ``pairing='random'``: a fresh random disjoint pairing per layer (6-qubit,
    40-layer case);
  - ``pairing='brickwork'``: nearest-neighbour CZ on a line, bonds (0,1),
    (2,3), ... on even layers and (1,2), (3,4), ... on odd layers (14-qubit,
    60-layer case).
"""
import numpy as np

from .algebra.algebra import kraus_to_ptm
from .bases import general
from .operations import qubits as gates

__all__ = ['haar_unitary', 'noisy_layers', 'layer_slices', 'damping_ptm', 'circuit_bytes',
           'random_kraus', 'mixed_qudit_program']

_B1 = (general(2),)
GATE_TIME = 20e-9


def haar_unitary(d, rng):
    """Haar-distributed unitary: QR of a complex Ginibre matrix with the
    phase correction of Mezzadri, 'How to generate random matrices from the
    classical compact groups' (2007)."""
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r) / np.abs(np.diagonal(r))
    return q * ph


def damping_ptm(t1, t2, t=GATE_TIME):
    g, lam = gates.t1t2_rates(t, t1, t2)
    return gates.amp_phase_damping(g, lam).ptm


def noisy_layers(n_qubits, n_layers, seed, pairing='brickwork'):
    """Returns (circuit, meta) -- see the module docstring."""
    rng = np.random.default_rng(seed)
    t1 = rng.uniform(20e-6, 40e-6, size=n_qubits)
    t2 = np.array([rng.uniform(10e-6, t1[q]) for q in range(n_qubits)])
    damp = [damping_ptm(t1[q], t2[q]) for q in range(n_qubits)]
    cz = gates.cphase(np.pi).ptm
    circuit = []
    layer_starts = []
    for layer in range(n_layers):
        layer_starts.append(len(circuit))
        for q in range(n_qubits):
            u = haar_unitary(2, rng)
            circuit.append((kraus_to_ptm(u[None], _B1, _B1), (q,)))
        if pairing == 'random':
            perm = rng.permutation(n_qubits)
            pairs = [tuple(sorted((int(perm[2 * i]), int(perm[2 * i + 1]))))
                     for i in range(n_qubits // 2)]
        elif pairing == 'brickwork':
            start = layer % 2
            pairs = [(q, q + 1) for q in range(start, n_qubits - 1, 2)]
        else:
            raise ValueError(pairing)
        for pair in pairs:
            circuit.append((cz, pair))
        for q in range(n_qubits):
            circuit.append((damp[q], (q,)))
    meta = {'t1': t1, 't2': t2, 'layer_starts': layer_starts, 'seed': seed,
            'n_qubits': n_qubits, 'n_layers': n_layers, 'pairing': pairing}
    return circuit, meta


def layer_slices(meta, n_gates):
    starts = list(meta['layer_starts']) + [n_gates]
    return [slice(a, b) for a, b in zip(starts[:-1], starts[1:])]


def circuit_bytes(circuit, n_qubits, elem_bytes, dim=4):
    """Algorithmic HBM bytes of applying ``circuit`` gate by gate to an
    n-qubit full-basis vector: (N_in + N_out) * s per gate (PTM bytes are
    negligible and ignored)."""
    n = dim ** n_qubits
    return len(circuit) * 2 * n * elem_bytes


def random_kraus(dim, n_kraus, rng):
    """n_kraus Kraus operators of a random CPTP map on a dim-level system:
    a Haar-like random isometry dim -> n_kraus * dim (QR of a complex
    Gaussian matrix) cut into n_kraus blocks, so sum_k K_k^dag K_k = 1."""
    z = rng.standard_normal((n_kraus * dim, dim)) + 1j * rng.standard_normal((n_kraus * dim, dim))
    q, r = np.linalg.qr(z)
    q = q * (np.diagonal(r) / np.abs(np.diagonal(r)))
    return q.reshape(n_kraus, dim, dim)


def mixed_qudit_program(dims, n_steps, seed):
    """Mixed qubit/qutrit workload (10 qudits, 6 of them qutrits, in the
    benchmark configuration).  Qutrits live in the Gell-Mann basis, qubits
    in the full qubit basis until measured, after which they are reduced to
    their diagonal subbasis.  Each step is a dict with
      pair    : the two qudits of a random 2-qudit CPTP map (4 Kraus ops)
      kraus   : its Kraus operators on the pair's Hilbert space
      measure : a qubit whose populations are read out, then whose basis is
                reduced to subbasis([0, 1]) (a 0/1 selection PTM)
      keep    : three qudits for partial_trace
      take    : per-qudit index lists (for those three qudits) for multitake
    The PTMs depend on the bases at each point and are built by the runner
    with kraus_to_ptm."""
    rng = np.random.default_rng(seed)
    n = len(dims)
    qubits = [q for q in range(n) if dims[q] == 2]
    steps = []
    for _ in range(n_steps):
        pair = tuple(int(x) for x in rng.choice(n, 2, replace=False))
        D = dims[pair[0]] * dims[pair[1]]
        keep = tuple(sorted(int(x) for x in rng.choice(n, 3, replace=False)))
        steps.append({
            'pair': pair,
            'kraus': random_kraus(D, 4, rng),
            'measure': int(rng.choice(qubits)),
            'keep': keep,
            'take': [sorted(int(x) for x in rng.choice(dims[q] ** 2, 2, replace=False))
                     for q in keep],
        })
    return steps


# ---------------------------------------------------------------------------
# Differentiable, batched noisy circuits (torch, on the device)
#
# A layer is: RX(a) then RZ(b) on every qubit, CZ on the brickwork bonds of
# that layer, then amplitude + phase damping of every qubit for GATE_TIME.
# Gates are fused per layer before they reach the device: each bond becomes
# one 2-qubit PTM (D (x) D) CZ (R (x) R), every qubit left out of a bond one
# 1-qubit PTM D R, so a layer costs n/2 (+1) passes over the state instead
# of 2.5 n.

def brickwork_pairs(n, layer):
    return [(q, q + 1) for q in range(layer % 2, n - 1, 2)]


def layer_targets(n, layer):
    """Targets of the fused gates of one layer, in fused_layer_gates order:
    the brickwork bonds, then every qubit left out of a bond."""
    pairs = brickwork_pairs(n, layer)
    paired = set(q for pq in pairs for q in pq)
    return [tuple(pq) for pq in pairs] + [(q,) for q in range(n) if q not in paired]


def brickwork_window_plan(n):
    """Dense blocks of two consecutive brickwork layers (the gates of
    layer_targets(n, 0) followed by those of layer_targets(n, 1), indexed in
    that order): [(block qubits, source gate indices in application order),
    ...] in application order.

    Block k covers qubits 2k, 2k+1, 2k+2 and holds the first layer's bond
    (2k, 2k+1) and the second layer's bond (2k+1, 2k+2) (plus the single-qubit
    gates at the ends); the blocks run from the highest k down, so the
    second-layer bond (2k+1, 2k+2) comes after the first-layer bond
    (2k+2, 2k+3) of block k+1 that shares its qubit.  ceil((n-1)/2) blocks,
    one fewer than the greedy fusion pass for odd n."""
    ta, tb = layer_targets(n, 0), layer_targets(n, 1)
    ia = {t: i for i, t in enumerate(ta)}
    ib = {t: len(ta) + i for i, t in enumerate(tb)}
    blocks = []
    for k in reversed(range(max(1, n // 2))):
        qs = [q for q in (2 * k, 2 * k + 1, 2 * k + 2) if q < n]
        src = []
        for t in ((2 * k, 2 * k + 1), (2 * k + 2,)):
            if t in ia:
                src.append(ia[t])
        for t in ((2 * k + 1, 2 * k + 2),):
            if t in ib:
                src.append(ib[t])
        for q in qs:
            if (q,) in ib and (q == 0 or q == n - 1):
                src.append(ib[(q,)])
        blocks.append((tuple(qs), src))
    return blocks


def fused_layer_gates(n, layer, rot, damp):
    """Gates of one layer for a batch: rot, damp [B, n, 4, 4] PTMs."""
    from .autograd import Gate, cz_ptm, kron_ptm
    cz = cz_ptm(rot.device).to(rot.dtype)
    # one view per qubit (unbind: a single stack in the backward instead of
    # a zero-filled full-size gradient per selected qubit)
    rq, dq = rot.unbind(1), damp.unbind(1)
    gates = []
    for t in layer_targets(n, layer):
        if len(t) == 2:
            q, q2 = t
            pre = kron_ptm(rq[q], rq[q2])
            post = kron_ptm(dq[q], dq[q2])
            gates.append(Gate(post @ cz @ pre, t))
        else:
            q = t[0]
            gates.append(Gate(dq[q] @ rq[q], t))
    return gates


_ROT_TERMS = {}


def _rotation_terms(device):
    """M [9, 16] with PTM(RZ(b) RX(a)) = sum over i, j of f_j(b) f_i(a)
    M[3 j + i] for f = (1, cos, sin).  A one-qubit rotation's PTM is a
    first-order trigonometric polynomial of its angle (its Kraus operator's
    entries are cos(t/2), sin(t/2) or e^(+-it/2), and the PTM is quadratic
    in them), so P(t) = A + B cos t + C sin t with A = (P(0) + P(pi)) / 2,
    B = (P(0) - P(pi)) / 2, C = P(pi/2) - A, read off the Kraus-built PTMs
    (autograd.rx_ptm / rz_ptm) once per device."""
    M = _ROT_TERMS.get(device)
    if M is None:
        import torch
        from .autograd import rx_ptm, rz_ptm
        t = torch.tensor([0.0, np.pi, np.pi / 2], dtype=torch.float64, device=device)

        def terms(P):
            A = (P[0] + P[1]) / 2
            return [A, (P[0] - P[1]) / 2, P[2] - A]
        X, Z = terms(rx_ptm(t)), terms(rz_ptm(t))
        M = torch.stack([(Z[j] @ X[i]).reshape(16) for j in range(3) for i in range(3)])
        from .autograd import shared_constant
        _ROT_TERMS[device] = shared_constant(M)
    return M


def rotation_ptms(angles):
    """angles [..., 2] (RX angle, RZ angle) -> PTMs [..., 4, 4] of RZ RX,
    as one real product with the nine trigonometric terms (no complex
    Kraus arithmetic per angle; equal to rz_ptm(b) @ rx_ptm(a) to rounding)."""
    import torch
    shp = angles.shape[:-1]
    a = angles.reshape(-1, 2)
    one = torch.ones_like(a[:, 0])
    fx = torch.stack([one, torch.cos(a[:, 0]), torch.sin(a[:, 0])], -1)
    fz = torch.stack([one, torch.cos(a[:, 1]), torch.sin(a[:, 1])], -1)
    coef = (fz[:, :, None] * fx[:, None, :]).reshape(-1, 9)
    M = _rotation_terms(angles.device).to(angles.dtype)
    return (coef @ M).reshape(*shp, 4, 4)


def damping_ptms(t1, t2, t=GATE_TIME * 1e6):
    """t1, t2 [..., n] in microseconds -> PTMs [..., n, 4, 4]."""
    from .autograd import damping_ptm
    shp = t1.shape
    p = damping_ptm(t1.reshape(-1), t2.reshape(-1), t)
    return p.reshape(*shp, 4, 4)


class TrainableCircuit:
    """12 qubits, 40 layers, 96 trainable parameters per parameter set:
    RX/RZ angles shared by the layers of a 5-layer block, three blocks used
    cyclically (block (layer // 5) mod 3: 3 x 12 x 2 = 72 angles), plus T1
    and T2 of every qubit (24).  (Eight independent blocks plus T1/T2 would
    be 216 parameters, not the 96 of the benchmark; the cyclic three-block
    split is the choice made here.)  The loss is KL(target || simulated) over
    the 2^n populations; the target comes from the same circuit with
    seeded, perturbed parameters."""

    def __init__(self, n=12, layers=40, block=5, n_blocks=3):
        self.n, self.layers, self.block, self.n_blocks = n, layers, block, n_blocks
        self.n_params = n_blocks * n * 2 + 2 * n

    def initial_params(self, sets, seed, device, dtype):
        import torch
        g = np.random.default_rng(seed)
        ang = g.uniform(0, 2 * np.pi, size=(sets, self.n_blocks, self.n, 2))
        t1 = g.uniform(20, 40, size=(sets, self.n))
        t2 = np.array([[g.uniform(10, t1[s, q]) for q in range(self.n)] for s in range(sets)])
        mk = lambda a: torch.tensor(a, dtype=dtype, device=device)
        return mk(ang), mk(t1), mk(t2)

    def gates(self, ang, t1, t2, fuse_qudits=0):
        """The circuit's gates (per layer: workloads.fused_layer_gates);
        with fuse_qudits >= 3 composed into dense blocks of up to that many
        qubits (fusion.fuse_batched: differentiable torch products)."""
        rot = rotation_ptms(ang)                   # [S, blocks, n, 4, 4]
        damp = damping_ptms(t1, t2)                # [S, n, 4, 4]
        out = []
        for layer in range(self.layers):
            b = (layer // self.block) % self.n_blocks
            out += fused_layer_gates(self.n, layer, rot[:, b], damp)
        if fuse_qudits and fuse_qudits > 2:
            from .autograd import Gate
            from .fusion import fuse_batched
            out = [Gate(m, q) for m, q in fuse_batched(out, [4] * self.n, fuse_qudits)]
        return out

    def populations(self, ang, t1, t2, dtype, checkpoint_every=8, fuse_qudits=0):
        import torch
        from .autograd import run_circuit, diagonal
        from .bases import general
        S = ang.shape[0]
        N = 4 ** self.n
        x0 = torch.zeros((S, N), dtype=dtype, device=ang.device)
        x0[:, 0] = 1
        gates = self.gates(ang, t1, t2, fuse_qudits)
        x = run_circuit(x0, gates, (4,) * self.n, checkpoint_every)
        return diagonal(x, [general(2)] * self.n)

    @staticmethod
    def kl(target, p, eps=1e-30):
        import torch
        t = target.clamp_min(eps)
        return (target * (torch.log(t) - torch.log(p.clamp_min(eps)))).sum(-1)


def random_circuit_batch(n_circuits, seed, n_range=(4, 11), layer_range=(20, 80)):
    """The batched sweep: per circuit a qubit count ~ U{4..11}, a layer count
    ~ U{20..80}, RX/RZ angles ~ U(0, 2 pi) for every layer and qubit, and
    T1 ~ U(20, 40) us, T2 ~ U(10, T1) us per qubit."""
    g = np.random.default_rng(seed)
    out = []
    for _ in range(n_circuits):
        n = int(g.integers(n_range[0], n_range[1] + 1))
        L = int(g.integers(layer_range[0], layer_range[1] + 1))
        t1 = g.uniform(20, 40, size=n)
        t2 = np.array([g.uniform(10, t1[q]) for q in range(n)])
        out.append({'n': n, 'layers': L, 'angles': g.uniform(0, 2 * np.pi, size=(L, n, 2)),
                    't1': t1, 't2': t2})
    return out
