"""Gate fusion (quantumsim_amd.fusion) against gate-by-gate application.

The fused list must give the same vector as the original gate list, applied
with the oracle's einsum restatement of PauliVectorNumpy.apply_ptm; this
pins the ordering rules (gates sharing a qudit keep their order, gates on
disjoint qudits may be reordered) and the embedding of a gate into a larger
block (a transposed or misplaced axis changes the result).
"""
import numpy as np
import pytest

from oracle import bases as ob
from oracle.pauli_vector import PauliVectorNumpy
from quantumsim_amd.fusion import embed_ptm, fuse
from quantumsim_amd.workloads import noisy_layers


def _apply_all(dims_basis, gates, x0):
    pv = PauliVectorNumpy(dims_basis, x0.copy(), force=True)
    for p, q in gates:
        pv.apply_ptm(np.asarray(p), *q)
    return pv.to_pv()


def _random_program(dims, n_gates, rng, max_k=3):
    n = len(dims)
    out = []
    for _ in range(n_gates):
        k = int(rng.integers(1, max_k + 1))
        qs = tuple(int(q) for q in rng.choice(n, size=k, replace=False))
        shape = [dims[q] for q in qs] * 2
        out.append((rng.standard_normal(shape) / np.sqrt(np.prod(shape[:k])), qs))
    return out


@pytest.mark.parametrize('max_q', [1, 2, 3, 4])
@pytest.mark.parametrize('seed', [0, 1, 2])
def test_fuse_random_mixed(max_q, seed):
    rng = np.random.default_rng(seed)
    bl = [ob.general(2), ob.general(3), ob.general(2), ob.general(2), ob.general(3)]
    dims = [b.dim_pauli for b in bl]
    gates = _random_program(dims, 40, rng)
    x0 = rng.standard_normal(dims)
    ref = _apply_all(bl, gates, x0)
    fused = fuse(gates, dims, max_qudits=max_q)
    assert sum(f.n_source for f in fused) == len(gates)
    for f in fused:
        # a block only exceeds the limit when a single gate already does
        assert len(f.qudits) <= max(max_q, 3)
        assert list(f.qudits) == sorted(f.qudits)
    got = _apply_all(bl, fused, x0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_fuse_noisy_layers_counts():
    n = 8
    circuit, _ = noisy_layers(n, 6, seed=3, pairing='brickwork')
    dims = [4] * n
    f2 = fuse(circuit, dims, 2)
    f4 = fuse(circuit, dims, 4)
    # a brickwork layer is n 1q unitaries, the CZ bonds, n dampings: with
    # blocks of 2 qubits every 1q gate joins a bond, leaving about one dense
    # 2q PTM per bond
    assert len(f2) < len(circuit) / 3
    assert len(f4) < len(f2)
    rng = np.random.default_rng(5)
    x0 = rng.standard_normal(dims)
    bl = [ob.general(2)] * n
    ref = _apply_all(bl, circuit, x0)
    for f in (f2, f4):
        np.testing.assert_allclose(_apply_all(bl, f, x0), ref, rtol=0, atol=1e-12)


def test_embed_ptm_identity_and_order():
    rng = np.random.default_rng(7)
    dims = [4, 9, 4]
    p = rng.standard_normal([9, 4, 9, 4])        # acts on (1, 0) in that order
    e = embed_ptm(p, (1, 0), (0, 1, 2), dims)
    # brute force: apply p to a random vector directly, and e as a matrix
    x = rng.standard_normal(dims)
    y_ref = np.einsum('abcd,dcz->baz', p, x)      # p[out1,out0,in1,in0] x[in0,in1,z]
    y = (e @ x.reshape(-1)).reshape(dims)
    np.testing.assert_allclose(y, y_ref, atol=1e-13)


def test_fuse_rejects_nonsquare():
    with pytest.raises(ValueError):
        fuse([(np.zeros((3, 4)), (0,))], [4, 4])


def test_fuse_batched_matches_host_fusion_per_batch_entry():
    """The torch composition along fuse_plan equals fuse() of each batch
    entry's PTMs; shared [1, D, D] PTMs broadcast; sequential application of
    the blocks equals the gate list (oracle)."""
    import torch
    from quantumsim_amd.fusion import fuse_batched, fuse_plan

    class G:
        def __init__(self, ptm, targets):
            self.ptm, self.targets = ptm, targets

    rng = np.random.default_rng(7)
    dims = [4, 4, 4, 4, 4]
    B = 3
    prog = []
    for t in [(0, 1), (2,), (2, 3), (4,), (1, 2), (3, 4), (0,), (1,), (2, 1), (4, 3)]:
        D = 4 ** len(t)
        shared = len(prog) % 3 == 0
        prog.append(G(torch.tensor(rng.standard_normal((1 if shared else B, D, D)) / D),
                      t))
    for k in (2, 3):
        blocks = fuse_batched(prog, dims, k)
        assert [b for _, b in blocks] == [q for q, _ in fuse_plan([g.targets for g in prog],
                                                                   dims, k)]
        for b in range(B):
            host = [(g.ptm[min(b, g.ptm.shape[0] - 1)].numpy().reshape([4] * (2 * len(g.targets))),
                     g.targets) for g in prog]
            ref = fuse(host, dims, k)
            assert len(ref) == len(blocks)
            for (m, q), f in zip(blocks, ref):
                assert q == f.qudits
                got = m[min(b, m.shape[0] - 1)].numpy()
                np.testing.assert_allclose(got, f.ptm.reshape(got.shape), atol=1e-14)
            x0 = rng.standard_normal([4] * 5)
            bl = [ob.general(2)] * 5
            seq = _apply_all(bl, host, x0)
            fused = _apply_all(bl, [(m[min(b, m.shape[0] - 1)].numpy().reshape([4] * (2 * len(q))),
                                     q) for m, q in blocks], x0)
            np.testing.assert_allclose(seq, fused, atol=1e-12)


def test_fuse_batched_gradient_reaches_sources():
    import torch
    from quantumsim_amd.fusion import fuse_batched

    class G:
        def __init__(self, ptm, targets):
            self.ptm, self.targets = ptm, targets

    rng = np.random.default_rng(8)
    a = torch.tensor(rng.standard_normal((2, 16, 16)), requires_grad=True)
    b = torch.tensor(rng.standard_normal((2, 16, 16)), requires_grad=True)
    blocks = fuse_batched([G(a, (0, 1)), G(b, (1, 2))], [4, 4, 4], 3)
    assert len(blocks) == 1 and blocks[0][1] == (0, 1, 2)

    def f(a_, b_):
        m, _ = fuse_batched([G(a_, (0, 1)), G(b_, (1, 2))], [4, 4, 4], 3)[0]
        return (m ** 2).sum()
    assert torch.autograd.gradcheck(f, (a, b), eps=1e-6, atol=1e-6)


@pytest.mark.parametrize('n', list(range(2, 13)))
def test_brickwork_window_plan(n):
    """The right-to-left tiling of two brickwork layers: every gate once,
    blocks of at most 3 qubits, n // 2 blocks, and the blocks applied in
    order equal the gates applied in order (oracle)."""
    from quantumsim_amd.workloads import brickwork_window_plan, layer_targets
    rng = np.random.default_rng(n)
    tg = layer_targets(n, 0) + layer_targets(n, 1)
    plan = brickwork_window_plan(n)
    assert len(plan) == max(1, n // 2)
    assert sorted(i for _, s in plan for i in s) == list(range(len(tg)))
    gates = [(rng.standard_normal([4] * (2 * len(t))) / 4 ** len(t), t) for t in tg]
    blocks = []
    for qs, src in plan:
        assert len(qs) <= 3 and all(set(tg[i]) <= set(qs) for i in src)
        m = np.eye(4 ** len(qs))
        for i in src:
            m = embed_ptm(gates[i][0], gates[i][1], qs, [4] * n) @ m
        blocks.append((m.reshape([4] * (2 * len(qs))), qs))
    x0 = rng.standard_normal([4] * n)
    bl = [ob.general(2)] * n
    np.testing.assert_allclose(_apply_all(bl, blocks, x0), _apply_all(bl, gates, x0), atol=1e-12)
