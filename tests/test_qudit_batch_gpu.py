"""Batched and differentiable paths for qudits on the GPU.

* qsv_apply_ptm_batched with PTMs that are neither inline nor dense
  64 x 64 (two qutrits: 81 x 81; qutrit + qubit: 36 x 36; three qudits:
  324 x 324) against the oracle's apply_ptm per state (float64 within
  1e-12, float32 within 1e-5);
* qsv_ptm_grad above 64 x 64 against numpy's einsum, and the adjoint of a
  qutrit-pair apply through torch.autograd.gradcheck;
* CircuitSet: circuits of uneven qudit shapes and depths (qutrits, qubits,
  mixed; shared gate prefixes) run together -- final states against the
  oracle per circuit, populations against the oracle's diagonal, and every
  PTM's gradient against torch's autograd through a float64 einsum chain on
  the CPU.
"""
import math

import numpy as np
import pytest
import torch

from quantumsim_amd import bases
from quantumsim_amd.autograd import apply, _ptm_grad_raw
from quantumsim_amd.circuit_set import CircuitSet
from oracle.pauli_vector import PauliVectorNumpy
from oracle import bases as ob

pytestmark = pytest.mark.gpu
DEV = 'cuda'
Q3, Q2 = bases.general(3), bases.general(2)


def _obases(bs):
    return [ob.Basis(np.array(b.vectors), list(b.labels)) for b in bs]


def _orthogonal(rng, D):
    q, r = np.linalg.qr(rng.standard_normal((D, D)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 1e-5)])
@pytest.mark.parametrize('targets', [(0, 1), (3, 0), (1, 2), (2, 1), (0, 1, 2), (1,), (2, 3)])
def test_batched_apply_qudit_sizes_match_oracle(targets, dtype, tol):
    bs = [Q3, Q3, Q2, Q3]
    shape = tuple(b.dim_pauli for b in bs)                   # (9, 9, 4, 9)
    rng = np.random.default_rng(sum(targets) + 7 * len(targets))
    B = 5
    D = math.prod(shape[q] for q in targets)
    x = rng.standard_normal((B, math.prod(shape)))
    P = rng.standard_normal((B, D, D)) / math.sqrt(D)
    got = apply(torch.tensor(x, dtype=dtype, device=DEV), torch.tensor(P, dtype=dtype, device=DEV),
                shape, targets).double().cpu().numpy()
    ob_ = _obases(bs)
    for b in range(B):
        xb = x[b] if dtype == torch.float64 else x[b].astype(np.float32).astype(np.float64)
        pb = P[b] if dtype == torch.float64 else P[b].astype(np.float32).astype(np.float64)
        o = PauliVectorNumpy(ob_, xb.reshape(shape).copy())
        o.apply_ptm(pb.reshape([shape[q] for q in targets] * 2), *targets)
        ref = o.to_pv().reshape(-1)
        err = np.abs(got[b] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < tol, (targets, b, err)


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 1e-5)])
@pytest.mark.parametrize('targets,dout', [((0, 1), (9, 9)), ((3, 1), (9, 9)),
                                          ((1, 2), (9, 4)), ((4, 0), (4, 9)),
                                          ((0, 1), (9, 4))])
def test_big_ptm_grad_matches_einsum(targets, dout, dtype, tol):
    """dP[b, o, i] = sum over the other indices of ybar[b, .., o, ..] x[b, .., i, ..]
    for PTMs above 64 entries a side (the float64 LDS kernel)."""
    rng = np.random.default_rng(3 + sum(targets))
    shape = (9, 9, 4, 9, 4)
    n, B = len(shape), 3
    oshape = list(shape)
    for t, d in zip(targets, dout):
        oshape[t] = d
    x = rng.standard_normal((B,) + shape)
    yb = rng.standard_normal([B] + oshape)
    if dtype == torch.float32:
        x, yb = x.astype(np.float32).astype(np.float64), yb.astype(np.float32).astype(np.float64)
    got = _ptm_grad_raw(torch.tensor(x.reshape(B, -1), dtype=dtype, device=DEV),
                        torch.tensor(yb.reshape(B, -1), dtype=dtype, device=DEV), shape,
                        targets, dout).double().cpu().numpy()
    k = len(targets)
    lx = [n + 1] + list(range(n))
    ly = [n + 1] + list(range(n))
    for j, t in enumerate(targets):
        ly[1 + t] = n + 2 + j
    ref = np.einsum(yb, ly, x, lx, [n + 1] + [n + 2 + j for j in range(k)] + list(targets))
    ref = ref.reshape(got.shape)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < tol, err


def test_qutrit_pair_apply_gradcheck():
    rng = np.random.default_rng(9)
    shape = (9, 4, 9)
    B = 2
    x = torch.tensor(rng.standard_normal((B, math.prod(shape))), device=DEV, requires_grad=True)
    P = torch.tensor(rng.standard_normal((B, 81, 81)) / 9, device=DEV, requires_grad=True)
    Ps = torch.tensor(rng.standard_normal((1, 36, 36)) / 6, device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: apply(a, b, shape, (2, 0)), (x, P),
                                    eps=1e-6, atol=1e-8, rtol=1e-6, fast_mode=True)
    assert torch.autograd.gradcheck(lambda a, b: apply(a, b, shape, (1, 2)), (x, Ps),
                                    eps=1e-6, atol=1e-8, rtol=1e-6, fast_mode=True)


def _circuits(rng):
    """Uneven shapes and depths; circuits 0, 1 share a gate prefix, 2 has
    the same shape but another structure, 5 has no gates.  PTM sizes 4, 9,
    36, 64 and 81 a side (random orthogonal matrices)."""
    mix = [Q3, Q3, Q2]
    def g(shape, targets):
        D = math.prod(shape[q] for q in targets)
        return (torch.tensor(_orthogonal(rng, D), requires_grad=True), targets)
    sm = (9, 9, 4)
    s4 = (4, 4, 4, 4)
    s3 = (9, 9, 9)
    c0 = [g(sm, (0, 1)), g(sm, (2,)), g(sm, (1, 2)), g(sm, (0,)), g(sm, (2, 0))]
    c1 = [g(sm, t) for _, t in c0[:3]]
    c2 = [g(sm, (2, 1)), g(sm, (0,)), g(sm, (0, 1))]
    c3 = [g(s4, (0, 1, 2)), g(s4, (3,)), g(s4, (1, 2)), g(s4, (3, 2, 1))]
    c4 = [g(s3, (0, 2)), g(s3, (1,)), g(s3, (1, 2)), g(s3, (2, 0))]
    return [(mix, c0), (mix, c1), (mix, c2), ([Q2] * 4, c3), ([Q3] * 3, c4), (mix, [])]


def _cpu_chain(bs, gates, x0):
    """The same circuit with torch einsums on the CPU (float64 autograd)."""
    shape = [b.dim_pauli for b in bs]
    n = len(shape)
    x = x0.reshape(shape)
    for P, tg in gates:
        k = len(tg)
        Pt = P.reshape([shape[q] for q in tg] * 2)
        lx = list(range(n))
        lout = list(range(n))
        lp = [n + j for j in range(k)] + list(tg)
        for j, q in enumerate(tg):
            lout[q] = n + j
        x = torch.einsum(Pt, lp, x, lx, lout)
    return x.reshape(-1)


@pytest.mark.parametrize('dtype,tol,gtol', [(torch.float64, 1e-12, 1e-10),
                                            (torch.float32, 2e-5, 2e-4)])
def test_circuit_set_matches_oracle(dtype, tol, gtol):
    rng = np.random.default_rng(21)
    circuits = _circuits(rng)
    cs = CircuitSet(circuits, dtype=dtype, checkpoint_every=2)
    shapes = [b[0] for b in cs.batches]
    assert (9, 9, 4) in shapes and (4, 4, 4, 4) in shapes and (9, 9, 9) in shapes
    # circuits 0 and 1 (and the empty 5) share one batch
    assert any(sorted(m) == [0, 1, 5] for _, _, m in cs.batches)
    states = cs.run()
    pops = cs.populations(states)
    weights = [rng.standard_normal(s.numel()) for s in states]
    loss = sum((s.double() * torch.tensor(w, device=DEV)).sum() for s, w in zip(states, weights))
    loss.backward()
    for c, (bs, gates) in enumerate(circuits):
        o = PauliVectorNumpy(_obases(bs))
        for P, tg in gates:
            o.apply_ptm(P.detach().numpy().reshape([o.dim_pauli[q] for q in tg] * 2), *tg)
        ref = o.to_pv().reshape(-1)
        got = states[c].detach().double().cpu().numpy()
        assert np.abs(got - ref).max() < tol * max(1.0, np.abs(ref).max()), c
        assert np.abs(pops[c].detach().cpu().numpy() - o.diagonal().reshape(-1)).max() < tol, c
        # gradients: the same loss through a CPU einsum chain
        leaves = [P.detach().clone().requires_grad_(True) for P, _ in gates]
        x0 = torch.zeros(int(np.prod(o.dim_pauli)), dtype=torch.float64)
        x0[0] = 1
        y = _cpu_chain(bs, [(L, tg) for L, (_, tg) in zip(leaves, gates)], x0)
        if leaves:
            ref_g = torch.autograd.grad((y * torch.tensor(weights[c])).sum(), leaves)
            for (P, tg), rg in zip(gates, ref_g):
                err = (P.grad - rg).abs().max() / max(rg.abs().max(), 1e-300)
                assert err < gtol, (c, tg, float(err))
