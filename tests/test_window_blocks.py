"""Window blocks built by qsv_window_blocks (quantumsim_amd.windows),
replayed on the CPU (qsv_emulate_window_blocks: the kernels' own per-thread
phases, every index checked) against the same blocks composed with torch
ops (workloads.fused_layer_gates + fusion.fuse_batched) and against torch
autograd through that composition."""
import numpy as np
import pytest
import torch

from quantumsim_amd.windows import WindowPlan, WBLOCK_DTYPE, emulate
from quantumsim_amd.workloads import (rotation_ptms, damping_ptms, fused_layer_gates,
                                      brickwork_window_plan, layer_targets)
from quantumsim_amd.autograd import Gate, cz_ptm
from quantumsim_amd.fusion import fuse_batched


def torch_blocks(rot, damp, keep, n, nw):
    """The torch composition of the batched path (CircuitBatch with
    block_kernel=False): [nw, B, D, D] per block position."""
    B = rot.shape[0]
    damp_w = damp[:, None].expand(B, nw, n, 4, 4).reshape(B * nw, n, 4, 4)
    kp = keep.reshape(B * nw).bool()[:, None, None]
    gates = []
    for par in (0, 1):
        rot_p = rot[:, par::2][:, :nw].reshape(B * nw, n, 4, 4)
        for gt in fused_layer_gates(n, par, rot_p, damp_w):
            ptm = gt.ptm
            if par == 1:
                ptm = torch.where(kp, ptm, torch.eye(ptm.shape[-1], dtype=ptm.dtype,
                                                     device=ptm.device))
            gates.append(Gate(ptm, gt.targets))
    out = []
    for m, q in fuse_batched(gates, [4] * n, 3, plan=brickwork_window_plan(n)):
        D = m.shape[-1]
        out.append(m.reshape(B, nw, D, D).transpose(0, 1))
    return out


def random_factors(n, B, nw, rng, device='cpu'):
    ang = torch.tensor(rng.uniform(0, 2 * np.pi, (B, 2 * nw, n, 2)), device=device,
                       requires_grad=True)
    t1 = torch.tensor(rng.uniform(20, 40, (B, n)), device=device, requires_grad=True)
    t2 = torch.tensor(rng.uniform(10, 20, (B, n)), device=device, requires_grad=True)
    keep = torch.tensor(rng.integers(0, 2, (B, nw)), dtype=torch.uint8, device=device)
    return ang, t1, t2, keep


def test_plan_covers_every_gate_once():
    for n in range(2, 13):
        plan = WindowPlan(n, 3, 2)
        ta, tb = layer_targets(n, 0), layer_targets(n, 1)
        seen = []
        for row, (qs, D, off) in zip(plan.table, plan.positions):
            assert D == 4 ** row['nq'] and list(row['qubits'][:row['nq']]) == list(qs)
            for f in range(row['nf']):
                slots = [row['f_slot0'][f]] + ([row['f_slot1'][f]] if row['f_kind'][f] == 2
                                               else [])
                seen.append((int(row['f_layer'][f]), tuple(qs[s] for s in slots)))
        assert sorted(seen) == sorted([(0, t) for t in ta] + [(1, t) for t in tb])
        assert plan.out_elems == sum(3 * 2 * D * D for _, D, _ in plan.positions)
    assert WBLOCK_DTYPE.itemsize == 96


@pytest.mark.parametrize('n', [2, 3, 4, 5, 6, 7, 8])
def test_emulated_kernels_match_torch_composition(n):
    rng = np.random.default_rng(n)
    B, nw = 3, 2
    ang, t1, t2, keep = random_factors(n, B, nw, rng)
    rot, damp = rotation_ptms(ang), damping_ptms(t1, t2)
    ref = torch_blocks(rot, damp, keep, n, nw)
    plan = WindowPlan(n, nw, B)
    cz = cz_ptm(torch.device('cpu')).reshape(16, 16)
    ups = [torch.tensor(rng.standard_normal(r.shape)) for r in ref]
    g_rot, g_damp = torch.autograd.grad(sum((r * u).sum() for r, u in zip(ref, ups)),
                                        [rot, damp])
    dout = np.concatenate([u.numpy().reshape(-1) for u in ups])
    out, drot, ddp = emulate(plan, rot.detach().numpy(), damp.detach().numpy(), keep.numpy(),
                             cz.numpy(), dout)
    for (qs, D, off), r in zip(plan.positions, ref):
        got = out[off:off + nw * B * D * D].reshape(nw, B, D, D)
        np.testing.assert_allclose(got, r.detach().numpy(), rtol=0, atol=1e-14)
    # every gradient slot written (the buffers start as NaN)
    assert not np.isnan(drot).any() and not np.isnan(ddp).any()
    scale = max(1.0, float(g_rot.abs().max()))
    np.testing.assert_allclose(drot, g_rot.numpy(), rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(ddp.sum(axis=(1, 2)), g_damp.numpy(), rtol=0,
                               atol=1e-13 * max(1.0, float(g_damp.abs().max())))
    # masked windows: zero gradients for their layer-1 rotations
    for b in range(B):
        for w in range(nw):
            if not keep[b, w]:
                assert np.all(drot[b, 2 * w + 1] == 0)


def test_table_validation():
    plan = WindowPlan(4, 2, 2)
    rng = np.random.default_rng(0)
    rot, damp = rng.standard_normal((2, 4, 4, 4, 4)), rng.standard_normal((2, 4, 4, 4))
    keep, cz = np.ones((2, 2), np.uint8), np.eye(16)
    bad = WindowPlan(4, 2, 2)
    bad.table[0]['qubits'][:2] = [3, 2]              # not ascending
    with pytest.raises(ValueError):
        emulate(bad, rot, damp, keep, cz)
    bad = WindowPlan(4, 2, 2)
    bad.table[1]['out_off'] = plan.out_elems         # output past the buffer
    with pytest.raises(ValueError):
        emulate(bad, rot, damp, keep, cz)
    with pytest.raises(ValueError):                  # too few rotation layers
        emulate(plan, rot[:, :3], damp, keep, cz)
