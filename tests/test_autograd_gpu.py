"""Differentiation and batching on the GPU.

* batched apply == one PauliVectorHIP per state (float64, bitwise);
* the adjoint kernels pass torch.autograd.gradcheck (central finite
  differences in float64);
* the checkpointed circuit gradient equals the gradient of the plain chain;
* the 96-parameter trainable circuit: autograd gradient vs central finite
  differences (h = 1e-6, float64) and vs finite differences of the numpy
  oracle run on the same PTMs;
* the circuit batcher: populations vs the oracle per circuit, gradients vs
  one-circuit-at-a-time runs, and HIP-graph replays vs the eager step.
"""
import numpy as np
import pytest
import torch

import quantumsim_amd as qa
from quantumsim_amd import bases
from quantumsim_amd import _native as nat
from quantumsim_amd.autograd import apply, run_circuit, run_schedule, Gate, diagonal
from quantumsim_amd.workloads import TrainableCircuit, random_circuit_batch
from quantumsim_amd.batch import CircuitBatch, plain_forward_backward
from oracle.pauli_vector import PauliVectorNumpy
from oracle import bases as ob

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_apply_batched_matches_single():
    rng = np.random.default_rng(0)
    n, B = 5, 6
    x = rng.standard_normal((B, 4 ** n))
    for targets in [(0,), (4,), (1, 3), (4, 0), (3, 4)]:
        D = 4 ** len(targets)
        P = rng.standard_normal((B, D, D))
        y = apply(torch.tensor(x, device=DEV), torch.tensor(P, device=DEV), (4,) * n,
                  targets).cpu().numpy()
        for b in range(B):
            v = qa.PauliVectorHIP([bases.general(2)] * n, x[b].reshape((4,) * n))
            v.apply_ptm(P[b].reshape((4,) * (2 * len(targets))), *targets)
            assert np.array_equal(y[b], v.to_pv().reshape(-1))


@pytest.mark.parametrize('targets', [(1,), (2,), (0, 2), (2, 0), (1, 2)])
def test_apply_gradcheck(targets):
    rng = np.random.default_rng(1)
    n, B = 3, 2
    D = 4 ** len(targets)
    x = torch.tensor(rng.standard_normal((B, 4 ** n)), device=DEV, requires_grad=True)
    P = torch.tensor(rng.standard_normal((B, D, D)), device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: apply(a, b, (4,) * n, targets), (x, P),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)
    Ps = torch.tensor(rng.standard_normal((1, D, D)), device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: apply(a, b, (4,) * n, targets), (x, Ps),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 1e-5)])
@pytest.mark.parametrize('targets', [(0, 1), (4, 5), (8, 9), (9, 8), (0, 9), (3,), (9,)])
def test_ptm_grad_kernel_matches_einsum(targets, dtype, tol):
    """qsv_ptm_grad (MFMA, float64 accumulation) on 10-qubit states, batch 3,
    every target position (vector and per-orbit paths) against numpy:
    dP[b, o, i] = sum over the other indices of ybar[b, .., o, ..] x[b, .., i, ..]."""
    from quantumsim_amd.autograd import _ptm_grad_raw
    rng = np.random.default_rng(5)
    n, B = 10, 3
    shape = (4,) * n
    x = rng.standard_normal((B,) + shape)
    yb = rng.standard_normal((B,) + shape)
    if dtype == torch.float32:
        x, yb = x.astype(np.float32).astype(np.float64), yb.astype(np.float32).astype(np.float64)
    got = _ptm_grad_raw(torch.tensor(x.reshape(B, -1), dtype=dtype, device=DEV),
                        torch.tensor(yb.reshape(B, -1), dtype=dtype, device=DEV), shape,
                        targets, tuple(4 for _ in targets)).double().cpu().numpy()
    k = len(targets)
    lx = [n + 1] + list(range(n))
    ly = [n + 1] + list(range(n))
    for j, t in enumerate(targets):
        ly[1 + t] = n + 2 + j
    ref = np.einsum(yb, ly, x, lx, [n + 1] + [n + 2 + j for j in range(k)] + list(targets))
    ref = ref.reshape(B, 4 ** k, 4 ** k)
    err = np.max(np.abs(got.reshape(ref.shape) - ref)) / np.max(np.abs(ref))
    print(targets, dtype, 'rel err %.3e' % err)
    assert err < tol


def test_checkpointed_circuit_equals_chain():
    rng = np.random.default_rng(2)
    n, B = 4, 3
    shape = (4,) * n
    x0 = torch.tensor(rng.standard_normal((B, 4 ** n)), device=DEV, requires_grad=True)
    specs = [(0, 1), (2,), (3, 1), (0,), (2, 3), (1,), (3,), (0, 2), (1, 2), (3,)]
    ptms = [torch.tensor(rng.standard_normal((B, 4 ** len(t), 4 ** len(t))) / 4, device=DEV,
                         requires_grad=True) for t in specs]
    w = torch.tensor(rng.standard_normal((B, 4 ** n)), device=DEV)
    y1 = run_circuit(x0, [Gate(p, t) for p, t in zip(ptms, specs)], shape, checkpoint_every=3)
    g1 = torch.autograd.grad((y1 * w).sum(), [x0] + ptms)
    y2 = x0
    for p, t in zip(ptms, specs):
        y2 = apply(y2, p, shape, t)
    g2 = torch.autograd.grad((y2 * w).sum(), [x0] + ptms)
    assert torch.equal(y1, y2)
    for a, b in zip(g1, g2):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-12)



@pytest.mark.parametrize('k', [1, 2, 3])
@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 2e-5)])
def test_run_schedule_matches_per_state_chains(k, dtype, tol):
    """A shrinking schedule (5 states, then 3, then 1; positions of 1, 2
    and 3 qubits, the last on the 64 x 64 MFMA kernels) against each state
    run alone through the gates of the segments it takes part in, values
    and gradients of x0 and of every step tensor."""
    rng = np.random.default_rng(5)
    n, B = 5, 5
    shape = (4,) * n
    positions = [(0, 1), (4,), (1, 2, 3)]
    steps = [torch.tensor(rng.standard_normal((3, B, 4 ** len(t), 4 ** len(t))) / 2 ** len(t),
                          dtype=dtype, device=DEV, requires_grad=True) for t in positions]
    segments = [(5, 3, [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1)]),
                (3, 1, [(2, 1), (0, 2)]),
                (1, 0, [(1, 2), (2, 2)])]
    x0 = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV,
                      requires_grad=True)
    w = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV)
    out = run_schedule(x0, steps, positions, segments, shape, k)
    g1 = torch.autograd.grad((out * w).sum(), [x0] + steps)
    rows = []
    for b in range(B):
        y = x0[b:b + 1]
        for active, _, gates in segments:
            if b < active:
                for p, s in gates:
                    y = apply(y, steps[p][s][b:b + 1], shape, positions[p])
        rows.append(y)
    ref = torch.cat(rows)
    g2 = torch.autograd.grad((ref * w).sum(), [x0] + steps)
    scale = lambda t: max(1.0, t.abs().max().item())
    assert (out - ref).abs().max().item() <= tol * scale(ref)
    for a, b in zip(g1, g2):
        assert (a - b).abs().max().item() <= tol * scale(b)
    # PTMs a state never meets get a zero gradient
    for grad, rows_unused in ((g1[1][2], 3), (g1[2][2], 1), (g1[3][2], 1), (g1[3][1], 3)):
        assert grad[rows_unused:].abs().max().item() == 0

def _kl_loss(tc, params, target, dtype):
    ang, t1, t2 = params
    p = tc.populations(ang, t1, t2, dtype, checkpoint_every=5)
    return tc.kl(target, p).sum()


def test_trainable_circuit_fd():
    tc = TrainableCircuit(n=4, layers=10)
    ang, t1, t2 = tc.initial_params(2, seed=3, device=DEV, dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang + 0.2, t1 * 1.1, t2 * 0.95, torch.float64)
    params = [ang.clone().requires_grad_(), t1.clone().requires_grad_(),
              t2.clone().requires_grad_()]
    loss = _kl_loss(tc, params, target, torch.float64)
    grads = torch.autograd.grad(loss, params)
    rng = np.random.default_rng(4)
    h = 1e-6
    for _ in range(12):
        which = int(rng.integers(3))
        idx = tuple(int(rng.integers(s)) for s in params[which].shape)
        vals = []
        for sgn in (1, -1):
            p2 = [t.detach().clone() for t in params]
            p2[which][idx] += sgn * h
            with torch.no_grad():
                vals.append(_kl_loss(tc, p2, target, torch.float64).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        an = grads[which][idx].item()
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 2e-8, (which, idx, fd, an)


def test_gradient_matches_oracle_fd():
    """Finite differences of the numpy oracle, run on the same PTMs, agree
    with the HIP adjoint (a small instance so numpy finishes quickly)."""
    tc = TrainableCircuit(n=4, layers=6)
    ang, t1, t2 = tc.initial_params(1, seed=5, device=DEV, dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang * 0.9, t1, t2, torch.float64)
    params = [ang.clone().requires_grad_(), t1.clone().requires_grad_(),
              t2.clone().requires_grad_()]
    grads = torch.autograd.grad(_kl_loss(tc, params, target, torch.float64), params)
    tgt = target[0].cpu().numpy()

    def oracle_loss(a, b, c):
        with torch.no_grad():
            gates = tc.gates(a, b, c)
        o = PauliVectorNumpy([ob.general(2)] * tc.n)
        for g in gates:
            k = len(g.targets)
            o.apply_ptm(g.ptm[0].cpu().numpy().reshape((4,) * (2 * k)), *g.targets)
        p = np.maximum(o.diagonal(), 1e-30)
        t = np.maximum(tgt, 1e-30)
        return float(np.sum(tgt * (np.log(t) - np.log(p))))

    h = 1e-6
    for which, idx in [(0, (0, 0, 1, 0)), (0, (0, 2, 3, 1)), (1, (0, 2)), (2, (0, 1))]:
        vals = []
        for sgn in (1, -1):
            p2 = [t.detach().clone() for t in params]
            p2[which][idx] += sgn * h
            vals.append(oracle_loss(*p2))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = grads[which][idx].item()
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 2e-8, (which, idx, fd, an)


def _oracle_circuit_pops(c):
    from quantumsim_amd.workloads import rotation_ptms, damping_ptms, fused_layer_gates
    n = c['n']
    with torch.no_grad():
        rot = rotation_ptms(torch.tensor(c['angles'][None], device=DEV))
        damp = damping_ptms(torch.tensor(c['t1'][None], device=DEV),
                            torch.tensor(c['t2'][None], device=DEV))
        o = PauliVectorNumpy([ob.general(2)] * n)
        for l in range(c['layers']):
            for g in fused_layer_gates(n, l, rot[:, l], damp):
                k = len(g.targets)
                o.apply_ptm(g.ptm[0].cpu().numpy().reshape((4,) * (2 * k)), *g.targets)
    return o.diagonal()


def test_circuit_batch_matches_oracle_and_single():
    circuits = random_circuit_batch(10, seed=7, n_range=(3, 5), layer_range=(2, 6))
    cb = CircuitBatch(circuits, dtype=torch.float64, checkpoint_every=4)
    w = cb.random_weights(8)
    loss, pops, grads = cb.step(w)
    for gi, g in enumerate(cb.groups):
        for b, i in enumerate(g.members):
            ref = _oracle_circuit_pops(circuits[i])
            assert np.max(np.abs(pops[gi][b].detach().cpu().numpy() - ref)) < 1e-12
            _, _, g1 = plain_forward_backward(circuits[i], w[gi][b:b + 1],
                                              dtype=torch.float64, checkpoint_every=4)
            L = circuits[i]['layers']
            assert torch.allclose(grads[3 * gi][b][:L], g1[0][0], rtol=1e-10, atol=1e-12)
            assert torch.allclose(grads[3 * gi + 1][b], g1[1][0], rtol=1e-10, atol=1e-12)
            assert torch.allclose(grads[3 * gi + 2][b], g1[2][0], rtol=1e-10, atol=1e-12)


def test_circuit_batch_float32_and_graph_replay():
    circuits = random_circuit_batch(12, seed=9, n_range=(3, 6), layer_range=(3, 8))
    cb64 = CircuitBatch(circuits, dtype=torch.float64, checkpoint_every=4)
    w = cb64.random_weights(10)
    _, pops64, grads64 = cb64.step(w)
    cb = CircuitBatch(circuits, dtype=torch.float32, checkpoint_every=4)
    loss_e, pops_e, grads_e = cb.step(w)
    for a, b in zip(pops_e, pops64):
        assert (a - b).abs().max().item() < 1e-5           # float32 state vs float64
    cb.capture(w)
    for rep in range(3):
        torch.randn(1000, device=DEV).sum().item()         # eager work between replays
        loss_g, pops_g, grads_g = cb.replay()
        torch.cuda.synchronize()
        for a, b in zip(pops_g, pops_e):
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-14), rep
        for a, b in zip(grads_g, grads_e):
            assert torch.allclose(a, b, rtol=1e-6, atol=1e-9), rep


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_persistent_kernel_matches_batched_launches(dtype, tol):
    """Circuits of 4-9 qubits: the narrow ones run in the persistent
    whole-circuit kernels (one forward + one adjoint launch), the wide ones
    in batched launches; populations and every parameter gradient equal the
    all-batched-launch run."""
    from quantumsim_amd import persistent
    circuits = random_circuit_batch(16, seed=21, n_range=(4, 9), layer_range=(3, 9))
    cb = CircuitBatch(circuits, dtype=dtype, checkpoint_every=4)
    assert cb.small and len(cb.small) < len(cb.groups)
    assert persistent.max_qubits(dtype) >= 6
    ref = CircuitBatch(circuits, dtype=dtype, checkpoint_every=4, persistent_kernel=False)
    w = cb.random_weights(3)
    _, pops, grads = cb.step(w)
    _, pops_r, grads_r = ref.step(w)
    for a, b in zip(pops, pops_r):
        assert (a - b).abs().max().item() < tol
    for a, b in zip(grads, grads_r):
        scale = b.abs().max().item() + 1e-30
        assert (a - b).abs().max().item() < (1e-10 if dtype == torch.float64 else 1e-4) * scale


# ------------------------------------------------ dense 64 x 64 (three-qubit) paths
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('targets', [(1, 2, 3), (2, 3, 4), (4, 0, 2), (0, 1, 2)])
def test_apply_batched_dense64_matches_single(targets, dtype):
    """Batched 64 x 64 apply (MFMA, per-state row-major PTMs gathered in the
    kernel) equals one PauliVectorHIP per state (host PTM in fragment
    order) bitwise; a shared PTM (stride 0) too."""
    rng = np.random.default_rng(31)
    n, B = 5, 5
    print('dense64 batched', targets, dtype, flush=True)
    x = rng.standard_normal((B, 4 ** n))
    P = rng.standard_normal((B, 64, 64)) / 8
    xt = torch.tensor(x, dtype=dtype, device=DEV)
    y = apply(xt, torch.tensor(P, dtype=dtype, device=DEV), (4,) * n, targets).cpu()
    ys = apply(xt, torch.tensor(P[:1], dtype=dtype, device=DEV), (4,) * n, targets).cpu()
    for b in range(B):
        for PP, yy in ((P[b], y), (P[0], ys)):
            v = qa.PauliVectorHIP([bases.general(2)] * n, x[b].reshape((4,) * n), dtype=dtype)
            v.apply_ptm(PP.reshape((4,) * 6), *targets)
            assert torch.equal(yy[b], v.to_torch().reshape(-1).cpu()), b


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 1e-5)])
@pytest.mark.parametrize('targets', [(1, 2, 3), (7, 8, 9), (9, 0, 4), (0, 1)])
def test_ptm_grad64_kernel_matches_einsum(targets, dtype, tol):
    """The wide gradient kernel (dense 3-qubit PTMs; a 16 x 16 case for
    comparison) on 10-qubit states, batch 3, against numpy."""
    from quantumsim_amd.autograd import _ptm_grad_raw
    rng = np.random.default_rng(6)
    n, B = 10, 3
    shape = (4,) * n
    x = rng.standard_normal((B,) + shape)
    yb = rng.standard_normal((B,) + shape)
    if dtype == torch.float32:
        x, yb = x.astype(np.float32).astype(np.float64), yb.astype(np.float32).astype(np.float64)
    print('grad64', targets, dtype, flush=True)
    got = _ptm_grad_raw(torch.tensor(x.reshape(B, -1), dtype=dtype, device=DEV),
                        torch.tensor(yb.reshape(B, -1), dtype=dtype, device=DEV), shape,
                        targets, tuple(4 for _ in targets)).double().cpu().numpy()
    k = len(targets)
    lx = [n + 1] + list(range(n))
    ly = [n + 1] + list(range(n))
    for j, t in enumerate(targets):
        ly[1 + t] = n + 2 + j
    ref = np.einsum(yb, ly, x, lx, [n + 1] + [n + 2 + j for j in range(k)] + list(targets))
    ref = ref.reshape(B, 4 ** k, 4 ** k)
    err = np.max(np.abs(got.reshape(ref.shape) - ref)) / np.max(np.abs(ref))
    print(targets, dtype, 'rel err %.3e' % err)
    assert err < tol


@pytest.mark.parametrize('n,targets,B', [(6, (1, 2, 3), 2), (8, (0, 1, 2), 3), (10, (5, 6, 7), 3),
                                         (11, (2, 3, 4), 2), (7, (4, 1, 0), 5)])
def test_ptm_grad64_direct_kernel_matches_einsum(n, targets, B):
    """The float32 64 x 64 gradient with direct HBM -> LDS loads (innermost
    axis a bystander, whole 32-orbit tiles), for the planner's block count
    and forced ones (1 .. 64 blocks per state, some with no tile), against
    numpy's einsum of the same float32 inputs."""
    from quantumsim_amd.autograd import _ptm_grad_raw
    rng = np.random.default_rng(n + B)
    shape = (4,) * n
    x = rng.standard_normal((B,) + shape).astype(np.float32).astype(np.float64)
    yb = rng.standard_normal((B,) + shape).astype(np.float32).astype(np.float64)
    xl = [n + 1] + list(range(n))
    yl = [n + 1] + list(range(n))
    for j, t in enumerate(targets):
        yl[1 + t] = n + 2 + j
    ref = np.einsum(yb, yl, x, xl, [n + 1] + [n + 2 + j for j in range(3)] + list(targets))
    ref = ref.reshape(B, 64, 64)
    L = nat.lib()
    xt = torch.tensor(x.reshape(B, -1), dtype=torch.float32, device=DEV)
    yt = torch.tensor(yb.reshape(B, -1), dtype=torch.float32, device=DEV)
    print('grad64 direct', n, targets, B, flush=True)
    got = _ptm_grad_raw(xt, yt, shape, targets, (4, 4, 4)).double().cpu().numpy()
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-5
    ws = L.qsv_ptm_grad_workspace_bytes(B, n, nat.i64(shape), 3, nat.i32(targets),
                                        nat.i32((4, 4, 4)))
    buf = torch.empty(ws, dtype=torch.uint8, device=DEV)
    for nb in (1, 3, 17, 64):
        out = torch.empty((B, 64, 64), dtype=torch.float32, device=DEV)
        rc = L.qsv_ptm_grad_tiled(nat.QSV_F32, xt.data_ptr(), yt.data_ptr(), B, n,
                                  nat.i64(shape), 3, nat.i32(targets), nat.i32((4, 4, 4)),
                                  out.data_ptr(), 0, buf.data_ptr(), ws,
                                  nat.ctypes.byref(nat.Tile(0, 0, nb, 0)),
                                  torch.cuda.current_stream().cuda_stream)
        nat.check(rc, 'grad')
        err = np.max(np.abs(out.double().cpu().numpy() - ref)) / np.max(np.abs(ref))
        assert err < 1e-5, (nb, err)


def test_apply_dense64_gradcheck():
    rng = np.random.default_rng(32)
    n, B = 4, 2
    x = torch.tensor(rng.standard_normal((B, 4 ** n)), device=DEV, requires_grad=True)
    P = torch.tensor(rng.standard_normal((B, 64, 64)) / 8, device=DEV, requires_grad=True)
    for targets in [(1, 2, 3), (3, 0, 1)]:
        assert torch.autograd.gradcheck(lambda a, b: apply(a, b, (4,) * n, targets), (x, P),
                                        eps=1e-6, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_fused_batch_equals_unfused(dtype):
    """Circuits of 8-9 qubits (batched launches) with uneven depths: the
    two-layer windows fused into 3-qubit blocks (identity PTMs for circuits
    ending inside a window) give the populations and gradients of the
    gate-by-gate batch; HIP-graph replays equal the eager step."""
    circuits = random_circuit_batch(6, seed=41, n_range=(8, 9), layer_range=(3, 8))
    assert len(set(c['layers'] for c in circuits)) > 2
    fused = CircuitBatch(circuits, dtype=dtype, checkpoint_every=4, fuse_qudits=3)
    plain = CircuitBatch(circuits, dtype=dtype, checkpoint_every=4, fuse_qudits=0)
    assert not fused.small
    w = fused.random_weights(5)
    _, pf, gf = fused.step(w)
    _, pp, gp = plain.step(w)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    for a, b in zip(pf, pp):
        assert (a - b).abs().max().item() < tol
    for a, b in zip(gf, gp):
        scale = b.abs().max().item() + 1e-30
        assert (a - b).abs().max().item() < (1e-10 if dtype == torch.float64 else 1e-4) * scale
    fused.capture(w)
    # capture() tunes the launch shapes: an eager step with the tuned shapes
    # is the reference for the replays (float32 gradients of dense blocks
    # sum in float32 within a wave, so other block counts round differently)
    _, pe, ge = fused.step(w)
    for a, b in zip(ge, gf):
        scale = b.abs().max().item() + 1e-30
        assert (a - b).abs().max().item() < (1e-12 if dtype == torch.float64 else 1e-5) * scale
    # eager step and graph replays (float32 dL/dP partial sums may be
    # partitioned differently by the tuned launch shapes)
    for rep in range(3):
        torch.randn(1000, device=DEV).sum().item()
        _, pg, gg = fused.replay()
        torch.cuda.synchronize()
        for a, b in zip(pg, pe):
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-14), rep
        for a, b in zip(gg, ge):
            scale = b.abs().max().item() + 1e-30
            assert (a - b).abs().max().item() < (1e-11 if dtype == torch.float64
                                                 else 1e-5) * scale, rep


def test_fused_batch_matches_oracle():
    """One fused 8-qubit circuit vs the oracle applying the unfused layer
    PTMs (float64)."""
    circuits = random_circuit_batch(2, seed=43, n_range=(8, 8), layer_range=(5, 7))
    cb = CircuitBatch(circuits, dtype=torch.float64, checkpoint_every=4)
    with torch.no_grad():
        pops = cb.populations()
    for gi, g in enumerate(cb.groups):
        for b, i in enumerate(g.members):
            ref = _oracle_circuit_pops(circuits[i])
            assert np.max(np.abs(pops[gi][b].cpu().numpy() - ref)) < 1e-12


def test_trainable_circuit_fused_gradient():
    """The 12-qubit trainable circuit (reduced depth and sets) fused into
    3-qubit blocks: loss and gradients equal the unfused run (float64)."""
    tc = TrainableCircuit(n=8, layers=10)
    ang, t1, t2 = tc.initial_params(3, seed=8, device=DEV, dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang + 0.1, t1, t2, torch.float64)
    out = []
    for fq in (0, 3):
        params = [ang.clone().requires_grad_(), t1.clone().requires_grad_(),
                  t2.clone().requires_grad_()]
        p = tc.populations(*params, torch.float64, checkpoint_every=4, fuse_qudits=fq)
        loss = tc.kl(target, p).sum()
        out.append((loss.item(), torch.autograd.grad(loss, params)))
    assert abs(out[0][0] - out[1][0]) < 1e-12
    for a, b in zip(out[0][1], out[1][1]):
        assert torch.allclose(a, b, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 1e-5)])
@pytest.mark.parametrize('n,targets', [(6, (1, 2, 3)), (6, (0, 1, 2)), (6, (3, 0, 2)),
                                       (10, (4, 5, 6)), (10, (7, 8, 9)), (5, (2, 3, 4)),
                                       (9, (0, 1, 2)), (8, (3, 4, 5)), (10, (0, 4, 8)),
                                       (7, (6, 1, 3)), (6, (3, 4, 5)), (9, (6, 7, 8)),
                                       (8, (1, 4, 7)), (10, (2, 5, 9)), (7, (6, 3, 4))])
def test_fused_backward_matches_separate_kernels(n, targets, dtype, tol):
    """One-pass backward of dense 64 x 64 blocks (qsv_apply_ptm_batched_backward)
    vs the transposed apply (bitwise: the same k-steps of four entries in
    the same order; float32 states on whole 32-orbit tiles run the
    direct-to-LDS k_bwd64d (innermost axis a bystander) or k_bwd64q (the
    last target innermost)) and the PTM gradient kernel; per-state and
    shared PTMs; geometries where it does not apply fall back to the
    separate kernels."""
    from quantumsim_amd.autograd import _backward_raw, _apply_raw, _ptm_grad_raw
    rng = np.random.default_rng(n + sum(targets))
    B = 3
    shape = (4,) * n
    print('fused backward', n, targets, dtype, flush=True)
    x = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV)
    gy = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV)
    for P in (torch.tensor(rng.standard_normal((B, 64, 64)) / 8, dtype=dtype, device=DEV),
              torch.tensor(rng.standard_normal((1, 64, 64)) / 8, dtype=dtype, device=DEV)):
        gx, gp = _backward_raw(x, gy, P, shape, targets, True, fused=True)
        gx2 = _apply_raw(gy, P.transpose(1, 2), shape, targets, [4, 4, 4])
        gp2 = _ptm_grad_raw(x, gy, shape, targets, [4, 4, 4])
        assert torch.equal(gx, gx2)
        err = ((gp.double() - gp2.double()).abs().max() / gp2.double().abs().max()).item()
        assert err < tol, err
        xn = x.double().cpu().numpy().reshape((B,) + shape)
        yn = gy.double().cpu().numpy().reshape((B,) + shape)
        lx = [n + 1] + list(range(n))
        ly = [n + 1] + list(range(n))
        for j, t in enumerate(targets):
            ly[1 + t] = n + 2 + j
        ref = np.einsum(yn, ly, xn, lx, [n + 1] + [n + 2 + j for j in range(3)] + list(targets))
        err = np.max(np.abs(gp.double().cpu().numpy() - ref.reshape(B, 64, 64))) / np.max(np.abs(ref))
        assert err < tol, err
        # xbar = P^T ybar against numpy (the transposed apply above may run
        # the same family of direct kernels)
        Pn = P.double().cpu().numpy().reshape((P.shape[0],) + (4,) * 6)
        if P.shape[0] == 1:
            Pn = np.broadcast_to(Pn, (B,) + (4,) * 6)
        lp = [n + 1] + [n + 2 + j for j in range(3)] + list(targets)
        xref = np.einsum(Pn, lp, yn, ly, [n + 1] + list(range(n))).reshape(B, -1)
        err = np.max(np.abs(gx.double().cpu().numpy() - xref)) / np.max(np.abs(xref))
        assert err < tol, err


def test_fused_backward_misaligned_states_fall_back():
    """States that are not 16-byte aligned (a view one float in) take the
    two separate kernels, and the geometry stays eligible for aligned
    states afterwards; both agree with the aligned one-pass result."""
    from quantumsim_amd.autograd import _backward_raw, _FUSED_GEOMETRY
    n, targets, B = 7, (1, 2, 3), 2
    shape = (4,) * n
    g = torch.Generator().manual_seed(11)
    N = 4 ** n
    xs = torch.randn(B * N + 1, generator=g).to(DEV)
    ys = torch.randn(B * N + 1, generator=g).to(DEV)
    P = (torch.randn(B, 64, 64, generator=g) / 8).to(DEV)
    x_al, y_al = xs[:B * N].view(B, N), ys[:B * N].view(B, N)
    x_mis = xs[1:].view(B, N)
    y_mis = ys[1:].view(B, N)
    assert x_mis.data_ptr() % 16 and y_mis.data_ptr() % 16
    # the same values at both alignments
    x_mis.copy_(x_al.clone())
    y_mis.copy_(y_al.clone())
    x_al = x_mis.clone()
    y_al = y_mis.clone()
    gx1, gp1 = _backward_raw(x_mis, y_mis, P, shape, targets, True)
    assert _FUSED_GEOMETRY.get((torch.float32, shape, targets)) is not False
    gx2, gp2 = _backward_raw(x_al, y_al, P, shape, targets, True)
    assert _FUSED_GEOMETRY.get((torch.float32, shape, targets)) is True
    assert torch.equal(gx1, gx2)
    err = ((gp1.double() - gp2.double()).abs().max() / gp2.double().abs().max()).item()
    assert err < 1e-5, err


@pytest.mark.parametrize('n,targets', [(6, (3, 4, 5)), (8, (5, 6, 7)), (10, (0, 4, 9)),
                                       (9, (2, 7, 8)), (8, (1, 2, 3))])
def test_batched_apply_dense64_matches_einsum(n, targets):
    """Float32 batched 64 x 64 applies with per-state PTMs against a float64
    einsum: the direct-to-LDS kernel where the PTM's lowest index is the
    innermost axis (k_apply64q), k_fused64 elsewhere."""
    from quantumsim_amd.autograd import _apply_raw
    rng = np.random.default_rng(3 * n + targets[0])
    B = 3
    shape = (4,) * n
    print('dense64 apply', n, targets, flush=True)
    x = rng.standard_normal((B, 4 ** n))
    P = rng.standard_normal((B, 64, 64)) / 8
    y = _apply_raw(torch.tensor(x, dtype=torch.float32, device=DEV),
                   torch.tensor(P, dtype=torch.float32, device=DEV), shape, targets, [4, 4, 4])
    xs = x.reshape((B,) + shape)
    lx = [n + 1] + list(range(n))
    lp = [n + 1] + [n + 2 + j for j in range(3)] + list(targets)
    lo = [n + 1] + [n + 2 + targets.index(q) if q in targets else q for q in range(n)]
    ref = np.einsum(P.reshape((B,) + (4,) * 6), lp, xs, lx, lo).reshape(B, -1)
    err = np.max(np.abs(y.double().cpu().numpy() - ref)) / np.max(np.abs(ref))
    assert err < 1e-5, err


def test_circuit_batch_checkpoint_interval_from_budget():
    """The interval comes from the memory budget (every state kept when it
    fits, the longest interval when nothing does), and it changes only the
    memory: populations bitwise, gradients to 1e-12, graph replays too."""
    circuits = random_circuit_batch(8, seed=21, n_range=(7, 9), layer_range=(3, 11))
    keep_all = CircuitBatch(circuits, dtype=torch.float64, checkpoint_budget=1e15)
    recompute = CircuitBatch(circuits, dtype=torch.float64, checkpoint_budget=0)
    assert keep_all.k == 1 and recompute.k == 16
    assert keep_all.kept_bytes(1) > recompute.kept_bytes(16)
    w = keep_all.random_weights(3)
    _, p1, g1 = keep_all.step(w)
    _, p2, g2 = recompute.step(w)
    for a, b in zip(p1, p2):
        assert torch.equal(a, b)
    for a, b in zip(g1, g2):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-14)
    keep_all.capture(w, autotune=False)
    for _ in range(2):
        _, pr, gr = keep_all.replay()
        torch.cuda.synchronize()
        for a, b in zip(pr, p1):
            assert torch.equal(a, b)
        for a, b in zip(gr, g1):
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-14)


def test_circuit_batch_fp64_trace_and_renormalize():
    """Float64 traces of every circuit (the sum of its float64 populations,
    what the reference's renormalize returns) and populations divided by
    them; the gradient goes through the normalisation."""
    circuits = random_circuit_batch(10, seed=31, n_range=(4, 9), layer_range=(3, 9))
    for dtype, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        cb = CircuitBatch(circuits, dtype=dtype)
        raw = CircuitBatch(circuits, dtype=dtype, renormalize=False)
        w = cb.random_weights(5)
        _, p, g = cb.step(w)
        q = raw.populations()
        for a, b, t in zip(p, q, cb.traces):
            assert t.dtype == torch.float64 and t.shape == (a.shape[0],)
            assert torch.allclose(t, b.detach().sum(-1), rtol=0, atol=1e-15)
            # trace-preserving channels: only the float32 states' rounding moves it
            assert (t - 1).abs().max().item() < (1e-13 if dtype == torch.float64 else 1e-5)
            assert (a.detach().sum(-1) - 1).abs().max().item() < 1e-14
            assert torch.allclose(a.detach(), b.detach() / t[:, None], rtol=1e-14, atol=0)
        loss = sum(((b / b.sum(-1, keepdim=True)) * wt).sum() for b, wt in zip(q, w))
        g2 = torch.autograd.grad(loss, raw.parameters())
        for a, b in zip(g, g2):
            assert (a - b).abs().max().item() <= tol * max(1.0, b.abs().max().item())


def test_circuit_batch_side_stream_equals_one_stream():
    """The narrow circuits' persistent kernels on a second stream give the
    one-stream results, eagerly and in graph replays (joined streams)."""
    circuits = random_circuit_batch(14, seed=41, n_range=(4, 9), layer_range=(3, 12))
    one = CircuitBatch(circuits, dtype=torch.float32, overlap_narrow=False)
    two = CircuitBatch(circuits, dtype=torch.float32)
    assert one.side is None and two.side is not None and two.small
    w = one.random_weights(4)
    _, p1, g1 = one.step(w)
    _, p2, g2 = two.step(w)
    for a, b in zip(p1, p2):
        assert torch.equal(a, b)
    for a, b in zip(g1, g2):
        assert torch.equal(a, b)
    two.capture(w, autotune=False)
    for _ in range(3):
        _, pr, gr = two.replay()
        torch.cuda.synchronize()
        for a, b in zip(pr, p1):
            assert torch.equal(a, b)
        for a, b in zip(gr, g1):
            assert torch.equal(a, b)


def test_capture_step_trainable_circuit():
    """A differentiable step (trainable circuit in 3-qubit blocks, KL loss,
    gradients of every parameter) captured with capture_step: replays equal
    the eager step, and follow parameter updates made in place."""
    from quantumsim_amd.autograd import capture_step
    tc = TrainableCircuit(n=6, layers=8)
    ang, t1, t2 = tc.initial_params(4, seed=8, device=DEV, dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang + 0.1, t1, t2, torch.float32)
    params = [ang.requires_grad_(), t1.requires_grad_(), t2.requires_grad_()]

    def step():
        p = tc.populations(*params, torch.float32, checkpoint_every=4, fuse_qudits=3)
        return torch.autograd.grad(tc.kl(target, p).sum(), params)

    graph, stream, out = capture_step(step)
    for rep in range(3):
        with torch.no_grad():
            ang.add_(0.01)                      # in place: the graph reads the same memory
        with torch.cuda.stream(stream):
            graph.replay()
        torch.cuda.synchronize()
        ref = step()
        for a, b in zip(out, ref):
            assert torch.equal(a, b), rep


def test_batched_apply_large_batch_equals_per_state():
    """A batch whose states exceed the nontemporal threshold together (16
    float32 states of 11 qubits: 0.5 GB read and written) but not one by one:
    the batched launch (nontemporal accesses) matches per-state launches
    (cached accesses) for per-state 16 x 16 PTMs on middle and inner qubits
    (direct and LDS-tiled kernels) and dense 64 x 64 blocks (MFMA)."""
    from quantumsim_amd.autograd import _apply_raw
    n, B = 11, 16
    shape = [4] * n
    g = torch.Generator().manual_seed(5)
    x = torch.randn(B, 4 ** n, generator=g).to(DEV, torch.float32)
    for targets in ([4, 5], [9, 10], [3, 4, 5], [8, 9, 10]):
        D = 4 ** len(targets)
        ptm = (torch.randn(B, D, D, generator=g, dtype=torch.float64) / D).to(DEV)
        dims_out = [4] * len(targets)
        y = _apply_raw(x, ptm, shape, targets, dims_out)
        for b in range(B):
            yb = _apply_raw(x[b:b + 1].contiguous(), ptm[b:b + 1], shape, targets, dims_out)
            scale = float(yb.abs().max())
            err = float((y[b] - yb[0]).abs().max())
            assert err <= 1e-6 * scale, (targets, b, err, scale)
