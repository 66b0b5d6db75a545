"""The benchmark workloads at their own sizes, checked on the GPU.

* 14 qubits, 60 layers (the plain-method circuit): float64 trace exactly
  preserved, float32 populations against the float64 run, fused blocks
  against gate by gate.
* The batched sweep: one circuit per width 4..11 (plus an 11-qubit circuit
  at each end of the 20..80-layer range) in one batch, float64 populations
  against the oracle (1e-12) and float32 renormalised populations against
  the oracle (1e-5 absolute), gradients against central differences; and
  the full 512-circuit float32 batch itself: graph replays equal the eager
  step, float64 traces within 1e-5 of 1, renormalised populations summing
  to 1, the deepest circuit of every width against the oracle.
* The differentiable workload: 12 qubits, 40 layers, 96 parameters,
  checkpoints every 8 blocks, float64: gradient entries against central
  differences (h = 1e-6), populations against the oracle, and a
  directional derivative against central differences of the oracle's own
  loss; the 64-set float32 step captured in a HIP graph against eager runs
  and against float64.

The oracle's circuits are built here from Kraus operators with the
oracle's kraus_to_ptm (quantumsim/algebra/algebra.py:11-30) and the
Kraus sets of quantumsim/operations/qubits.py (RX :38-53, RZ :74-90,
cphase :109-123, amplitude / phase damping :182-214), so they pin the
product's PTM construction and block composition as well as its kernels.
"""
import numpy as np
import pytest
import torch

import quantumsim_amd as qa
from quantumsim_amd import bases
from quantumsim_amd.batch import CircuitBatch
from quantumsim_amd.fusion import apply_circuit
from quantumsim_amd.workloads import noisy_layers, random_circuit_batch, TrainableCircuit
from oracle.pauli_vector import PauliVectorNumpy
from oracle import bases as ob, algebra as oa

pytestmark = pytest.mark.gpu
DEV = 'cuda'
QB = [ob.general(2)]
GATE_US = 0.02                      # 20 ns gate, relaxation times in microseconds


# ------------------------------------------------------------ the oracle side
def _ptm1(kraus):
    return oa.kraus_to_ptm(np.asarray(kraus, dtype=complex), QB, QB).reshape(4, 4)


def _rx(a):
    c, s = np.cos(a / 2), np.sin(a / 2)
    return _ptm1([[[c, -1j * s], [-1j * s, c]]])


def _rz(b):
    return _ptm1([np.diag([np.exp(-0.5j * b), np.exp(0.5j * b)])])


def _damp(t1, t2, t=GATE_US):
    g = 1 - np.exp(-t / t1)
    lam = 1 - np.exp(-2 * t * (1 / t2 - 0.5 / t1))
    amp = _ptm1([[[1, 0], [0, np.sqrt(1 - g)]], [[0, np.sqrt(g)], [0, 0]]])
    ph = _ptm1([[[1, 0], [0, np.sqrt(1 - lam)]], [[0, 0], [0, np.sqrt(lam)]]])
    return ph @ amp


_CZ = oa.kraus_to_ptm(np.diag([1, 1, 1, -1]).astype(complex)[None], QB * 2,
                      QB * 2).reshape(16, 16)


def oracle_populations(n, layer_angles, t1, t2):
    """Populations of the layered circuit on n qubits from |0..0>: in layer
    l, RX(a_q) then RZ(b_q) on every qubit, CZ on the bonds (q, q+1) for q =
    l mod 2, l mod 2 + 2, ..., then amplitude + phase damping of every
    qubit.  A bond's gates are applied as one 16x16 PTM
    (D (x) D) CZ (R (x) R), a qubit outside every bond as D R.
    layer_angles: (L, n, 2) = (RX angle, RZ angle)."""
    D = [_damp(t1[q], t2[q]) for q in range(n)]
    o = PauliVectorNumpy(QB * n, force=True)
    for layer, ang in enumerate(layer_angles):
        R = [_rz(ang[q, 1]) @ _rx(ang[q, 0]) for q in range(n)]
        bonds = [(q, q + 1) for q in range(layer % 2, n - 1, 2)]
        paired = {q for b in bonds for q in b}
        for q, r in bonds:
            P = np.kron(D[q], D[r]) @ _CZ @ np.kron(R[q], R[r])
            o.apply_ptm(P.reshape(4, 4, 4, 4), q, r)
        for q in range(n):
            if q not in paired:
                o.apply_ptm(D[q] @ R[q], q)
    return o.diagonal()


def _circuit_pops(c):
    return oracle_populations(c['n'], c['angles'], c['t1'], c['t2'])


def _trainable_layer_angles(tc, ang):
    """(L, n, 2) angles of TrainableCircuit's layers: block
    (layer // block) mod n_blocks of one parameter set's [blocks, n, 2]."""
    return np.stack([ang[(l // tc.block) % tc.n_blocks] for l in range(tc.layers)])


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


# ------------------------------------------------ 14 qubits, 60 layers (plain)
def test_14q_60_layer_circuit_full_size():
    """The plain-method circuit at 14 qubits (4^14 entries): the gates are
    CPTP, so the float64 trace stays 1; the float32 run's populations are
    within 1e-5 of the float64 run's; dense 3-qubit blocks (MFMA kernel)
    equal gate-by-gate application to 1e-12.  (The 10-qubit slice of the
    same circuit is compared with the oracle in test_workloads_gpu.)"""
    n = 14
    circuit, _ = noisy_layers(n, 60, 1, 'brickwork')
    v = qa.PauliVectorHIP([bases.general(2)] * n, force=True)
    for p, q in circuit:
        v.apply_ptm(p, *q)
    d64 = v.diagonal()
    assert abs(v.trace() - 1) < 1e-12
    del v
    v32 = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype='float32')
    for p, q in circuit:
        v32.apply_ptm(p, *q)
    err32 = maxabs(v32.diagonal(), d64)
    print('14q float32 populations vs float64: %.3e' % err32)
    assert err32 < 1e-5
    assert abs(v32.renormalize() - 1) < 1e-5
    del v32
    vf = qa.PauliVectorHIP([bases.general(2)] * n, force=True)
    napp = apply_circuit(vf, circuit, 3)
    errf = maxabs(vf.diagonal(), d64)
    print('14q fused <= 3 qubits: %d applications, populations vs gate by gate %.3e'
          % (napp, errf))
    assert napp < len(circuit) // 5
    assert errf < 1e-12


# ------------------------------------------------ the batched sweep
def _width_sweep():
    """One circuit per width 4..11 at depths spread over 20..80 layers, plus
    an 11-qubit circuit at 80 and one at 20 layers (seeded as the sweep
    draws them: angles ~ U(0, 2 pi), T1 ~ U(20, 40) us, T2 ~ U(10, T1))."""
    depth = {4: 20, 5: 80, 6: 37, 7: 63, 8: 29, 9: 71, 10: 46, 11: 54}
    g = np.random.default_rng(77)
    out = []
    for n, L in list(depth.items()) + [(11, 80), (11, 20)]:
        t1 = g.uniform(20, 40, size=n)
        t2 = np.array([g.uniform(10, t1[q]) for q in range(n)])
        out.append({'n': n, 'layers': L, 'angles': g.uniform(0, 2 * np.pi, size=(L, n, 2)),
                    't1': t1, 't2': t2})
    return out


def test_batch_every_width_against_oracle():
    """Widths 4..11 and depths 20..80 in one batch (narrow circuits in the
    persistent kernels, wide ones in batched 64x64 MFMA launches): float64
    populations within 1e-12 of the oracle, float32 renormalised
    populations within 1e-5 absolute, float64 traces within 1e-12 of 1."""
    circuits = _width_sweep()
    refs = [_circuit_pops(c) for c in circuits]
    cb64 = CircuitBatch(circuits, dtype=torch.float64, renormalize=False)
    cb32 = CircuitBatch(circuits, dtype=torch.float32)
    assert cb64.small and len(cb64.small) < len(cb64.groups)     # both paths used
    with torch.no_grad():
        p64 = cb64.populations()
        p32 = cb32.populations()
    worst64 = worst32 = 0.0
    for gi, g in enumerate(cb64.groups):
        assert (cb64.traces[gi] - 1).abs().max().item() < 1e-12
        for b, i in enumerate(g.members):
            e64 = maxabs(p64[gi][b].cpu().numpy(), refs[i])
            e32 = maxabs(p32[gi][b].cpu().numpy(), refs[i])
            print('%2dq %2d layers: float64 %.2e  float32 %.2e' % (g.n, g.layers[b], e64, e32))
            worst64, worst32 = max(worst64, e64), max(worst32, e32)
    assert worst64 < 1e-12
    assert worst32 < 1e-5


def test_batch_every_width_gradient_central_differences():
    """Gradients of loss = sum_c <w_c, p_c> for the width sweep in float64:
    angles of the first and last layer of the 11-qubit 80-layer circuit, a
    T1 of a 9-qubit circuit and a T2 of a persistent-kernel circuit against
    central differences (h = 1e-6) of the batch's own float64 forward."""
    circuits = _width_sweep()
    cb = CircuitBatch(circuits, dtype=torch.float64)
    w = cb.random_weights(5)
    _, _, grads = cb.step(w)
    by_n = {g.n: gi for gi, g in enumerate(cb.groups)}
    g11 = cb.groups[by_n[11]]
    b80 = g11.layers.index(80)
    picks = [(by_n[11], 0, (b80, 0, 3, 0)), (by_n[11], 0, (b80, 79, 10, 1)),
             (by_n[9], 1, (0, 4)), (by_n[5], 2, (0, 2)), (by_n[7], 0, (0, 30, 6, 0))]
    h = 1e-6

    def loss():
        with torch.no_grad():
            return sum((p * ww).sum() for p, ww in zip(cb.populations(), w)).item()
    for gi, which, idx in picks:
        t = cb.parameters()[3 * gi + which]
        vals = []
        for sgn in (1, -1):
            with torch.no_grad():
                t[idx] += sgn * h
            vals.append(loss())
            with torch.no_grad():
                t[idx] -= sgn * h
        fd = (vals[0] - vals[1]) / (2 * h)
        an = grads[3 * gi + which][idx].item()
        print('group %d param %d %s: autograd %.10e  central difference %.10e'
              % (gi, which, idx, an, fd))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-8


def test_batch_512_circuits_full_size():
    """The 512-circuit sweep (4..11 qubits, 20..80 layers) in float32 with
    float64 readout, as benchmarked: graph replays equal the eager step,
    every circuit's float64 trace is within 1e-5 of 1, renormalised
    populations sum to 1, populations are within 1e-5 of the float64 batch
    and, for the deepest circuit of every width, of the oracle."""
    circuits = random_circuit_batch(512, seed=2024)
    cb = CircuitBatch(circuits, dtype=torch.float32, checkpoint_every=4)
    w = cb.random_weights(7)
    loss_e, pops_e, grads_e = cb.step(w)
    traces_e = [t.clone() for t in cb.traces]
    pops_e = [p.detach().clone() for p in pops_e]
    grads_e = [g.clone() for g in grads_e]
    cb.capture(w, autotune=False)
    for rep in range(2):
        torch.randn(1000, device=DEV).sum().item()          # eager work between replays
        loss_g, pops_g, grads_g = cb.replay()
        torch.cuda.synchronize()
        for a, b in zip(pops_g, pops_e):
            assert torch.equal(a, b), rep
        for a, b in zip(grads_g, grads_e):
            assert torch.equal(a, b), rep
    for t, p in zip(traces_e, pops_e):
        assert (t - 1).abs().max().item() < 1e-5
        assert (p.sum(-1) - 1).abs().max().item() < 1e-12
    del cb, grads_e, grads_g, pops_g
    torch.cuda.empty_cache()
    cb64 = CircuitBatch(circuits, dtype=torch.float64, checkpoint_every=16,
                        requires_grad=False)
    with torch.no_grad():
        p64 = cb64.populations()
    err = max(maxabs(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(pops_e, p64))
    print('512 circuits: float32 vs float64 populations %.3e' % err)
    assert err < 1e-5
    worst = 0.0
    for gi, g in enumerate(cb64.groups):
        i = g.members[0]                                       # the deepest of this width
        ref = _circuit_pops(circuits[i])
        e64 = maxabs(p64[gi][0].cpu().numpy(), ref)
        e32 = maxabs(pops_e[gi][0].cpu().numpy(), ref)
        print('%2dq %2d layers vs oracle: float64 %.2e  float32 %.2e'
              % (g.n, g.layers[0], e64, e32), flush=True)
        assert e64 < 1e-12
        worst = max(worst, e32)
    assert worst < 1e-5


# ------------------------------------------------ the differentiable workload
def _trainable_setup(sets, dtype=torch.float64):
    tc = TrainableCircuit(n=12, layers=40)
    assert tc.n_params == 96
    ang, t1, t2 = tc.initial_params(sets, seed=11, device=DEV, dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang + 0.1, t1 * 1.05, t2 * 0.97, dtype)
    return tc, [ang, t1, t2], target


def _trainable_loss(tc, params, target, dtype):
    p = tc.populations(*params, dtype, checkpoint_every=8, fuse_qudits=3)
    return tc.kl(target, p).sum(), p


def test_trainable_12q_gradient_central_differences_and_oracle():
    """12 qubits, 40 layers, 96 parameters per set, KL loss, dense 3-qubit
    blocks, a checkpoint every 8 blocks, float64, 4 parameter sets: six
    gradient entries (angles, T1, T2, different sets) against central
    differences (h = 1e-6); set 0's populations against the oracle."""
    tc, base, target = _trainable_setup(4)
    params = [t.clone().requires_grad_() for t in base]
    loss, pops = _trainable_loss(tc, params, target, torch.float64)
    grads = torch.autograd.grad(loss, params)
    h = 1e-6
    picks = [(0, (0, 0, 0, 0)), (0, (3, 1, 5, 1)), (0, (2, 2, 11, 0)), (1, (0, 4)),
             (1, (3, 9)), (2, (1, 7))]
    for which, idx in picks:
        vals = []
        for sgn in (1, -1):
            p2 = [t.detach().clone() for t in params]
            p2[which][idx] += sgn * h
            with torch.no_grad():
                vals.append(_trainable_loss(tc, p2, target, torch.float64)[0].item())
        fd = (vals[0] - vals[1]) / (2 * h)
        an = grads[which][idx].item()
        print('param %d %s: autograd %.10e  central difference %.10e' % (which, idx, an, fd))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 2e-8
    ang, t1, t2 = [t.detach().cpu().numpy() for t in base]
    ref = oracle_populations(12, _trainable_layer_angles(tc, ang[0]), t1[0], t2[0])
    err = maxabs(pops[0].detach().cpu().numpy(), ref)
    print('12q populations vs oracle: %.3e' % err)
    assert err < 1e-12


def test_trainable_12q_directional_derivative_vs_oracle():
    """The oracle's own loss (its circuit built from Kraus operators, run by
    the numpy backend) differentiated by central differences (h = 1e-6)
    along a random direction through all 96 parameters of set 0, against
    the HIP adjoint's gradient projected on that direction."""
    tc, base, target = _trainable_setup(1)
    params = [t.clone().requires_grad_() for t in base]
    loss, _ = _trainable_loss(tc, params, target, torch.float64)
    grads = torch.autograd.grad(loss, params)
    rng = np.random.default_rng(3)
    dirs = [rng.standard_normal(t.shape) for t in base]      # radians and microseconds
    an = sum(float((g.cpu().numpy() * d).sum()) for g, d in zip(grads, dirs))
    tgt = np.maximum(target[0].cpu().numpy(), 1e-30)
    t_raw = target[0].cpu().numpy()
    h = 1e-6

    def oracle_loss(sgn):
        ang, t1, t2 = [t.detach().cpu().numpy()[0] + sgn * h * d[0]
                       for t, d in zip(base, dirs)]
        p = oracle_populations(12, _trainable_layer_angles(tc, ang), t1, t2)
        return float(np.sum(t_raw * (np.log(tgt) - np.log(np.maximum(p, 1e-30)))))
    fd = (oracle_loss(1) - oracle_loss(-1)) / (2 * h)
    print('directional derivative: HIP adjoint %.10e  oracle central difference %.10e'
          % (an, fd))
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-8


def test_trainable_12q_64_sets_float32_graph():
    """The benchmarked step: 64 parameter sets, float32 states: replays of
    the step captured in a HIP graph equal the eager step; populations
    within 1e-5 of the float64 run and gradients within 1e-3 (relative to
    each tensor's largest entry) of it."""
    from quantumsim_amd.autograd import capture_step
    sets = 64
    tc, base, target = _trainable_setup(sets, torch.float32)
    params = [t.clone().requires_grad_() for t in base]

    def step():
        loss, p = _trainable_loss(tc, params, target, torch.float32)
        return (loss, p) + tuple(torch.autograd.grad(loss, params))
    eager = [t.detach().clone() for t in step()]
    graph, stream, outs = capture_step(step)
    for rep in range(2):
        torch.randn(1000, device=DEV).sum().item()
        with torch.cuda.stream(stream):
            graph.replay()
        stream.synchronize()
        for a, b in zip(outs, eager):
            assert torch.equal(a, b), rep
    pops32, g32 = eager[1], eager[2:]
    del graph, outs
    torch.cuda.empty_cache()
    sub = 16                                    # float64 comparison on the first 16 sets
    p64s = [t[:sub].clone().requires_grad_() for t in base]
    loss64, pops64 = _trainable_loss(tc, p64s, target[:sub].double(), torch.float64)
    g64 = torch.autograd.grad(loss64, p64s)
    err = (pops32[:sub].double() - pops64).abs().max().item()
    print('64 sets float32: populations vs float64 %.3e' % err)
    assert err < 1e-5
    for a, b in zip(g32, g64):
        rel = (a[:sub].double() - b).abs().max().item() / b.abs().max().item()
        print('gradient relative error vs float64: %.3e' % rel)
        assert rel < 1e-3
