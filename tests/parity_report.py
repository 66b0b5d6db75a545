"""Measured error of every result of the benchmark workloads, as a table.

Each result is compared with the float64 numpy oracle on the same seeded
inputs.  Where the oracle cannot finish at full size (the 14-qubit circuit,
the 512-circuit batch, the 12-qubit 64-set gradient) the full-size result
is compared with the float64 device run, which the same table pins to the
oracle on the instances the oracle can run, and with size-independent
properties (trace preservation, finite differences).

Run on a GPU from the repository root:

    python tests/parity_report.py --out parity.md

Every row: workload, result, element type, what it is compared with, the
maximum absolute error (relative where the row says so) and the bar the
tests hold it to.  Not collected by pytest (the same checks with their
assertions live in the test_*_gpu.py files).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ROWS = []
T0 = time.time()


def row(workload, result, dtype, vs, err, bar):
    ROWS.append((workload, result, dtype, vs, err, bar))
    print('[%6.1f s] %-34s %-34s %-4s %-26s %.3e  (bar %s)'
          % (time.time() - T0, workload, result, dtype, vs, err, bar), flush=True)


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64))))


# ------------------------------------------------------------ plain circuits
def plain_circuits():
    import quantumsim_amd as qa
    from quantumsim_amd import bases
    from quantumsim_amd.fusion import apply_circuit
    from quantumsim_amd.workloads import noisy_layers
    from oracle.pauli_vector import PauliVectorNumpy
    from oracle import bases as ob

    def run(circuit, n, dtype):
        v = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype=dtype)
        for p, q in circuit:
            v.apply_ptm(p, *q)
        return v

    for n, layers, pairing, seed, name in [(6, 40, 'random', 0, '6q 40 layers (CPU config)'),
                                           (10, 60, 'brickwork', 1, '10q 60-layer brickwork')]:
        circuit, _ = noisy_layers(n, layers, seed, pairing)
        o = PauliVectorNumpy([ob.general(2)] * n, force=True)
        for p, q in circuit:
            o.apply_ptm(p, *q)
        d_ref, t_ref, pv_ref = o.diagonal(), o.trace(), o.to_pv()
        v = run(circuit, n, 'float64')
        row(name, 'Pauli vector', 'f64', 'oracle', maxabs(v.to_pv(), pv_ref), '1e-12')
        row(name, 'diagonal (%d probabilities)' % 2 ** n, 'f64', 'oracle',
            maxabs(v.diagonal(), d_ref), '1e-12')
        row(name, 'trace', 'f64', 'oracle', abs(v.trace() - t_ref), '1e-12')
        v32 = run(circuit, n, 'float32')
        row(name, 'diagonal', 'f32', 'oracle', maxabs(v32.diagonal(), d_ref), '1e-5')
        row(name, 'trace (float64 sum)', 'f32', 'oracle', abs(v32.trace() - t_ref), '1e-5')
        row(name, 'renormalize() return', 'f32', 'oracle', abs(v32.renormalize() - t_ref),
            '1e-5')
        for k in (3, 4):
            vf = qa.PauliVectorHIP([bases.general(2)] * n, force=True)
            apply_circuit(vf, circuit, k)
            row(name, 'fused <= %d qubits (MFMA), vector' % k, 'f64', 'oracle',
                maxabs(vf.to_pv(), pv_ref), '1e-12')
            vf32 = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype='float32')
            apply_circuit(vf32, circuit, k)
            row(name, 'fused <= %d qubits (MFMA), diagonal' % k, 'f32', 'oracle',
                maxabs(vf32.diagonal(), d_ref), '1e-5')

    # the 14-qubit circuit itself: too large for the oracle (~1 s per gate)
    n = 14
    circuit, _ = noisy_layers(n, 60, 1, 'brickwork')
    v = run(circuit, n, 'float64')
    d64 = v.diagonal()
    row('14q 60-layer brickwork', 'trace (CPTP: exactly 1)', 'f64', '1', abs(v.trace() - 1),
        '1e-12')
    del v
    v32 = run(circuit, n, 'float32')
    row('14q 60-layer brickwork', 'diagonal', 'f32', 'f64 device run',
        maxabs(v32.diagonal(), d64), '1e-5')
    del v32
    vf = qa.PauliVectorHIP([bases.general(2)] * n, force=True)
    apply_circuit(vf, circuit, 3)
    row('14q 60-layer brickwork', 'fused <= 3 qubits, diagonal', 'f64', 'f64 gate by gate',
        maxabs(vf.diagonal(), d64), '1e-12')


# ------------------------------------------------------------ mixed qudits
def mixed_qudits(full_size):
    import quantumsim_amd as qa
    from quantumsim_amd import bases
    from quantumsim_amd.algebra import kraus_to_ptm, dm_to_pv
    from quantumsim_amd.workloads import mixed_qudit_program
    from oracle.pauli_vector import PauliVectorNumpy, multitake
    from oracle import bases as ob

    FULL = {3: bases.gell_mann(3), 2: bases.general(2)}
    DIAG2 = bases.general(2).subbasis([0, 1])
    SEL = np.zeros((2, 4))
    SEL[0, 0] = SEL[1, 1] = 1

    def obs(bs):
        return [ob.Basis(np.array(b.vectors), list(b.labels)) for b in bs]

    def initial(dims, rng):
        pv = np.ones(())
        for d in dims:
            a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
            rho = a @ a.conj().T
            rho /= np.trace(rho)
            pv = np.multiply.outer(pv, dm_to_pv(rho, [FULL[d]]))
        return pv

    cases = [([3, 2, 3, 2, 3], 12, 5, 11, '3 qutrits + 2 qubits, 12 steps')]
    if full_size:
        cases.append(([3, 3, 2, 3, 2, 3, 3, 2, 3, 2], 3, 6, 12,
                      '6 qutrits + 4 qubits (full), 3 steps'))
    for dims, steps, seed, iseed, name in cases:
        rng = np.random.default_rng(iseed)
        pv0 = initial(dims, rng)
        bs = [FULL[d] for d in dims]
        v = qa.PauliVectorHIP(bs, pv0.copy(), force=True)
        o = PauliVectorNumpy(obs(bs), pv0.copy(), force=True)
        err = {'pv': 0.0, 'meas_prob': 0.0, 'partial_trace': 0.0, 'trace': 0.0,
               'diagonal': 0.0}
        exact = {'reduction': 0, 'multitake': 0}
        for step in mixed_qudit_program(dims, steps, seed=seed):
            a, b = step['pair']
            b_in = (v.bases[a], v.bases[b])
            b_out = tuple(FULL[v.bases[q].dim_hilbert] for q in (a, b))
            ptm = kraus_to_ptm(step['kraus'], b_in, b_out)
            v.apply_ptm(ptm, a, b)
            o.apply_ptm(ptm, a, b)
            for q, bb in zip((a, b), b_out):
                v.bases[q] = bb
                o.bases[q] = obs([bb])[0]
            err['pv'] = max(err['pv'], maxabs(v.to_pv(), o.to_pv()))
            q = step['measure']
            err['meas_prob'] = max(err['meas_prob'], maxabs(v.meas_prob(q), o.meas_prob(q)))
            if v.bases[q].dim_pauli == 4:
                before = v.to_pv()
                v.apply_ptm(SEL, q)
                o.apply_ptm(SEL, q)
                v.bases[q] = DIAG2
                o.bases[q] = obs([DIAG2])[0]
                idx = [list(range(s)) for s in before.shape]
                idx[q] = [0, 1]
                exact['reduction'] += int(np.count_nonzero(v.to_pv() != multitake(before, idx)))
            keep = step['keep']
            err['partial_trace'] = max(err['partial_trace'],
                                       maxabs(v.partial_trace(*keep).to_pv(),
                                              o.partial_trace(*keep).to_pv()))
            cur = v.to_pv()
            idx = [list(range(s)) for s in cur.shape]
            for qq, lst in zip(keep, step['take']):
                idx[qq] = [i for i in lst if i < cur.shape[qq]] or [0]
            exact['multitake'] += int(np.count_nonzero(v.multitake(*idx) != multitake(cur, idx)))
            err['trace'] = max(err['trace'], abs(v.trace() - o.trace()))
            err['diagonal'] = max(err['diagonal'], maxabs(v.diagonal(), o.diagonal()))
            print('   %s: step on %s done' % (name, step['pair']), flush=True)
        for k, e in err.items():
            row(name, k + ' (max over steps)', 'f64', 'oracle', e, '1e-12')
        for k, c in exact.items():
            row(name, k + ': entries differing', 'f64', 'numpy gather', float(c), '0 (bit-exact)')
        del v, o


# ------------------------------------------------------------ batched sweep
def batched(n_circuits):
    import torch
    from quantumsim_amd.batch import CircuitBatch
    from quantumsim_amd.workloads import (random_circuit_batch, rotation_ptms, damping_ptms,
                                          fused_layer_gates)
    from oracle.pauli_vector import PauliVectorNumpy
    from oracle import bases as ob

    circuits = random_circuit_batch(n_circuits, seed=2024)
    name = '%d circuits, 4-11 qubits' % n_circuits
    cb32 = CircuitBatch(circuits, dtype=torch.float32, checkpoint_every=16)
    w = cb32.random_weights(7)
    _, p32, g32 = cb32.step(w)
    torch.cuda.synchronize()
    p32 = [p.detach().cpu().numpy() for p in p32]
    g32 = [g.detach().cpu().numpy() for g in g32]
    del cb32
    cb64 = CircuitBatch(circuits, dtype=torch.float64, checkpoint_every=16)
    _, p64, g64 = cb64.step(w)
    torch.cuda.synchronize()
    p64 = [p.detach().cpu().numpy() for p in p64]
    g64 = [g.detach().cpu().numpy() for g in g64]
    groups = [(g.n, list(g.members)) for g in cb64.groups]
    del cb64
    torch.cuda.empty_cache()
    row(name, 'populations (all circuits)', 'f32', 'f64 device run',
        max(maxabs(a, b) for a, b in zip(p32, p64)), '1e-5')
    rel = max(float(np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-300))
              for a, b in zip(g32, g64))
    row(name, 'parameter gradients (relative)', 'f32', 'f64 device run', rel, '1e-4')

    # the oracle on one circuit of every width up to 10 qubits
    def oracle_pops(c):
        n = c['n']
        with torch.no_grad():
            rot = rotation_ptms(torch.tensor(c['angles'][None], device='cuda'))
            damp = damping_ptms(torch.tensor(c['t1'][None], device='cuda'),
                                torch.tensor(c['t2'][None], device='cuda'))
            o = PauliVectorNumpy([ob.general(2)] * n)
            for layer in range(c['layers']):
                for g in fused_layer_gates(n, layer, rot[:, layer], damp):
                    k = len(g.targets)
                    o.apply_ptm(g.ptm[0].cpu().numpy().reshape((4,) * (2 * k)), *g.targets)
        return o.diagonal()

    worst64 = worst32 = 0.0
    checked = []
    for gi, (n, members) in enumerate(groups):
        if n > 10:
            continue
        b = 0                                   # one circuit of this width
        c = circuits[members[b]]
        ref = oracle_pops(c)
        worst64 = max(worst64, maxabs(p64[gi][b], ref))
        worst32 = max(worst32, maxabs(p32[gi][b], ref))
        checked.append('%dq/%dL' % (n, c['layers']))
        print('   oracle checked %s' % checked[-1], flush=True)
    row(name, 'populations, %d circuits (%s)' % (len(checked), ' '.join(checked)), 'f64',
        'oracle', worst64, '1e-12')
    row(name, 'populations, same circuits', 'f32', 'oracle', worst32, '1e-5')


# ------------------------------------------------------------ differentiation
def differentiable(sets):
    import torch
    from quantumsim_amd.workloads import TrainableCircuit
    from oracle.pauli_vector import PauliVectorNumpy
    from oracle import bases as ob

    # full size: 12 qubits, 40 layers, 96 parameters x `sets`
    tc = TrainableCircuit(n=12, layers=40)
    ang, t1, t2 = tc.initial_params(sets, seed=11, device='cuda', dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang + 0.1, t1 * 1.05, t2 * 0.97, torch.float64)

    def loss(params, dtype):
        p = tc.populations(*params, dtype, checkpoint_every=8, fuse_qudits=3)
        return tc.kl(target, p).sum()

    params = [ang.clone().requires_grad_(), t1.clone().requires_grad_(),
              t2.clone().requires_grad_()]
    g64 = torch.autograd.grad(loss(params, torch.float64), params)
    g32 = torch.autograd.grad(loss(params, torch.float32), params)
    name = '12q 40 layers, 96 params x %d sets' % sets
    rel = max(float((a.double() - b).abs().max() / (b.abs().max() + 1e-300))
              for a, b in zip(g32, g64))
    row(name, 'gradient (relative)', 'f32', 'f64 device run', rel, '1e-4')
    h = 1e-6
    worst = 0.0
    picks = [(0, (0, 0, 0, 0)), (0, (sets - 1, 1, 5, 1)), (0, (3 % sets, 2, 11, 0)),
             (1, (0, 4)), (1, (sets - 1, 9)), (2, (0, 7))]
    for which, idx in picks:
        vals = []
        for sgn in (1, -1):
            p2 = [t.detach().clone() for t in params]
            p2[which][idx] += sgn * h
            with torch.no_grad():
                vals.append(loss(p2, torch.float64).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        an = g64[which][idx].item()
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    row(name, 'gradient, 6 entries (relative)', 'f64', 'central FD h=1e-6', worst, '1e-6')

    # the oracle's own finite differences on a small instance
    tc4 = TrainableCircuit(n=4, layers=6)
    a4, b4, c4 = tc4.initial_params(1, seed=5, device='cuda', dtype=torch.float64)
    with torch.no_grad():
        tgt = tc4.populations(a4 * 0.9, b4, c4, torch.float64)
    p4 = [a4.clone().requires_grad_(), b4.clone().requires_grad_(), c4.clone().requires_grad_()]
    g4 = torch.autograd.grad(tc4.kl(tgt, tc4.populations(*p4, torch.float64, 5)).sum(), p4)
    tn = tgt[0].cpu().numpy()

    def oracle_loss(a, b, c):
        with torch.no_grad():
            gates = tc4.gates(a, b, c)
        o = PauliVectorNumpy([ob.general(2)] * tc4.n)
        for g in gates:
            k = len(g.targets)
            o.apply_ptm(g.ptm[0].cpu().numpy().reshape((4,) * (2 * k)), *g.targets)
        p = np.maximum(o.diagonal(), 1e-30)
        t = np.maximum(tn, 1e-30)
        return float(np.sum(tn * (np.log(t) - np.log(p))))

    worst = 0.0
    for which, idx in [(0, (0, 0, 1, 0)), (0, (0, 2, 3, 1)), (1, (0, 2)), (2, (0, 1))]:
        vals = []
        for sgn in (1, -1):
            p2 = [t.detach().clone() for t in p4]
            p2[which][idx] += sgn * h
            vals.append(oracle_loss(*p2))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = g4[which][idx].item()
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    row('4q 6 layers, 32 params', 'gradient, 4 entries (relative)', 'f64',
        'oracle central FD', worst, '1e-6')


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--out', default=None, help='also write the markdown table here')
    ap.add_argument('--circuits', type=int, default=512)
    ap.add_argument('--param-sets', type=int, default=16,
                    help='parameter sets of the 12-qubit gradient (independent of each other; '
                         'float64 with 8-gate checkpoints needs ~4.4 GB per set)')
    ap.add_argument('--no-full-mixed', action='store_true',
                    help='skip the full-size 10-qudit mixed instance')
    args = ap.parse_args()
    import torch
    if not torch.cuda.is_available():
        sys.exit('parity_report.py needs a GPU')
    plain_circuits()
    mixed_qudits(not args.no_full_mixed)
    batched(args.circuits)
    differentiable(args.param_sets)
    lines = ['| workload | result | type | compared with | max error | bar |',
             '|---|---|---|---|---|---|']
    for w, r, d, vs, e, bar in ROWS:
        lines.append('| %s | %s | %s | %s | %.2e | %s |' % (w, r, d, vs, e, bar))
    table = '\n'.join(lines)
    print(table)
    if args.out:
        with open(args.out, 'w') as fh:
            fh.write('Command: `python tests/parity_report.py`\n\n' + table + '\n')


if __name__ == '__main__':
    main()
