"""Workloads for rocprofv3 runs (kernel trace, FETCH_SIZE / WRITE_SIZE).

    python -m quantumsim_amd.profiling apply2q --qubits 14 --dtype float64 --reps 3
    python -m quantumsim_amd.profiling fused --qubits 14 --dtype float64 --reps 3

Every run first copies a buffer of known size (2 * 4^n * element bytes of
traffic, torch's copy kernel) so that HBM counters can be scaled against a
plain copy, then runs the kernels under study.
  apply2q : the 13 two-qubit PTM applications of one brickwork layer pair
  apply1q : one-qubit PTM applications on qubits 0, n/2, n-1
  fused   : dense 3- and 4-qubit PTMs (MFMA kernel) on middle and inner qubits
  diagonal: the float64 population readout (2^n-entry gather) of the state
  batched : the 512-circuit batch tuned and captured, then a marker kernel
            (k_scale of 4,099 elements) and graph replays of forward + adjoint;
            --block-kernel 0 builds the window blocks with torch products
  differentiable: the 12-qubit, 40-layer, 96-parameter circuit over 64
            parameter sets (float32, 3-qubit blocks, checkpoints every 8
            applications): one warm-up step, the marker, then forward +
            adjoint steps
Both step workloads print the algorithmic bytes of one step by kernel
family (quantumsim_amd.ledger) before the marker.
"""
import argparse
import os
import sys

import numpy as np


MARKER_ELEMS = 4099


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('what', choices=['apply2q', 'apply1q', 'fused', 'batched', 'diagonal',
                                     'differentiable'])
    ap.add_argument('--qubits', type=int, default=14)
    ap.add_argument('--dtype', default='float64')
    ap.add_argument('--reps', type=int, default=3)
    ap.add_argument('--block-kernel', type=int, default=1, choices=[0, 1],
                    help='batched: window blocks by the block kernel (1) or torch products (0)')
    a = ap.parse_args(argv)
    import torch
    import quantumsim_amd as qa
    from quantumsim_amd.ledger import Ledger
    from quantumsim_amd import bases
    from quantumsim_amd.workloads import noisy_layers

    n = a.qubits
    N = 4 ** n
    torch.cuda.set_device(0)
    dt = torch.float64 if a.dtype == 'float64' else torch.float32
    src = torch.ones(N, dtype=dt, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    print('copy reference: %d bytes read, %d bytes written' % (N * src_bytes(dt), N * src_bytes(dt)),
          flush=True)
    if a.what == 'differentiable':
        from quantumsim_amd.workloads import TrainableCircuit
        from quantumsim_amd import _native as nat
        tc = TrainableCircuit(n=12, layers=40)
        ang, t1, t2 = tc.initial_params(64, seed=11, device='cuda', dtype=torch.float64)
        with torch.no_grad():
            target = tc.populations(ang + 0.1, t1 * 1.05, t2 * 0.97, torch.float32)
        params = [ang.requires_grad_(), t1.requires_grad_(), t2.requires_grad_()]

        def step():
            p = tc.populations(*params, torch.float32, checkpoint_every=8, fuse_qudits=3)
            return torch.autograd.grad(tc.kl(target, p).sum(), params)
        with Ledger(timed=False) as led:
            step()
        torch.cuda.synchronize()
        _print_ledger(led)
        m = torch.ones(MARKER_ELEMS, dtype=torch.float64, device='cuda')
        nat.check(nat.lib().qsv_scale(nat.QSV_F64, m.data_ptr(), m.numel(), 1.0,
                                      torch.cuda.current_stream().cuda_stream), 'scale')
        torch.cuda.synchronize()
        for _ in range(a.reps):
            step()
        torch.cuda.synchronize()
        print('done', flush=True)
        return 0
    if a.what == 'batched':
        from quantumsim_amd.batch import CircuitBatch
        from quantumsim_amd.workloads import random_circuit_batch
        # the benchmark's memory budget: kept states may take 60 % of the
        # free device memory (every state is kept)
        cb = CircuitBatch(random_circuit_batch(512, seed=2024), dtype=torch.float32,
                          checkpoint_budget=0.6 * torch.cuda.mem_get_info()[0],
                          block_kernel=bool(a.block_kernel))
        w = cb.random_weights(7)
        led = Ledger(timed=False)
        cb.capture(w, autotune=True, ledger=led)   # tuning + warm-up + capture
        torch.cuda.synchronize()
        _print_ledger(led)
        # marker: one k_scale launch (not used by the batch); the measured
        # replays follow it
        from quantumsim_amd import _native as nat
        m = torch.ones(MARKER_ELEMS, dtype=torch.float64, device='cuda')
        nat.check(nat.lib().qsv_scale(nat.QSV_F64, m.data_ptr(), m.numel(), 1.0,
                                      torch.cuda.current_stream().cuda_stream), 'scale')
        torch.cuda.synchronize()
        for _ in range(a.reps):
            cb.replay()
        torch.cuda.synchronize()
        print('done', flush=True)
        return 0
    v = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype=dt)
    v.to_torch().copy_(torch.randn((4,) * n, dtype=dt, device='cuda'))
    if a.what == 'diagonal':
        for _ in range(a.reps):
            v.diagonal(get_data=False)
        torch.cuda.synchronize()
        print('done: %d diagonal readouts' % a.reps, flush=True)
        return 0
    circuit, _ = noisy_layers(n, 2, seed=1, pairing='brickwork')
    if a.what == 'apply2q':
        gates = [(p, q) for p, q in circuit if len(q) == 2]
    elif a.what == 'apply1q':
        p1 = circuit[0][0]
        gates = [(p1, (q,)) for q in (0, n // 2, n - 1)]
    else:
        rng = np.random.default_rng(0)
        gates = [(rng.standard_normal([4] * 6) / 64, (5, 6, 7)),
                 (rng.standard_normal([4] * 6) / 64, (n - 3, n - 2, n - 1)),
                 (rng.standard_normal([4] * 8) / 256, (5, 6, 7, 8)),
                 (rng.standard_normal([4] * 8) / 256, (n - 4, n - 3, n - 2, n - 1))]
    for _ in range(a.reps):
        for p, q in gates:
            v.apply_ptm(p, *q)
    torch.cuda.synchronize()
    print('done: %d launches per rep' % len(gates), flush=True)
    return 0


def _print_ledger(led):
    """Algorithmic bytes of one step by kernel family (the step the measured
    launches repeat)."""
    for fam, row in sorted(led.rows.items()):
        print('ledger %-20s launches %6d  algorithmic bytes %d' % (fam, row[0], row[1]),
              flush=True)


def src_bytes(dt):
    import torch
    return 8 if dt == torch.float64 else 4


if __name__ == '__main__':
    sys.exit(main())
