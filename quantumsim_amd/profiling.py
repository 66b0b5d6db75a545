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
  backward: the backward of one dense 64 x 64 block over --batch float32
            states (--qubits, --targets): five warm-up rounds of both, the
            marker, then twice --reps one-pass backwards
            (qsv_apply_ptm_batched_backward) alternating with --reps of the
            two kernels it replaces (qsv_ptm_grad and the transposed
            qsv_apply_ptm_batched)
"""
import argparse
import os
import sys

import numpy as np


MARKER_ELEMS = 4099


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('what', choices=['apply2q', 'apply1q', 'fused', 'batched', 'diagonal',
                                     'differentiable', 'backward'])
    ap.add_argument('--targets', default='5,6,7', help='backward: the block\'s three qubits')
    ap.add_argument('--batch', type=int, default=64, help='backward: states in the batch')
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
    if a.what == 'backward':
        from quantumsim_amd import _native as nat
        from quantumsim_amd.autograd import _apply_raw, _fused_backward_raw, _ptm_grad_raw
        tg = tuple(int(t) for t in a.targets.split(','))
        B, shape = a.batch, (4,) * n
        g = torch.Generator(device='cuda').manual_seed(7)
        x = torch.randn(B, N, device='cuda', generator=g)
        gy = torch.randn(B, N, device='cuda', generator=g)
        P = torch.randn(B, 64, 64, device='cuda', generator=g) / 8
        Pt = P.transpose(1, 2).contiguous()
        gx = torch.empty_like(x)
        gp = torch.empty(B, 64, 64, device='cuda')

        def one_pass():
            if _fused_backward_raw(x, gy, P, shape, tg, gx_out=gx, gp_out=gp) is None:
                raise SystemExit('the one-pass backward does not take this geometry')

        def two_kernels():
            _ptm_grad_raw(x, gy, shape, tg, [4, 4, 4], out=gp)
            _apply_raw(gy, Pt, shape, tg, [4, 4, 4], out=gx)
        # warm-up of both (the device's clocks rise over the first
        # milliseconds of heavy work: the variant timed first would lose)
        for _ in range(5):
            one_pass()
            two_kernels()
        torch.cuda.synchronize()
        print('backward: %d qubits, targets %s, %d float32 states; one pass reads x and '
              'ybar and writes xbar (%d bytes), the two kernels move 4 N s per state '
              '(%d bytes)' % (n, tg, B, 3 * B * N * 4, 4 * B * N * 4), flush=True)
        m = torch.ones(MARKER_ELEMS, dtype=torch.float64, device='cuda')
        nat.check(nat.lib().qsv_scale(nat.QSV_F64, m.data_ptr(), m.numel(), 1.0,
                                      torch.cuda.current_stream().cuda_stream), 'scale')
        torch.cuda.synchronize()
        ms = []
        for fn in (one_pass, two_kernels, one_pass, two_kernels):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(a.reps):
                fn()
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_time(e1) / a.reps)
        print('event-timed per call (kernels and launch gaps), alternating: one pass %.3f / '
              '%.3f ms, two kernels %.3f / %.3f ms' % (ms[0], ms[2], ms[1], ms[3]), flush=True)
        print('done', flush=True)
        return 0
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
