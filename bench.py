#!/usr/bin/env python3
"""Benchmark of the Pauli-vector hot path on one MI355X.

Workload: the 14-qubit (4^14 = 268,435,456-entry) density matrix in the full
qubit Pauli basis driven by the seeded 60-layer noisy brickwork circuit.
One step = the 13 two-qubit PTM applications of one even+odd brickwork layer
pair, bonds (0,1), (2,3), ..., (12,13), then (1,2), ..., (11,12): every
target position from outermost to innermost.  The state stays resident in
HBM; PTMs travel in the kernel arguments.

Printed JSON line (stdout):
  value      = 2-qubit PTM applications per second (float64 state)
  roofline   = the 2-qubit apply kernel against 8 TB/s HBM; algorithmic bytes
               per launch = (N_in + N_out) * 8 + 16*16*8 (PTM)
  cpu_baseline = the float64 numpy oracle (same einsum structure as the
               reference backend) applying the same 2-qubit PTMs to a
               14-qubit state on this host's cores, bounded sample.
Extra fields: float32 rate, per-position kernel GB/s, 1-qubit rate, the
rate of whole circuit layers, the diagonal readout time, the whole 60-layer
circuit gate by gate and fused into dense 2-/3-qubit PTMs, the batched
forward + adjoint of 512 circuits and the differentiable 12-qubit step.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# dense float32 MFMA peak of MI355X (TFLOP/s): the bound of the 64 x 64
# contractions next to HBM (16-21 flops per byte against a ridge of 19.7)
MFMA_F32_PEAK_TFLOPS = 157.3
# share of the free device memory the batched step's kept states may take
BATCH_BUDGET_FRACTION = 0.6


def _cpu_threads():
    """Host threads of the numpy baseline: this process's CPU affinity,
    capped by OMP_NUM_THREADS when the environment already sets it."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get('OMP_NUM_THREADS', '')
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return n


# BLAS threads are fixed when numpy first loads OpenBLAS: set them before
# anything imports numpy (torch does)
CPU_THREADS = _cpu_threads()
os.environ.setdefault('OPENBLAS_NUM_THREADS', str(CPU_THREADS))
os.environ.setdefault('OMP_NUM_THREADS', str(CPU_THREADS))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--cpu-runs', type=int, default=5,
                    help='timed runs of each CPU baseline sample (the median is reported)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--qubits', type=int, default=14)
    ap.add_argument('--no-extras', action='store_true',
                    help='skip the batched and differentiable measurements')
    ap.add_argument('--circuits', type=int, default=512)
    ap.add_argument('--param-sets', type=int, default=64)
    ap.add_argument('--fuse-qudits', type=int, default=3,
                    help='dense block size of the batched and differentiable paths '
                         '(3: 64 x 64 MFMA blocks; 0: gate by gate)')
    return ap.parse_args()


def timed(fn, reps, stream):
    """Average seconds per call of fn() over reps calls, HIP events on the
    stream the kernels are launched on."""
    import torch
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / 1e3 / reps


def load_traffic():
    """HBM bytes per launch of the 2-qubit apply kernel from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled)."""
    path = os.path.join(ROOT, 'profiles', 'apply_ptm_2q_14q_f64_pmc.txt')
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith('traffic_bytes_per_launch:'):
                return float(line.split(':', 1)[1])
    return None


def _median_seconds(fn, runs):
    """Median wall seconds of ``runs`` calls of fn() (after one warm-up call
    made by the caller)."""
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return statistics.median(times), times


def cpu_baseline(n, step_gates, circuit, runs, gpu_slice_pops):
    """The float64 numpy oracle (the reference backend's einsum structure,
    quantumsim/pauli_vectors/numpy.py:50-80) on this host, OPENBLAS / OMP
    threads set to the cores used.  Each sample: one warm-up, then the
    median of ``runs`` timed runs.
      value: the benchmark step (13 two-qubit PTMs on the 14-qubit state)
             -> 2-qubit applications per second;
      cfg0:  the 6-qubit, 40-layer circuit (random pairing) in full;
      slice10: the 60-layer 14-qubit circuit restricted to qubits 0-9 (its
             gates on those qubits), also checked against the GPU's float64
             populations of the same slice."""
    import numpy as np
    from oracle.pauli_vector import PauliVectorNumpy
    from oracle import bases as ob
    from quantumsim_amd.workloads import noisy_layers
    QB = [ob.general(2)]
    # 14 qubits: the benchmark step
    o = PauliVectorNumpy(QB * n, force=True)
    o._data = np.random.default_rng(0).standard_normal([4] * n)
    p0, q0 = step_gates[0]
    o.apply_ptm(p0, *q0)                                         # warm-up

    def step():
        for ptm, qs in step_gates:
            o.apply_ptm(ptm, *qs)
    med, times = _median_seconds(step, runs)
    del o
    # the 6-qubit, 40-layer circuit
    c0, _ = noisy_layers(6, 40, 0, 'random')

    def run_c0():
        v = PauliVectorNumpy(QB * 6)
        for ptm, qs in c0:
            v.apply_ptm(ptm, *qs)
        return v.diagonal(), v.trace()
    run_c0()
    med0, _ = _median_seconds(run_c0, runs)
    # the 10-qubit slice
    sl = [(p, q) for p, q in circuit if max(q) < 10]
    res_sl = {}

    def run_slice():
        v = PauliVectorNumpy(QB * 10)
        for ptm, qs in sl:
            v.apply_ptm(ptm, *qs)
        res_sl['pops'] = v.diagonal()
    PauliVectorNumpy(QB * 10).apply_ptm(sl[0][0], *sl[0][1])   # warm-up
    med10, _ = _median_seconds(run_slice, runs)
    err = float(np.max(np.abs(res_sl['pops'] - gpu_slice_pops)))
    threads = {k: os.environ.get(k, 'unset') for k in ('OPENBLAS_NUM_THREADS', 'OMP_NUM_THREADS')}
    return {
        'value': len(step_gates) / med, 'unit': '2q-PTM apps/s', 'cores': CPU_THREADS,
        'kind': 'port',
        'sample': 'the benchmark step (%d two-qubit PTM applications, bonds %s) on the '
                  '%d-qubit float64 state: one warm-up application, then the median of %d '
                  'timed steps; numpy einsum(optimize=True) as in the reference backend; '
                  'threads %s' % (len(step_gates), [q for _, q in step_gates], n, runs,
                                  threads),
        'step_seconds': [round(x, 3) for x in times],
        'threads': threads,
        'cfg0_6q_40_layers': {'gates': len(c0), 'median_s': round(med0, 4),
                              'gates_per_s': round(len(c0) / med0, 1)},
        'slice_10q_60_layers': {'gates': len(sl), 'median_s': round(med10, 3),
                                'gates_per_s': round(len(sl) / med10, 1),
                                'gpu_f64_populations_max_abs_err': float('%.3e' % err)},
    }


def bench_batched(args):
    """512 circuits (4..11 qubits, 20..80 layers, float32 states, float64
    readout): one HIP-graph replay of the batched forward + adjoint against
    one-circuit-at-a-time eager runs of a stratified subsample."""
    import numpy as np
    import torch
    from quantumsim_amd import persistent
    from quantumsim_amd.batch import CircuitBatch
    from quantumsim_amd.ledger import Ledger
    from quantumsim_amd.workloads import random_circuit_batch
    from quantumsim_amd.autotune import Autotuner, set_tuner
    circuits = random_circuit_batch(args.circuits, seed=2024)
    tuner = Autotuner(path=None)                 # tuned in this run, not read from disk
    set_tuner(tuner)
    t0 = time.perf_counter()
    torch.cuda.reset_peak_memory_stats()
    free0 = torch.cuda.mem_get_info()[0]
    # checkpoint interval chosen from an explicit memory budget: this
    # process owns the device, so the kept states may take 60 % of the free
    # memory (every state is kept when they fit)
    budget = BATCH_BUDGET_FRACTION * free0
    cb = CircuitBatch(circuits, dtype=torch.float32, fuse_qudits=args.fuse_qudits,
                      checkpoint_budget=budget)
    w = cb.random_weights(7)
    # the warm-up step before capture (tuned shapes) is counted: every
    # library launch of one step, each timed alone (device idle before and
    # after it: the two streams would otherwise overlap inside the events)
    led = Ledger(timed=True, isolate=True)
    cb.capture(w, autotune=True, ledger=led)
    per_kernel = led.summary(HBM_PEAK_GBS, MFMA_F32_PEAK_TFLOPS)
    alg_bytes = led.total_bytes()
    t_setup = time.perf_counter() - t0
    peak_gb = torch.cuda.max_memory_allocated() / 1e9
    stream = cb.stream
    reps = 3
    # CUDAGraph.replay() launches on the current stream: make it the batch's
    # stream, the one the events are recorded on (otherwise the events time
    # only the host-side launch of the replays)
    with torch.cuda.stream(stream):
        sec = timed(lambda: cb.graph.replay(), reps, stream)
    log('batched: %d circuits fwd+adjoint %.3f s per step (setup+capture %.1f s)'
        % (len(circuits), sec, t_setup))
    # plain baseline: each circuit alone, eager launches, same kernels.
    # Every sampled circuit's engine is built and warmed up outside the
    # timed region; only its step is timed (the earlier form, which built a
    # fresh engine inside the timed call, is kept as *_incl_setup).  For
    # each circuit the faster of gate by gate and dense blocks (and, for
    # circuits that fit one, of the persistent kernel) is the baseline.
    by_n = {}
    for i, c in enumerate(circuits):
        by_n.setdefault(c['n'], []).append(i)
    sample = [by_n[n][j] for n in sorted(by_n) for j in range(min(2, len(by_n[n])))]
    lds_n = persistent.max_qubits(torch.float32)
    weights = np.random.default_rng(7)
    per, per_setup, choice = [], [], {}
    for i in sample:
        c = circuits[i]
        wt = torch.tensor(weights.standard_normal((1, 2 ** c['n'])), dtype=torch.float64,
                          device='cuda')
        best = None
        for fq in sorted({0, args.fuse_qudits}):
            for pk in ((False, True) if c['n'] <= lds_n else (False,)):
                b = CircuitBatch([c], dtype=torch.float32, persistent_kernel=pk,
                                 fuse_qudits=fq)
                b.step([wt])                                    # warm-up
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                b.step([wt])
                torch.cuda.synchronize()
                t = time.perf_counter() - t1
                del b
                if best is None or t < best[0]:
                    best = (t, fq, pk)
        per.append(best[0])
        choice[i] = best[1:]
        # construction inside the timed call, as first measured
        fq, pk = best[1:]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        CircuitBatch([c], dtype=torch.float32, persistent_kernel=pk, fuse_qudits=fq).step([wt])
        torch.cuda.synchronize()
        per_setup.append(time.perf_counter() - t1)

    def weighted(ts):
        # per-width averages weighted by the widths' share of the batch
        mean_by_n = {}
        for i, t in zip(sample, ts):
            mean_by_n.setdefault(circuits[i]['n'], []).append(t)
        return sum(len(by_n[n]) * statistics.mean(v) for n, v in mean_by_n.items())
    plain_sec, plain_sec_setup = weighted(per), weighted(per_setup)
    log('batched: plain loop %.3f s (%.3f s with engine construction) for %d circuits'
        % (plain_sec, plain_sec_setup, len(circuits)))
    # probability error of the float32 step against the same batch run with
    # float64 states (the oracle comparison of these circuits is in the
    # GPU tests)
    cb.replay()
    torch.cuda.synchronize()
    trace_dev = max((t - 1).abs().max().item() for t in cb.g_traces)
    cb64 = CircuitBatch(circuits, dtype=torch.float64, checkpoint_every=16,
                        fuse_qudits=args.fuse_qudits, requires_grad=False)
    with torch.no_grad():
        p64 = cb64.populations()
    prob_err = max((a - b).abs().max().item() for a, b in zip(cb.g_pops, p64))
    del cb64, p64
    res = {'circuits': len(circuits), 'fwd_adjoint_s': round(sec, 4),
           'circuits_per_s': round(len(circuits) / sec, 1),
           'plain_circuits_per_s': round(len(circuits) / plain_sec, 2),
           'speedup_vs_plain': round(plain_sec / sec, 2),
           'plain_circuits_per_s_incl_setup': round(len(circuits) / plain_sec_setup, 2),
           'speedup_vs_plain_incl_setup': round(plain_sec_setup / sec, 2),
           'plain_sample': len(sample),
           'plain_timing': 'per sampled circuit: engine built and warmed up outside the '
                           'timed region, one eager step timed; the fastest of gate by gate '
                           '/ dense blocks / persistent kernel per circuit; per-width means '
                           'weighted by the batch',
           'plain_choice_counts': {'%s' % (k,): sum(1 for v in choice.values() if v == k)
                                   for k in set(choice.values())},
           'roofline': {'bound': 'hbm', 'alg_bytes_per_step': alg_bytes,
                        'achieved': round(alg_bytes / sec / 1e9, 1), 'peak': HBM_PEAK_GBS,
                        'unit': 'GB/s', 'frac': round(alg_bytes / sec / 1e9 / HBM_PEAK_GBS, 4),
                        'per_kernel': per_kernel,
                        'note': 'alg bytes: every library launch of one step ((N_in + N_out) '
                                's per apply, 2 N s + dP per PTM gradient, 3 N s + dP per '
                                'one-pass backward, trajectories of the persistent kernels); '
                                'achieved = bytes / graph-replay step time; per_kernel ms: '
                                'events around each launch of one eager step, the device idle '
                                'before and after it; the persistent kernels keep their states '
                                'in LDS, so their HBM fraction is not their bound; mfma_frac: '
                                'the dense families\' MFMA flops / time against the float32 '
                                'MFMA peak'},
           'dtype': 'f32 state, f64 readout',
           'checkpoint_every': cb.k, 'kept_state_gb': round(cb.kept_bytes(cb.k) / 1e9, 1),
           'checkpoint_budget_gb': round(budget / 1e9, 1),
           'peak_alloc_gb': round(peak_gb, 1), 'free_gb_at_start': round(free0 / 1e9, 1),
           'setup_s': round(t_setup, 1),
           'max_prob_err_vs_f64': float('%.3e' % prob_err), 'fuse_qudits': args.fuse_qudits,
           'readout': 'float64 populations / float64 trace per circuit',
           'max_trace_dev_f32': float('%.3e' % trace_dev),
           'autotune': tuner.summary()}
    log('batched: %s' % res)
    del cb
    torch.cuda.empty_cache()
    return res


def bench_circuit(args, circuit, n):
    """The whole seeded 60-layer circuit on the 14-qubit state: gate by gate
    against gates fused into dense PTMs of up to 2 and 3 qubits (the 3-qubit
    blocks run on the MFMA kernel).  Fusion is host work done once per
    circuit, outside the timed region (timed separately)."""
    import torch
    import quantumsim_amd as qa
    from quantumsim_amd import bases
    from quantumsim_amd.fusion import fuse
    stream = torch.cuda.current_stream()
    plans = {1: [(p, tuple(q)) for p, q in circuit]}
    fuse_s = {}
    for k in (2, 3):
        t0 = time.perf_counter()
        plans[k] = [(f.ptm, f.qudits) for f in fuse(circuit, [4] * n, k)]
        fuse_s[k] = time.perf_counter() - t0
    res = {'gates': len(circuit), 'layers': 60}
    for dtype in ('float64', 'float32'):
        v = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype=dtype)
        row = {}
        for k, gates in plans.items():
            def run(gates=gates):
                for p, q in gates:
                    v.apply_ptm(p, *q)
            run()
            sec = timed(run, 1, stream)
            row['fuse%d' % k if k > 1 else 'gate_by_gate'] = {
                'applications': len(gates), 'ms': round(sec * 1e3, 2),
                'gates_per_s': round(len(circuit) / sec, 1)}
        res[dtype[:1] + dtype[-2:]] = row
        del v
        torch.cuda.empty_cache()
    res['fuse_host_s'] = {'fuse2': round(fuse_s[2], 3), 'fuse3': round(fuse_s[3], 3)}
    res['speedup_fuse3_vs_gate_by_gate'] = {
        d: round(res[d]['gate_by_gate']['ms'] / res[d]['fuse3']['ms'], 2) for d in ('f64', 'f32')}
    log('circuit: %s' % res)
    return res


def bench_differentiable(args):
    """12 qubits, 40 layers, 96 parameters x 64 parameter sets, float32
    states: KL loss forward + adjoint (checkpoint every 8 gates), timed
    eagerly and as replays of the step captured in a HIP graph."""
    import torch
    from quantumsim_amd.workloads import TrainableCircuit
    from quantumsim_amd.autograd import capture_step
    from quantumsim_amd.ledger import Ledger
    tc = TrainableCircuit(n=12, layers=40)
    sets = args.param_sets
    free0 = torch.cuda.mem_get_info()[0]
    ang, t1, t2 = tc.initial_params(sets, seed=11, device='cuda', dtype=torch.float64)
    with torch.no_grad():
        target = tc.populations(ang + 0.1, t1 * 1.05, t2 * 0.97, torch.float32)
    params = [ang.requires_grad_(), t1.requires_grad_(), t2.requires_grad_()]

    def step():
        p = tc.populations(*params, torch.float32, checkpoint_every=8,
                           fuse_qudits=args.fuse_qudits)
        loss = tc.kl(target, p).sum()
        return torch.autograd.grad(loss, params)

    step()
    torch.cuda.synchronize()
    sec_eager = timed(step, 2, torch.cuda.current_stream())
    # the same step captured once in a HIP graph (on a stream of its own,
    # default priority) and replayed: no host launch gaps
    graph, stream, g_out = capture_step(step)
    with torch.cuda.stream(stream):
        sec = timed(lambda: graph.replay(), 3, stream)
    eager_out = step()
    torch.cuda.synchronize()
    graph_vs_eager = max(((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item()
                         for a, b in zip(g_out, eager_out))
    del graph, g_out, eager_out
    torch.cuda.empty_cache()
    with Ledger(timed=True, isolate=True) as led:
        step()
    per_kernel = led.summary(HBM_PEAK_GBS, MFMA_F32_PEAK_TFLOPS)
    alg_bytes = led.total_bytes()
    torch.cuda.empty_cache()
    with torch.no_grad():
        first = [p.detach()[:1] for p in params]
        n_gates = len(tc.gates(*first))
        n_apps = len(tc.gates(*first, fuse_qudits=args.fuse_qudits))
    res = {'qubits': 12, 'layers': 40, 'params_per_set': tc.n_params, 'param_sets': sets,
           'fwd_adjoint_s': round(sec, 4), 'param_sets_per_s': round(sets / sec, 2),
           'timing': 'HIP-graph replays', 'fwd_adjoint_s_eager': round(sec_eager, 4),
           'graph_vs_eager_max_rel_diff': float('%.3e' % graph_vs_eager),
           'gates': n_gates, 'applications': n_apps, 'fuse_qudits': args.fuse_qudits,
           'checkpoint_every': 8, 'free_gb_at_start': round(free0 / 1e9, 1),
           'roofline': {'bound': 'hbm', 'alg_bytes_per_step': alg_bytes,
                        'achieved': round(alg_bytes / sec / 1e9, 1), 'peak': HBM_PEAK_GBS,
                        'unit': 'GB/s', 'frac': round(alg_bytes / sec / 1e9 / HBM_PEAK_GBS, 4),
                        'per_kernel': per_kernel}}
    log('differentiable: %s' % res)
    return res


def main():
    args = parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: there is no multi-GPU path '
              '(--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import quantumsim_amd as qa
    from quantumsim_amd import bases
    from quantumsim_amd.workloads import noisy_layers, layer_slices

    n = args.qubits
    N = 4 ** n
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    circuit, meta = noisy_layers(n, 60, seed=1, pairing='brickwork')
    twoq = [(p, q) for p, q in circuit[:len(circuit)] if len(q) == 2]
    step_gates = [g for g in twoq[:(n // 2) + (n - 1) // 2]]
    assert len(step_gates) == n - 1
    pairs = [q for _, q in step_gates]

    res = {}
    for dtype, es in (('float64', 8), ('float32', 4)):
        v = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype=dtype)
        v.to_torch().copy_(torch.randn((4,) * n, dtype=v.dtype, device=v.device))

        def step():
            for p, q in step_gates:
                v.apply_ptm(p, *q)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sec_step = timed(step, args.steps, stream)
        wall = time.perf_counter() - t0
        apps = len(step_gates) / sec_step
        per_launch = sec_step / len(step_gates)
        alg_bytes = 2 * N * es + 256 * es
        res[dtype] = dict(apps=apps, sec_step=sec_step, per_launch=per_launch,
                          gbs=alg_bytes / per_launch / 1e9, wall=wall, alg_bytes=alg_bytes)
        # per-position kernel rates
        pos = {}
        for label, qs in (('outer(0,1)', (0, 1)), ('middle(6,7)', (n // 2 - 1, n // 2)),
                          ('inner(%d,%d)' % (n - 2, n - 1), (n - 2, n - 1))):
            p = step_gates[0][0]
            t = timed(lambda: v.apply_ptm(p, *qs), 10, stream)
            pos[label] = round(alg_bytes / t / 1e9, 1)
        res[dtype]['positions_GBs'] = pos
        # one-qubit applications
        oneq = [(p, q) for p, q in circuit[:3 * n] if len(q) == 1][:n]
        t1 = timed(lambda: [v.apply_ptm(p, *q) for p, q in oneq], 5, stream) / len(oneq)
        res[dtype]['oneq_apps'] = 1.0 / t1
        res[dtype]['oneq_GBs'] = (2 * N * es + 16 * es) / t1 / 1e9
        # whole circuit layers (1q Haar + CZ + damping per layer)
        sl = layer_slices(meta, len(circuit))
        two_layers = circuit[sl[0].start:sl[1].stop]
        tl = timed(lambda: [v.apply_ptm(p, *q) for p, q in two_layers], 3, stream) / 2
        res[dtype]['layer_ms'] = tl * 1e3
        res[dtype]['layer_gates'] = len(two_layers) / 2
        # diagonal readout (gather of 2^n entries, float64 out): per call,
        # host work included; the first call builds the cached gather plan
        v.diagonal(get_data=False)
        td = timed(lambda: v.diagonal(get_data=False), 10, stream)
        res[dtype]['diag_us'] = td * 1e6
        del v
        torch.cuda.empty_cache()
        log('%s: %.1f 2q apps/s, %.1f GB/s per launch, positions %s, 1q %.1f apps/s, '
            'layer %.2f ms, diag %.1f us' % (dtype, apps, res[dtype]['gbs'], pos,
                                              res[dtype]['oneq_apps'], tl * 1e3, td * 1e6))

    # device-to-device copy ceiling: same bytes moved as one apply launch
    src = torch.empty(N, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    src.fill_(1.0)
    tc = timed(lambda: dst.copy_(src), 10, stream)
    copy_gbs = 2 * N * 8 / tc / 1e9
    del src, dst
    torch.cuda.empty_cache()
    log('copy ceiling %.1f GB/s' % copy_gbs)

    r64 = res['float64']
    out = {
        'metric': '2q-PTM apps/s + HBM GB/s (14q DM); batched fwd+grad circuits/s vs plain',
        'value': round(r64['apps'], 2),
        'unit': '2q-PTM applications/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(r64['sec_step'] * 1e3, 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic (seeded noisy brickwork circuit, random state)',
        'config': {'workload': '14-qubit noisy circuit, plain method: 2-qubit PTM applications '
                               'on the 4^14-entry Pauli vector, all 13 brickwork bonds per step',
                   'qubits': n, 'entries': N, 'gates_per_step': len(step_gates)},
        'roofline': {'bound': 'hbm', 'achieved': round(r64['gbs'], 1), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(r64['gbs'] / HBM_PEAK_GBS, 4),
                     'traffic': load_traffic(),
                     'kernel': '2-qubit apply: k_apply_small (outer/middle bonds), k_apply_lds (inner bonds)',
                     'alg_bytes_per_launch': r64['alg_bytes'],
                     'avg_launch_ms': round(r64['per_launch'] * 1e3, 4)},
        'f32': {'apps_per_s': round(res['float32']['apps'], 2),
                'GBs': round(res['float32']['gbs'], 1),
                'frac': round(res['float32']['gbs'] / HBM_PEAK_GBS, 4)},
        'positions_GBs': {'f64': r64['positions_GBs'],
                          'f32': res['float32']['positions_GBs']},
        'oneq': {'f64_apps_per_s': round(r64['oneq_apps'], 1),
                 'f64_GBs': round(r64['oneq_GBs'], 1),
                 'f32_apps_per_s': round(res['float32']['oneq_apps'], 1)},
        'layer': {'gates_per_layer': r64['layer_gates'],
                  'f64_ms': round(r64['layer_ms'], 3),
                  'f32_ms': round(res['float32']['layer_ms'], 3)},
        'copy_ceiling_GBs': round(copy_gbs, 1),
        'diagonal_us': {'f64': round(r64['diag_us'], 1),
                        'f32': round(res['float32']['diag_us'], 1)},
    }
    if not args.no_extras:
        out['circuit'] = bench_circuit(args, circuit, n)
        out['batched'] = bench_batched(args)
        out['differentiable'] = bench_differentiable(args)
    if not args.no_cpu:
        # the GPU's float64 populations of the 10-qubit slice, for the
        # numpy check that comes with the CPU timing
        v10 = qa.PauliVectorHIP([bases.general(2)] * 10)
        for p, q in circuit:
            if max(q) < 10:
                v10.apply_ptm(p, *q)
        pops10 = v10.diagonal()
        del v10
        log('cpu baseline (numpy oracle, median of %d runs, %d threads)...'
            % (args.cpu_runs, CPU_THREADS))
        out['cpu_baseline'] = cpu_baseline(n, step_gates, circuit, args.cpu_runs, pops10)
        log('cpu baseline: %s' % out['cpu_baseline'])
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
