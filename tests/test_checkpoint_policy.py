"""Host logic of the batched engine's checkpoint policy (no GPU needed):
the runs of layers with a constant set of running circuits, the bytes the
forward keeps for the adjoint, and the interval chosen from a budget."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from quantumsim_amd.batch import CircuitBatch, choose_checkpoint_interval
from quantumsim_amd.workloads import brickwork_window_plan, layer_targets


def _batch(depths_by_n, window=2, fuse_qudits=3, dtype=torch.float32):
    """A CircuitBatch shell with only the fields the host planning reads."""
    cb = CircuitBatch.__new__(CircuitBatch)
    cb.window, cb.fuse_qudits, cb.dtype, cb.small = window, fuse_qudits, dtype, []
    cb.groups = [SimpleNamespace(n=n, layers=sorted(d, reverse=True), members=list(range(len(d))))
                 for n, d in sorted(depths_by_n.items())]
    return cb


def _per_circuit_apps(cb, n, depth):
    """Applications one circuit of depth `depth` goes through, counted from
    the layer structure alone."""
    if cb.window == 2:
        return (depth + 1) // 2 * len(brickwork_window_plan(n))
    return sum(len(layer_targets(n, l % 2)) for l in range(depth))


@pytest.mark.parametrize('window,fq', [(2, 3), (1, 0)])
@pytest.mark.parametrize('seed', [0, 1, 2])
def test_segments_cover_every_circuit_once(window, fq, seed):
    rng = np.random.default_rng(seed)
    depths = {n: list(rng.integers(1, 30, size=int(rng.integers(1, 12)))) for n in (8, 9, 10)}
    cb = _batch(depths, window, fq)
    for g in cb.groups:
        segs = cb._segments(g)
        # contiguous, starting at layer 0, ending past the deepest circuit
        assert segs[0][0] == 0
        for a, b in zip(segs, segs[1:]):
            assert a[1] == b[0] and a[3] == b[2]
        assert segs[-1][3] == 0 and segs[-1][1] >= g.layers[0]
        for first, end, active, nxt in segs:
            assert end > first and (end - first) % window == 0
            # the running prefix: circuits deeper than the first layer
            assert active == sum(1 for d in g.layers if d > first)
            # nobody finishes inside the run (only at its last window)
            assert all(d >= end - window + 1 for d in g.layers[:active])
            assert nxt == sum(1 for d in g.layers if d > end)
        # every circuit runs exactly its own windows
        windows = sum(active * (end - first) // window for first, end, active, _ in segs)
        assert windows == sum((d + window - 1) // window for d in g.layers)


@pytest.mark.parametrize('window,fq', [(2, 3), (1, 0)])
def test_kept_bytes_every_state(window, fq):
    """k = 1 keeps every state each circuit passes through; on top come at
    most one stored output per segment, the final states and the adjoint's
    three buffers."""
    depths = {8: [5, 9, 3, 14], 10: [7, 2, 11]}
    cb = _batch(depths, window, fq)
    es = 4
    passed = extra = 0
    for g in cb.groups:
        S = 4 ** g.n * es
        passed += sum(_per_circuit_apps(cb, g.n, d) for d in g.layers) * S
        extra += (len(cb._segments(g)) + 1) * len(g.layers) * S
    extra += 3 * len(depths[10]) * 4 ** 10 * es
    assert passed <= cb.kept_bytes(1) <= passed + extra


def test_kept_bytes_decreases_with_interval():
    cb = _batch({9: [40, 33, 21, 8], 11: [30, 29, 12]})
    kept = [cb.kept_bytes(k) for k in (1, 2, 4, 8, 16)]
    assert kept[0] > kept[1] > kept[2] > kept[3]
    # past the run length a longer interval keeps fewer checkpoints but
    # recomputes a longer segment at once: here 16 states of the 3-circuit
    # 11-qubit run outweigh the checkpoints saved
    assert kept[4] > kept[3]
    # float64 states double every figure
    cb64 = _batch({9: [40, 33, 21, 8], 11: [30, 29, 12]}, dtype=torch.float64)
    assert cb64.kept_bytes(4) == 2 * cb.kept_bytes(4)


def test_choose_checkpoint_interval():
    cb = _batch({10: [40, 35, 30, 25, 20]})
    k1, k4 = cb.kept_bytes(1), cb.kept_bytes(4)
    assert choose_checkpoint_interval(cb.kept_bytes, k1) == 1
    assert choose_checkpoint_interval(cb.kept_bytes, k1 - 1) == 2
    assert choose_checkpoint_interval(cb.kept_bytes, k4) == 4
    assert choose_checkpoint_interval(cb.kept_bytes, 0) == 16
    assert choose_checkpoint_interval(lambda k: 100 // k, 30, (1, 3, 5)) == 5


def test_run_schedule_rejects_inconsistent_arguments():
    """Argument checks run on the host before any launch."""
    from quantumsim_amd.autograd import run_schedule
    shape = (4, 4, 4)
    x0 = torch.zeros((3, 64))
    steps = [torch.zeros((2, 3, 16, 16)), torch.zeros((2, 3, 4, 4))]
    targets = [(0, 1), (2,)]
    ok = [(3, 1, [(0, 0), (1, 0)]), (1, 0, [(0, 1), (1, 1)])]
    bad = {
        'prefix does not start with every state': [(2, 0, [(0, 0)])],
        'segments do not chain': [(3, 1, [(0, 0)]), (2, 0, [(0, 1)])],
        'states left running': [(3, 1, [(0, 0)])],
        'continuing above running': [(3, 4, [(0, 0)])],
        'PTM reused': [(3, 0, [(0, 0), (0, 0)])],
        'step out of range': [(3, 0, [(0, 2)])],
        'position out of range': [(3, 0, [(2, 0)])],
    }
    for what, segs in bad.items():
        with pytest.raises(ValueError):
            run_schedule(x0, steps, targets, segs, shape)
    with pytest.raises(ValueError):          # wrong element type
        run_schedule(x0, [s.double() for s in steps], targets, ok, shape)
    with pytest.raises(ValueError):          # wrong PTM size for the targets
        run_schedule(x0, steps, [(0,), (2,)], ok, shape)
    with pytest.raises(ValueError):          # state size does not match the shape
        run_schedule(torch.zeros((3, 65)), steps, targets, ok, shape)
    with pytest.raises(ValueError):          # batch size differs
        run_schedule(torch.zeros((2, 64)), steps, targets, ok, shape)


def test_stream_balance():
    """Batched groups go to the stream with less state traffic so far,
    largest first; the largest stays on the main stream."""
    cb = _batch({8: [40] * 60, 9: [40] * 60, 10: [40] * 60, 11: [40] * 60})
    side = [cb.groups[g].n for g in cb._balance_streams()]
    assert side == [8, 9, 10]                  # 11 qubits: 4x the 10-qubit traffic
    cb = _batch({10: [40] * 10, 11: [40] * 10})
    assert [cb.groups[g].n for g in cb._balance_streams()] == [10]
    cb = _batch({11: [40] * 10})
    assert cb._balance_streams() == []           # one group: nothing beside it
    cb = _batch({9: [10] * 4, 10: [10] * 4})
    cb.small = [0]                               # persistent groups are not balanced
    assert [cb.groups[g].n for g in cb._balance_streams()] == []
