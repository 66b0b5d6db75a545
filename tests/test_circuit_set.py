"""Host-side planning of CircuitSet (no GPU): grouping by qudit shape and
gate-structure prefix, run_schedule segments, argument checks."""
import numpy as np
import pytest
import torch

from quantumsim_amd.circuit_set import plan_batches, segments_of, CircuitSet
from quantumsim_amd.bases import general


def test_plan_batches_groups_by_shape_and_prefix():
    shapes = [(9, 9), (4, 4, 4), (9, 9), (9, 9), (4, 4, 4), (9, 4)]
    a, b = ((0,), (0, 1), (1,)), ((1,), (0,))
    structures = [a[:2], ((0, 1, 2),), a, b, (), a[:1]]
    got = plan_batches(shapes, structures)
    # shapes in (rank, dims) order; members by decreasing depth
    assert got == [((9, 4), a[:1], [5]),
                   ((9, 9), a, [2, 0]),
                   ((9, 9), b, [3]),
                   ((4, 4, 4), ((0, 1, 2),), [1, 4])]


def test_segments_of():
    assert segments_of([5, 5, 3, 0]) == [(4, 3, []), (3, 2, [(p, 0) for p in range(3)]),
                                         (2, 0, [(3, 0), (4, 0)])]
    assert segments_of([2]) == [(1, 0, [(0, 0), (1, 0)])]
    assert segments_of([0, 0]) == [(2, 0, [])]
    # every position once, the running set shrinking to zero
    d = [7, 7, 6, 3, 3, 1]
    segs = segments_of(d)
    assert [g for s in segs for g in s[2]] == [(p, 0) for p in range(7)]
    assert segs[0][0] == len(d) and segs[-1][1] == 0
    for (a0, c0, _), (a1, _, _) in zip(segs, segs[1:]):
        assert c0 == a1


def test_circuit_set_argument_checks():
    q3 = general(3)
    good = torch.eye(9, dtype=torch.float64)
    with pytest.raises(ValueError, match='bad targets'):
        CircuitSet([([q3, q3], [(good, (2,))])], device='cpu')
    with pytest.raises(ValueError, match='bad targets'):
        CircuitSet([([q3, q3], [(torch.eye(81), (0, 0))])], device='cpu')
    with pytest.raises(ValueError, match='PTM of shape'):
        CircuitSet([([q3, general(2)], [(torch.eye(81), (0, 1))])], device='cpu')
    with pytest.raises(ValueError, match='dtype'):
        CircuitSet([], dtype=torch.float16, device='cpu')
    # flat [D, D] or per-qudit (9, 4, 9, 4) forms
    cs = CircuitSet([([q3, general(2)], [(torch.eye(36).reshape(9, 4, 9, 4), (0, 1)),
                                         (torch.eye(9), (0,))])],
                    device='cpu')
    assert len(cs) == 1 and cs.batches == [((9, 4), ((0, 1), (0,)), [0])]


def test_circuit_set_size_limit():
    q3 = general(3)
    with pytest.raises(ValueError, match='1024 entries a side'):
        CircuitSet([([q3] * 4, [(torch.eye(6561), (0, 1, 2, 3))])], device='cpu')
