"""Persistent whole-circuit kernels (qsv_circuits_forward / _adjoint),
replayed on the CPU by qsv_emulate_circuits:

* populations of random circuits of different widths (1-2-qubit gates on
  random targets, random PTMs) equal the oracle applying the same PTMs gate
  by gate and reading its diagonal;
* dL/dP for L = sum_c <w_c, p_c> equals central finite differences of the
  oracle's L (exact up to rounding: L is linear in each PTM);
* the table builder rejects bad targets, PTM offsets and shared slots.
"""
import numpy as np
import pytest

from oracle import bases as ob
from oracle.pauli_vector import PauliVectorNumpy
from quantumsim_amd.persistent import CircuitTable, emulate


def _random_program(rng, widths, depth):
    circuits, pool, off = [], [], 0
    for n in widths:
        gates = []
        for _ in range(depth):
            k = 1 if n == 1 else int(rng.integers(1, 3))
            targets = tuple(int(q) for q in rng.choice(n, size=k, replace=False))
            D = 4 ** k
            pool.append(rng.standard_normal(D * D) / D ** 0.5)
            gates.append((targets, off))
            off += D * D
        circuits.append((n, gates))
    return circuits, np.concatenate(pool)


def _oracle_pops(circuits, pool):
    out = []
    for n, gates in circuits:
        o = PauliVectorNumpy([ob.general(2)] * n, force=True)
        for targets, off in gates:
            k = len(targets)
            D = 4 ** k
            o.apply_ptm(pool[off:off + D * D].reshape([4] * (2 * k)), *targets)
        out.append(o.diagonal())
    return np.concatenate(out)


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_forward_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    circuits, pool = _random_program(rng, [1, 2, 3, 4, 3], 7)
    table = CircuitTable(circuits, pool.size)
    pops, _ = emulate(table, pool)
    ref = _oracle_pops(circuits, pool)
    assert np.allclose(pops, ref, rtol=1e-12, atol=1e-12)


def test_adjoint_matches_finite_differences():
    rng = np.random.default_rng(7)
    circuits, pool = _random_program(rng, [2, 3], 5)
    table = CircuitTable(circuits, pool.size)
    w = rng.standard_normal(table.pop_elems)
    _, dpool = emulate(table, pool, w)
    h = 1e-4
    for idx in rng.choice(pool.size, 25, replace=False):
        p1, p2 = pool.copy(), pool.copy()
        p1[idx] += h
        p2[idx] -= h
        fd = (w @ _oracle_pops(circuits, p1) - w @ _oracle_pops(circuits, p2)) / (2 * h)
        assert abs(dpool[idx] - fd) < 1e-9 * max(1.0, abs(fd)), idx


def test_table_rejects_bad_input():
    with pytest.raises(ValueError):
        CircuitTable([(3, [((3,), 0)])], 16)            # target out of range
    with pytest.raises(ValueError):
        CircuitTable([(3, [((1, 1), 0)])], 256)         # repeated target
    with pytest.raises(ValueError):
        CircuitTable([(3, [((0, 1), 10)])], 256)        # PTM beyond the pool
    with pytest.raises(ValueError):                      # two gates sharing PTM entries
        CircuitTable([(2, [((0,), 0), ((1,), 8)])], 32)
    with pytest.raises(ValueError):                      # ... across circuits too
        CircuitTable([(2, [((0,), 0)]), (2, [((0, 1), 0)])], 256)
    CircuitTable([(2, [((0,), 0), ((1,), 16)])], 32)     # adjacent slots are fine


def test_wide_circuits_and_swizzle():
    """6- and 7-qubit circuits (LDS swizzle over the full 4^7 range, every
    target position) against the oracle, plus the adjoint at 5 qubits."""
    rng = np.random.default_rng(11)
    circuits, pool = _random_program(rng, [6, 7], 6)
    table = CircuitTable(circuits, pool.size)
    pops, _ = emulate(table, pool)
    assert np.allclose(pops, _oracle_pops(circuits, pool), rtol=1e-12, atol=1e-12)
    circuits, pool = _random_program(rng, [5], 4)
    table = CircuitTable(circuits, pool.size)
    w = rng.standard_normal(table.pop_elems)
    _, dpool = emulate(table, pool, w)
    h = 1e-4
    for idx in rng.choice(pool.size, 8, replace=False):
        p1, p2 = pool.copy(), pool.copy()
        p1[idx] += h
        p2[idx] -= h
        fd = (w @ _oracle_pops(circuits, p1) - w @ _oracle_pops(circuits, p2)) / (2 * h)
        assert abs(dpool[idx] - fd) < 1e-9 * max(1.0, abs(fd))
