"""The npz circuit file: save/load round trip (bases table, rectangular
PTMs with basis changes, initial state), rejection of foreign files, and a
replay through the oracle that equals applying the gates directly."""
import numpy as np
import pytest

from quantumsim_amd import bases, circuitio
from quantumsim_amd.workloads import noisy_layers
from oracle import bases as ob
from oracle.pauli_vector import PauliVectorNumpy


def _oracle_run(bs, gates, pv=None):
    ob_map = lambda b: ob.Basis(np.asarray(b.vectors), list(b.labels))
    o = PauliVectorNumpy([ob_map(b) for b in bs], None if pv is None else pv.copy(), force=True)
    for g in gates:
        o.apply_ptm(np.asarray(g[0]), *g[1])
        if len(g) > 2:
            for q, b in zip(g[1], g[2]):
                o.bases[q] = ob_map(b)
    return o


def test_round_trip_noisy_circuit(tmp_path):
    circuit, _ = noisy_layers(4, 5, seed=2, pairing='brickwork')
    bs = [bases.general(2)] * 4
    path = str(tmp_path / 'c.npz')
    circuitio.save(path, bs, circuit)
    bs2, gates2, pv2 = circuitio.load(path)
    assert pv2 is None and len(gates2) == len(circuit)
    assert all(np.array_equal(b.vectors, bases.general(2).vectors) for b in bs2)
    for (p, q), (p2, q2, outs) in zip(circuit, gates2):
        assert np.array_equal(np.asarray(p), p2) and tuple(q) == q2
    ref = _oracle_run(bs, circuit).to_pv()
    got = _oracle_run(bs2, gates2).to_pv()
    assert np.array_equal(ref, got)


def test_basis_changes_and_initial_state(tmp_path):
    rng = np.random.default_rng(0)
    full, cls = bases.general(2), bases.general(2).subbasis([0, 1])
    sel = np.zeros((2, 4))
    sel[0, 0] = sel[1, 1] = 1                   # projection onto the classical subbasis
    pv = rng.standard_normal((4, 4, 9))
    bs = [full, full, bases.gell_mann(3)]
    gates = [(rng.standard_normal((4, 4, 4, 4)), (0, 1)),
             (sel, (1,), [cls]),
             (rng.standard_normal((9, 9)), (2,))]
    path = str(tmp_path / 'c.npz')
    circuitio.save(path, bs, gates, initial_pv=pv)
    bs2, gates2, pv2 = circuitio.load(path)
    assert np.array_equal(pv2, pv)
    assert gates2[1][2][0].dim_pauli == 2 and gates2[2][2][0].dim_pauli == 9
    a = _oracle_run(bs, gates, pv)
    b = _oracle_run(bs2, gates2, pv2)
    assert np.array_equal(a.to_pv(), b.to_pv())
    assert [x.dim_pauli for x in b.bases] == [4, 2, 9]


def test_rejects_foreign_files(tmp_path):
    path = str(tmp_path / 'x.npz')
    np.savez(path, format=np.array('something-else'))
    with pytest.raises(ValueError):
        circuitio.load(path)
    with pytest.raises(ValueError):             # qudit 1 does not exist
        circuitio.save(str(tmp_path / 'y.npz'), [bases.general(2)], [(np.eye(4), (0, 1))])
    with pytest.raises(ValueError):             # 2-dimensional PTM on 2 qudits
        circuitio.save(str(tmp_path / 'y.npz'), [bases.general(2)] * 2, [(np.eye(4), (0, 1))])
