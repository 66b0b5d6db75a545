"""End-to-end parity of the benchmark circuits: the seeded noisy circuits
run gate by gate through PauliVectorHIP and through the float64 oracle.

* 6 qubits, 40 layers, random CZ pairing: diagonal (64 probabilities) and
  trace within 1e-12 in float64; diagonal within 1e-5 absolute in float32.
* the 14-qubit, 60-layer brickwork circuit restricted to qubits 0-9 (the
  10-qubit slice), checked the same way.
* the same circuits fused into dense 2-, 3- and 4-qubit PTMs (MFMA kernels
  for 3 and 4) against the gate-by-gate oracle.
"""
import numpy as np
import pytest

import quantumsim_amd as qa
from quantumsim_amd import bases
from quantumsim_amd.workloads import noisy_layers
from quantumsim_amd.fusion import apply_circuit
from oracle.pauli_vector import PauliVectorNumpy
from oracle import bases as ob

pytestmark = pytest.mark.gpu


def _run(circuit, n, dtype):
    v = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype=dtype)
    for ptm, qs in circuit:
        v.apply_ptm(ptm, *qs)
    return v


def _run_oracle(circuit, n):
    o = PauliVectorNumpy([ob.general(2)] * n, force=True)
    for ptm, qs in circuit:
        o.apply_ptm(ptm, *qs)
    return o


@pytest.mark.parametrize('n,layers,pairing,seed', [(6, 40, 'random', 0),
                                                   (10, 60, 'brickwork', 1)])
def test_circuit_parity(n, layers, pairing, seed):
    circuit, _ = noisy_layers(n, layers, seed, pairing)
    o = _run_oracle(circuit, n)
    d_ref, t_ref = o.diagonal(), o.trace()
    v = _run(circuit, n, 'float64')
    d64 = v.diagonal()
    err64 = np.max(np.abs(d64 - d_ref))
    print('n=%d fp64 max |diag err| = %.3e' % (n, err64))
    assert err64 < 1e-12
    assert abs(v.trace() - t_ref) < 1e-12
    assert np.max(np.abs(v.to_pv() - o.to_pv())) < 1e-12
    v32 = _run(circuit, n, 'float32')
    err32 = np.max(np.abs(v32.diagonal() - d_ref))
    print('n=%d fp32 max |diag err| = %.3e' % (n, err32))
    assert err32 < 1e-5
    assert abs(v32.trace() - t_ref) < 1e-5
    assert abs(v32.renormalize() - t_ref) < 1e-5


@pytest.mark.parametrize('max_qudits', [2, 3, 4])
@pytest.mark.parametrize('n,layers,pairing,seed', [(6, 40, 'random', 0),
                                                   (10, 60, 'brickwork', 1)])
def test_fused_circuit_parity(n, layers, pairing, seed, max_qudits):
    circuit, _ = noisy_layers(n, layers, seed, pairing)
    o = _run_oracle(circuit, n)
    d_ref = o.diagonal()
    v = qa.PauliVectorHIP([bases.general(2)] * n, force=True)
    napp = apply_circuit(v, circuit, max_qudits)
    assert napp < len(circuit)
    err64 = np.max(np.abs(v.to_pv() - o.to_pv()))
    print('n=%d fused<=%d: %d applications for %d gates, fp64 max |pv err| = %.3e'
          % (n, max_qudits, napp, len(circuit), err64))
    assert err64 < 1e-12
    assert abs(v.trace() - o.trace()) < 1e-12
    v32 = qa.PauliVectorHIP([bases.general(2)] * n, force=True, dtype='float32')
    apply_circuit(v32, circuit, max_qudits)
    err32 = np.max(np.abs(v32.diagonal() - d_ref))
    print('n=%d fused<=%d fp32 max |diag err| = %.3e' % (n, max_qudits, err32))
    assert err32 < 1e-5
