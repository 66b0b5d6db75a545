"""Circuit tools on the GPU: a subbasis-optimised circuit (rectangular
PTMs, bases that shrink and grow) run on PauliVectorHIP gives the full
circuit's populations, and so does its replay from the npz circuit file."""
import numpy as np
import pytest

import quantumsim_amd as qa
from quantumsim_amd import bases, circuitio
from quantumsim_amd.subbases import optimal_subbases
from quantumsim_amd.operations import qubits as gates
from quantumsim_amd.workloads import noisy_layers

pytestmark = pytest.mark.gpu
FULL = bases.general(2)


def _run_full(n, circuit, dtype='float64'):
    v = qa.PauliVectorHIP([FULL] * n, force=True, dtype=dtype)
    for p, q in circuit:
        v.apply_ptm(np.asarray(p), *q)
    return v.diagonal()


def _run_opt(init, opt, dtype='float64'):
    v = qa.PauliVectorHIP(init, force=True, dtype=dtype)
    for p, q, outs in opt:
        v.apply_ptm(p, *q)
        for qq, b in zip(q, outs):
            v.bases[qq] = b
    return v.diagonal()


@pytest.mark.parametrize('dtype,tol', [('float64', 1e-12), ('float32', 1e-5)])
def test_optimised_circuits(dtype, tol, tmp_path):
    n = 8
    circuit, _ = noisy_layers(n, 8, seed=6, pairing='brickwork')
    X, cnot = gates.rotate_x(np.pi).ptm, gates.cnot().ptm
    classical = [(X, (0,)), (cnot, (0, 1)), (X, (3,)), (cnot, (3, 7)), (cnot, (1, 5)),
                 (gates.phase_damping(0.2).ptm, (5,)), (cnot, (7, 2))]
    for circ in (circuit, classical):
        init, opt, fin = optimal_subbases([FULL] * n, circ, initial=[[0]] * n,
                                          final=[[0, 1]] * n)
        ref = _run_full(n, circ, dtype)
        got = _run_opt(init, opt, dtype)
        assert np.max(np.abs(got - ref)) < tol
        path = str(tmp_path / 'c.npz')
        circuitio.save(path, init, opt)
        v = circuitio.run(path, dtype=dtype)
        assert np.max(np.abs(v.diagonal() - ref)) < tol
