"""Circuit tools on the GPU, against the oracle: a subbasis-optimised
circuit (rectangular PTMs, bases that shrink and grow) run on
PauliVectorHIP, and its replay from the npz circuit file, give the plain
circuit's populations as the float64 oracle computes them in full bases;
the reference compiler's ZZ-parity known answer (leaky qutrit CPhases,
quantumsim/test/test_compiler.py:268-314) runs optimised on the GPU with
the oracle's populations."""
import numpy as np
import pytest

import quantumsim_amd as qa
from quantumsim_amd import bases, circuitio
from quantumsim_amd.subbases import optimal_subbases
from quantumsim_amd.operations import qubits as gates
from quantumsim_amd.workloads import noisy_layers
from oracle.pauli_vector import PauliVectorNumpy
from oracle import bases as ob

pytestmark = pytest.mark.gpu
FULL = bases.general(2)


def _ob(b):
    return ob.Basis(np.asarray(b.vectors), list(b.labels))


def _oracle_full(n, circuit):
    o = PauliVectorNumpy([ob.general(2)] * n, force=True)
    for p, q in circuit:
        o.apply_ptm(np.asarray(p), *q)
    return o.diagonal()


def _run_opt(init, opt, dtype='float64'):
    v = qa.PauliVectorHIP(init, force=True, dtype=dtype)
    for p, q, outs in opt:
        v.apply_ptm(p, *q)
        for qq, b in zip(q, outs):
            v.bases[qq] = b
    return v


@pytest.mark.parametrize('dtype,tol', [('float64', 1e-12), ('float32', 1e-5)])
def test_optimised_circuits_against_oracle(dtype, tol, tmp_path):
    n = 8
    circuit, _ = noisy_layers(n, 8, seed=6, pairing='brickwork')
    X, cnot = gates.rotate_x(np.pi).ptm, gates.cnot().ptm
    classical = [(X, (0,)), (cnot, (0, 1)), (X, (3,)), (cnot, (3, 7)), (cnot, (1, 5)),
                 (gates.phase_damping(0.2).ptm, (5,)), (cnot, (7, 2))]
    for circ in (circuit, classical):
        init, opt, fin = optimal_subbases([FULL] * n, circ, initial=[[0]] * n,
                                          final=[[0, 1]] * n)
        ref = _oracle_full(n, circ)
        v = _run_opt(init, opt, dtype)
        assert [b.dim_pauli for b in v.bases] == [b.dim_pauli for b in fin]
        assert np.max(np.abs(v.diagonal() - ref)) < tol
        path = str(tmp_path / 'c.npz')
        circuitio.save(path, init, opt)
        w = circuitio.run(path, dtype=dtype)
        assert np.max(np.abs(w.diagonal() - ref)) < tol


def test_zz_parity_optimised_on_gpu():
    """Three qutrits (general(3), 9 elements each) through the optimised
    ZZ-parity circuit: the vector shrinks to 2 x 4 x 1 elements at the
    start and ends in the reference's subbases; populations equal the
    oracle's plain full-basis run to 1e-12."""
    from test_subbases import zz_parity_circuit, _random_hermitian
    from oracle import algebra as oa
    full, gl, initial, final = zz_parity_circuit()
    init, opt, fin = optimal_subbases([full] * 3, gl, initial=initial, final=final)
    dm = _random_hermitian(27, 85)
    pv = oa.dm_to_pv(dm, [_ob(full)] * 3)
    mask = np.zeros(pv.shape, dtype=bool)
    mask[np.ix_(*initial)] = True
    o = PauliVectorNumpy([_ob(full)] * 3, np.where(mask, pv, 0.0))
    for p, q in gl:
        o.apply_ptm(np.asarray(p), *q)
    v = qa.PauliVectorHIP(init, np.ascontiguousarray(pv[np.ix_(*initial)]))
    assert v.data_shape == (2, 2, 1)
    for p, q, outs in opt:
        v.apply_ptm(p, *q)
        for qq, b in zip(q, outs):
            v.bases[qq] = b
    assert v.bases[1] == full.subbasis([0, 1, 2, 6])
    assert np.max(np.abs(v.diagonal() - o.diagonal())) < 1e-12
