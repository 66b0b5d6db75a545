"""Subbasis optimisation (quantumsim_amd.subbases, the behaviour of the
reference compiler's optimal_bases) on gate lists: the optimised circuit,
run in its subbases with rectangular PTMs, gives the same populations as
the plain circuit in full bases (oracle), and shrinks the vector where the
circuit allows it."""
import numpy as np
import pytest

from quantumsim_amd import bases
from quantumsim_amd.subbases import optimal_subbases
from quantumsim_amd.operations import qubits as gates
from quantumsim_amd.workloads import noisy_layers
from oracle import bases as ob
from oracle.pauli_vector import PauliVectorNumpy


def _ob(b):
    return ob.Basis(np.asarray(b.vectors), list(b.labels))


def _run_full(n, circuit):
    o = PauliVectorNumpy([ob.general(2)] * n, force=True)
    for p, q in circuit:
        o.apply_ptm(np.asarray(p), *q)
    return o.diagonal()


def _run_opt(init_bases, opt):
    o = PauliVectorNumpy([_ob(b) for b in init_bases], force=True)
    sizes = [o.to_pv().size]
    for p, q, outs in opt:
        o.apply_ptm(p, *q)
        for qq, b in zip(q, outs):
            o.bases[qq] = _ob(b)
        sizes.append(o.to_pv().size)
    return o.diagonal(), max(sizes)


FULL = bases.general(2)


def test_noisy_circuit_same_populations():
    n = 4
    circuit, _ = noisy_layers(n, 6, seed=4, pairing='brickwork')
    init, opt, fin = optimal_subbases([FULL] * n, circuit, initial=[[0]] * n,
                                      final=[[0, 1]] * n)
    assert [b.dim_pauli for b in init] == [1] * n            # |0><0| only at the start
    assert [b.dim_pauli for b in fin] == [2] * n             # populations only at the end
    d_ref = _run_full(n, circuit)
    d_opt, _ = _run_opt(init, opt)
    assert np.allclose(d_opt, d_ref, atol=1e-13)


def test_classical_circuit_stays_small():
    """Bit flips, CNOTs and dephasing keep every qubit in the classical
    subbasis: the vector stays 2^n instead of 4^n."""
    n = 5
    X = gates.rotate_x(np.pi).ptm
    cnot = gates.cnot().ptm
    deph = gates.phase_damping(0.3).ptm
    circuit = [(X, (0,)), (cnot, (0, 1)), (X, (2,)), (cnot, (2, 3)), (deph, (3,)),
               (cnot, (1, 4)), (X, (0,)), (cnot, (4, 2))]
    init, opt, fin = optimal_subbases([FULL] * n, circuit, initial=[[0]] * n,
                                      final=[[0, 1]] * n)
    d_ref = _run_full(n, circuit)
    d_opt, biggest = _run_opt(init, opt)
    assert np.allclose(d_opt, d_ref, atol=1e-13)
    assert biggest <= 2 ** n
    assert all(b.dim_pauli <= 2 for b in fin)


def test_hadamard_needs_the_full_basis_only_where_it_must():
    n = 3
    H = gates.hadamard().ptm
    circuit = [(H, (0,)), (gates.cphase(np.pi).ptm, (0, 1)), (H, (0,)),
               (gates.amp_damping(0.1).ptm, (2,))]
    init, opt, fin = optimal_subbases([FULL] * n, circuit, initial=[[0]] * n,
                                      final=[[0, 1]] * n)
    d_ref = _run_full(n, circuit)
    d_opt, biggest = _run_opt(init, opt)
    assert np.allclose(d_opt, d_ref, atol=1e-13)
    assert biggest < 4 ** n
    assert init[2].dim_pauli == 1


def test_rejects_non_full_ptms():
    with pytest.raises(ValueError):
        optimal_subbases([FULL] * 2, [(np.eye(2), (0,))])
