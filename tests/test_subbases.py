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


# ------------------------------------------------------------------ the
# reference compiler's known answers (quantumsim/test/test_compiler.py),
# restated on gate lists: which subbases the optimised circuit ends up in,
# and that optimised and plain circuits give the same populations.
def _kraus(kraus, full):
    from quantumsim_amd.algebra import kraus_to_ptm
    k = int(round(np.log(np.asarray(kraus).shape[-1]) / np.log(full.dim_hilbert)))
    return kraus_to_ptm(np.asarray(kraus, dtype=complex), (full,) * k, (full,) * k)


def _rx(angle, d):
    """RX on levels 0, 1 of a d-level system, level 2 untouched
    (quantumsim/operations/qubits.py:38-53, qutrits.py:41-56)."""
    s, c = np.sin(angle / 2), np.cos(angle / 2)
    u = np.eye(d, dtype=complex)
    u[:2, :2] = [[c, -1j * s], [-1j * s, c]]
    return u


def _cphase_qutrit(angle, leakage):
    """The 'legacy' qutrit CPhase (quantumsim/operations/qutrits.py:201-222):
    exp(-i angle (1 - arcsin(sqrt(leakage)) / pi) G), G coupling |11> and
    |02> (indices 4 and 2 of the 9 two-qutrit levels); the second qutrit
    leaks."""
    from scipy.linalg import expm
    g = np.zeros((9, 9))
    g[2, 4] = g[4, 2] = 1
    return expm(-1j * angle * (1 - np.arcsin(np.sqrt(leakage)) / np.pi) * g)


def _random_hermitian(dim, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
    return (a + a.conj().T) / 2


def _oracle_full_run(full_bases, init_supports, gates, dm):
    """The plain circuit on the oracle: the state given by ``dm`` restricted
    to the initial supports (the reference's from_dm into subbases), padded
    to the full bases, every PTM applied in the full bases."""
    from oracle import algebra as oa
    obs = [_ob(b) for b in full_bases]
    pv = oa.dm_to_pv(dm, obs)
    mask = np.zeros(pv.shape, dtype=bool)
    mask[np.ix_(*init_supports)] = True
    o = PauliVectorNumpy(obs, np.where(mask, pv, 0.0), force=True)
    for p, q in gates:
        o.apply_ptm(np.asarray(p), *q)
    return o.diagonal()


def _oracle_opt_run(full_bases, init_supports, init, opt, dm):
    from oracle import algebra as oa
    pv = oa.dm_to_pv(dm, [_ob(b) for b in full_bases])[np.ix_(*init_supports)]
    o = PauliVectorNumpy([_ob(b) for b in init], np.ascontiguousarray(pv), force=True)
    for p, q, outs in opt:
        o.apply_ptm(p, *q)
        for qq, b in zip(q, outs):
            o.bases[qq] = _ob(b)
    return o.diagonal()


def test_compile_cnot_shapes():
    """test_compiler.py:108-120: CNOT in the full bases stays 4x4x4x4, on
    two classical qubits (subbasis [0, 1]) it is 2x2x2x2, with the control
    in |0><0| (subbasis [0]) and a full target it is 1x4x1x4."""
    cnot = gates.cnot().ptm
    for initial, shape in (([None, None], (4, 4, 4, 4)), ([[0, 1], [0, 1]], (2, 2, 2, 2)),
                           ([[0], None], (1, 4, 1, 4))):
        _, opt, _ = optimal_subbases([FULL] * 2, [(cnot, (0, 1))], initial=initial)
        assert opt[0][0].shape == shape, (initial, opt[0][0].shape)


@pytest.mark.parametrize('d', [2, 3])
def test_compile_single_qubit_chain(d):
    """test_compiler.py:122-164: two RX(pi/5) fused into one PTM equal RX(2
    pi/5) and act identically on a Pauli vector; RX(pi) twice keeps the
    classical subbasis [0, 1] whether it is given at the input or asked for
    at the output."""
    from quantumsim_amd.fusion import fuse
    full = bases.general(d)
    obs = [_ob(full)]
    angle = np.pi / 5
    rx1 = _kraus(_rx(angle, d), full)
    rx2 = _kraus(_rx(2 * angle, d), full)
    fused = fuse([(rx1, (0,)), (rx1, (0,))], [d * d], 2)
    assert len(fused) == 1
    np.testing.assert_allclose(fused[0].ptm, rx2, atol=1e-14)
    from oracle import algebra as oa
    pv = oa.dm_to_pv(_random_hermitian(d, 487), obs)
    o0 = PauliVectorNumpy(obs, pv.copy())
    o0.apply_ptm(rx1, 0)
    o0.apply_ptm(rx1, 0)
    o1 = PauliVectorNumpy(obs, pv.copy())
    o1.apply_ptm(fused[0].ptm, 0)
    np.testing.assert_allclose(o1.to_pv(), o0.to_pv(), atol=1e-14)
    sub = full.subbasis([0, 1])
    rxpi = _kraus(_rx(np.pi, d), full)
    chain = [(rxpi, (0,)), (rxpi, (0,))]
    init, opt, fin = optimal_subbases([full], chain, initial=[[0, 1]])
    assert init[0] == sub and fin[0] == sub
    assert all(outs[0] == sub for _, _, outs in opt)
    init, opt, fin = optimal_subbases([full], chain, final=[[0, 1]])
    assert init[0] == sub and fin[0] == sub


def zz_parity_circuit():
    """The ZZ-parity check of test_compiler.py:268-314 on three qutrits:
    RX(-pi/2) on the ancilla 2, leaky CPhases (0, 2) and (2, 1), RX(pi/2) on
    the ancilla, RX(pi) on both data qutrits; the data qutrits start
    classical (subbasis [0, 1]), the ancilla in |0><0|; the readout keeps
    the data qutrits whole and the ancilla's populations (subbasis
    [0, 1, 2])."""
    full = bases.general(3)
    gl = [(_kraus(_rx(-np.pi / 2, 3), full), (2,)),
          (_kraus(_cphase_qutrit(np.pi, 0.1), full), (0, 2)),
          (_kraus(_cphase_qutrit(np.pi, 0.25), full), (2, 1)),
          (_kraus(_rx(np.pi / 2, 3), full), (2,)),
          (_kraus(_rx(np.pi, 3), full), (0,)),
          (_kraus(_rx(np.pi, 3), full), (1,))]
    initial = [[0, 1], [0, 1], [0]]
    final = [None, None, [0, 1, 2]]
    return full, gl, initial, final


def test_compile_zz_parity_leakage_subbases():
    """The leaking qutrit 1 ends in subbasis [0, 1, 2, 6] (|0><0|, |1><1|,
    |2><2| and Y20), qutrit 0, which does not leak, in [0, 1], the measured
    ancilla in [0, 1, 2] (test_compiler.py:293-303); the optimised circuit's
    populations equal the plain circuit's (test_compiler.py:305-314)."""
    full, gl, initial, final = zz_parity_circuit()
    init, opt, fin = optimal_subbases([full] * 3, gl, initial=initial, final=final)
    assert [b for b in init] == [full.subbasis([0, 1]), full.subbasis([0, 1]),
                                 full.subbasis([0])]
    assert fin[0] == full.subbasis([0, 1])
    assert fin[1] == full.subbasis([0, 1, 2, 6])
    assert list(fin[1].labels) == [full.labels[i] for i in (0, 1, 2, 6)]
    assert fin[2] == full.subbasis([0, 1, 2])
    dm = _random_hermitian(27, 85)
    d_ref = _oracle_full_run([full] * 3, initial, gl, dm)
    d_opt = _oracle_opt_run([full] * 3, initial, init, opt, dm)
    assert np.allclose(d_opt, d_ref, atol=1e-12)
