"""Density matrix <-> Pauli vector and expectation values on the device
(quantumsim_amd.algebra.device) against the oracle's conversions
(oracle/algebra.py, pinned in test_oracle.py) and Hilbert-space brute force
tr(O rho), for qubits, qutrits and mixed dimensions (which the reference's
uniform-dimension conversion does not cover)."""
import numpy as np
import pytest
import torch

import quantumsim_amd as qa
from quantumsim_amd import bases
from quantumsim_amd.algebra import device as dalg
from oracle import algebra as oa, bases as ob, brute

pytestmark = pytest.mark.gpu


def _oracle_bases(bs):
    out = []
    for b in bs:
        if b.dim_pauli == b.dim_hilbert ** 2 and np.allclose(b.vectors, ob.general(b.dim_hilbert).vectors):
            out.append(ob.general(b.dim_hilbert))
        else:
            out.append(ob.Basis(np.asarray(b.vectors), list(b.labels)))
    return out


BASIS_SETS = {
    'qubits3': lambda: [bases.general(2)] * 3,
    'gellmann2': lambda: [bases.gell_mann(2)] * 2,
    'qutrits2': lambda: [bases.general(3), bases.gell_mann(3)],
    'mixed': lambda: [bases.general(2), bases.general(3), bases.gell_mann(2)],
}


@pytest.mark.parametrize('name', list(BASIS_SETS))
def test_dm_pv_round_trip(name):
    bs = BASIS_SETS[name]()
    dims = [b.dim_hilbert for b in bs]
    rng = np.random.default_rng(3)
    rho = brute.random_density_matrix(int(np.prod(dims)), rng)
    pv = dalg.dm_to_pv(rho, bs).cpu().numpy()
    # the oracle's conversion (uniform dimension only), or brute force
    if len(set(dims)) == 1:
        ref = oa.dm_to_pv(rho, _oracle_bases(bs))
        assert np.allclose(pv, ref, atol=1e-13)
    # brute force: pv_x = Re tr(rho P_x) with P_x the Kronecker product
    for idx in list(np.ndindex(*pv.shape))[:40]:
        P = np.array([[1.0]])
        for q, x in enumerate(idx):
            P = np.kron(P, np.asarray(bs[q].vectors)[x])
        assert abs(pv[idx] - np.trace(rho @ P).real) < 1e-13
    back = dalg.pv_to_dm(pv, bs).cpu().numpy()
    assert np.allclose(back, rho, atol=1e-13)


@pytest.mark.parametrize('name', list(BASIS_SETS))
def test_from_dm_to_dm_methods(name):
    bs = BASIS_SETS[name]()
    D = int(np.prod([b.dim_hilbert for b in bs]))
    rho = brute.random_density_matrix(D, np.random.default_rng(4))
    v = qa.PauliVectorHIP.from_dm(rho, bs)
    assert np.allclose(v.to_dm(), rho, atol=1e-13)
    assert abs(v.trace() - 1.0) < 1e-13
    v32 = qa.PauliVectorHIP.from_dm(rho, bs, dtype='float32')
    assert np.allclose(v32.to_dm(), rho, atol=1e-6)


def test_exp_value_pauli_strings_and_dense():
    rng = np.random.default_rng(5)
    bs = [bases.general(2)] * 3
    rho = brute.random_density_matrix(8, rng)
    v = qa.PauliVectorHIP.from_dm(rho, bs)
    S = dalg.SIGMA
    for op in ['XYZ', 'IZI', 'zzx', 'III']:
        O = np.array([[1.0]])
        for ch in op:
            O = np.kron(O, S[ch.upper()])
        ref = np.trace(O @ rho)
        got = v.exp_value(op)
        assert abs(got - ref) < 1e-12, op
    A = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    assert abs(v.exp_value(A) - np.trace(A @ rho)) < 1e-12
    assert abs(v.exp_value(A.reshape([2] * 6)) - np.trace(A @ rho)) < 1e-12
    with pytest.raises(ValueError):
        v.exp_value('XY')
    with pytest.raises(ValueError):
        v.exp_value('XQZ')
