"""The reference's operation tests restated on PauliVectorHIP.

* quantumsim test/test_operations_qubits.py:13-141 (rotations, Euler
  rotation, Hadamard, CNOT on PauliVector states, checked through
  meas_prob / to_pv), with the operation library's numpy PTMs
  (quantumsim_amd.operations.qubits) and the same PTMs as device operations
  (quantumsim_amd.operations.device.PTMOperation, PTMs never leaving the GPU);
* test/test_operations.py:153-195 (a gate sequence applied one by one equals
  its product: here the sequence composed on the host into one dense 3-qubit
  PTM, applied on the MFMA kernel, and composed on the device with
  PTMOperation.compose).
"""
import numpy as np
import pytest
import torch

import quantumsim_amd as qa
from quantumsim_amd import bases
from quantumsim_amd.operations import qubits as lib
from quantumsim_amd.operations.device import PTMOperation
from quantumsim_amd.fusion import fuse

pytestmark = pytest.mark.gpu

QB = (bases.general(2),)


def _ops(kind):
    """The library's operations as numpy PTMOps or as device PTMOperations."""
    if kind == 'host':
        return lambda op: op
    return lambda op: PTMOperation(torch.tensor(op.ptm, device='cuda'), op.bases_in,
                                   op.bases_out)


KINDS = ['host', 'device']


@pytest.mark.parametrize('kind', KINDS)
@pytest.mark.parametrize('rot', ['x', 'y'])
def test_rotate_xy(rot, kind):
    wrap = _ops(kind)
    make = lib.rotate_x if rot == 'x' else lib.rotate_y
    dm = qa.PauliVectorHIP(QB * 3)
    r90, r180, r360 = wrap(make(0.5 * np.pi)), wrap(make(np.pi)), wrap(make(2 * np.pi))
    r90(dm, 1)
    r180(dm, 2)
    assert np.allclose(dm.meas_prob(0), (1, 0))
    assert np.allclose(dm.meas_prob(1), (0.5, 0.5))
    assert np.allclose(dm.meas_prob(2), (0, 1))
    r180(dm, 1)
    assert np.allclose(dm.meas_prob(1), (0.5, 0.5))
    r90(dm, 1)
    assert np.allclose(dm.meas_prob(1), (1, 0))
    r360(dm, 0)
    assert np.allclose(dm.meas_prob(0), (1, 0))


@pytest.mark.parametrize('kind', KINDS)
def test_rotate_z(kind):
    wrap = _ops(kind)
    sqrt2 = np.sqrt(2)
    dm = qa.PauliVectorHIP(QB)
    r90, r180, r360 = (wrap(lib.rotate_z(0.5 * np.pi)), wrap(lib.rotate_z(np.pi)),
                       wrap(lib.rotate_z(2 * np.pi)))
    r90(dm, 0)
    assert np.allclose(dm.to_pv(), [1, 0, 0, 0])
    r180(dm, 0)
    assert np.allclose(dm.to_pv(), [1, 0, 0, 0])
    sup = qa.PauliVectorHIP(QB, np.array([0.5, 0.5, sqrt2, 0]))   # a hand-made Hadamard
    r180(sup, 0)
    assert np.allclose(sup.to_pv(), [0.5, 0.5, -sqrt2, 0])
    r90(sup, 0)
    assert np.allclose(sup.to_pv(), [0.5, 0.5, 0, -sqrt2])
    r180(sup, 0)
    assert np.allclose(sup.to_pv(), [0.5, 0.5, 0, sqrt2])
    r360(sup, 0)
    assert np.allclose(sup.to_pv(), [0.5, 0.5, 0, sqrt2])


@pytest.mark.parametrize('kind', KINDS)
def test_rotate_euler_and_hadamard(kind):
    wrap = _ops(kind)
    dm = qa.PauliVectorHIP(QB * 2)
    wrap(lib.rotate_euler(0, 0.5 * np.pi, 0))(dm, 0)
    assert np.allclose(dm.meas_prob(0), (0.5, 0.5))
    wrap(lib.rotate_euler(0, 0.5 * np.pi, 0))(dm, 1)
    assert np.allclose(dm.meas_prob(1), (0.5, 0.5))
    dm = qa.PauliVectorHIP(QB * 2)
    h = wrap(lib.hadamard())
    h(dm, 1)
    assert np.allclose(dm.meas_prob(0), (1, 0))
    assert np.allclose(dm.meas_prob(1), (0.5, 0.5))
    h(dm, 1)
    assert np.allclose(dm.meas_prob(1), (1, 0))


@pytest.mark.parametrize('kind', KINDS)
def test_cnot(kind):
    cnot = _ops(kind)(lib.cnot())
    s = qa.PauliVectorHIP.from_dm(np.diag([0.25, 0, 0.75, 0, 0, 0, 0, 0]), QB * 3)
    for q, p in ((0, (1, 0)), (1, (0.25, 0.75)), (2, (1, 0))):
        assert np.allclose(s.meas_prob(q), p)
    cnot(s, 0, 1)
    for q, p in ((0, (1, 0)), (1, (0.25, 0.75)), (2, (1, 0))):
        assert np.allclose(s.meas_prob(q), p)
    cnot(s, 1, 2)
    for q, p in ((0, (1, 0)), (1, (0.25, 0.75)), (2, (0.25, 0.75))):
        assert np.allclose(s.meas_prob(q), p)


def _sequence():
    return [(lib.rotate_x(np.pi / 2), (0,)), (lib.rotate_y(0.3333), (1,)),
            (lib.cphase(), (0, 2)), (lib.cphase(), (1, 2)), (lib.rotate_x(-np.pi / 2), (0,))]


def _random_dm(seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


@pytest.mark.parametrize('dtype,tol', [('float64', 1e-12), ('float32', 1e-6)])
def test_chain_equals_dense_3q_ptm(dtype, tol):
    """test_operations.py:153-195: applying the sequence one gate at a time
    equals applying its 3-qubit PTM (composed on the host, MFMA kernel) and
    its product composed on the device."""
    dm = _random_dm(93)
    seq = _sequence()
    s1 = qa.PauliVectorHIP.from_dm(dm, QB * 3, dtype=dtype)
    for op, qs in seq:
        op(s1, *qs)
    blocks = fuse([(op.ptm, qs) for op, qs in seq], [4, 4, 4], 3)
    assert len(blocks) == 1 and blocks[0].qudits == (0, 1, 2)
    s2 = qa.PauliVectorHIP.from_dm(dm, QB * 3, dtype=dtype)
    s2.apply_ptm(blocks[0].ptm, 0, 1, 2)
    assert np.max(np.abs(s1.to_pv() - s2.to_pv())) < tol
    # device composition: every gate embedded in the 3-qubit space, then chained
    from quantumsim_amd.fusion import embed_ptm
    total = None
    for op, qs in seq:
        e = PTMOperation(torch.tensor(embed_ptm(op.ptm, qs, (0, 1, 2), [4, 4, 4]).reshape(
            (4,) * 6), device='cuda'), QB * 3)
        total = e if total is None else total.compose(e)
    s3 = qa.PauliVectorHIP.from_dm(dm, QB * 3, dtype=dtype)
    total(s3, 0, 1, 2)
    assert np.max(np.abs(s1.to_pv() - s3.to_pv())) < tol
