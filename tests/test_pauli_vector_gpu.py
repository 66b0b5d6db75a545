"""GPU parity of PauliVectorHIP against the oracle and the reference's own
known answers.

The first class restates quantumsim's test/test_pauli_vectors.py (the
backend-parametrised drop-in tests) against the new class; the rest compare
every operation with the float64 oracle on seeded inputs (tolerance 1e-12
relative to the largest magnitude for float64, bit-exact for gathers).
"""
import itertools
import math
import warnings

import numpy as np
import pytest
from scipy.stats import unitary_group

import quantumsim_amd as qa
from quantumsim_amd import bases
from quantumsim_amd.algebra import kraus_to_ptm, ptm_convert_basis
from oracle.pauli_vector import PauliVectorNumpy, multitake
from oracle import bases as ob

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.shape != b.shape:
        return np.inf
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def random_density_matrix(dim, seed):
    # test_pauli_vectors.py:51-57
    rng = np.random.RandomState(seed)
    diag = rng.rand(dim)
    diag /= np.sum(diag)
    unitary = random_unitary_matrix(dim, seed + 1)
    return unitary @ np.diag(diag) @ unitary.conj().T


def random_unitary_matrix(dim, seed):
    return unitary_group.rvs(dim, random_state=np.random.RandomState(seed))


def _oracle_bases(bs):
    """The oracle's own basis objects for a list of product bases."""
    out = []
    for b in bs:
        out.append(ob.Basis(np.array(b.vectors), list(b.labels)))
    return out


# ------------------------------------------------- reference drop-in tests
class TestReferenceSuite:
    """quantumsim test/test_pauli_vectors.py:65-307 with PauliVectorHIP."""

    def test_create_trivial(self):
        dm = qa.PauliVectorHIP([])
        assert dm.to_pv() == pytest.approx(1)
        assert dm.diagonal() == pytest.approx(1)

    @pytest.mark.parametrize('dm_dims', [[2] * 9, [3, 2, 2]])
    def test_create(self, dm_dims):
        target_shape = [d ** 2 for d in dm_dims]
        data = np.random.default_rng(5).random(target_shape)
        bs = [bases.general(d) for d in dm_dims]
        for expansion in (None, data):
            dm = qa.PauliVectorHIP(bs, expansion)
            assert dm.n_qubits == len(dm_dims)
            assert dm.dim_hilbert == pytest.approx(dm_dims)
            assert dm.dim_pauli == pytest.approx(target_shape)
            assert dm.size == np.prod(np.array(dm_dims) ** 2)
            if expansion is not None:
                np.testing.assert_array_equal(dm.to_pv(), expansion)
        wrong = int(np.prod(dm_dims))
        with pytest.raises(ValueError):
            qa.PauliVectorHIP(bs, data.reshape(wrong, wrong))
        for shape in (target_shape, (wrong, wrong)):
            with pytest.raises(ValueError):
                qa.PauliVectorHIP(bases=bs, pv=np.zeros(shape, dtype=complex))
        with pytest.raises(ValueError):
            qa.PauliVectorHIP(bases=[bases.general(2)] * 16)

    @pytest.mark.parametrize('bs', [
        (bases.general(2),) * 3,
        (bases.general(2).subbasis([0, 1, 2]), bases.general(2).subbasis([0, 1]),
         bases.general(2))])
    def test_create_from_dm_general(self, bs):
        dm = np.zeros((8, 8), dtype=complex)
        dm[0, 0], dm[1, 1] = 0.25, 0.75
        dm[0, 1], dm[1, 0] = 0.5 - 0.33j, 0.5 + 0.33j
        s = qa.PauliVectorHIP.from_dm(dm, bs)
        pv = s.to_pv()
        assert pv.shape == tuple(b.dim_pauli for b in bs)
        assert all(s.bases[i] == bs[i] for i in range(3))
        assert pv[0, 0, 0] == 0.25 and pv[0, 0, 1] == 0.75
        assert pv[0, 0, 2] == pytest.approx(0.5 * 2 ** 0.5)
        assert pv[0, 0, 3] == pytest.approx(0.33 * 2 ** 0.5)

    def test_apply_random_ptm1q_pv1q(self):
        unitary = random_unitary_matrix(2, 45)
        dm_before = random_density_matrix(2, seed=256)
        dm_after = unitary @ dm_before @ unitary.conj().T
        b = (bases.general(2),)
        pv0 = qa.PauliVectorHIP.from_dm(dm_before, b)
        pv1 = qa.PauliVectorHIP.from_dm(dm_after, b)
        assert pv0.meas_prob(0) != pytest.approx(pv1.meas_prob(0))
        ptm = kraus_to_ptm(unitary.reshape(1, 2, 2), b, b)
        b_gm = (bases.gell_mann(2),)
        ptm2 = ptm_convert_basis(ptm, b, b, b_gm, b_gm)
        assert np.allclose(ptm2[0, 1:], 0) and ptm2[0, 0] == pytest.approx(1)
        pv0.apply_ptm(ptm, 0)
        assert rel_err(pv0.to_pv(), pv1.to_pv()) < TOL64

    @pytest.mark.parametrize('qubit', [0, 1, 2])
    def test_apply_1q_random_ptm(self, qubit):
        unity = np.identity(2)
        unitary = random_unitary_matrix(2, 45)
        args = [unity, [0, 3], unity, [1, 4], unity, [2, 5]]
        args[2 * qubit] = unitary
        u3 = np.einsum(*args, optimize=True).reshape(8, 8)
        dm_before = random_density_matrix(8, seed=46)
        dm_after = u3 @ dm_before @ u3.conj().T
        b = (bases.general(2),) * 3
        pv0 = qa.PauliVectorHIP.from_dm(dm_before, b)
        pv1 = qa.PauliVectorHIP.from_dm(dm_after, b)
        for q in range(3):
            if q != qubit:
                assert pv0.meas_prob(q) == pytest.approx(pv1.meas_prob(q))
            else:
                assert pv0.meas_prob(q) != pytest.approx(pv1.meas_prob(q))
        ptm = kraus_to_ptm(unitary.reshape(1, 2, 2), b[:1], b[:1])
        pv0.apply_ptm(ptm, qubit)
        assert rel_err(pv0.to_pv(), pv1.to_pv()) < TOL64

    @pytest.mark.parametrize('qubit', [0, 1, 2])
    def test_apply_1q_projecting_ptm(self, qubit):
        # test_pauli_vectors.py:193-209 (whose final comparison lacks an
        # assert; it is asserted here)
        unitary = random_unitary_matrix(2, 45)
        b_in = (bases.general(2),)
        b_out = (bases.general(2).subbasis([1]),)
        pv = qa.PauliVectorHIP.from_dm(random_density_matrix(8, seed=46), b_in * 3)
        pv_before = pv.to_pv()
        ix_out = [0, 1, 2]
        ix_out[qubit] = 3
        ptm = kraus_to_ptm(unitary.reshape(1, 2, 2), b_in, b_out)
        pv_ref = np.einsum(ptm, [3, qubit], pv_before, [0, 1, 2], ix_out)
        pv.apply_ptm(ptm, qubit)
        assert rel_err(pv.to_pv(), pv_ref) < TOL64

    @pytest.mark.parametrize('qubits', [(0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)])
    def test_apply_2q_random_ptm(self, qubits):
        unity = np.identity(2)
        unitary = random_unitary_matrix(4, 45).reshape(2, 2, 2, 2)
        decay = np.array([[0.9, 0], [0, 0.7]])
        op = decay @ unitary
        q0, q1 = qubits
        (q2,) = tuple(q for q in range(3) if q not in qubits)
        op3q = np.einsum(op, [q0, q1, q0 + 3, q1 + 3], unity, [q2, q2 + 3],
                         optimize=True).reshape(8, 8)
        dm_before = random_density_matrix(8, seed=46)
        dm_after = op3q @ dm_before @ op3q.conj().T
        assert dm_after.trace() < 1.
        b = (bases.general(2),) * 3
        pv0 = qa.PauliVectorHIP.from_dm(dm_before, b)
        pv1 = qa.PauliVectorHIP.from_dm(dm_after, b)
        ptm = kraus_to_ptm(op.reshape(1, 4, 4), b[:2], b[:2])
        pv0.apply_ptm(ptm, *qubits)
        assert rel_err(pv0.to_pv(), pv1.to_pv()) < TOL64

    @pytest.mark.parametrize('bs', [
        (bases.general(2),) * 3,
        (bases.PauliBasis(bases.general(2).vectors[[0, 2, 1, 3]], ['0', 'X10', '1', 'Y10']),) * 3,
        (bases.general(2), bases.general(2).subbasis([0, 1]),
         bases.general(2).subbasis([0, 1, 2]))])
    def test_diagonal_meas_prob(self, bs):
        diag = np.array([0.25, 0, 0.75, 0, 0, 0, 0, 0])
        s = qa.PauliVectorHIP.from_dm(np.diag(diag), bs)
        assert np.allclose(s.diagonal(), diag)
        assert s.trace() == pytest.approx(1)
        assert np.allclose(s.meas_prob(0), (1, 0))
        assert np.allclose(s.meas_prob(1), (0.25, 0.75))
        diag = np.array([0.25, 0.5, 0, 0, 0, 0, 0, 0])
        s = qa.PauliVectorHIP.from_dm(np.diag(diag), bs)
        assert np.allclose(s.diagonal(), diag)
        assert s.trace() == pytest.approx(0.75)
        assert np.allclose(s.meas_prob(0), (0.75, 0.))
        assert np.allclose(s.meas_prob(1), (0.75, 0.))
        assert np.allclose(s.meas_prob(2), (0.25, 0.5))

    @pytest.mark.parametrize('dim1,dim2', [[2, 2], [2, 3], [3, 3]])
    def test_get_diagonal(self, dim1, dim2):
        dm = qa.PauliVectorHIP([bases.general(dim1), bases.general(dim2)])
        ref = np.zeros(dim1 * dim2)
        ref[0] = 1.
        assert np.array_equal(dm.diagonal(), ref)

    @pytest.mark.parametrize('bs', [
        (bases.general(2),) * 3,
        (bases.general(2).subbasis([0, 1]), bases.general(2).subbasis([0, 1, 2]),
         bases.general(2).subbasis([0, 1]))])
    def test_diagonal_indicated(self, bs):
        dm = np.array([[min(i, j) * 10 + max(i, j) for i in range(1, 9)]
                       for j in range(1, 9)])
        s = qa.PauliVectorHIP.from_dm(dm, bs)
        assert np.allclose(s.diagonal(), np.diag(dm))


# ----------------------------------------------------- oracle parity, f64
BASIS_SETS = {
    'qubits3': lambda: [bases.general(2)] * 3,
    'qubits5': lambda: [bases.general(2)] * 5,
    'mixed': lambda: [bases.general(3), bases.general(2), bases.gell_mann(2),
                      bases.general(3)],
    'gm_qutrits': lambda: [bases.gell_mann(3)] * 3,
    'subbases': lambda: [bases.general(2).subbasis([0, 1]), bases.general(3),
                         bases.general(2).subbasis([0]), bases.general(2)],
}


def _targets_all(n, k):
    return list(itertools.permutations(range(n), k))


@pytest.mark.parametrize('name', list(BASIS_SETS))
@pytest.mark.parametrize('k', [1, 2, 3])
def test_apply_parity_all_positions(name, k, rng):
    bs = BASIS_SETS[name]()
    n = len(bs)
    dims = [b.dim_pauli for b in bs]
    pv = rng.standard_normal(dims)
    tlist = _targets_all(n, k)
    if len(tlist) > 24:
        tlist = [tlist[i] for i in rng.choice(len(tlist), 24, replace=False)]
    for targets in tlist:
        din = [dims[t] for t in targets]
        # square and rectangular PTMs (basis reduction / expansion)
        for dout in (din, [max(1, d // 2) for d in din], [d + 1 for d in din]):
            ptm = rng.standard_normal(list(dout) + din)
            v = qa.PauliVectorHIP(bs, pv.copy())
            o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
            v.apply_ptm(ptm, *targets)
            o.apply_ptm(ptm, *targets)
            err = rel_err(v.to_pv(), o.to_pv())
            assert err < TOL64, (name, targets, dout, err)


@pytest.mark.parametrize('n', [4, 6, 9])
@pytest.mark.parametrize('k', [3, 4])
def test_apply_fused_dense(n, k, rng):
    # 64 x 64 and 256 x 256 PTMs (test_operations.py:174-195 applies 64 x 64)
    # on the MFMA kernels: outer, inner (vector accesses) and scattered
    # targets; n = 9 gives several tiles per wave
    bs = [bases.general(2)] * n
    pv = rng.standard_normal([4] * n)
    for targets in [tuple(range(k)), tuple(range(n - k, n)), tuple(rng.permutation(n)[:k])]:
        ptm = rng.standard_normal([4] * (2 * k)) / 4 ** (k / 2)
        o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
        o.apply_ptm(ptm, *targets)
        v = qa.PauliVectorHIP(bs, pv.copy())
        v.apply_ptm(ptm, *targets)
        assert rel_err(v.to_pv(), o.to_pv()) < TOL64, (n, targets)
        v32 = qa.PauliVectorHIP(bs, pv.copy(), dtype='float32')
        v32.apply_ptm(ptm, *targets)
        assert rel_err(v32.to_pv(), o.to_pv()) < 1e-5, (n, targets)


def test_apply_fused_dense_ququarts(rng):
    # a Pauli dimension of 16 (d = 4 qudit) inside 64- and 256-wide PTMs
    bs = [bases.general(2), bases.general(4), bases.general(2), bases.general(2),
          bases.general(4)]
    dims = [b.dim_pauli for b in bs]
    pv = rng.standard_normal(dims)
    for targets in [(1, 2), (3, 4), (0, 1, 2), (2, 3, 4), (4, 0)]:
        din = [dims[t] for t in targets]
        ptm = rng.standard_normal(din + din) / np.sqrt(np.prod(din))
        o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
        o.apply_ptm(ptm, *targets)
        for dt, tol in (('float64', TOL64), ('float32', 1e-5)):
            v = qa.PauliVectorHIP(bs, pv.copy(), dtype=dt)
            v.apply_ptm(ptm, *targets)
            assert rel_err(v.to_pv(), o.to_pv()) < tol, (targets, dt)


def test_apply_qutrit_pair_81(rng):
    bs = [bases.gell_mann(3), bases.general(2), bases.general(3)]
    pv = rng.standard_normal([9, 4, 9])
    for targets in [(0, 2), (2, 0), (0, 1), (1, 2)]:
        din = [bs[t].dim_pauli for t in targets]
        ptm = rng.standard_normal(din + din)
        v = qa.PauliVectorHIP(bs, pv.copy())
        o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
        v.apply_ptm(ptm, *targets)
        o.apply_ptm(ptm, *targets)
        assert rel_err(v.to_pv(), o.to_pv()) < TOL64


def test_basis_reduction_bit_exact(rng):
    # a 0/1 selection PTM (general(2) -> subbasis([0, 1])) is a gather:
    # results must equal the input entries exactly
    bs = [bases.general(2)] * 4
    pv = rng.standard_normal([4] * 4)
    sel = np.zeros((2, 4))
    sel[0, 0] = sel[1, 1] = 1
    for q in range(4):
        v = qa.PauliVectorHIP(bs, pv.copy())
        v.apply_ptm(sel, q)
        idx = [list(range(4))] * 4
        idx[q] = [0, 1]
        assert np.array_equal(v.to_pv(), multitake(pv, idx))
        o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
        o.apply_ptm(sel, q)
        assert np.array_equal(v.to_pv(), o.to_pv())


@pytest.mark.parametrize('name', list(BASIS_SETS))
def test_readouts_parity(name, rng):
    bs = BASIS_SETS[name]()
    dims = [b.dim_pauli for b in bs]
    pv = rng.standard_normal(dims)
    v = qa.PauliVectorHIP(bs, pv.copy())
    o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
    assert rel_err(v.diagonal(), o.diagonal()) < TOL64
    assert abs(v.trace() - o.trace()) < TOL64 * max(1, abs(o.trace()))
    assert isinstance(v.trace(), np.float64)
    for q in range(len(bs)):
        mp = v.meas_prob(q)
        assert isinstance(mp, np.ndarray) and mp.dtype == np.float64
        assert rel_err(mp, o.meas_prob(q)) < TOL64
    n = len(bs)
    for keep in [(0,), (n - 1,), (0, n - 1), tuple(range(n)), (1, 2)]:
        pt = v.partial_trace(*keep)
        po = o.partial_trace(*keep)
        assert all(a is bs[q] for a, q in zip(pt.bases, keep))
        assert rel_err(pt.to_pv(), po.to_pv()) < TOL64


def test_diagonal_is_bit_exact_gather_for_general(rng):
    bs = [bases.general(2), bases.general(3), bases.general(2).subbasis([0, 2, 3])]
    pv = rng.standard_normal([b.dim_pauli for b in bs])
    v = qa.PauliVectorHIP(bs, pv.copy())
    o = PauliVectorNumpy(_oracle_bases(bs), pv.copy())
    assert np.array_equal(v.diagonal(), o.diagonal())
    # the device form is the same flat vector, no host copy
    d = v.diagonal(get_data=False)
    assert d.is_cuda and d.dim() == 1
    assert np.array_equal(d.cpu().numpy(), o.diagonal())


def test_partial_trace_quirks(rng):
    # numpy.py:86-95: bases follow caller order, data axes stay ascending
    bs = [bases.general(2), bases.general(3)]
    v = qa.PauliVectorHIP(bs, rng.standard_normal((4, 9)))
    with pytest.raises(ValueError):
        v.partial_trace(1, 0)
    b2 = [bases.general(2), bases.gell_mann(2)]
    pv = rng.standard_normal((4, 4))
    v2 = qa.PauliVectorHIP(b2, pv)
    pt = v2.partial_trace(1, 0)
    assert pt.bases[0] is b2[1]
    assert np.array_equal(pt.to_pv(), pv)
    with pytest.raises(ValueError):
        v2.partial_trace(2)


def test_renormalize_and_copy(rng):
    b = [bases.general(2)] * 3
    pv = np.abs(rng.standard_normal((4, 4, 4)))
    v = qa.PauliVectorHIP(b, pv.copy())
    o = PauliVectorNumpy(_oracle_bases(b), pv.copy())
    c = v.copy()
    tr = v.renormalize()
    assert abs(tr - o.renormalize()) < TOL64 * abs(tr)
    assert rel_err(v.to_pv(), o.to_pv()) < TOL64
    assert v.trace() == pytest.approx(1, abs=1e-14)
    assert np.array_equal(c.to_pv(), pv)                  # the copy is independent
    z = qa.PauliVectorHIP(b, np.zeros((4, 4, 4)))
    with pytest.warns(UserWarning):
        assert z.renormalize() == 0


def test_apply_errors(rng):
    v = qa.PauliVectorHIP([bases.general(2)] * 3)
    with pytest.raises(ValueError):
        v.apply_ptm(np.eye(4), 0, 1)                      # ndim mismatch
    with pytest.raises(ValueError):
        v.apply_ptm(np.eye(4), 3)                         # no such qubit
    with pytest.raises(ValueError):
        v.apply_ptm(np.eye(9), 0)                         # wrong input dim
    with pytest.raises(ValueError):
        v.apply_ptm(np.eye(16).reshape(4, 4, 4, 4), 1, 1)
    with pytest.raises(ValueError):
        v.meas_prob(5)
    assert np.array_equal(v.diagonal(), np.eye(8)[0])    # state untouched


def test_default_state_is_index_zero():
    # numpy.py:39-41 (for gell_mann this is (I/sqrt d)^n, not |0..0>)
    v = qa.PauliVectorHIP([bases.gell_mann(2), bases.general(3)])
    ref = np.zeros((4, 9))
    ref[0, 0] = 1
    assert np.array_equal(v.to_pv(), ref)


def test_float32_state(rng):
    b = [bases.general(2)] * 6
    pv = rng.standard_normal([4] * 6)
    v = qa.PauliVectorHIP(b, pv.copy(), dtype='float32')
    o = PauliVectorNumpy(_oracle_bases(b), pv.astype(np.float32).astype(np.float64))
    for targets in [(0,), (5,), (2, 3), (5, 0), (1, 3, 4)]:
        ptm = rng.standard_normal([4] * (2 * len(targets))) / 4 ** len(targets)
        v.apply_ptm(ptm, *targets)
        o.apply_ptm(ptm, *targets)
    assert v.to_pv().dtype == np.float32
    assert rel_err(v.to_pv(), o.to_pv()) < 1e-5
    assert v.diagonal().dtype == np.float64


# ------------------------------------- full-size properties (14 qubits)
@pytest.mark.parametrize('dtype,tol', [('float64', 1e-12), ('float32', 2e-6)])
def test_14q_orthogonal_round_trip(dtype, tol, rng):
    """A unitary channel's PTM in an orthonormal real basis is orthogonal:
    applying R then R^T must restore the vector, at every position class
    (outer, middle, innermost) -- a size-independent check at full size."""
    n = 14
    b = [bases.general(2)] * n
    v = qa.PauliVectorHIP(b, force=True, dtype=dtype)
    import torch
    x0 = torch.randn((4,) * n, dtype=torch.float64, device=v.device).to(v.dtype)
    v.to_torch().copy_(x0)
    norm0 = torch.linalg.vector_norm(x0.double()).item()
    U = random_unitary_matrix(4, 7)
    R = kraus_to_ptm(U[None], (bases.general(2),) * 2, (bases.general(2),) * 2)
    Rm = R.reshape(16, 16)
    assert np.allclose(Rm @ Rm.T, np.eye(16), atol=1e-14)
    RT = Rm.T.reshape(4, 4, 4, 4)
    for targets in [(0, 1), (6, 7), (12, 13), (13, 0), (3, 10), (1, 2), (2, 3)]:
        v.apply_ptm(R, *targets)
        assert abs(torch.linalg.vector_norm(v.to_torch().double()).item() / norm0 - 1) < tol
        v.apply_ptm(RT, *targets)
    err = (v.to_torch().double() - x0.double()).abs().max().item()
    assert err < tol * 10, err
    r1 = kraus_to_ptm(random_unitary_matrix(2, 8)[None], (bases.general(2),),
                      (bases.general(2),))
    for q in (0, 7, 13):
        v.apply_ptm(r1, q)
        v.apply_ptm(np.ascontiguousarray(r1.T), q)
    err = (v.to_torch().double() - x0.double()).abs().max().item()
    assert err < tol * 10, err
    # the padded layout: strides of the outer axes are not powers of two
    if dtype == 'float64':
        assert v._pstr[0] != 4 ** (n - 1)


def test_two_streams_dense_ptms_match_one_stream(rng):
    """Vectors driven on two torch streams at once, with dense PTMs that are
    staged through a device workspace (64 x 64 on MFMA, 81 x 81 qutrit
    pairs in the table kernel): every result equals the same calls made
    one after another on one stream (each stream has its own workspace)."""
    import torch
    cases = [([bases.general(2)] * 6, [4] * 6, (1, 2, 4)),
             ([bases.general(3)] * 4, [9] * 4, (0, 2))]
    for bs, shape, targets in cases:
        d = math.prod(shape[t] for t in targets)
        pvs = [rng.standard_normal(shape) for _ in range(2)]
        ptms = [[rng.standard_normal([shape[t] for t in targets] * 2) / d for _ in range(4)]
                for _ in range(2)]
        ref = []
        for pv, seq in zip(pvs, ptms):
            v = qa.PauliVectorHIP(bs, pv)
            for p in seq:
                v.apply_ptm(p, *targets)
            ref.append(v.to_pv())
        vs = [qa.PauliVectorHIP(bs, pv) for pv in pvs]
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        for s in streams:
            s.wait_stream(torch.cuda.current_stream())
        for r in range(4):
            for v, s, seq in zip(vs, streams, ptms):
                with torch.cuda.stream(s):
                    v.apply_ptm(seq[r], *targets)
        torch.cuda.synchronize()
        for v, want in zip(vs, ref):
            assert np.array_equal(v.to_pv(), want), (shape, targets)


def test_side_stream_growing_subbasis_vector_matches_oracle(rng):
    """A vector made on the default stream and driven on a side stream with
    PTMs that grow its extent (a subbasis qubit mapped back to the full
    basis), so buffers allocated on the default stream are replaced while
    the side stream's kernels may still use them; the default stream then
    allocates and fills memory of the same sizes before anything is
    synchronised.  The result equals the oracle's."""
    import torch
    sub = bases.general(2).subbasis([0, 1])
    full = bases.general(2)
    bs = [sub, sub] + [full] * 5
    shape = [b.dim_pauli for b in bs]
    pv = rng.standard_normal(shape)
    grow = rng.standard_normal((4, 2))
    sq = [rng.standard_normal((4, 4, 4, 4)) / 4 for _ in range(3)]
    o = PauliVectorNumpy(_oracle_bases(bs), pv)
    for q in (0, 1):
        o.apply_ptm(grow, q)
    for p in sq:
        o.apply_ptm(p, 2, 0)
    want = o.to_pv()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    v = qa.PauliVectorHIP(bs, pv)
    with torch.cuda.stream(side):
        for q in (0, 1):
            v.apply_ptm(grow, q)          # extent grows: the old buffers are dropped
        for p in sq:
            v.apply_ptm(p, 2, 0)
    junk = [torch.full((4 ** 7,), float('nan'), dtype=torch.float64, device=v.device)
            for _ in range(6)]
    side.synchronize()
    torch.cuda.synchronize()
    del junk
    assert rel_err(v.to_pv(), want) < TOL64
