"""Pins the CPU oracle (oracle/) before the tests trust it as their reference.

Each check is chosen so a plausible mistake (wrong element order, a dropped
conjugate, transposed PTM, mislabelled axis) fails it: the reference's own
known-answer values, Hilbert-space brute force, closed forms and
invariants.  Reference test locations are cited per test.
"""
import itertools
import math
import warnings

import numpy as np
import pytest

from oracle import bases as ob, algebra as oa, brute
from oracle.pauli_vector import PauliVectorNumpy, multitake

SQ2 = np.sqrt(2)


# ------------------------------------------------------------------ bases
def test_general2_elements():
    # library.py:10-53: [|0><0|, |1><1|, X10, Y10]
    v = ob.general(2).vectors
    r = np.sqrt(0.5)     # the reference's constant (library.py:7)
    exp = np.array([[[1, 0], [0, 0]], [[0, 0], [0, 1]],
                    [[0, r], [r, 0]], [[0, -1j * r], [1j * r, 0]]])
    assert np.array_equal(v, exp)


def test_general3_order():
    # test_compiler.py:301 subbasis([0,1,2,6]) = {0,1,2,Y20};
    # test_operations_qutrits.py:61-93: X10 at index 3, Y10 at 4
    b = ob.general(3)
    assert b.labels == ['0', '1', '2', 'X10', 'Y10', 'X20', 'Y20', 'X21', 'Y21']
    y20 = b.vectors[6]
    assert y20[2, 0] == 1j * np.sqrt(0.5) and y20[0, 2] == -1j * np.sqrt(0.5)


def test_gell_mann2_is_pauli_over_sqrt2():
    # qubits.py:5 zips ['I','X','Y','Z'] with gell_mann(2).vectors
    v = ob.gell_mann(2).vectors * SQ2
    assert np.allclose(v[0], np.eye(2))
    assert np.allclose(v[1], [[0, 1], [1, 0]])
    assert np.allclose(v[2], [[0, -1j], [1j, 0]])
    assert np.allclose(v[3], [[1, 0], [0, -1]])


@pytest.mark.parametrize('fac', [ob.general, ob.gell_mann])
@pytest.mark.parametrize('d', [2, 3, 4])
def test_orthonormal_hermitian(fac, d):
    v = fac(d).vectors
    gram = np.einsum('xab,yba->xy', v, v)
    assert np.allclose(gram, np.eye(d * d), atol=1e-14)
    assert np.allclose(v, np.conj(np.transpose(v, (0, 2, 1))))


# ---------------------------------------------------------------- algebra
def test_amp_damping_ptm_kat():
    # test_operations.py:22-37 (gell_mann basis, p = 0.5)
    p = 0.5
    K = np.array([[[1, 0], [0, np.sqrt(1 - p)]], [[0, np.sqrt(p)], [0, 0]]])
    b = (ob.gell_mann(2),)
    exp = np.array([[1, 0, 0, 0], [0, np.sqrt(1 - p), 0, 0],
                    [0, 0, np.sqrt(1 - p), 0], [p, 0, 0, 1 - p]])
    assert np.allclose(oa.kraus_to_ptm(K, b, b), exp, atol=1e-14)


def test_cz_ptm_properties():
    # test_operations.py:42-51
    b = (ob.gell_mann(2),) * 2
    ptm = oa.kraus_to_ptm(np.diag([1, 1, 1, -1]), b, b).reshape(16, 16)
    assert np.all(ptm.round(3) <= 1) and np.all(ptm.round(3) >= -1)
    assert np.isclose(ptm[0].sum(), 1) and np.isclose(ptm[:, 0].sum(), 1)


def test_from_dm_kat():
    # test_pauli_vectors.py:109-125
    dm = np.zeros((8, 8), complex)
    dm[0, 0], dm[1, 1] = 0.25, 0.75
    dm[0, 1], dm[1, 0] = 0.5 - 0.33j, 0.5 + 0.33j
    for bases in [(ob.general(2),) * 3,
                  (ob.general(2).subbasis([0, 1, 2]), ob.general(2).subbasis([0, 1]),
                   ob.general(2))]:
        pv = oa.dm_to_pv(dm, bases)
        assert pv.shape == tuple(b.dim_pauli for b in bases)
        assert pv[0, 0, 0] == 0.25 and pv[0, 0, 1] == 0.75
        assert np.isclose(pv[0, 0, 2], 0.5 * SQ2) and np.isclose(pv[0, 0, 3], 0.33 * SQ2)


@pytest.mark.parametrize('dims', [[2, 2, 2], [3, 2], [3, 3]])
def test_dm_pv_round_trip(dims, rng):
    # pv_to_dm(dm_to_pv(rho)) == rho in full bases (needs uniform d)
    if len(set(dims)) > 1:
        pytest.skip('dm_to_pv assumes a uniform Hilbert dimension (algebra.py:46)')
    D = int(np.prod(dims))
    rho = brute.random_density_matrix(D, rng)
    for fac in (ob.general, ob.gell_mann):
        bases = [fac(d) for d in dims]
        assert np.allclose(oa.pv_to_dm(oa.dm_to_pv(rho, bases), bases), rho, atol=1e-13)


def test_kraus_to_ptm_is_cptp_in_gell_mann(rng):
    # trace preservation <=> row 0 of the PTM in gell_mann is e_0
    for d in (2, 3):
        X = rng.standard_normal((4 * d, d)) + 1j * rng.standard_normal((4 * d, d))
        q, _ = np.linalg.qr(X)
        K = q.reshape(4, d, d)
        b = (ob.gell_mann(d),)
        ptm = oa.kraus_to_ptm(K, b, b)
        e0 = np.zeros(d * d)
        e0[0] = 1
        assert np.allclose(ptm[0], e0, atol=1e-13)


def test_ptm_convert_basis_matches_direct(rng):
    U = brute.random_unitary(2, rng)
    bg, bm = (ob.general(2),), (ob.gell_mann(2),)
    p_g = oa.kraus_to_ptm(U, bg, bg)
    p_m = oa.kraus_to_ptm(U, bm, bm)
    assert np.allclose(oa.ptm_convert_basis(p_g, bg, bg, bm, bm), p_m, atol=1e-14)


# ----------------------------------------------------- backend vs brute force
def _random_kraus(dims_t, rng, n_kraus=2):
    D = int(np.prod(dims_t))
    X = rng.standard_normal((n_kraus * D, D)) + 1j * rng.standard_normal((n_kraus * D, D))
    q, _ = np.linalg.qr(X)
    return q.reshape(n_kraus, D, D) * 0.9   # trace-decreasing CP map


CASES = [
    ([2, 2, 2], (1,)), ([2, 2, 2], (0, 2)), ([2, 2, 2], (2, 0)), ([2, 2, 2], (1, 0)),
    ([3, 2, 2], (0, 1)), ([3, 2, 2], (2, 0)), ([2, 3, 2], (1,)), ([2, 2, 2], (2, 0, 1)),
    ([2, 3, 2], (0, 2, 1)),
]


@pytest.mark.parametrize('fac', [ob.general, ob.gell_mann])
@pytest.mark.parametrize('dims,targets', CASES)
def test_oracle_apply_matches_hilbert(dims, targets, fac, rng):
    bases = [fac(d) for d in dims]
    D = int(np.prod(dims))
    rho = brute.random_density_matrix(D, rng)
    pv = np.asarray(_dm_to_pv_mixed(rho, bases, dims))
    K = _random_kraus([dims[t] for t in targets], rng)
    tb = [bases[t] for t in targets]
    ptm = oa.kraus_to_ptm(K, tb, tb)
    v = PauliVectorNumpy(bases, pv.copy())
    v.apply_ptm(ptm, *targets)
    rho2 = brute.apply_channel(rho, brute.embed_kraus(K, targets, dims))
    assert np.allclose(v.to_pv(), _dm_to_pv_mixed(rho2, bases, dims), atol=1e-13)


def _dm_to_pv_mixed(rho, bases, dims):
    """Re tr(P_x rho) for qudits of unequal dimension (index loops)."""
    n = len(dims)
    out = np.zeros([b.dim_pauli for b in bases])
    for xs in itertools.product(*[range(b.dim_pauli) for b in bases]):
        P = np.array([[1.0 + 0j]])
        for b, x in zip(bases, xs):
            P = np.kron(P, b.vectors[x])
        out[xs] = np.trace(P @ rho).real
    return out


def test_planted_transposed_ptm_fails(rng):
    # a mistake planted in the oracle (PTM transposed) must fail brute force
    dims, targets = [2, 2], (0, 1)
    bases = [ob.general(2)] * 2
    rho = brute.random_density_matrix(4, rng)
    K = _random_kraus([2, 2], rng)
    ptm = oa.kraus_to_ptm(K, bases, bases)
    bad = ptm.reshape(16, 16).T.reshape(4, 4, 4, 4)
    v = PauliVectorNumpy(bases, _dm_to_pv_mixed(rho, bases, dims))
    v.apply_ptm(bad, *targets)
    rho2 = brute.apply_channel(rho, brute.embed_kraus(K, targets, dims))
    assert not np.allclose(v.to_pv(), _dm_to_pv_mixed(rho2, bases, dims), atol=1e-6)


def test_planted_swapped_targets_fails(rng):
    dims = [2, 2, 2]
    bases = [ob.general(2)] * 3
    rho = brute.random_density_matrix(8, rng)
    K = _random_kraus([2, 2], rng)
    ptm = oa.kraus_to_ptm(K, bases[:2], bases[:2])
    v = PauliVectorNumpy(bases, _dm_to_pv_mixed(rho, bases, dims))
    v.apply_ptm(ptm, 2, 0)       # applied with the targets swapped
    rho2 = brute.apply_channel(rho, brute.embed_kraus(K, (0, 2), dims))
    assert not np.allclose(v.to_pv(), _dm_to_pv_mixed(rho2, bases, dims), atol=1e-6)


# --------------------------------------------------------------- readouts
READ_BASES = [
    lambda: [ob.general(2)] * 3,
    lambda: [ob.gell_mann(2)] * 3,
    lambda: [ob.general(2), ob.general(2).subbasis([0, 1]), ob.general(2).subbasis([0, 1, 2])],
    lambda: [ob.general(3), ob.gell_mann(2), ob.general(2)],
    lambda: [ob.gell_mann(3), ob.general(2).subbasis([1, 3])],
]


@pytest.mark.parametrize('mk', READ_BASES)
def test_oracle_readouts_match_hilbert(mk, rng):
    bases = mk()
    dims = [b.dim_hilbert for b in bases]
    D = int(np.prod(dims))
    rho = brute.random_density_matrix(D, rng)
    pv = _dm_to_pv_mixed(rho, bases, dims)
    # what a truncated basis can represent
    rho_r = sum(pv[xs] * _kron_elems(bases, xs)
                for xs in itertools.product(*[range(b.dim_pauli) for b in bases]))
    v = PauliVectorNumpy(bases, pv)
    assert np.allclose(v.diagonal(), np.diag(rho_r).real, atol=1e-13)
    assert np.isclose(v.trace(), np.trace(rho_r).real, atol=1e-13)
    for q in range(len(dims)):
        red = brute.partial_trace(rho_r, dims, [q])
        assert np.allclose(v.meas_prob(q), np.diag(red).real, atol=1e-13)
    for keep in [(0,), (0, 1), (1,), (0, len(dims) - 1)]:
        keep = tuple(sorted(set(keep)))
        pt = v.partial_trace(*keep)
        red = brute.partial_trace(rho_r, dims, keep)
        kb = [bases[q] for q in keep]
        assert np.allclose(pt.to_pv(), _dm_to_pv_mixed(red, kb, [dims[q] for q in keep]),
                           atol=1e-13)


def _kron_elems(bases, xs):
    P = np.array([[1.0 + 0j]])
    for b, x in zip(bases, xs):
        P = np.kron(P, b.vectors[x])
    return P


def test_diagonal_meas_prob_kat():
    # test_pauli_vectors.py:239-271
    for bases in [[ob.general(2)] * 3,
                  [ob.general(2), ob.general(2).subbasis([0, 1]),
                   ob.general(2).subbasis([0, 1, 2])]]:
        diag = np.array([0.25, 0, 0.75, 0, 0, 0, 0, 0])
        s = PauliVectorNumpy(bases, oa.dm_to_pv(np.diag(diag), bases))
        assert np.allclose(s.diagonal(), diag)
        assert np.isclose(s.trace(), 1)
        assert np.allclose(s.meas_prob(0), (1, 0))
        assert np.allclose(s.meas_prob(1), (0.25, 0.75))
        diag = np.array([0.25, 0.5, 0, 0, 0, 0, 0, 0])
        s = PauliVectorNumpy(bases, oa.dm_to_pv(np.diag(diag), bases))
        assert np.isclose(s.trace(), 0.75)
        assert np.allclose(s.meas_prob(2), (0.25, 0.5))


def test_diagonal_indicated_kat():
    # test_pauli_vectors.py:285-307: entries 10*min(i,j) + max(i,j)
    dm = np.array([[min(i, j) * 10 + max(i, j) for i in range(1, 9)] for j in range(1, 9)])
    for bases in [[ob.general(2)] * 3,
                  [ob.general(2).subbasis([0, 1]), ob.general(2).subbasis([0, 1, 2]),
                   ob.general(2).subbasis([0, 1])]]:
        s = PauliVectorNumpy(bases, oa.dm_to_pv(dm, bases))
        assert np.allclose(s.diagonal(), np.diag(dm))


def test_default_state_and_trivial():
    # test_pauli_vectors.py:66-69, 273-283
    assert PauliVectorNumpy([]).to_pv() == 1
    for d1, d2 in [(2, 2), (2, 3), (3, 3)]:
        s = PauliVectorNumpy([ob.general(d1), ob.general(d2)])
        ref = np.zeros(d1 * d2)
        ref[0] = 1
        assert np.array_equal(s.diagonal(), ref)


def test_partial_trace_axis_order_quirk(rng):
    # numpy.py:86-95: data axes ascending, bases in caller order
    bases = [ob.general(2), ob.general(3)]
    s = PauliVectorNumpy(bases, rng.standard_normal((4, 9)))
    with pytest.raises(ValueError):
        s.partial_trace(1, 0)
    b2 = [ob.general(2), ob.gell_mann(2)]
    s2 = PauliVectorNumpy(b2, rng.standard_normal((4, 4)))
    pt = s2.partial_trace(1, 0)
    assert pt.bases[0] is b2[1] and np.array_equal(pt.to_pv(), s2.to_pv())


def test_renormalize_semantics():
    b = [ob.general(2)]
    s = PauliVectorNumpy(b, np.array([0.5, 0.25, 0.1, 0.0]))
    tr = s.renormalize()
    assert tr == 0.75 and np.isclose(s.trace(), 1)
    z = PauliVectorNumpy(b, np.zeros(4))
    with pytest.warns(UserWarning):
        assert z.renormalize() == 0


def test_multitake_matches_loops(rng):
    x = rng.standard_normal((4, 9, 4))
    idx = [[0, 1], [0, 1, 2, 8], [3, -1, 0]]
    got = multitake(x, idx)
    for a, i in enumerate(idx[0]):
        for b, j in enumerate(idx[1]):
            for c, k in enumerate(idx[2]):
                exp = 0.0 if k < 0 else x[i, j, k]
                assert got[a, b, c] == exp
