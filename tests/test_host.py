"""CPU-only checks of the product's host side: the C ABI loads and exports
what include/qsv.h declares, argument validation returns errors without GPU
work, the index arithmetic is exact, and the product's bases/algebra agree
with the oracle's independent restatement."""
import ctypes
import os
import re

import numpy as np
import pytest

import quantumsim_amd as qa
from quantumsim_amd import _native as nat
from quantumsim_amd.algebra import kraus_to_ptm, ptm_convert_basis, dm_to_pv, pv_to_dm
from quantumsim_amd.bases import general, gell_mann
from quantumsim_amd.pauli_vectors.hip import _BasisPlan
from oracle import bases as ob, algebra as oa, brute

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, 'include')):
        if fn.endswith('.h'):
            with open(os.path.join(ROOT, 'include', fn)) as fh:
                src = fh.read()
            names |= set(re.findall(r'\b(qsv_[a-z0-9_]+)\s*\(', src))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nat.library_path)
    declared = _declared_symbols()
    assert declared, 'no declarations found'
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared <= set(nat.exported_symbols())
    assert nat.lib().qsv_abi_version() >= 100


def test_fastdiv_selftest():
    assert nat.lib().qsv_selftest_fastdiv(20000, 12345) == 0


def test_apply_argument_errors_are_value_errors():
    L = nat.lib()
    ptm = np.eye(4)
    p = ptm.ctypes.data_as(nat._c_f64p)
    cases = [
        dict(rank=2, dims=[4, 4], k=1, targets=[2], dout=[4]),    # no such qubit
        dict(rank=2, dims=[4, 4], k=2, targets=[1, 1], dout=[4, 4]),  # duplicate
        dict(rank=2, dims=[4, 4], k=0, targets=[0], dout=[4]),    # no target
        dict(rank=2, dims=[4, 0], k=1, targets=[0], dout=[4]),    # bad dim
    ]
    for c in cases:
        rc = L.qsv_apply_ptm(nat.QSV_F64, 16, 32, c['rank'], nat.i64(c['dims']), c['k'],
                             nat.i32(c['targets']), nat.i32(c['dout']), p, None, 0, None)
        assert rc == -1
        with pytest.raises(ValueError):
            nat.check(rc, 'apply_ptm')
    # a 81 x 81 PTM needs a workspace; none given -> error before any launch
    rc = L.qsv_apply_ptm(nat.QSV_F64, 16, 32, 2, nat.i64([9, 9]), 2, nat.i32([0, 1]),
                         nat.i32([9, 9]), np.eye(81).ctypes.data_as(nat._c_f64p), None, 0,
                         None)
    assert rc == -1 and b'workspace' in L.qsv_last_error()
    assert L.qsv_apply_ptm_workspace_bytes(nat.QSV_F64, 2, nat.i32([4, 4]),
                                           nat.i32([4, 4])) == 0
    assert L.qsv_apply_ptm_workspace_bytes(nat.QSV_F64, 2, nat.i32([9, 9]),
                                           nat.i32([9, 9])) > 81 * 81 * 8


def test_other_argument_errors():
    L = nat.lib()
    assert L.qsv_axis_sum(nat.QSV_F64, 16, 1, 65, 1, 16, 16, 1 << 20, None) == -1
    assert L.qsv_scale(7, 16, 4, 1.0, None) == -1
    assert L.qsv_multitake(nat.QSV_F64, nat.QSV_F32, 16, 32, 1, nat.i64([4]), nat.i64([2]),
                           16, nat.i32([0]), None) == -1


@pytest.mark.parametrize('d', [2, 3, 4])
def test_product_bases_equal_oracle(d):
    for fp, fo in [(general, ob.general), (gell_mann, ob.gell_mann)]:
        assert np.array_equal(fp(d).vectors, fo(d).vectors)
        assert np.array_equal(fp(d).computational_basis_vectors,
                              fo(d).computational_basis_vectors)


def test_basis_index_hints():
    # pauli_basis.py:38-46
    g = general(3)
    assert g.computational_basis_indices == {0: 0, 1: 1, 2: 2}
    assert gell_mann(2).computational_basis_indices == {0: None, 1: None}
    assert gell_mann(2).trace_index == 0
    assert g.subbasis([0, 1, 2, 6]).labels == ['0', '1', '2', 'Y20']
    assert g.subbasis([1, 3]).computational_basis_indices == {0: None, 1: 0, 2: None}


def test_product_algebra_equals_oracle(rng):
    for d, n in [(2, 1), (2, 2), (3, 1), (3, 2)]:
        D = d ** n
        X = rng.standard_normal((3 * D, D)) + 1j * rng.standard_normal((3 * D, D))
        K = np.linalg.qr(X)[0].reshape(3, D, D)
        for fp, fo in [(general, ob.general), (gell_mann, ob.gell_mann)]:
            bp, bo = (fp(d),) * n, (fo(d),) * n
            assert np.allclose(kraus_to_ptm(K, bp, bp), oa.kraus_to_ptm(K, bo, bo),
                               atol=1e-14)
        rho = brute.random_density_matrix(D, rng)
        bp, bo = (general(d),) * n, (ob.general(d),) * n
        assert np.allclose(dm_to_pv(rho, bp), oa.dm_to_pv(rho, bo), atol=1e-14)
        pv = dm_to_pv(rho, bp)
        assert np.allclose(pv_to_dm(pv, bp), oa.pv_to_dm(pv, bo), atol=1e-14)
        bm, bom = (gell_mann(d),) * n, (ob.gell_mann(d),) * n
        p = kraus_to_ptm(K, bp, bp)
        assert np.allclose(ptm_convert_basis(p, bp, bp, bm, bm),
                           oa.ptm_convert_basis(p, bo, bo, bom, bom), atol=1e-13)


def test_basis_plans():
    gp = _BasisPlan(general(2))
    assert gp.take == [0, 1] and gp.transform is None
    assert list(gp.tvec) == [1, 1, 0, 0]
    sb = _BasisPlan(general(2).subbasis([1, 2]))
    assert sb.take == [-1, 0] and sb.transform is None
    gm = _BasisPlan(gell_mann(3))
    assert gm.take == [0, 4, 8] and gm.transform.shape == (3, 3)
    assert np.allclose(gm.transform.sum(axis=0) * np.sqrt(3), [3, 0, 0])


def test_constructor_validation_without_gpu():
    b = [general(2)] * 16
    with pytest.raises(ValueError):
        qa.PauliVectorHIP(b)                       # size guard, pauli_vector.py:33-37
    with pytest.raises(ValueError):
        qa.PauliVectorHIP([general(2)] * 2, np.zeros((4, 4), complex))
    with pytest.raises(ValueError):
        qa.PauliVectorHIP([general(2)] * 2, np.zeros((2, 2, 2, 2)))
    with pytest.raises(ValueError):
        qa.PauliVectorHIP([general(2)] * 2, [[0.0] * 4] * 4)


def test_t1t2_rates_closed_form():
    from quantumsim_amd.operations.qubits import t1t2_rates
    t, t1, t2 = 20e-9, 30e-6, 17e-6
    g, lam = t1t2_rates(t, t1, t2)
    assert np.isclose(np.sqrt(1 - g) * np.sqrt(1 - lam), np.exp(-t / t2), rtol=1e-14)
    assert np.isclose(1 - g, np.exp(-t / t1), rtol=1e-14)


def test_autotune_cache_roundtrip(tmp_path):
    """The tuner's cache file: entries survive save/load, power-of-two batch
    buckets, summary sums (no GPU needed)."""
    from quantumsim_amd.autotune import Autotuner, _bucket
    assert [_bucket(b) for b in (1, 2, 3, 4, 5, 64, 65)] == [1, 2, 4, 4, 8, 64, 128]
    path = str(tmp_path / 'tiles.json')
    t = Autotuner(path=path)
    t.entries['apply|float32|64|4x4|0,1|4,4|dev'] = {'kernel': 2, 'tile_orbits': 256,
                                                     'us': 10.0, 'default_us': 12.5}
    t.entries['grad|float32|64|4x4|0,1|4,4|dev'] = {'grad_blocks': 32, 'us': 5.0,
                                                    'default_us': 5.5}
    t.save()
    t2 = Autotuner(path=path)
    assert t2.entries == t.entries
    s = t2.summary()
    assert s['apply'] == {'geometries': 1, 'default_us': 12.5, 'tuned_us': 10.0}
    assert s['grad']['geometries'] == 1
