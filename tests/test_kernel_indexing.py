"""CPU replay of the apply kernels' index arithmetic (qsv_emulate_apply_ptm).

For every geometry the library plans the same kernel path the GPU call would
take (direct, LDS-tiled or table), replays every thread's global and LDS
accesses with bounds checks and exactly-once output coverage, and computes
the result in float64; the result must match the oracle.  This pins the
planners and tile decompositions before any GPU run."""
import itertools

import numpy as np
import pytest

from quantumsim_amd import _native as nat
from oracle.pauli_vector import PauliVectorNumpy
from oracle import bases as ob


def _padded(shape, pads):
    """element strides with pads[q] extra elements per step of axis q"""
    st, acc = [0] * len(shape), 1
    for q in reversed(range(len(shape))):
        st[q] = acc + pads[q]
        acc = st[q] * shape[q]
    return st, acc


def emulate(dtype, pv, ptm, targets, pads_in=None, pads_out=None):
    """Run the CPU replay; with pads, the vectors are laid out with padded
    strides (padding filled with NaN, which must never reach a result)."""
    L = nat.lib()
    k = len(targets)
    dims = pv.shape
    out_shape = list(dims)
    for j, q in enumerate(targets):
        out_shape[q] = ptm.shape[j]
    path = (nat.ctypes.c_int32 * 1)()
    ptm = np.ascontiguousarray(ptm, dtype=np.float64)
    if pads_in is None:
        pv = np.ascontiguousarray(pv, dtype=np.float64)
        out = np.zeros(out_shape)
        rc = L.qsv_emulate_apply_ptm(dtype, pv.ctypes.data_as(nat._c_f64p),
                                     out.ctypes.data_as(nat._c_f64p), len(dims), nat.i64(dims),
                                     None, None, k, nat.i32(targets), nat.i32(ptm.shape[:k]),
                                     ptm.ctypes.data_as(nat._c_f64p), path)
        nat.check(rc, 'emulate')
        return out, path[0]
    si, ni = _padded(dims, pads_in)
    so, no = _padded(out_shape, pads_out)
    buf_in = np.full(ni, np.nan)
    np.lib.stride_tricks.as_strided(buf_in, dims, [8 * x for x in si])[...] = pv
    buf_out = np.full(no, np.nan)
    rc = L.qsv_emulate_apply_ptm(dtype, buf_in.ctypes.data_as(nat._c_f64p),
                                 buf_out.ctypes.data_as(nat._c_f64p), len(dims), nat.i64(dims),
                                 nat.i64(si), nat.i64(so), k, nat.i32(targets),
                                 nat.i32(ptm.shape[:k]), ptm.ctypes.data_as(nat._c_f64p), path)
    nat.check(rc, 'emulate')
    out = np.lib.stride_tricks.as_strided(buf_out, out_shape, [8 * x for x in so]).copy()
    return out, path[0]


def oracle_apply(pv, ptm, targets):
    o = PauliVectorNumpy([ob.Basis(np.eye(1)[None].repeat(d, 0).reshape(d, 1, 1), [''] * d)
                          for d in pv.shape], pv.copy())
    o.apply_ptm(ptm, *targets)
    return o.to_pv()


GEOMS = []
for n in (1, 2, 3, 5, 7):
    for k in (1, 2):
        if k > n:
            continue
        for t in itertools.permutations(range(n), k):
            GEOMS.append(([4] * n, t))
GEOMS += [([2] * 6, (5, 2, 0)), ([2] * 6, (0, 1, 2, 3)), ([2] * 7, (6, 5, 4, 3)),
          ([9, 4, 4, 9], (3,)), ([9, 4, 4, 9], (0, 3)), ([4, 9, 4, 3], (2, 3)),
          ([3, 4, 2, 4, 4], (4, 3)), ([1, 4, 4, 1], (3, 2)), ([4, 4, 4, 4, 4, 4], (4,)),
          # dense 64 x 64 and 256 x 256 PTMs (MFMA path)
          ([4] * 5, (0, 1, 2)), ([4] * 5, (4, 2, 0)), ([4] * 6, (1, 2, 3, 4)),
          ([4] * 4, (3, 0, 1, 2)), ([2] * 8, (7, 0, 3, 1, 5, 2)), ([4, 4, 4, 4, 3], (0, 1, 2)),
          # innermost axis = lowest PTM index: vector accesses; short tails
          ([4] * 5, (2, 3, 4)), ([4] * 6, (2, 3, 4, 5)), ([4] * 3, (0, 1, 2)),
          ([4] * 4, (0, 1, 2, 3)), ([4] * 4, (1, 2, 3)), ([16, 4, 4], (1, 2)),
          ([4, 4, 16], (1, 2)), ([4, 16, 4], (0, 1, 2)), ([9, 4, 4, 4], (1, 2, 3)),
          ([4, 4, 4, 9], (0, 1, 2)), ([4] * 5, (4, 0, 1))]


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets', GEOMS)
def test_emulated_kernels_match_oracle(dims, targets, dtype, rng):
    pv = rng.standard_normal(dims)
    din = [dims[t] for t in targets]
    douts = [din]
    if np.prod(din) <= 16:
        douts += [[max(1, d // 2) for d in din], [min(d + 1, 16) for d in din]]
    for dout in douts:
        if np.prod(dout) > 16 and np.prod(din) <= 16:
            continue
        ptm = rng.standard_normal(list(dout) + din)
        got, path = emulate(dtype, pv, ptm, targets)
        ref = oracle_apply(pv, ptm, targets)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13), (dims, targets, dout, path)


def test_paths_taken():
    """The 14-qubit geometry at CPU-friendly size: inner targets go to the
    LDS-tiled kernel, outer ones to the direct kernel."""
    rng = np.random.default_rng(0)
    pv = rng.standard_normal([4] * 8)
    ptm = rng.standard_normal([4] * 4)
    assert emulate(nat.QSV_F64, pv, ptm, (6, 7))[1] == 2
    assert emulate(nat.QSV_F64, pv, ptm, (7, 0))[1] == 2
    assert emulate(nat.QSV_F64, pv, ptm, (0, 1))[1] == 1
    assert emulate(nat.QSV_F64, pv, ptm, (2, 3))[1] == 1
    big = rng.standard_normal([9] * 4)
    assert emulate(nat.QSV_F64, big, rng.standard_normal([9] * 4), (1, 2))[1] == 3
    # MFMA path; bits 4-7 / 8-11: elements per load / store access along the
    # PTM index, bits 12-15: orbits per lane vector
    assert emulate(nat.QSV_F64, pv, rng.standard_normal([4] * 6), (1, 2, 3))[1] == 4 | 16 | 256 | 2 << 12
    assert emulate(nat.QSV_F32, pv, rng.standard_normal([4] * 6), (1, 2, 3))[1] == 4 | 16 | 256 | 4 << 12
    assert emulate(nat.QSV_F32, pv, rng.standard_normal([4] * 8), (0, 7, 3, 4))[1] == 4 | 16 | 256 | 1 << 12
    assert emulate(nat.QSV_F64, pv, rng.standard_normal([4] * 6), (5, 6, 7))[1] == 4 | 32 | 512 | 1 << 12
    assert emulate(nat.QSV_F32, pv, rng.standard_normal([4] * 8), (4, 5, 6, 7))[1] == 4 | 64 | 1024 | 1 << 12
    assert emulate(nat.QSV_F32, pv, rng.standard_normal([4] * 6), (7, 6, 5))[1] == 4 | 16 | 256 | 1 << 12


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets', [([4] * 6, (0, 1)), ([4] * 6, (5, 4)), ([4] * 6, (2,)),
                                          ([4] * 6, (0, 5)), ([4] * 5, (1, 2, 3)),
                                          ([4] * 5, (4, 0, 2, 1)), ([3, 4, 2, 4], (3, 1)),
                                          ([4] * 5, (2, 3, 4)), ([4] * 6, (1, 3, 4, 5))])
def test_emulated_kernels_padded_layout(dims, targets, dtype, rng):
    """Padded strides on the outer axes (as the backend lays out large
    vectors): same results, no access to padding (NaN-filled), every logical
    output written once."""
    n = len(dims)
    pads = [4 if q < 3 else 0 for q in range(n)]      # multiples of the vector width
    pv = rng.standard_normal(dims)
    din = [dims[t] for t in targets]
    ptm = rng.standard_normal(din + din)
    got, path = emulate(dtype, pv, ptm, targets, pads, pads)
    assert not np.isnan(got).any()
    assert np.allclose(got, oracle_apply(pv, ptm, targets), rtol=1e-13, atol=1e-13), path
    odd = [3 if q < 2 else 0 for q in range(n)]         # not a multiple: scalar accesses
    got, path = emulate(dtype, pv, ptm, targets, odd, [0] * n)
    assert np.allclose(got, oracle_apply(pv, ptm, targets), rtol=1e-13, atol=1e-13), path


def emulate_grad(dtype, x, ybar, targets, dout, grad_blocks=0):
    """Replay of qsv_ptm_grad (one state) on the host: returns (dP, vn)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    ybar = np.ascontiguousarray(ybar, dtype=np.float64)
    k = len(targets)
    din = [x.shape[t] for t in targets]
    out = np.zeros(int(np.prod(dout)) * int(np.prod(din)))
    vn = (nat.ctypes.c_int32 * 1)()
    rc = nat.lib().qsv_emulate_ptm_grad(dtype, x.ctypes.data_as(nat._c_f64p),
                                        ybar.ctypes.data_as(nat._c_f64p), x.ndim,
                                        nat.i64(x.shape), k, nat.i32(targets), nat.i32(dout),
                                        out.ctypes.data_as(nat._c_f64p), grad_blocks, vn)
    nat.check(rc, 'emulate_ptm_grad')
    return out.reshape(list(dout) + din), vn[0]


GRAD_GEOMS = [([4] * 5, (1, 2)), ([4] * 5, (3, 4)), ([4] * 5, (4, 3)), ([4] * 5, (0, 4)),
              ([4] * 6, (2,)), ([4] * 6, (5,)), ([4] * 3, (0, 1)), ([4, 4], (0, 1)),
              ([9, 4, 4, 9], (1, 2)), ([9, 4, 4, 3], (0,)), ([4, 9, 4], (2,)),
              ([2] * 7, (3, 1)), ([4] * 7, (0, 1))]


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets', GRAD_GEOMS)
def test_emulated_grad_matches_oracle(dims, targets, dtype, rng):
    """dP[o, i] = sum over all other indices of ybar[.., o, ..] x[.., i, ..]:
    the oracle's einsum, against the lane-by-lane replay of the MFMA
    gradient kernel (square and rectangular PTMs)."""
    n = len(dims)
    din = [dims[t] for t in targets]
    for dout in (din, [max(1, d - 1) for d in din]):
        oshape = list(dims)
        for t, d in zip(targets, dout):
            oshape[t] = d
        x = rng.standard_normal(dims)
        ybar = rng.standard_normal(oshape)
        got, vn = emulate_grad(dtype, x, ybar, targets, dout)
        xl = list(range(n))
        yl = list(range(n))
        outl = [n + j for j in range(len(targets))] + [t for t in targets]
        for j, t in enumerate(targets):
            yl[t] = n + j
        ref = np.einsum(ybar, yl, x, xl, outl)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (dims, targets, dout, vn)


def test_grad_vector_width():
    rng = np.random.default_rng(3)
    x = rng.standard_normal([4] * 5)
    # innermost axis a bystander: 16-byte orbit vectors
    assert emulate_grad(nat.QSV_F64, x, x, (1, 2), [4, 4])[1] == 2
    assert emulate_grad(nat.QSV_F32, x, x, (1, 2), [4, 4])[1] == 4
    # innermost axis a target: per-orbit lanes
    assert emulate_grad(nat.QSV_F32, x, x, (3, 4), [4, 4])[1] == 1


def emulate_tiled(dtype, pv, ptm, targets, tile):
    """Replay of the batched path's kernel for one state with launch shape
    `tile`; returns (out, path) or None when the tile is infeasible."""
    k = len(targets)
    dims = list(pv.shape)
    dout = list(ptm.shape[:k])
    oshape = list(dims)
    for t, d in zip(targets, dout):
        oshape[t] = d
    out = np.zeros(oshape)
    x = np.ascontiguousarray(pv, dtype=np.float64)
    P = np.ascontiguousarray(ptm, dtype=np.float64)
    path = (nat.ctypes.c_int32 * 1)()
    rc = nat.lib().qsv_emulate_apply_ptm_tiled(dtype, x.ctypes.data_as(nat._c_f64p),
                                               out.ctypes.data_as(nat._c_f64p), len(dims),
                                               nat.i64(dims), k, nat.i32(targets), nat.i32(dout),
                                               P.ctypes.data_as(nat._c_f64p),
                                               nat.ctypes.byref(tile), path)
    if rc == -3:
        return None
    nat.check(rc, 'emulate_apply_ptm_tiled')
    return out, path[0]


TILE_CANDIDATES = [(1, 0), (2, 0), (2, 16), (2, 64), (2, 128), (2, 256), (2, 512), (2, 1024)]


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets', [([4] * 5, (3, 4)), ([4] * 5, (4, 3)), ([4] * 5, (1, 2)),
                                          ([4] * 5, (0, 1)), ([4] * 6, (5,)), ([4] * 6, (2,)),
                                          ([4] * 4, (0, 3)), ([4] * 3, (1, 2))])
def test_emulated_tile_candidates_match_oracle(dims, targets, dtype, rng):
    """Every launch shape the autotuner may pick computes the same result
    (the replay checks each access and the exactly-once coverage)."""
    pv = rng.standard_normal(dims)
    din = [dims[t] for t in targets]
    ptm = rng.standard_normal(din + din)
    ref = oracle_apply(pv, ptm, targets)
    feasible = 0
    for kern, orbits in TILE_CANDIDATES:
        r = emulate_tiled(dtype, pv, ptm, targets, nat.Tile(kern, orbits, 0, 0))
        if r is None:
            assert kern == 2
            continue
        feasible += 1
        out, path = r
        assert path == (2 if kern == 2 else 1)
        assert np.allclose(out, ref, rtol=1e-13, atol=1e-13), (dims, targets, kern, orbits)
    assert feasible >= 1


def test_tile_validation():
    pv = np.zeros([4] * 3)
    ptm = np.zeros([4, 4])
    for bad in (nat.Tile(3, 0, 0, 0), nat.Tile(0, -1, 0, 0), nat.Tile(0, 0, 300, 0),
                nat.Tile(0, 0, 0, 1), nat.Tile(0, 5000, 0, 0)):
        with pytest.raises(ValueError):
            emulate_tiled(nat.QSV_F64, pv, ptm, (1,), bad)


@pytest.mark.parametrize('blocks', [1, 2, 3, 7, 64, 256])
def test_emulated_grad_block_counts(blocks, rng):
    dims, targets = [4] * 6, (2, 3)
    x = rng.standard_normal(dims)
    y = rng.standard_normal(dims)
    ref, _ = emulate_grad(nat.QSV_F64, x, y, targets, [4, 4])
    got, _ = emulate_grad(nat.QSV_F64, x, y, targets, [4, 4], grad_blocks=blocks)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


# ------------------------------------------------ dense 64 x 64 batched paths
@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets', [([4] * 5, (1, 2, 3)), ([4] * 5, (2, 3, 4)),
                                          ([4] * 5, (0, 1, 2)), ([4] * 6, (4, 1, 3)),
                                          ([4] * 3, (0, 1, 2)), ([4] * 4, (3, 2, 1)),
                                          ([2] * 8, (1, 3, 4, 6, 7, 0))])
def test_emulated_batched_dense64_matches_oracle(dims, targets, dtype, rng):
    """The batched path's MFMA kernel for dense 64 x 64 PTMs (row-major PTM
    gathered into fragment order per state), replayed on the CPU."""
    pv = rng.standard_normal(dims)
    din = [dims[t] for t in targets]
    ptm = rng.standard_normal(din + din)
    ref = oracle_apply(pv, ptm, targets)
    out, path = emulate_tiled(dtype, pv, ptm, targets, nat.Tile(0, 0, 0, 0))
    assert path & 7 == 4, path
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12), (dims, targets, path)
    # a launch shape is refused for dense PTMs
    with pytest.raises(ValueError, match='no launch shape'):
        emulate_tiled(dtype, pv, ptm, targets, nat.Tile(2, 64, 0, 0))


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets,dout', [
    ([4] * 5, (1, 2, 3), [4, 4, 4]), ([4] * 5, (2, 3, 4), [4, 4, 4]),
    ([4] * 5, (4, 0, 2), [4, 4, 4]), ([4] * 4, (0, 1, 2), [4, 2, 4]),
    ([9, 4, 4, 9], (0, 1), [9, 4]), ([9, 4, 4, 9], (3, 2), [9, 2]),
    ([4] * 3, (0, 1, 2), [4, 4, 4])])
def test_emulated_grad64_matches_oracle(dims, targets, dout, dtype, rng):
    """The wide (17..64) PTM gradient kernel, replayed: loads in bounds and
    aligned, every orbit once, the MFMA tile maps of both element types."""
    n = len(dims)
    oshape = list(dims)
    for t, d in zip(targets, dout):
        oshape[t] = d
    x = rng.standard_normal(dims)
    ybar = rng.standard_normal(oshape)
    xl = list(range(n))
    yl = list(range(n))
    outl = [n + j for j in range(len(targets))] + [t for t in targets]
    for j, t in enumerate(targets):
        yl[t] = n + j
    ref = np.einsum(ybar, yl, x, xl, outl)
    for blocks in (0, 1, 3, 64):
        got, vn = emulate_grad(dtype, x, ybar, targets, dout, blocks)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (dims, targets, dout, blocks, vn)


def test_grad64_load_modes():
    """Wide gradient: orbit vectors when the innermost axis is a bystander,
    runs of four entries (16-byte loads along the PTM index) when it is a
    target, scalar loads otherwise."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal([4] * 5)
    # float64 orbit vectors go to the LDS-staged kernel (512)
    for dtype, vec in ((nat.QSV_F64, 2 | 512), (nat.QSV_F32, 4)):
        assert emulate_grad(dtype, x, x, (1, 2, 3), [4, 4, 4])[1] == vec
        assert emulate_grad(dtype, x, x, (2, 3, 4), [4, 4, 4])[1] == 1 | 256
        assert emulate_grad(dtype, x, x, (4, 2, 3), [4, 4, 4])[1] == 1
        assert emulate_grad(dtype, x, x, (0, 1, 4), [4, 4, 4])[1] == 1 | 256


def emulate_backward(dtype, x, ybar, ptm, targets, grad_blocks=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ybar = np.ascontiguousarray(ybar, dtype=np.float64)
    P = np.ascontiguousarray(ptm, dtype=np.float64).reshape(64, 64)
    xbar = np.zeros_like(x)
    dptm = np.zeros(64 * 64)
    rc = nat.lib().qsv_emulate_apply_ptm_batched_backward(
        dtype, x.ctypes.data_as(nat._c_f64p), ybar.ctypes.data_as(nat._c_f64p),
        xbar.ctypes.data_as(nat._c_f64p), x.ndim, nat.i64(x.shape), len(targets),
        nat.i32(targets), P.ctypes.data_as(nat._c_f64p), dptm.ctypes.data_as(nat._c_f64p),
        grad_blocks)
    if rc == -3:
        return None
    nat.check(rc, 'emulate_backward')
    return xbar, dptm.reshape(64, 64)


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets', [([4] * 5, (1, 2, 3)), ([4] * 6, (0, 1, 2)),
                                          ([4] * 5, (3, 0, 2)), ([4] * 4, (0, 1, 2)),
                                          ([4] * 6, (4, 1, 2))])
def test_emulated_fused_backward_matches_oracle(dims, targets, dtype, rng):
    """The fused backward of a dense 64 x 64 block (xbar = P^T ybar and dP
    in one pass), replayed: both parts against numpy.  Float32 takes only
    the direct kernel's geometries (whole 32-orbit tiles): QSV_ENOTSUP on
    the others."""
    n = len(dims)
    x = rng.standard_normal(dims)
    ybar = rng.standard_normal(dims)
    P = rng.standard_normal((4,) * 6)
    if dtype == nat.QSV_F32 and not emulate_grad(dtype, x, ybar, targets, [4, 4, 4])[1] & 2048:
        assert emulate_backward(dtype, x, ybar, P, targets) is None
        return
    ref_xbar = oracle_apply(ybar, np.ascontiguousarray(
        P.reshape(64, 64).T).reshape((4,) * 6), targets)
    xl, yl = list(range(n)), list(range(n))
    for j, t in enumerate(targets):
        yl[t] = n + j
    ref_dp = np.einsum(ybar, yl, x, xl, [n + j for j in range(3)] + list(targets)).reshape(64, 64)
    for blocks in (0, 1, 3):
        r = emulate_backward(dtype, x, ybar, P, targets, blocks)
        assert r is not None
        xbar, dp = r
        assert np.allclose(xbar, ref_xbar, rtol=1e-12, atol=1e-12), (dims, targets)
        assert np.allclose(dp, ref_dp, rtol=1e-12, atol=1e-12), (dims, targets, blocks)


@pytest.mark.parametrize('dims,targets', [([4] * 6, (1, 2, 3)), ([4] * 6, (2, 3, 4)),
                                          ([4] * 7, (0, 1, 2)), ([4] * 6, (0, 2, 4)),
                                          ([4] * 6, (3, 1, 0)), ([4, 4, 4, 4, 2, 4], (0, 1, 2))])
def test_emulated_fused_backward_direct_matches_oracle(dims, targets, rng):
    """Float32 fused backward on whole 32-orbit tiles with 16-byte orbit
    vectors (the geometries of the direct-to-LDS gradient, which the fused
    backward then runs as k_bwd64d), replayed: the gradient half as for
    k_ptm_grad64d, the adjoint apply's 8-byte operand reads (entry and
    orbits checked in the current stage), every xbar store in bounds and
    each xbar entry written once; both results against numpy for several
    block counts."""
    n = len(dims)
    x = rng.standard_normal(dims)
    ybar = rng.standard_normal(dims)
    P = rng.standard_normal((4,) * 6)
    assert emulate_grad(nat.QSV_F32, x, ybar, targets, [4, 4, 4])[1] & 2048
    ref_xbar = oracle_apply(ybar, np.ascontiguousarray(
        P.reshape(64, 64).T).reshape((4,) * 6), targets)
    xl, yl = list(range(n)), list(range(n))
    for j, t in enumerate(targets):
        yl[t] = n + j
    ref_dp = np.einsum(ybar, yl, x, xl, [n + j for j in range(3)] + list(targets)).reshape(64, 64)
    for blocks in (0, 1, 3, 64):
        xbar, dp = emulate_backward(nat.QSV_F32, x, ybar, P, targets, blocks)
        assert np.allclose(xbar, ref_xbar, rtol=1e-12, atol=1e-12), (dims, targets, blocks)
        assert np.allclose(dp, ref_dp, rtol=1e-12, atol=1e-12), (dims, targets, blocks)


@pytest.mark.parametrize('dims,targets', [([4] * 6, (3, 4, 5)), ([4] * 6, (0, 2, 5)),
                                          ([4] * 7, (4, 5, 6)), ([4] * 7, (1, 3, 6)),
                                          ([2, 4, 4, 4, 4, 4], (3, 4, 5)),
                                          ([4, 4, 4, 4, 2, 4, 4, 4], (5, 6, 7))])
def test_emulated_fused_backward_quad_matches_oracle(dims, targets, rng):
    """Float32 fused backward of blocks whose last target is the innermost
    axis (entries in contiguous runs of four; k_bwd64q), replayed: every
    direct load in bounds with its four entries contiguous, every operand
    read of both contractions finding its (entry, orbit) in the current
    stage, every xbar store in bounds and each xbar entry written once;
    both results against numpy for several block counts."""
    n = len(dims)
    x = rng.standard_normal(dims)
    ybar = rng.standard_normal(dims)
    P = rng.standard_normal((4,) * 6)
    assert emulate_grad(nat.QSV_F32, x, ybar, targets, [4, 4, 4])[1] == 1 | 256  # quad
    ref_xbar = oracle_apply(ybar, np.ascontiguousarray(
        P.reshape(64, 64).T).reshape((4,) * 6), targets)
    xl, yl = list(range(n)), list(range(n))
    for j, t in enumerate(targets):
        yl[t] = n + j
    ref_dp = np.einsum(ybar, yl, x, xl, [n + j for j in range(3)] + list(targets)).reshape(64, 64)
    for blocks in (0, 1, 3, 64):
        xbar, dp = emulate_backward(nat.QSV_F32, x, ybar, P, targets, blocks)
        assert np.allclose(xbar, ref_xbar, rtol=1e-12, atol=1e-12), (dims, targets, blocks)
        assert np.allclose(dp, ref_dp, rtol=1e-12, atol=1e-12), (dims, targets, blocks)
    # float64 keeps the separate kernels there
    assert emulate_backward(nat.QSV_F64, x, ybar, P, targets) is None


@pytest.mark.parametrize('dims,targets', [([4] * 6, (3, 4, 5)), ([4] * 6, (0, 2, 5)),
                                          ([4] * 7, (4, 5, 6)), ([4] * 7, (1, 3, 6)),
                                          ([2, 4, 4, 4, 4, 4], (3, 4, 5)),
                                          ([4, 4, 4, 4, 2, 4, 4, 4], (5, 6, 7))])
def test_emulated_apply_direct_quad_matches_oracle(dims, targets, rng):
    """Float32 64 x 64 apply of the batched path when the PTM's lowest index
    is the innermost axis (k_apply64q), replayed: every direct load of x in
    bounds with its four entries contiguous, every 16-byte operand read
    finding its (entry, orbit) in the current stage, every y store in
    bounds and each y entry written once; y against the oracle for several
    block counts."""
    x = rng.standard_normal(dims)
    P = rng.standard_normal((4,) * 6)
    ref = oracle_apply(x, P, targets)
    L = nat.lib()
    Pf = np.ascontiguousarray(P.reshape(64, 64))
    xc = np.ascontiguousarray(x)
    for blocks in (0, 1, 3, 64):
        y = np.zeros_like(xc)
        rc = L.qsv_emulate_apply_ptm_direct(nat.QSV_F32, xc.ctypes.data_as(nat._c_f64p),
                                            y.ctypes.data_as(nat._c_f64p), x.ndim,
                                            nat.i64(dims), 3, nat.i32(targets),
                                            Pf.ctypes.data_as(nat._c_f64p), blocks)
        nat.check(rc, 'emulate_apply_ptm_direct')
        assert np.allclose(y, ref, rtol=1e-12, atol=1e-12), (dims, targets, blocks)


def test_apply_direct_not_applicable():
    """Float64, partial tiles or an innermost bystander: QSV_ENOTSUP."""
    L = nat.lib()
    P = np.zeros(64 * 64)
    for dtype, dims, targets in ((nat.QSV_F64, [4] * 6, (3, 4, 5)),
                                 (nat.QSV_F32, [4] * 5, (2, 3, 4)),
                                 (nat.QSV_F32, [4] * 6, (1, 2, 3))):
        x = np.zeros(dims)
        y = np.zeros(dims)
        rc = L.qsv_emulate_apply_ptm_direct(dtype, x.ctypes.data_as(nat._c_f64p),
                                            y.ctypes.data_as(nat._c_f64p), len(dims),
                                            nat.i64(dims), 3, nat.i32(targets),
                                            P.ctypes.data_as(nat._c_f64p), 0)
        assert rc == -3, (dtype, dims, targets)


def test_fused_backward_not_applicable():
    """Partial 32-orbit tiles in float32, an innermost target in float64, or
    a lowest PTM index off the innermost axis: QSV_ENOTSUP, callers fall
    back to the separate kernels."""
    x = np.zeros([4] * 5)
    P = np.zeros((4,) * 6)
    assert emulate_backward(nat.QSV_F32, x, x, P, (2, 3, 4)) is None   # 16 orbits
    assert emulate_backward(nat.QSV_F64, x, x, P, (0, 1, 4)) is None
    x = np.zeros([4] * 6)
    assert emulate_backward(nat.QSV_F32, x, x, P, (5, 3, 4)) is None   # lowest index not innermost


# ------------------------------------------- PTMs above 64 x 64 (qudit pairs)
BIG_GEOMS = [([9, 9, 4], (0, 1), [9, 9]), ([4, 9, 9], (2, 1), [9, 9]),
             ([9, 4, 9], (0, 2), [9, 4]), ([9, 4, 2, 4], (0, 1, 2), [9, 4, 2]),
             ([4, 9, 9, 2], (1, 2), [4, 9])]


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets,dout', BIG_GEOMS)
def test_emulated_batched_table_matches_oracle(dims, targets, dout, dtype, rng):
    """The batched path for PTMs that are neither inline nor dense 64 x 64
    (two qutrits: 81 x 81, qutrit + qubit pairs, rectangular outputs): the
    table kernel with offsets decoded on the device, replayed (the decoded
    offsets are checked against the planner's)."""
    pv = rng.standard_normal(dims)
    din = [dims[t] for t in targets]
    ptm = rng.standard_normal(list(dout) + din)
    ref = oracle_apply(pv, ptm, targets)
    out, path = emulate_tiled(dtype, pv, ptm, targets, nat.Tile(0, 0, 0, 0))
    assert path == 3
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12), (dims, targets, dout)
    with pytest.raises(ValueError, match='no launch shape'):
        emulate_tiled(dtype, pv, ptm, targets, nat.Tile(2, 64, 0, 0))


@pytest.mark.parametrize('dtype', [nat.QSV_F64, nat.QSV_F32])
@pytest.mark.parametrize('dims,targets,dout', BIG_GEOMS)
def test_emulated_big_grad_matches_oracle(dims, targets, dout, dtype, rng):
    """The gradient kernel for PTMs above 64 entries on a side, replayed:
    decoded offsets, loads in bounds, every (entry, orbit) staged once, any
    number of blocks per state; against numpy's einsum."""
    n = len(dims)
    oshape = list(dims)
    for t, d in zip(targets, dout):
        oshape[t] = d
    x = rng.standard_normal(dims)
    ybar = rng.standard_normal(oshape)
    xl = list(range(n))
    yl = list(range(n))
    outl = [n + j for j in range(len(targets))] + [t for t in targets]
    for j, t in enumerate(targets):
        yl[t] = n + j
    ref = np.einsum(ybar, yl, x, xl, outl)
    for blocks in (0, 1, 3, 32):
        got, vn = emulate_grad(dtype, x, ybar, targets, dout, blocks)
        assert vn & 1024, vn
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (dims, targets, dout, blocks)


def test_big_grad_limits():
    """Four qubits (256 x 256 = 65536 entries) exceed the gradient's 8192."""
    x = np.zeros([4] * 5)
    with pytest.raises(ValueError, match='8192'):
        emulate_grad(nat.QSV_F64, x, x, (0, 1, 2, 3), [4, 4, 4, 4])
    with pytest.raises(ValueError, match='grad_blocks'):
        emulate_grad(nat.QSV_F64, np.zeros([9, 9, 2]), np.zeros([9, 9, 2]), (0, 1), [9, 9], 65)


@pytest.mark.parametrize('dims,targets', [([4] * 6, (1, 2, 3)), ([4] * 6, (2, 3, 4)),
                                          ([4] * 7, (0, 1, 2)), ([4] * 6, (0, 2, 4)),
                                          ([4] * 6, (3, 1, 0)), ([4, 4, 4, 4, 2, 4], (0, 1, 2))])
def test_emulated_grad64_direct_matches_oracle(dims, targets, rng):
    """Float32 64 x 64 gradient with 16-byte orbit vectors (the innermost axis
    a bystander) and whole 32-orbit tiles: the direct-to-LDS kernel, replayed
    (every direct load in bounds, its four orbits consecutive, its LDS slot;
    every MFMA operand read finding its (entry, orbit) in the current stage;
    every partial entry written once), for several block counts."""
    n = len(dims)
    x = rng.standard_normal(dims)
    ybar = rng.standard_normal(dims)
    xl = list(range(n))
    yl = list(range(n))
    outl = [n + j for j in range(len(targets))] + [t for t in targets]
    for j, t in enumerate(targets):
        yl[t] = n + j
    ref = np.einsum(ybar, yl, x, xl, outl)
    for blocks in (0, 1, 3, 7, 64):
        got, vn = emulate_grad(nat.QSV_F32, x, ybar, targets, [4, 4, 4], blocks)
        assert vn & 2048 and vn & 0xFF == 4, vn
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (dims, targets, blocks)
    # float64 states keep the register/LDS-staged kernels
    assert not emulate_grad(nat.QSV_F64, x, ybar, targets, [4, 4, 4])[1] & 2048


def test_grad64_direct_needs_whole_tiles(rng):
    """16 orbits (five qubits) or an innermost target: not the direct kernel."""
    x = rng.standard_normal([4] * 5)
    assert not emulate_grad(nat.QSV_F32, x, x, (1, 2, 3), [4, 4, 4])[1] & 2048
    x = rng.standard_normal([4] * 6)
    assert not emulate_grad(nat.QSV_F32, x, x, (3, 4, 5), [4, 4, 4])[1] & 2048
