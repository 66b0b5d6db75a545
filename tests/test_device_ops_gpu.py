"""The device-side operation pipeline on the GPU.

* PTMs handed to PauliVectorHIP.apply_ptm as device tensors
  (qsv_apply_ptm_dev_strided) give bitwise the results of the host-PTM
  path: inline sizes, offset-table sizes and the dense 64 x 64 / 256 x 256
  MFMA kernels (PTM packed into fragment order on the device), on
  contiguous and on padded (12-qubit float64) layouts;
* Kraus-defined operations on qubits and qutrits, defined in other bases
  than the vector's, applied with PTMOperation.__call__: float64 within
  1e-12 of the oracle (numpy PTMs applied to the oracle vector), float32
  within 1e-5, and the caller's bases update;
* gradients through Kraus parameters (PTMOperation.apply_to + the
  differentiable diagonal) against central finite differences (h = 1e-6,
  float64).
"""
import numpy as np
import pytest
import torch

import quantumsim_amd as qa
from quantumsim_amd.bases import general, gell_mann
from quantumsim_amd.autograd import diagonal
from quantumsim_amd.operations.device import PTMOperation
from oracle import algebra as oa, bases as ob
from oracle.pauli_vector import PauliVectorNumpy

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def _ob(bs):
    return [ob.Basis(np.array(b.vectors), list(b.labels)) for b in bs]


def _random_channel(D, n_kraus, rng):
    a = rng.standard_normal((n_kraus * D, D)) + 1j * rng.standard_normal((n_kraus * D, D))
    q, _ = np.linalg.qr(a)
    return q.reshape(n_kraus, D, D)


def _pair(n, dtype, rng):
    pv = rng.standard_normal((4,) * n)
    a = qa.PauliVectorHIP([general(2)] * n, pv, force=True, dtype=dtype)
    b = qa.PauliVectorHIP([general(2)] * n, pv, force=True, dtype=dtype)
    return a, b


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('n,targets,dout', [
    (6, (2,), None), (6, (5, 0), None), (6, (1,), (2,)),           # inline PTMs
    (6, (0, 3, 5), (4, 2, 4)),                                      # offset table (32 x 64)
    (6, (1, 2, 3), None), (6, (5, 4, 3), None), (7, (0, 3, 6), None),  # 64 x 64 MFMA
    (6, (2, 3, 4, 5), None), (6, (0, 2, 4, 5), None),               # 256 x 256 MFMA
])
def test_device_ptm_bitwise_equals_host_ptm(n, targets, dout, dtype):
    rng = np.random.default_rng(n * 100 + len(targets))
    k = len(targets)
    dout = dout or (4,) * k
    P = rng.standard_normal(tuple(dout) + (4,) * k)
    print('device PTM', n, targets, dout, dtype, flush=True)
    a, b = _pair(n, dtype, rng)
    a.apply_ptm(P, *targets)
    b.apply_ptm(torch.tensor(P, device=DEV), *targets)
    assert torch.equal(a.to_torch(), b.to_torch())


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('n,targets', [(6, (1, 2, 3)), (7, (0, 3, 6)), (6, (5, 4, 3)),
                                       (6, (2, 3, 4, 5)), (6, (0, 2, 4, 5))])
def test_dense_host_ptm_c_abi_equals_staged_path(n, targets, dtype):
    """Dense 64 x 64 / 256 x 256 host PTMs: PauliVectorHIP stages them
    through pinned memory and applies them as device PTMs; the C ABI's
    host-PTM entry point (qsv_apply_ptm: fragment-ordered copy into the
    workspace) gives the same bits."""
    from quantumsim_amd import _native as nat
    rng = np.random.default_rng(n * 7 + len(targets))
    k = len(targets)
    P = np.ascontiguousarray(rng.standard_normal((4,) * (2 * k)))
    pv = rng.standard_normal((4,) * n)
    print('host PTM via C ABI', n, targets, dtype, flush=True)
    v = qa.PauliVectorHIP([general(2)] * n, pv, force=True, dtype=dtype)
    v.apply_ptm(P, *targets)
    x = torch.tensor(pv, dtype=dtype, device=DEV).reshape(-1)
    y = torch.empty_like(x)
    L = nat.lib()
    code = nat.dtype_code(dtype)
    ws_bytes = nat.check(L.qsv_apply_ptm_workspace_bytes(code, k, nat.i32([4] * k),
                                                         nat.i32([4] * k)), 'ws')
    ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=DEV)
    rc = L.qsv_apply_ptm(code, x.data_ptr(), y.data_ptr(), n, nat.i64([4] * n), k,
                         nat.i32(targets), nat.i32([4] * k), P.ctypes.data_as(nat._c_f64p),
                         ws.data_ptr(), ws_bytes, torch.cuda.current_stream().cuda_stream)
    nat.check(rc, 'apply_ptm')
    torch.cuda.synchronize()
    assert torch.equal(v.to_torch().reshape(-1), y)


@pytest.mark.parametrize('targets', [(0, 1, 2), (9, 10, 11), (0, 5, 11), (0, 1)])
def test_device_ptm_padded_layout(targets):
    """12 qubits float64 (128 MiB): the outer axes' strides are padded."""
    rng = np.random.default_rng(12)
    k = len(targets)
    P = rng.standard_normal((4,) * (2 * k)) / 4 ** k
    print('padded device PTM', targets, flush=True)
    a, b = _pair(12, torch.float64, rng)
    a.apply_ptm(P, *targets)
    b.apply_ptm(torch.tensor(P, device=DEV), *targets)
    assert torch.equal(a.to_torch(), b.to_torch())


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_kraus_operations_match_oracle(dtype, tol):
    """Qubit and qutrit channels defined in Gell-Mann bases, applied to a
    vector in general bases (converted on the device), vs the oracle."""
    rng = np.random.default_rng(31)
    bases = [general(3), general(2), general(3), general(2)]
    dims = [b.dim_pauli for b in bases]
    pv = rng.standard_normal(dims) / np.sqrt(np.prod(dims))
    v = qa.PauliVectorHIP(bases, pv, dtype=dtype)
    ref = PauliVectorNumpy(_ob(bases), pv.copy())
    ref_bases = list(bases)
    program = [((0, 1), [gell_mann(3), gell_mann(2)]), ((3,), [gell_mann(2)]),
               ((2, 0), [gell_mann(3), general(3)]), ((1, 3), [general(2), gell_mann(2)]),
               ((2,), [gell_mann(3)])]
    for qs, b_def in program:
        D = int(np.prod([b.dim_hilbert for b in b_def]))
        K = _random_channel(D, 2, rng)
        op = PTMOperation.from_kraus(torch.as_tensor(K, device=DEV), b_def)
        op(v, *qs)
        # oracle: the same channel's PTM built and converted by the oracle
        P_def = oa.kraus_to_ptm(K, _ob(b_def), _ob(b_def))
        P_vec = oa.ptm_convert_basis(P_def, _ob(b_def), _ob(b_def),
                                     _ob([ref_bases[q] for q in qs]), _ob(b_def))
        ref.apply_ptm(P_vec, *qs)
        for q, b in zip(qs, b_def):
            ref_bases[q] = b
    assert [b == r for b, r in zip(v.bases, ref_bases)] == [True] * 4
    np.testing.assert_allclose(v.to_pv(), ref.to_pv(), atol=tol, rtol=0)


def test_kraus_operation_reduces_basis_on_device():
    """A channel whose output basis is a subbasis (measurement-like
    projection) shrinks the vector; diagonal as in the oracle."""
    rng = np.random.default_rng(41)
    bases = [general(2)] * 3
    pv = rng.standard_normal((4, 4, 4)) / 8
    v = qa.PauliVectorHIP(bases, pv)
    ref = PauliVectorNumpy(_ob(bases), pv.copy())
    proj = [np.diag([1.0, 0.0]).astype(complex), np.diag([0.0, 1.0]).astype(complex)]
    op = PTMOperation.from_kraus(np.stack(proj), [general(2)], [general(2).subbasis([0, 1])])
    op(v, 1)
    ref.apply_ptm(op.ptm([general(2)], [general(2).subbasis([0, 1])]).numpy(), 1)
    ref.bases[1] = _ob([general(2).subbasis([0, 1])])[0]
    assert v.dim_pauli == (4, 2, 4)
    np.testing.assert_allclose(v.to_pv(), ref.to_pv(), atol=1e-14)
    np.testing.assert_allclose(v.diagonal(), ref.diagonal(), atol=1e-14)


def test_gradient_through_kraus_parameters():
    """Loss = sum_j w_j p_j over the populations of a 4-qubit state after
    parametrised amplitude-damped rotations defined in the Gell-Mann basis
    and a fixed CZ; d loss / d params vs central finite differences."""
    rng = np.random.default_rng(51)
    n, B = 4, 3
    w = torch.tensor(rng.standard_normal((B, 2 ** n)), device=DEV)
    cz = PTMOperation.from_kraus(np.diag([1, 1, 1, -1]).astype(complex), [gell_mann(2)] * 2)

    def kraus(th, g):
        c = torch.cos(th / 2).to(torch.complex128)
        s = torch.sin(th / 2).to(torch.complex128)
        rx = torch.stack([torch.stack([c, -1j * s]), torch.stack([-1j * s, c])])
        one = torch.ones((), dtype=torch.complex128, device=th.device)
        zero = torch.zeros((), dtype=torch.complex128, device=th.device)
        a0 = torch.stack([torch.stack([one, zero]),
                          torch.stack([zero, torch.sqrt(1 - g).to(torch.complex128)])])
        a1 = torch.stack([torch.stack([zero, torch.sqrt(g).to(torch.complex128)]),
                          torch.stack([zero, zero])])
        return torch.stack([a0 @ rx, a1 @ rx])

    def loss(params):
        x = torch.zeros((B, 4 ** n), dtype=torch.float64, device=DEV)
        x[:, 0] = 1.0
        bases = [general(2)] * n
        for layer in range(2):
            for q in range(n):
                th, g = params[layer, q, 0], params[layer, q, 1]
                op = PTMOperation.from_kraus(kraus(th, g), [gell_mann(2)])
                x, bases = op.apply_to(x, bases, (q,))
            for q in range(layer % 2, n - 1, 2):
                x, bases = cz.apply_to(x, bases, (q, q + 1))
        # back to general(2) for the population gather
        for q in range(n):
            ident = PTMOperation(torch.eye(4, dtype=torch.float64, device=DEV), [general(2)])
            x, bases = ident.apply_to(x, bases, (q,))
        return (diagonal(x, bases) * w).sum()

    p0 = np.stack([rng.uniform(0, 2 * np.pi, (2, n)), rng.uniform(0.01, 0.2, (2, n))], -1)
    params = torch.tensor(p0, device=DEV, requires_grad=True)
    L = loss(params)
    L.backward()
    g = params.grad.cpu().numpy()
    h = 1e-6
    fd = np.zeros_like(p0)
    with torch.no_grad():
        for idx in np.ndindex(p0.shape):
            pp, pm = p0.copy(), p0.copy()
            pp[idx] += h
            pm[idx] -= h
            fd[idx] = (loss(torch.tensor(pp, device=DEV)).item()
                       - loss(torch.tensor(pm, device=DEV)).item()) / (2 * h)
    np.testing.assert_allclose(g, fd, atol=1e-7, rtol=1e-6)
