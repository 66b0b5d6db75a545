"""qsv_ptm_grad_noalloc (partial tiles in the library's static device array,
no allocation) against qsv_ptm_grad with a caller workspace, through the C
ABI: two calls in a row on one stream, the second on a smaller input (the
array keeps the first call's tiles), in both element types, written and
accumulated; and a request above the array's size is refused before any
GPU work."""
import numpy as np
import pytest
import torch

from quantumsim_amd import _native as nat

pytestmark = pytest.mark.gpu


def _grad_ref(L, code, x, y, batch, dims, targets, dout, acc, dptm, stream):
    ws_bytes = nat.check(L.qsv_ptm_grad_workspace_bytes(
        batch, len(dims), nat.i64(dims), len(targets), nat.i32(targets), nat.i32(dout)),
        'ws')
    ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=x.device)
    rc = L.qsv_ptm_grad(code, x.data_ptr(), y.data_ptr(), batch, len(dims), nat.i64(dims),
                        len(targets), nat.i32(targets), nat.i32(dout), dptm.data_ptr(), acc,
                        ws.data_ptr(), ws_bytes, stream)
    nat.check(rc, 'ptm_grad')
    return ws


CASES = [
    # (element type, batch, Pauli dims, targets): the first of each pair is
    # the larger call
    (torch.float32, 6, [4] * 10, [3, 4, 5]),       # 64 x 64, orbit vectors
    (torch.float32, 3, [4] * 7, [1, 2]),           # 16 x 16, smaller input
    (torch.float64, 4, [4] * 9, [6, 7, 8]),        # 64 x 64, innermost target runs
    (torch.float64, 2, [4] * 6, [0]),              # 4 x 4, smaller input
]


@pytest.mark.parametrize('accumulate', [0, 1])
def test_noalloc_matches_workspace_entry_point(accumulate):
    L = nat.lib()
    dev = torch.device('cuda', torch.cuda.current_device())
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream or None
    g = torch.Generator(device='cpu').manual_seed(5)
    for dtype, batch, dims, targets in CASES:
        print('noalloc case', dtype, batch, dims, targets, 'accumulate', accumulate, flush=True)
        code = nat.dtype_code(dtype)
        N = int(np.prod(dims))
        D = 4 ** len(targets)
        dout = [4] * len(targets)
        x = torch.randn(batch, N, generator=g, dtype=torch.float64).to(dev, dtype)
        y = torch.randn(batch, N, generator=g, dtype=torch.float64).to(dev, dtype)
        base = torch.randn(batch, D, D, generator=g, dtype=torch.float64).to(dev, dtype)
        want = base.clone()
        ws = _grad_ref(L, code, x, y, batch, dims, targets, dout, accumulate, want, sp)
        got = base.clone()
        rc = L.qsv_ptm_grad_noalloc(code, x.data_ptr(), y.data_ptr(), batch, len(dims),
                                    nat.i64(dims), len(targets), nat.i32(targets),
                                    nat.i32(dout), got.data_ptr(), accumulate, sp)
        nat.check(rc, 'ptm_grad_noalloc')
        torch.cuda.synchronize()
        del ws
        # same blocks per state when the tiles fit the array: the same sums
        assert torch.equal(got, want), (dtype, dims, targets)
        # and against a float64 torch contraction of the same states
        xs = x.double().reshape(batch, *dims)
        ys = y.double().reshape(batch, *dims)
        k = len(targets)
        rest = [q for q in range(len(dims)) if q not in targets]
        xm = xs.permute(0, *[1 + q for q in rest], *[1 + q for q in targets]).reshape(batch, -1, D)
        ym = ys.permute(0, *[1 + q for q in rest], *[1 + q for q in targets]).reshape(batch, -1, D)
        ref = torch.bmm(ym.transpose(1, 2).contiguous(), xm)
        if accumulate:
            ref = ref + base.double()
        tol = 1e-12 if dtype == torch.float64 else 2e-5
        scale = max(1.0, ref.abs().max().item())
        assert (got.double() - ref).abs().max().item() <= tol * scale


def test_noalloc_refuses_requests_above_its_array():
    L = nat.lib()
    cap = L.qsv_ptm_grad_noalloc_scratch_bytes()
    assert cap == 16 << 20
    dims, targets = [4] * 6, [0, 1, 2]
    batch = cap // (64 * 64 * 8) + 1          # one state more than the smallest tiles fit
    # pointers are never dereferenced: the size check comes first
    rc = L.qsv_ptm_grad_noalloc(nat.QSV_F32, 256, 256, batch, len(dims), nat.i64(dims), 3,
                                nat.i32(targets), nat.i32([4, 4, 4]), 256, 0, None)
    assert rc == -1
    assert b'bytes of scratch' in L.qsv_last_error()
