"""The autotuner on the GPU: every candidate launch shape of the batched
apply and of the PTM gradient gives the default result, the tuned choice is
cached and used, and a tuned, graph-captured circuit batch replays the eager
step (the batcher tests in test_autograd_gpu.py run with tuning on)."""
import numpy as np
import pytest
import torch

from quantumsim_amd import _native as nat
from quantumsim_amd.autotune import (Autotuner, set_tuner, tuning, APPLY_CANDIDATES,
                                     GRAD_CANDIDATES)
from quantumsim_amd.autograd import _apply_raw, _ptm_grad_raw

pytestmark = pytest.mark.gpu
DEV = 'cuda'


@pytest.fixture
def tuner():
    t = Autotuner(path=None, reps=2)
    prev = set_tuner(t)
    yield t
    set_tuner(prev)


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-13), (torch.float32, 1e-6)])
@pytest.mark.parametrize('n,targets', [(6, (4, 5)), (6, (0, 1)), (7, (3,)), (5, (2, 3)),
                                       (4, (0, 3))])
def test_apply_candidates_agree(n, targets, dtype, tol, tuner):
    L = nat.lib()
    rng = np.random.default_rng(0)
    B, shape = 5, (4,) * n
    x = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV)
    D = 4 ** len(targets)
    P = torch.tensor(rng.standard_normal((B, D, D)), dtype=dtype, device=DEV)
    ref = _apply_raw(x, P, shape, targets, (4,) * len(targets))
    scale = ref.abs().max().item()
    ran = 0
    for kern, orbits in APPLY_CANDIDATES:
        tile = nat.Tile(kern, orbits, 0, 0)
        y = torch.empty_like(ref)
        rc = L.qsv_apply_ptm_batched_tiled(nat.dtype_code(dtype), x.data_ptr(), y.data_ptr(), B,
                                           n, nat.i64(shape), len(targets), nat.i32(targets),
                                           nat.i32((4,) * len(targets)), P.data_ptr(), D * D,
                                           nat.ctypes.byref(tile),
                                           torch.cuda.current_stream().cuda_stream)
        if rc == -3:
            continue
        nat.check(rc, 'apply')
        ran += 1
        assert (y - ref).abs().max().item() <= tol * scale, (kern, orbits)
    assert ran >= 1
    with tuning(tuner):
        y2 = _apply_raw(x, P, shape, targets, (4,) * len(targets))
    assert tuner.entries, 'nothing tuned'
    assert (y2 - ref).abs().max().item() <= tol * scale


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-13), (torch.float32, 1e-6)])
@pytest.mark.parametrize('n,targets', [(7, (5, 6)), (7, (1, 2)), (6, (3,))])
def test_grad_candidates_agree(n, targets, dtype, tol, tuner):
    L = nat.lib()
    rng = np.random.default_rng(1)
    B, shape = 3, (4,) * n
    x = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV)
    yb = torch.tensor(rng.standard_normal((B, 4 ** n)), dtype=dtype, device=DEV)
    k = len(targets)
    ref = _ptm_grad_raw(x, yb, shape, targets, (4,) * k)
    scale = ref.abs().max().item()
    ws_bytes = L.qsv_ptm_grad_workspace_bytes(B, n, nat.i64(shape), k, nat.i32(targets),
                                              nat.i32((4,) * k))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    for nb in GRAD_CANDIDATES:
        out = torch.empty_like(ref)
        rc = L.qsv_ptm_grad_tiled(nat.dtype_code(dtype), x.data_ptr(), yb.data_ptr(), B, n,
                                  nat.i64(shape), k, nat.i32(targets), nat.i32((4,) * k),
                                  out.data_ptr(), 0, ws.data_ptr(), ws_bytes,
                                  nat.ctypes.byref(nat.Tile(0, 0, nb, 0)),
                                  torch.cuda.current_stream().cuda_stream)
        nat.check(rc, 'grad')
        assert (out - ref).abs().max().item() <= tol * scale, nb
    with tuning(tuner):
        g2 = _ptm_grad_raw(x, yb, shape, targets, (4,) * k)
    assert any(key.startswith('grad') for key in tuner.entries)
    assert (g2 - ref).abs().max().item() <= tol * scale
