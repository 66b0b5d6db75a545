"""Differentiable, batched Pauli-vector circuits.

States are float tensors of shape [B, N] on the GPU (B circuits or
parameter sets of one qudit shape, N = prod dim_pauli, C order, qudit 0
most significant).  Gates carry per-state PTMs [B, D_out, D_in] (or one
shared PTM [1, D_out, D_in]) built from parameters with torch ops, so
autograd reaches the parameters through them.

* ``apply`` contracts a PTM with the states in one libqsv launch
  (qsv_apply_ptm_batched); its backward is the hand-written adjoint:
  x_bar = P^T y_bar (same kernel, transposed PTM) and
  P_bar = sum over orbits y_bar x^T (qsv_ptm_grad, float64 accumulation).
* ``run_circuit`` applies a gate list and keeps a checkpoint every
  ``checkpoint_every`` gates; backward recomputes each segment from its
  checkpoint and runs the adjoint through it, so memory is
  (#gates / k + k) states instead of #gates states.
* ``diagonal`` gathers the populations (qsv_multitake, float64 out); its
  backward scatters the incoming gradient onto the gathered entries.

No counterpart in the reference, which has no differentiation; the PTM
constructors follow quantumsim/operations/qubits.py:38-90,182-220 and
quantumsim/algebra/algebra.py:11-30.
"""
import math

import numpy as np
import torch

from . import _native as nat
from . import ledger as _ledger
from .autotune import get_tuner
from .bases import general

__all__ = ['apply', 'run_circuit', 'run_schedule', 'capture_step', 'diagonal', 'Gate', 'qubit_basis_tensor', 'kraus_ptm',
           'rx_ptm', 'rz_ptm', 'damping_ptm', 'cz_ptm', 'kron_ptm', 'chain_ptm']


def _stream():
    s = torch.cuda.current_stream().cuda_stream
    return s if s else None


class _Workspace:
    """Scratch of the PTM-gradient partial sums, one buffer per (device,
    stream): launches on one stream run in order and may share it, launches
    on two streams may overlap and must not.  Grown outside graph capture
    (the eager warm-up sizes it)."""
    _bufs = {}

    @classmethod
    def get(cls, device, nbytes):
        key = (device, torch.cuda.current_stream(device).cuda_stream)
        t = cls._bufs.get(key)
        if t is None or t.numel() < nbytes:
            t = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            cls._bufs[key] = t
        return t


def _apply_raw(x, ptm, shape, targets, dims_out, out=None):
    """y[b] = ptm[b or 0] applied to x[b] on `targets` (no autograd); into
    ``out`` (contiguous, x's shape and type) when given."""
    B = x.shape[0]
    n_out = math.prod(shape[q] if q not in targets else dims_out[targets.index(q)]
                      for q in range(len(shape)))
    if out is None:
        y = torch.empty((B, n_out), dtype=x.dtype, device=x.device)
    else:
        assert out.shape == (B, n_out) and out.dtype == x.dtype and out.is_contiguous()
        y = out
    ptm = ptm.to(dtype=x.dtype).contiguous()
    pstride = 0 if ptm.shape[0] == 1 else ptm.shape[1] * ptm.shape[2]
    tile = get_tuner().apply_tile(x.dtype, B, shape, targets, dims_out)
    led = _ledger.active()
    ev = led.begin() if led else None
    rc = nat.lib().qsv_apply_ptm_batched_tiled(
        nat.dtype_code(x.dtype), x.data_ptr(), y.data_ptr(), B, len(shape), nat.i64(shape),
        len(targets), nat.i32(targets), nat.i32(dims_out), ptm.data_ptr(), pstride,
        None if tile is None else nat.ctypes.byref(tile), _stream())
    nat.check(rc, 'apply')
    if led:
        dense = ptm.shape[1] == 64 and ptm.shape[2] == 64
        led.end(ev, _ledger.APPLY_DENSE64 if dense else _ledger.APPLY_SMALL,
                x.element_size() * (B * (x.shape[1] + n_out) + ptm.numel()),
                2 * 64 * B * n_out if dense else 0)
    return y


def _ptm_grad_raw(x, ybar, shape, targets, dims_out, out=None):
    """dL/dP[b] = sum over orbits of ybar[b] x[b]^T for the states x [B, N];
    into ``out`` (contiguous, at least B PTMs of x's type; rows past B are
    left alone) when given."""
    B = x.shape[0]
    L = nat.lib()
    k = len(targets)
    din = math.prod(shape[q] for q in targets)
    dout = math.prod(dims_out)
    ws = nat.check(L.qsv_ptm_grad_workspace_bytes(B, len(shape), nat.i64(shape), k,
                                                  nat.i32(targets), nat.i32(dims_out)),
                   'ptm_grad')
    buf = _Workspace.get(x.device, ws)
    if out is None:
        out = torch.empty((B, dout, din), dtype=x.dtype, device=x.device)
    else:
        assert (out.shape[0] >= B and tuple(out.shape[1:]) == (dout, din)
                and out.dtype == x.dtype and out.is_contiguous())
    tile = get_tuner().grad_tile(x.dtype, B, shape, targets, dims_out)
    led = _ledger.active()
    ev = led.begin() if led else None
    rc = L.qsv_ptm_grad_tiled(nat.dtype_code(x.dtype), x.data_ptr(), ybar.data_ptr(), B,
                              len(shape), nat.i64(shape), k, nat.i32(targets), nat.i32(dims_out),
                              out.data_ptr(), 0, buf.data_ptr(), ws,
                              None if tile is None else nat.ctypes.byref(tile), _stream())
    nat.check(rc, 'ptm_grad')
    if led:
        wide = din > 16 or dout > 16
        led.end(ev, _ledger.GRAD_DENSE64 if wide else _ledger.GRAD_SMALL,
                x.element_size() * B * (x.shape[1] + ybar.shape[1] + dout * din),
                2 * B * x.shape[1] * dout if wide else 0)
    return out


# (dtype, shape, targets) -> whether qsv_apply_ptm_batched_backward takes
# the geometry (learned from its first call; QSV_ENOTSUP otherwise)
_FUSED_GEOMETRY = {}


def _fused_backward_raw(x, gy, ptm, shape, targets, gx_out=None, gp_out=None):
    """(dL/dx, dL/dP) of y = P x for square dense 64 x 64 PTMs in one pass
    (qsv_apply_ptm_batched_backward: x and gy read once, dL/dx written), or
    None where it does not apply (then the transposed apply and the PTM
    gradient kernel do the work).  ``gx_out`` (contiguous [B, N]) and
    ``gp_out`` (contiguous, at least B PTMs) receive the results when
    given.  ptm: [B or 1, 64, 64], contiguous, x's type."""
    key = (x.dtype, tuple(shape), tuple(targets))
    if ptm.shape[-1] != 64 or ptm.shape[-2] != 64 or _FUSED_GEOMETRY.get(key) is False:
        return None
    B = x.shape[0]
    gx = torch.empty_like(gy) if gx_out is None else gx_out
    if any(t.data_ptr() % 16 for t in (x, gy, gx)):
        return None
    L = nat.lib()
    d = [shape[q] for q in targets]
    pstride = 0 if ptm.shape[0] == 1 else 64 * 64
    ws = nat.check(L.qsv_ptm_grad_workspace_bytes(B, len(shape), nat.i64(shape), len(targets),
                                                  nat.i32(targets), nat.i32(d)), 'backward')
    buf = _Workspace.get(x.device, ws)
    gp = torch.empty((B, 64, 64), dtype=x.dtype, device=x.device) if gp_out is None else gp_out
    assert gp.shape[0] >= B and tuple(gp.shape[1:]) == (64, 64) and gp.is_contiguous()
    led = _ledger.active()
    ev = led.begin() if led else None
    rc = L.qsv_apply_ptm_batched_backward(
        nat.dtype_code(x.dtype), x.data_ptr(), gy.data_ptr(), gx.data_ptr(), B, len(shape),
        nat.i64(shape), len(targets), nat.i32(targets), ptm.data_ptr(), pstride, gp.data_ptr(),
        0, buf.data_ptr(), ws, _stream())
    if rc == -3:                           # not applicable: the separate kernels
        _FUSED_GEOMETRY[key] = False
        return None
    nat.check(rc, 'backward')
    _FUSED_GEOMETRY[key] = True
    if led:
        led.end(ev, _ledger.BWD_DENSE64,
                x.element_size() * (B * (x.shape[1] + gy.shape[1] + gx.shape[1] + 64 * 64)
                                    + ptm.numel()),
                2 * 64 * B * (x.shape[1] + gx.shape[1]))
    return gx, gp


def _backward_raw(x, gy, ptm, shape, targets, need_ptm_grad, fused=None):
    """(dL/dx, dL/dP or None) of y = P x on the states x [B, N] for the
    incoming gy [B, N]: the transposed apply and the PTM gradient kernel.

    ``fused`` (default: float32 states) uses the one-pass backward of square
    dense 64 x 64 PTMs instead where it applies (qsv_apply_ptm_batched_backward:
    three state streams instead of four).  For float32 those are the
    direct-to-LDS kernels (innermost axis a bystander, or the PTM's lowest
    index on it), measured on MI355X at 7-26 % less kernel time than the two
    kernels on 11-12-qubit batches; float64 keeps the two kernels by default (its
    one-pass kernel runs at two waves per SIMD and is slower).
    """
    d = [shape[q] for q in targets]
    if fused is None:
        fused = x.dtype == torch.float32
    if fused and need_ptm_grad:
        r = _fused_backward_raw(x, gy, ptm.to(dtype=x.dtype).contiguous(), shape, targets)
        if r is not None:
            return r
    gp = _ptm_grad_raw(x, gy, shape, targets, d) if need_ptm_grad else None
    gx = _apply_raw(gy, ptm.transpose(1, 2), shape, targets, d)
    return gx, gp


class _Apply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ptm, shape, targets, dims_out):
        ctx.save_for_backward(x, ptm)
        ctx.geo = (shape, targets, dims_out)
        return _apply_raw(x.contiguous(), ptm, shape, targets, dims_out)

    @staticmethod
    def backward(ctx, gy):
        x, ptm = ctx.saved_tensors
        shape, targets, dims_out = ctx.geo
        gy = gy.contiguous()
        gx = gp = None
        square = tuple(dims_out) == tuple(shape[q] for q in targets)
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1] and square:
            gx, gp = _backward_raw(x.contiguous(), gy, ptm, shape, targets, True)
            if ptm.shape[0] == 1 and x.shape[0] != 1:
                gp = gp.sum(0, keepdim=True)
            return gx, gp.to(ptm.dtype), None, None, None
        out_shape = tuple(shape[q] if q not in targets else dims_out[targets.index(q)]
                          for q in range(len(shape)))
        if ctx.needs_input_grad[0]:
            gx = _apply_raw(gy, ptm.transpose(1, 2), out_shape, targets,
                            [shape[q] for q in targets])
        if ctx.needs_input_grad[1]:
            gp = _ptm_grad_raw(x.contiguous(), gy, shape, targets, dims_out)
            if ptm.shape[0] == 1 and x.shape[0] != 1:
                gp = gp.sum(0, keepdim=True)
            gp = gp.to(ptm.dtype)
        return gx, gp, None, None, None


def apply(x, ptm, shape, targets, dims_out=None):
    """Differentiable batched PTM application.  x: [B, N]; ptm: [B or 1,
    D_out, D_in] (PTM axis order (out..., in...) flattened)."""
    targets = tuple(int(t) for t in targets)
    if dims_out is None:
        dims_out = tuple(shape[q] for q in targets)
    if ptm.dim() == 2:
        ptm = ptm.unsqueeze(0)
    return _Apply.apply(x, ptm, tuple(shape), targets, tuple(dims_out))


class Gate:
    """One gate of a circuit: square PTMs [B or 1, D, D] on ``targets``."""
    __slots__ = ('ptm', 'targets')

    def __init__(self, ptm, targets):
        self.ptm = ptm if ptm.dim() == 3 else ptm.unsqueeze(0)
        self.targets = tuple(int(t) for t in targets)


class _Circuit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, meta, *ptms):
        shape, targets_list, k = meta
        ckpts = [x0]
        x = x0
        with torch.no_grad():
            for j, (ptm, tg) in enumerate(zip(ptms, targets_list)):
                x = _apply_raw(x, ptm, shape, tg, [shape[q] for q in tg])
                if (j + 1) % k == 0 and j + 1 < len(ptms):
                    ckpts.append(x)
        ctx.meta = meta
        ctx.n_ckpt = len(ckpts)
        ctx.save_for_backward(*ckpts, *ptms)
        return x

    @staticmethod
    def backward(ctx, gy):
        shape, targets_list, k = ctx.meta
        saved = ctx.saved_tensors
        ckpts, ptms = saved[:ctx.n_ckpt], saved[ctx.n_ckpt:]
        G = len(ptms)
        grads = [None] * G
        gy = gy.contiguous()
        for seg in reversed(range(ctx.n_ckpt)):
            lo, hi = seg * k, min(G, (seg + 1) * k)
            xs = [ckpts[seg]]
            for j in range(lo, hi - 1):
                tg = targets_list[j]
                xs.append(_apply_raw(xs[-1], ptms[j], shape, tg, [shape[q] for q in tg]))
            for j in reversed(range(lo, hi)):
                tg = targets_list[j]
                gy, g = _backward_raw(xs[j - lo], gy, ptms[j], shape, tg,
                                      ctx.needs_input_grad[2 + j])
                if g is not None:
                    if ptms[j].shape[0] == 1 and gy.shape[0] != 1:
                        g = g.sum(0, keepdim=True)
                    grads[j] = g.to(ptms[j].dtype)
        return (gy, None, *grads)


def capture_step(fn, warmup=1, stream=None):
    """Record one call of ``fn`` (a forward + backward built from this
    module's operations and torch ops, with no host synchronisation) in a
    HIP graph.  ``fn`` runs ``warmup`` times eagerly on ``stream`` first (a
    new stream of default priority if None), which sizes every buffer and
    scratch area, then the eager blocks are handed back and the call is
    captured.  Returns (graph, stream, outputs): ``graph.replay()`` on
    ``stream`` recomputes ``outputs`` in place."""
    if stream is None:
        stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        for _ in range(max(1, int(warmup))):
            fn()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        outputs = fn()
    torch.cuda.synchronize()
    return graph, stream, outputs


def run_circuit(x0, gates, shape, checkpoint_every=8):
    """Apply ``gates`` (list of Gate, square PTMs) to states x0 [B, N];
    differentiable in x0 and every gate PTM, with checkpointing."""
    ptms = [g.ptm for g in gates]
    meta = (tuple(shape), [g.targets for g in gates], max(1, int(checkpoint_every)))
    return _Circuit.apply(x0, meta, *ptms)


class _Schedule(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, meta, *steps):
        shape, targets, segments, k = meta
        dims = [[shape[q] for q in t] for t in targets]
        finals = torch.empty_like(x0)
        kept, cks = [], []
        x = x0
        with torch.no_grad():
            for active, nxt, gates in segments:
                ck = [x]
                for j, (pos, w) in enumerate(gates):
                    x = _apply_raw(x, steps[pos][w], shape, targets[pos], dims[pos])
                    if (j + 1) % k == 0 and j + 1 < len(gates):
                        ck.append(x)
                kept.append(len(ck))
                cks += ck
                if nxt < active:
                    finals[nxt:active].copy_(x[nxt:])
                x = x[:nxt]
        ctx.meta = meta
        ctx.kept = kept
        ctx.save_for_backward(*cks, *steps)
        return finals

    @staticmethod
    def backward(ctx, gf):
        shape, targets, segments, k = ctx.meta
        dims = [[shape[q] for q in t] for t in targets]
        saved = ctx.saved_tensors
        n_ck = sum(ctx.kept)
        cks, steps = saved[:n_ck], saved[n_ck:]
        need = ctx.needs_input_grad[2:]
        gf = gf.contiguous()
        # every (position, step) PTM is used by exactly one gate: its rows
        # of the running prefix are written in place, the rest stay zero
        grads = [torch.zeros_like(t) if nd else None for t, nd in zip(steps, need)]
        # transposed PTMs for the adjoint applies, one copy per position that
        # needs them (float32 64 x 64 blocks run the one-pass backward)
        steps_t = {}

        def transposed(pos):
            if pos not in steps_t:
                steps_t[pos] = steps[pos].transpose(-1, -2).contiguous()
            return steps_t[pos]
        fused = gf.dtype == torch.float32
        bufs = [torch.empty_like(gf), torch.empty_like(gf)]
        cur = 0
        base = n_ck
        for si in reversed(range(len(segments))):
            active, nxt, gates = segments[si]
            base -= ctx.kept[si]
            ck = cks[base:base + ctx.kept[si]]
            # adjoint of the segment's output: the continuing rows are in
            # place from the later segments; the finished rows come from gf
            if nxt < active:
                bufs[cur][nxt:active].copy_(gf[nxt:active])
            G = len(gates)
            for c in reversed(range(len(ck))):
                lo, hi = c * k, min(G, (c + 1) * k)
                xs = [ck[c]]
                for j in range(lo, hi - 1):
                    pos, w = gates[j]
                    xs.append(_apply_raw(xs[-1], steps[pos][w], shape, targets[pos], dims[pos]))
                for j in reversed(range(lo, hi)):
                    pos, w = gates[j]
                    gy = bufs[cur][:active]
                    done = None
                    if fused and grads[pos] is not None:
                        done = _fused_backward_raw(xs[j - lo], gy, steps[pos][w], shape,
                                                   targets[pos], gx_out=bufs[1 - cur][:active],
                                                   gp_out=grads[pos][w])
                    if done is None:
                        if grads[pos] is not None:
                            _ptm_grad_raw(xs[j - lo], gy, shape, targets[pos], dims[pos],
                                          out=grads[pos][w])
                        _apply_raw(gy, transposed(pos)[w], shape, targets[pos], dims[pos],
                                   out=bufs[1 - cur][:active])
                    cur = 1 - cur
        gx = bufs[cur] if ctx.needs_input_grad[0] else None
        return (gx, None, *grads)


def run_schedule(x0, steps, targets, segments, shape, checkpoint_every=1):
    """Apply a gate schedule to a batch of states whose running set shrinks:
    circuits that share their gate structure but not their depth, sorted by
    decreasing depth, so those still running are a prefix of the batch.

    x0: states [B, N].  steps: per gate position p, the PTMs of every step
    of that position, [n_steps_p, B, D_p, D_p] (contiguous, x0's type; step
    w's PTM of state b is steps[p][w, b]).  targets: the qudits of each
    position.  segments: [(active, continuing, [(p, w), ...])] in order;
    each applies its gates to the first ``active`` states, after which the
    states ``continuing``..``active`` - 1 are final.  Returns the final
    states [B, N]; differentiable in x0 and every steps tensor, with a
    checkpoint every ``checkpoint_every`` gates of a segment.

    Each (p, w) must appear in at most one gate.  Against chaining
    run_circuit calls over slices, the adjoint keeps the running prefix in
    place (two ping-pong buffers; finished states' gradients are copied in
    as the prefix grows) and writes dL/dP straight into one zeroed gradient
    per position, so it makes no per-gate slice gradients, stacks or
    concatenations.
    """
    B = x0.shape[0]
    if x0.dim() != 2 or x0.shape[1] != math.prod(shape) or not x0.is_contiguous():
        raise ValueError('x0 must be contiguous [B, {}]'.format(math.prod(shape)))
    if len(steps) != len(targets):
        raise ValueError('one targets tuple per step position')
    for t, tg in zip(steps, targets):
        D = math.prod(shape[q] for q in tg)
        if (t.dim() != 4 or t.shape[1] != B or tuple(t.shape[2:]) != (D, D)
                or t.dtype != x0.dtype or not t.is_contiguous()):
            raise ValueError('steps tensors must be contiguous [n_steps, {}, D, D] of the '
                             'states\' type'.format(B))
    # the running prefix starts with every state, shrinks from segment to
    # segment and ends empty, so every state's final value is written once
    prev = B
    seen = set()
    for active, cont, gates in segments:
        if active != prev or not 0 <= cont <= active:
            raise ValueError('segments must start with all {} states and hand each '
                             'segment the previous one\'s continuing states'.format(B))
        prev = cont
        for p, w in gates:
            if not (0 <= p < len(steps) and 0 <= w < steps[p].shape[0]):
                raise ValueError('gate ({}, {}) outside the steps tensors'.format(p, w))
            if (p, w) in seen:
                raise ValueError('each (position, step) PTM drives one gate')
            seen.add((p, w))
    if prev != 0:
        raise ValueError('the last segment must leave no state running')
    meta = (tuple(shape), [tuple(int(q) for q in t) for t in targets],
            [(int(a), int(c), [(int(p), int(w)) for p, w in gs]) for a, c, gs in segments],
            max(1, int(checkpoint_every)))
    return _Schedule.apply(x0, meta, *steps)


_DIAG_TABLES = {}


def _scatter_index(shape, take, device):
    """(source flat index, gathered position) of every gathered entry that
    exists in the basis; device tensors made once, outside graph capture."""
    flat = np.zeros([len(t) for t in take], dtype=np.int64)
    valid = np.ones_like(flat, dtype=bool)
    stride = 1
    for q in reversed(range(len(shape))):
        t = np.asarray(take[q])
        sl = [None] * len(shape)
        sl[q] = slice(None)
        flat = flat + (np.where(t < 0, 0, t) * stride)[tuple(sl)]
        valid = valid & (t >= 0)[tuple(sl)]
        stride *= shape[q]
    v = valid.reshape(-1)
    return (torch.from_numpy(flat.reshape(-1)[v]).to(device),
            torch.from_numpy(np.flatnonzero(v)).to(device))


class _Diagonal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, shape, take, offsets, take_dims, idx_dev, scatter):
        B = x.shape[0]
        n = math.prod(take_dims)
        out = torch.empty((B, n), dtype=torch.float64, device=x.device)
        rank = len(shape) + 1
        in_dims = [B] + list(shape)
        out_dims = [B] + list(take_dims)
        rc = nat.lib().qsv_multitake(nat.dtype_code(x.dtype), nat.QSV_F64, x.data_ptr(),
                                     out.data_ptr(), rank, nat.i64(in_dims), nat.i64(out_dims),
                                     idx_dev.data_ptr(), nat.i32(offsets), _stream())
        nat.check(rc, 'diagonal')
        ctx.geo = (shape, x.dtype, B, scatter)
        return out

    @staticmethod
    def backward(ctx, g):
        shape, dtype, B, (src, pos) = ctx.geo
        gx = torch.zeros((B, math.prod(shape)), dtype=dtype, device=g.device)
        gx.index_add_(1, src, g.index_select(1, pos).to(dtype))
        return gx, None, None, None, None, None, None


def diagonal(x, bases):
    """Populations <j|rho_b|j> of every state, float64 [B, prod d_hilbert].
    Only bases whose populations are a gather (general(d) and its subbases)
    are supported here."""
    take = []
    for b in bases:
        c = np.real(np.asarray(b.computational_basis_vectors, dtype=complex))
        row = []
        for r in c:
            nz = np.flatnonzero(r)
            if nz.size == 0:
                row.append(-1)
            elif nz.size == 1 and r[nz[0]] == 1.0:
                row.append(int(nz[0]))
            else:
                raise NotImplementedError('differentiable diagonal needs gather bases')
        take.append(row)
    shape = tuple(b.dim_pauli for b in bases)
    B = x.shape[0]
    key = (tuple(tuple(t) for t in take), B, x.device)
    hit = _DIAG_TABLES.get(key)
    if hit is None:
        # axis 0 of the gather is the batch axis (index list 0..B-1)
        offsets, flat_all = [0], list(range(B))
        for t in take:
            offsets.append(len(flat_all))
            flat_all.extend(t)
        idx_dev = torch.tensor(flat_all, dtype=torch.int32, device=x.device)
        src, pos = _scatter_index(shape, take, x.device)
        # tables are shared by every caller of this geometry, on any stream
        hit = (offsets, shared_constant(idx_dev), (src, shared_constant(pos)))
        _DIAG_TABLES[key] = hit
    offsets, idx_dev, scatter = hit
    return _Diagonal.apply(x.contiguous(), shape, take, offsets, [len(t) for t in take],
                           idx_dev, scatter)


# ----------------------------------------------------------- PTM builders
_BASIS_CACHE = {}


def shared_constant(t):
    """Make a freshly built device constant safe to read from any stream:
    wait until the stream that built it has finished (once per constant;
    never during graph capture, where the eager warm-up has built it)."""
    if t.is_cuda and not torch.cuda.is_current_stream_capturing():
        torch.cuda.current_stream(t.device).synchronize()
    return t


def qubit_basis_tensor(device, dtype=torch.complex128):
    """general(2) basis matrices as a [4, 2, 2] complex tensor."""
    key = (device, dtype)
    t = _BASIS_CACHE.get(key)
    if t is None:
        t = torch.tensor(np.asarray(general(2).vectors), dtype=dtype, device=device)
        _BASIS_CACHE[key] = shared_constant(t)
    return t


def kraus_ptm(kraus, basis):
    """PTM[n, x, y] = sum_k Re tr(P_x K P_y K^dag) for batched Kraus sets
    kraus [n, K, d, d] in the basis [X, d, d] (quantumsim/algebra/algebra.py:11-30)."""
    kp = torch.einsum('nkab,ybc,nkdc->nyad', kraus, basis, kraus.conj())
    return torch.einsum('xda,nyad->nxy', basis, kp).real


def rx_ptm(theta):
    """RX(theta) = [[c, -is], [-is, c]] (quantumsim/operations/qubits.py:38-53)."""
    c = torch.cos(theta / 2).to(torch.complex128)
    s = torch.sin(theta / 2).to(torch.complex128)
    k = torch.stack([torch.stack([c, -1j * s], -1), torch.stack([-1j * s, c], -1)], -2)
    return kraus_ptm(k.unsqueeze(-3), qubit_basis_tensor(theta.device))


def rz_ptm(theta):
    """RZ(theta) = diag(e^{-i theta/2}, e^{i theta/2}) (qubits.py:74-90)."""
    e = torch.exp(-0.5j * theta.to(torch.complex128))
    z = torch.zeros_like(e)
    k = torch.stack([torch.stack([e, z], -1), torch.stack([z, e.conj()], -1)], -2)
    return kraus_ptm(k.unsqueeze(-3), qubit_basis_tensor(theta.device))


def damping_ptm(t1, t2, t):
    """Amplitude then phase damping for a wait t under T1, T2
    (gamma = 1 - e^{-t/T1}, 1/T_phi = 1/T2 - 1/(2 T1),
    lambda = 1 - e^{-2t/T_phi}); Kraus sets of qubits.py:182-214."""
    g = 1 - torch.exp(-t / t1)
    lam = 1 - torch.exp(-2 * t * (1 / t2 - 0.5 / t1))
    one = torch.ones_like(g)
    zero = torch.zeros_like(g)
    cd = lambda m: m.to(torch.complex128)
    ka = torch.stack([
        torch.stack([torch.stack([one, zero], -1), torch.stack([zero, torch.sqrt(1 - g)], -1)], -2),
        torch.stack([torch.stack([zero, torch.sqrt(g)], -1), torch.stack([zero, zero], -1)], -2),
    ], -3)
    kp = torch.stack([
        torch.stack([torch.stack([one, zero], -1), torch.stack([zero, torch.sqrt(1 - lam)], -1)], -2),
        torch.stack([torch.stack([zero, zero], -1), torch.stack([zero, torch.sqrt(lam)], -1)], -2),
    ], -3)
    basis = qubit_basis_tensor(t1.device)
    return kraus_ptm(cd(kp), basis) @ kraus_ptm(cd(ka), basis)


def cz_ptm(device):
    """CZ = cphase(pi) as a [1, 16, 16] PTM (qubits.py:109-123)."""
    key = ('cz', device)
    t = _BASIS_CACHE.get(key)
    if t is None:
        from .operations.qubits import cphase
        t = torch.tensor(cphase(np.pi).ptm.reshape(16, 16), dtype=torch.float64,
                         device=device).unsqueeze(0)
        _BASIS_CACHE[key] = shared_constant(t)
    return t


def kron_ptm(a, b):
    """Two-qubit PTM of independent one-qubit PTMs a [n, 4, 4] (first
    target) and b [n, 4, 4] (second target), PTM index order (out0 out1,
    in0 in1)."""
    return torch.einsum('nac,nbd->nabcd', a, b).reshape(a.shape[0], 16, 16)


def chain_ptm(*ptms):
    """PTM of applying ptms[0], then ptms[1], ... (batched matrix product)."""
    out = ptms[0]
    for p in ptms[1:]:
        out = p @ out
    return out
