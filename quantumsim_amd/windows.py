"""Dense blocks of brickwork windows built on the device in one launch.

A window is two consecutive layers of the layered noisy circuits of the
batched path (workloads.fused_layer_gates: per layer RX then RZ on every
qubit, CZ on the layer's brickwork bonds, amplitude + phase damping); its
gates are tiled into blocks of up to three qubits
(workloads.brickwork_window_plan).  ``window_blocks`` builds every block of
every window and circuit of a group from the per-qubit factors -- rotation
PTMs [B, 2 nw, n, 4, 4], damping PTMs [B, n, 4, 4] and the CZ PTM -- in ONE
kernel launch (qsv_window_blocks), and its backward maps the blocks'
gradients to those factors in one more (qsv_window_blocks_backward).  The
same products composed with torch ops (workloads.fused_layer_gates +
fusion.fuse_batched) take hundreds of small launches per group, forward and
backward.

No counterpart in the reference, which builds each operation's PTM on the
host (quantumsim/operations/operation.py:132-175, algebra.py:11-30).
"""
import numpy as np
import torch

from . import _native as nat
from .workloads import brickwork_window_plan, layer_targets

__all__ = ['WBLOCK_DTYPE', 'WindowPlan', 'window_blocks', 'emulate']

WBLOCK_DTYPE = np.dtype([('nq', '<i4'), ('qubits', '<i4', 3), ('nf', '<i4'),
                         ('reserved', '<i4'), ('out_off', '<i8'), ('f_kind', '<i4', 4),
                         ('f_layer', '<i4', 4), ('f_slot0', '<i4', 4), ('f_slot1', '<i4', 4)])
assert WBLOCK_DTYPE.itemsize == 96


class WindowPlan:
    """The block table of an n-qubit group with ``n_windows`` windows and
    ``batch`` circuits: per block position its qubits, factors (a lone qubit
    of layer l = kind 1, a bond of layer l = kind 2, in application order)
    and the offset of its [n_windows, batch, D, D] output."""

    def __init__(self, n, n_windows, batch, device=None):
        self.n, self.nw, self.batch = n, n_windows, batch
        ta, tb = layer_targets(n, 0), layer_targets(n, 1)
        gates = ta + tb
        plan = brickwork_window_plan(n)
        self.table = np.zeros(len(plan), WBLOCK_DTYPE)
        self.positions = []              # (block qubits, D, offset)
        off = 0
        for p, (qs, src) in enumerate(plan):
            qs = tuple(qs)
            if list(qs) != sorted(qs) or not 1 <= len(qs) <= 3 or not 1 <= len(src) <= 4:
                raise ValueError('window plan position %d: qubits %s, %d gates'
                                 % (p, qs, len(src)))
            D = 4 ** len(qs)
            row = self.table[p]
            row['nq'] = len(qs)
            row['qubits'][:len(qs)] = qs
            row['nf'] = len(src)
            row['out_off'] = off
            for f, i in enumerate(src):
                t = gates[i]
                row['f_kind'][f] = len(t)
                row['f_layer'][f] = 0 if i < len(ta) else 1
                row['f_slot0'][f] = qs.index(t[0])
                row['f_slot1'][f] = qs.index(t[1]) if len(t) == 2 else 0
            self.positions.append((qs, D, off))
            off += n_windows * batch * D * D
        self.out_elems = off
        self.device = device
        self.table_dev = (torch.from_numpy(self.table.view(np.uint8).copy()).to(device)
                          if device is not None else None)

    def host_ptr(self):
        return self.table.ctypes.data_as(nat.ctypes.c_void_p)


def _stream():
    s = torch.cuda.current_stream().cuda_stream
    return s if s else None


class _WindowBlocks(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rot, damp, keep, plan, dtype, cz):
        rot = rot.detach().to(torch.float64).contiguous()
        damp = damp.detach().to(torch.float64).contiguous()
        out = torch.empty(plan.out_elems, dtype=dtype, device=rot.device)
        rc = nat.lib().qsv_window_blocks(
            nat.dtype_code(dtype), len(plan.table), plan.host_ptr(), plan.table_dev.data_ptr(),
            plan.batch, plan.nw, plan.n, rot.shape[1], rot.data_ptr(), damp.data_ptr(),
            keep.data_ptr(), cz.data_ptr(), out.data_ptr(), plan.out_elems, _stream())
        nat.check(rc, 'window_blocks')
        ctx.plan, ctx.dtype = plan, dtype
        ctx.save_for_backward(rot, damp, keep, cz)
        return tuple(out[off:off + plan.nw * plan.batch * D * D].view(plan.nw, plan.batch, D, D)
                     for _, D, off in plan.positions)

    @staticmethod
    def backward(ctx, *grads):
        rot, damp, keep, cz = ctx.saved_tensors
        plan = ctx.plan
        parts = []
        for g, (_, D, _) in zip(grads, plan.positions):
            n_el = plan.nw * plan.batch * D * D
            parts.append(torch.zeros(n_el, dtype=ctx.dtype, device=rot.device) if g is None
                         else g.to(ctx.dtype).reshape(-1))
        dout = torch.cat(parts)
        drot = torch.empty_like(rot)
        ddp = torch.empty((plan.batch, plan.nw, 2, plan.n, 4, 4), dtype=torch.float64,
                          device=rot.device)
        rc = nat.lib().qsv_window_blocks_backward(
            nat.dtype_code(ctx.dtype), len(plan.table), plan.host_ptr(),
            plan.table_dev.data_ptr(), plan.batch, plan.nw, plan.n, rot.shape[1],
            rot.data_ptr(), damp.data_ptr(), keep.data_ptr(), cz.data_ptr(), dout.data_ptr(),
            plan.out_elems, drot.data_ptr(), ddp.data_ptr(), _stream())
        nat.check(rc, 'window_blocks_backward')
        return drot, ddp.sum(dim=(1, 2)), None, None, None, None


def window_blocks(rot, damp, keep, plan, dtype, cz):
    """Blocks of every window: a tuple over block positions of [n_windows,
    batch, D, D] tensors in ``dtype`` (the layout run_schedule takes),
    differentiable with respect to ``rot`` [batch, >= 2 n_windows, n, 4, 4]
    and ``damp`` [batch, n, 4, 4] (float64).  ``keep`` [batch, n_windows]
    uint8: 0 where layer 2w + 1 is past the circuit's depth (the identity
    instead of its gates); ``cz``: the [16, 16] CZ PTM (float64)."""
    if rot.shape[0] != plan.batch or rot.shape[2] != plan.n or rot.shape[1] < 2 * plan.nw:
        raise ValueError('rotation PTMs of shape %s for %d circuits, %d windows, %d qubits'
                         % (tuple(rot.shape), plan.batch, plan.nw, plan.n))
    if tuple(damp.shape) != (plan.batch, plan.n, 4, 4):
        raise ValueError('damping PTMs of shape %s' % (tuple(damp.shape),))
    if keep.dtype != torch.uint8 or tuple(keep.shape) != (plan.batch, plan.nw) or \
            not keep.is_contiguous():
        raise ValueError('keep must be a contiguous uint8 [batch, n_windows] tensor')
    if cz.dtype != torch.float64 or cz.numel() != 256 or not cz.is_contiguous():
        raise ValueError('cz must be a contiguous float64 16 x 16 PTM')
    return _WindowBlocks.apply(rot, damp, keep, plan, dtype, cz)


def emulate(plan, rot, damp, keep, cz, dout=None):
    """CPU replay of both kernels on numpy float64 arrays: (blocks flat,
    drot or None, ddamp_part or None)."""
    f64 = nat._c_f64p
    rot = np.ascontiguousarray(rot, dtype=np.float64)
    damp = np.ascontiguousarray(damp, dtype=np.float64)
    keep = np.ascontiguousarray(keep, dtype=np.uint8)
    cz = np.ascontiguousarray(cz, dtype=np.float64)
    out = np.zeros(plan.out_elems)
    drot = ddp = None
    if dout is not None:
        dout = np.ascontiguousarray(dout, dtype=np.float64)
        drot = np.full(rot.shape, np.nan)
        ddp = np.full((plan.batch, plan.nw, 2, plan.n, 4, 4), np.nan)
    rc = nat.lib().qsv_emulate_window_blocks(
        len(plan.table), plan.host_ptr(), plan.batch, plan.nw, plan.n, rot.shape[1],
        rot.ctypes.data_as(f64), damp.ctypes.data_as(f64),
        keep.ctypes.data_as(nat.ctypes.c_void_p), cz.ctypes.data_as(f64), out.ctypes.data_as(f64),
        plan.out_elems, None if dout is None else dout.ctypes.data_as(f64),
        None if drot is None else drot.ctypes.data_as(f64),
        None if ddp is None else ddp.ctypes.data_as(f64))
    nat.check(rc, 'emulate_window_blocks')
    return out, drot, ddp
