"""Persistent whole-circuit kernels for many narrow circuits at once.

``qsv_circuits_forward`` / ``qsv_circuits_adjoint`` (csrc/qsv_circuits.hip)
run every circuit of a table in one launch each, a workgroup per circuit
with its Pauli vector in LDS (widths up to ``max_qubits()``: 7 qubits in
float32 on MI355X).  This module builds the tables and wraps the pair as a
``torch.autograd.Function`` over the PTM pool, so PTMs computed from
parameters by torch (rotation angles, T1/T2) get their gradients through
the adjoint kernel.

Table layout (include/qsv.h, ``qsv_circuit`` / ``qsv_gate``):
  circuit: n, n_gates, gate0, traj (element offset of the n_gates saved
           states), pop (offset of the 2^n float64 populations);
  gate:    k, q0, q1, reserved, ptm (element offset of its 4^k x 4^k PTM).
"""
import math

import numpy as np
import torch

from . import _native as nat
from . import ledger as _ledger

__all__ = ['CIRCUIT_DTYPE', 'GATE_DTYPE', 'CircuitTable', 'max_qubits', 'run', 'emulate']

CIRCUIT_DTYPE = np.dtype([('n', '<i4'), ('n_gates', '<i4'), ('gate0', '<i8'), ('traj', '<i8'),
                          ('pop', '<i8')])
GATE_DTYPE = np.dtype([('k', '<i4'), ('q0', '<i4'), ('q1', '<i4'), ('reserved', '<i4'),
                       ('ptm', '<i8')])
assert CIRCUIT_DTYPE.itemsize == 32 and GATE_DTYPE.itemsize == 24


def max_qubits(dtype=torch.float32):
    """Widest circuit the persistent kernels hold in LDS on this device."""
    return nat.check(nat.lib().qsv_circuits_max_qubits(nat.dtype_code(dtype)), 'circuits')


class CircuitTable:
    """Host + device tables for a list of circuits.

    ``circuits``: list of (n, gates) with gates = [(targets, ptm_offset), ...]
    in application order; ``ptm_elems``: size of the PTM pool they index.
    Every gate must own its PTM slot."""

    def __init__(self, circuits, ptm_elems, device=None):
        ng = sum(len(g) for _, g in circuits)
        self.circ = np.zeros(len(circuits), CIRCUIT_DTYPE)
        self.gates = np.zeros(max(ng, 1), GATE_DTYPE)
        gi = traj = pop = 0
        self.max_n = 1
        for c, (n, gates) in enumerate(circuits):
            if not 1 <= n <= 8:
                raise ValueError('circuit %d: %d qubits' % (c, n))
            self.circ[c] = (n, len(gates), gi, traj, pop)
            for targets, off in gates:
                k = len(targets)
                if k not in (1, 2) or any(not 0 <= q < n for q in targets) or \
                        len(set(targets)) != k:
                    raise ValueError('circuit %d: bad targets %s' % (c, targets))
                D = 4 ** k
                if off < 0 or off + D * D > ptm_elems:
                    raise ValueError('circuit %d: PTM offset %d outside the pool' % (c, off))
                self.gates[gi] = (k, targets[0], targets[1] if k == 2 else 0, 0, off)
                gi += 1
            traj += len(gates) * 4 ** n
            pop += 2 ** n
            self.max_n = max(self.max_n, n)
        # every gate owns its PTM slot: the adjoint writes (not adds) dL/dP
        # there, so slots that overlap would lose gradients
        used = sorted((int(r['ptm']), int(r['ptm']) + 16 ** int(r['k']))
                      for r in self.gates[:gi])
        for (a0, a1), (b0, _) in zip(used[:-1], used[1:]):
            if b0 < a1:
                raise ValueError('PTM slots [%d, %d) and [%d, ...) overlap: every gate '
                                 'needs a slot of its own' % (a0, a1, b0))
        self.n_gates = gi
        self.traj_elems = traj
        self.pop_elems = pop
        self.ptm_elems = ptm_elems
        self.device = device
        if device is not None:
            self.circ_dev = torch.from_numpy(self.circ.view(np.uint8).copy()).to(device)
            self.gates_dev = torch.from_numpy(self.gates.view(np.uint8).copy()).to(device)

    def pop_slices(self):
        """(offset, length) of every circuit's populations."""
        return [(int(r['pop']), 2 ** int(r['n'])) for r in self.circ]


def _stream():
    s = torch.cuda.current_stream().cuda_stream
    return s if s else None


class _Persistent(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pool, table, dtype):
        L = nat.lib()
        code = nat.dtype_code(dtype)
        pool_t = pool.detach().to(dtype).contiguous()
        traj = torch.empty(max(table.traj_elems, 1), dtype=dtype, device=pool.device)
        pops = torch.empty(max(table.pop_elems, 1), dtype=torch.float64, device=pool.device)
        led = _ledger.active()
        ev = led.begin() if led else None
        rc = L.qsv_circuits_forward(code, len(table.circ), table.circ_dev.data_ptr(),
                                    table.gates_dev.data_ptr(), pool_t.data_ptr(),
                                    traj.data_ptr(), pops.data_ptr(), table.max_n, _stream())
        nat.check(rc, 'circuits_forward')
        if led:
            es = traj.element_size()
            led.end(ev, _ledger.PERSISTENT_FWD,
                    es * (table.ptm_elems + table.traj_elems) + 8 * table.pop_elems)
        ctx.table, ctx.dtype = table, dtype
        ctx.save_for_backward(pool_t, traj)
        return pops[:table.pop_elems]

    @staticmethod
    def backward(ctx, dpops):
        pool_t, traj = ctx.saved_tensors
        table = ctx.table
        dpops = dpops.to(torch.float64).contiguous()
        dpool = torch.zeros(pool_t.shape, dtype=torch.float64, device=pool_t.device)
        led = _ledger.active()
        ev = led.begin() if led else None
        rc = nat.lib().qsv_circuits_adjoint(nat.dtype_code(ctx.dtype), len(table.circ),
                                            table.circ_dev.data_ptr(), table.gates_dev.data_ptr(),
                                            pool_t.data_ptr(), traj.data_ptr(), dpops.data_ptr(),
                                            dpool.data_ptr(), table.max_n, _stream())
        nat.check(rc, 'circuits_adjoint')
        if led:
            es = traj.element_size()
            led.end(ev, _ledger.PERSISTENT_ADJ,
                    es * (table.ptm_elems + table.traj_elems)
                    + 8 * (table.pop_elems + table.ptm_elems))
        return dpool, None, None


def run(pool, table, dtype=torch.float32):
    """Float64 populations of every circuit of ``table`` (flat, in table
    order) from the PTM pool ``pool`` (float64 on the device; differentiable)."""
    if pool.numel() != table.ptm_elems:
        raise ValueError('PTM pool has %d elements, the table expects %d'
                         % (pool.numel(), table.ptm_elems))
    return _Persistent.apply(pool, table, dtype)


def emulate(table, pool, dpops=None):
    """CPU replay of both kernels (float64): (populations, dL/dpool or None)."""
    pool = np.ascontiguousarray(pool, dtype=np.float64)
    traj = np.zeros(max(table.traj_elems, 1))
    pops = np.zeros(max(table.pop_elems, 1))
    dpool = np.zeros_like(pool) if dpops is not None else None
    dp = None if dpops is None else np.ascontiguousarray(dpops, dtype=np.float64)
    f64p = nat._c_f64p
    rc = nat.lib().qsv_emulate_circuits(
        len(table.circ), table.circ.ctypes.data_as(nat.ctypes.c_void_p),
        table.gates.ctypes.data_as(nat.ctypes.c_void_p), table.n_gates,
        pool.ctypes.data_as(f64p), pool.size, traj.ctypes.data_as(f64p), traj.size,
        pops.ctypes.data_as(f64p), pops.size,
        None if dp is None else dp.ctypes.data_as(f64p),
        None if dpool is None else dpool.ctypes.data_as(f64p), max(table.max_n, 1))
    nat.check(rc, 'emulate_circuits')
    return pops[:table.pop_elems], dpool
