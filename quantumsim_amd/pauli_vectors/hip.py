"""PauliVectorHIP: the Pauli-vector backend on one MI355X.

Drop-in for quantumsim's ``PauliVectorNumpy`` / ``PauliVectorCuda``
(quantumsim/pauli_vectors/numpy.py:8-120, cuda.py:55-425): same constructor,
methods, return types, error types and bases-list semantics.  The state is
a flat, C-order torch tensor on the GPU (float64 by default, float32 on
request); every operation on it runs in libqsv's HIP kernels, launched on
torch's current stream.  Readouts (diagonal, trace, meas_prob) are computed
and returned in float64 whatever the state's element type.
"""
import math
import os
import warnings
from collections import OrderedDict

import numpy as np
import torch

from .. import _native as nat
from .layout import physical_strides
from .pauli_vector import PauliVectorBase

__all__ = ['PauliVectorHIP']


def _resolve_dtype(dtype):
    if dtype is None:
        dtype = os.environ.get('QSV_DTYPE', 'float64')
    if isinstance(dtype, str):
        dtype = {'float64': torch.float64, 'f64': torch.float64, 'double': torch.float64,
                 'float32': torch.float32, 'f32': torch.float32,
                 'float': torch.float32}.get(dtype, dtype)
    if dtype in (np.float64, np.dtype(np.float64)):
        dtype = torch.float64
    if dtype in (np.float32, np.dtype(np.float32)):
        dtype = torch.float32
    if dtype not in (torch.float64, torch.float32):
        raise ValueError('PauliVectorHIP stores float64 or float32, got {}'.format(dtype))
    return dtype


def _resolve_device(device):
    if not torch.cuda.is_available():
        raise RuntimeError('PauliVectorHIP needs a ROCm GPU; none is visible')
    if device is None:
        return torch.device('cuda', torch.cuda.current_device())
    device = torch.device(device)
    if device.type != 'cuda':
        raise ValueError('PauliVectorHIP lives on a GPU, got device {}'.format(device))
    if device.index is None:
        device = torch.device('cuda', torch.cuda.current_device())
    return device


def _stream(device):
    return ctypes_void(torch.cuda.current_stream(device).cuda_stream)


def ctypes_void(v):
    return v if v else None


class _Scratch:
    """Grow-only device buffers (PTM tables, partial sums), one per device,
    stream and purpose.

    Calls on one stream use a buffer in stream order; calls issued on
    different streams get different buffers, so they cannot overwrite a
    table another stream's kernel is still reading, and a buffer that grows
    is handed back to the allocator on the stream it was used on."""

    _bufs = {}

    @classmethod
    def get(cls, device, purpose, nbytes):
        key = (device, torch.cuda.current_stream(device).cuda_stream, purpose)
        t = cls._bufs.get(key)
        if t is None or t.numel() < nbytes:
            t = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
            cls._bufs[key] = t
        return t


class _BasisPlan:
    """How one basis enters the diagonal (numpy.py:66-80) and the traces
    (numpy.py:86-107): ``take`` lists the Pauli indices gathered for that
    axis (-1 = a level the basis cannot represent, value 0); ``transform``
    is None when the gather alone gives the populations, else the
    d_hilbert x len(take) matrix C[j, x] = Re <j|P_x|j> applied after it.
    ``tvec`` is t[x] = Re tr P_x, the row vector that traces the axis out."""

    def __init__(self, basis):
        c = np.real(np.asarray(basis.computational_basis_vectors, dtype=complex))
        self.dim_hilbert = c.shape[0]
        self.dim_pauli = c.shape[1]
        unit = True
        take = []
        for row in c:
            nz = np.flatnonzero(row)
            if nz.size == 0:
                take.append(-1)
            elif nz.size == 1 and row[nz[0]] == 1.0:
                take.append(int(nz[0]))
            else:
                unit = False
                break
        if unit:
            self.take = take
            self.transform = None
        else:
            cols = [int(x) for x in np.flatnonzero(np.any(c != 0, axis=0))]
            self.take = cols
            self.transform = np.ascontiguousarray(c[:, cols])
        self.tvec = np.ascontiguousarray(c.sum(axis=0))


class _DiagPlan:
    def __init__(self, bases, device):
        self.axes = [_basis_plan(b) for b in bases]
        flat, offsets = [], []
        for a in self.axes:
            offsets.append(len(flat))
            flat.extend(a.take)
        self.offsets = offsets
        self.idx = torch.tensor(flat if flat else [0], dtype=torch.int32, device=device)
        self.take_dims = [len(a.take) for a in self.axes]
        self.dims_hilbert = [a.dim_hilbert for a in self.axes]


_basis_plans = OrderedDict()


def _basis_plan(basis):
    key = id(basis)
    hit = _basis_plans.get(key)
    if hit is not None and hit[0] is basis:
        _basis_plans.move_to_end(key)
        return hit[1]
    plan = _BasisPlan(basis)
    _basis_plans[key] = (basis, plan)
    if len(_basis_plans) > 512:
        _basis_plans.popitem(last=False)
    return plan


class PauliVectorHIP(PauliVectorBase):
    """Pauli vector on an AMD GPU.

    Parameters
    ----------
    bases : list of PauliBasis
    pv : numpy array, torch tensor or None
        Expansion of shape ``dim_pauli`` (float16/32/64).  ``None`` sets index
        (0, ..., 0) to 1, as quantumsim/pauli_vectors/numpy.py:39-41 does.
    force : bool
        Lift the 2**22-element size guard.
    dtype : float64 (default) or float32
        Element type of the state on the device.  ``QSV_DTYPE`` in the
        environment sets the default.
    device : torch device, default the current GPU
    """

    _diag_plans = OrderedDict()

    def __init__(self, bases, pv=None, *, force=False, dtype=None, device=None):
        super().__init__(bases, pv, force=force)
        self._dtype = _resolve_dtype(dtype)
        shape = self.dim_pauli
        if pv is not None:
            pshape = tuple(getattr(pv, 'shape', ()))
            if isinstance(pv, (np.ndarray, torch.Tensor)) and shape != pshape:
                raise ValueError(
                    '`bases` Pauli dimensionality should be the same as the '
                    'shape of `data` array.\n'
                    ' - bases shapes: {}\n - data shape: {}'.format(shape, pshape))
            if isinstance(pv, np.ndarray) and pv.dtype not in (np.float16, np.float32,
                                                               np.float64):
                raise ValueError('`pv` must have floating point data type, got {}'
                                 .format(pv.dtype))
            if isinstance(pv, torch.Tensor) and pv.dtype not in (torch.float16, torch.float32,
                                                                 torch.float64):
                raise ValueError('`pv` must have floating point data type, got {}'
                                 .format(pv.dtype))
        if pv is not None and not isinstance(pv, (np.ndarray, torch.Tensor)):
            raise ValueError('`pv` should be Numpy array, torch tensor or None, got type `{}`'
                             .format(type(pv)))
        self._device = _resolve_device(device)
        self._shape = tuple(shape)
        self._pstr, extent = physical_strides(self._shape, self._elem_bytes())
        store = torch.zeros(extent, dtype=self._dtype, device=self._device)
        view = torch.as_strided(store, self._shape, self._pstr)
        if isinstance(pv, np.ndarray):
            view.copy_(torch.from_numpy(np.ascontiguousarray(pv)).to(
                device=self._device, dtype=self._dtype))
        elif isinstance(pv, torch.Tensor):
            view.copy_(pv.detach().to(device=self._device, dtype=self._dtype))
        else:
            view[(0,) * len(self._shape)] = 1
        self._store = store           # owning storage of the state
        self._data = store            # physical (padded) extent
        self._spare = None            # second buffer (apply_ptm writes there)

    # ------------------------------------------------------------ plumbing
    @property
    def dtype(self):
        return self._dtype

    @property
    def device(self):
        return self._device

    @property
    def data_shape(self):
        """Shape of the stored expansion (equals dim_pauli unless an
        apply_ptm changed a dimension and the caller has not yet updated
        ``bases``, see quantumsim/operations/operation.py:549-551)."""
        return self._shape

    def to_torch(self):
        """The device tensor of the expansion (a strided view, not a copy)."""
        return torch.as_strided(self._store, self._shape, self._pstr)

    def _code(self):
        return nat.dtype_code(self._dtype)

    def _elem_bytes(self):
        return 8 if self._dtype == torch.float64 else 4

    def _used_here(self, *tensors):
        """Mark the vector's own buffers as used by work just queued on the
        current stream.  A buffer allocated on another stream would otherwise
        go back to that stream's pool when it is dropped (a spare replaced by
        a larger one, the vector deleted) and could be handed out again while
        this stream's kernels still read or write it; the caching allocator
        ignores the mark for a buffer's own allocation stream."""
        cur = torch.cuda.current_stream(self._device)
        for t in tensors:
            if t is not None:
                t.record_stream(cur)

    def _take_spare(self, n):
        sp = self._spare
        if sp is None or sp.numel() < n:
            sp = torch.empty(max(n, 1), dtype=self._dtype, device=self._device)
        self._spare = None
        return sp

    def _swap_in(self, store, extent, shape, strides):
        self._spare = self._store
        self._store = store
        self._data = store[:extent]
        self._shape = tuple(shape)
        self._pstr = tuple(strides)

    # ----------------------------------------------------------------- API
    def to_pv(self):
        return self.to_torch().contiguous().cpu().numpy()

    def apply_ptm(self, ptm, *qubits):
        """Contract ``ptm`` (shape (out..., in...)) with the listed qudits
        (quantumsim/pauli_vectors/numpy.py:50-64).  ``bases`` is left alone;
        the caller updates it (quantumsim/operations/operation.py:549-551)."""
        # a PTM already on this vector's device stays there (no host copy,
        # no synchronisation): the device-side operation pipeline
        on_device = (isinstance(ptm, torch.Tensor) and ptm.device.type == 'cuda'
                     and ptm.device == self._device)
        if isinstance(ptm, torch.Tensor) and not on_device:
            ptm = ptm.detach().cpu().numpy()
        if not on_device:
            ptm = np.asarray(ptm)
        k = len(qubits)
        if ptm.ndim != 2 * k:
            raise ValueError('{}-qubit PTM must have {} dimensions, got {}'
                             .format(k, 2 * k, ptm.ndim))
        if (ptm.is_complex() if on_device else np.iscomplexobj(ptm)):
            raise ValueError('PTM must be real, got dtype {}'.format(ptm.dtype))
        if k == 0:
            raise ValueError('apply_ptm needs at least one qubit')
        qubits = tuple(int(q) for q in qubits)
        for q in qubits:
            self._validate_qubit(q, 'qubit')
        if len(set(qubits)) != k:
            raise ValueError('qubits {} are not distinct'.format(qubits))
        if k > 8:
            raise NotImplementedError('Applying {}-qubit PTM is not implemented in the '
                                      'active backend.'.format(k))
        dims_out_t = ptm.shape[:k]
        dims_in_t = ptm.shape[k:]
        for q, d in zip(qubits, dims_in_t):
            if self._shape[q] != d:
                raise ValueError('PTM input dimension {} does not match dimension {} of '
                                 'qubit {}'.format(d, self._shape[q], q))
        if not on_device and math.prod(dims_in_t) in (64, 256) and dims_in_t == dims_out_t:
            # dense 3-/4-qubit PTMs from the host: staged through pinned
            # memory and copied asynchronously, then applied as a device PTM
            # (the MFMA kernels gather or pack it on the device), so the call
            # neither blocks on a pageable copy nor copies a host table
            # (torch's pinned allocator keeps the staging block until the
            # copy has run)
            host_np = np.ascontiguousarray(
                ptm, dtype=np.float64 if self._dtype == torch.float64 else np.float32)
            ptm = torch.from_numpy(host_np).pin_memory().to(self._device, non_blocking=True)
            on_device = True
        if on_device:
            ptm = ptm.detach().to(self._dtype).contiguous()
        else:
            ptm = np.ascontiguousarray(ptm, dtype=np.float64)
        new_shape = list(self._shape)
        for q, d in zip(qubits, dims_out_t):
            new_shape[q] = d
        new_str, extent = physical_strides(new_shape, self._elem_bytes())
        out = self._take_spare(extent)
        code = self._code()
        L = nat.lib()
        ws_bytes = L.qsv_apply_ptm_workspace_bytes(code, k, nat.i32(dims_in_t),
                                                   nat.i32(dims_out_t))
        nat.check(ws_bytes, 'apply_ptm')
        ws = _Scratch.get(self._device, 'ptm', ws_bytes).data_ptr() if ws_bytes else None
        if on_device:
            rc = L.qsv_apply_ptm_dev_strided(
                code, self._store.data_ptr(), out.data_ptr(), len(self._shape),
                nat.i64(self._shape), nat.i64(self._pstr), nat.i64(new_str), k,
                nat.i32(qubits), nat.i32(dims_out_t), ptm.data_ptr(), ws, ws_bytes,
                _stream(self._device))
        else:
            rc = L.qsv_apply_ptm_strided(
                code, self._store.data_ptr(), out.data_ptr(), len(self._shape),
                nat.i64(self._shape), nat.i64(self._pstr), nat.i64(new_str), k,
                nat.i32(qubits), nat.i32(dims_out_t), ptm.ctypes.data_as(nat._c_f64p), ws,
                ws_bytes, _stream(self._device))
        if rc != 0:
            self._spare = out
            nat.check(rc, 'apply_ptm')
        self._used_here(self._store, out)
        self._swap_in(out, extent, new_shape, new_str)

    # ------------------------------------------------------------ readouts
    def _diag_plan(self):
        key = tuple(id(b) for b in self.bases) + (self._device,)
        hit = self._diag_plans.get(key)
        if hit is not None and all(a is b for a, b in zip(hit[0], self.bases)):
            self._diag_plans.move_to_end(key)
            return hit[1]
        plan = _DiagPlan(self.bases, self._device)
        self._diag_plans[key] = (list(self.bases), plan)
        if len(self._diag_plans) > 256:
            self._diag_plans.popitem(last=False)
        return plan

    def _check_bases_match(self):
        if self.dim_pauli != self._shape:
            raise ValueError('`bases` Pauli dimensionality {} does not match the data '
                             'shape {}'.format(self.dim_pauli, self._shape))

    def _diagonal_dev(self):
        """float64 device tensor of the populations, shape dim_hilbert."""
        self._check_bases_match()
        plan = self._diag_plan()
        rank = len(self._shape)
        L = nat.lib()
        s = _stream(self._device)
        n_take = math.prod(plan.take_dims)
        cur = torch.empty(max(n_take, 1), dtype=torch.float64, device=self._device)
        rc = L.qsv_multitake_strided(self._code(), nat.QSV_F64, self._store.data_ptr(),
                                     cur.data_ptr(), rank, nat.i64(self._shape),
                                     nat.i64(self._pstr), nat.i64(plan.take_dims),
                                     plan.idx.data_ptr(), nat.i32(plan.offsets), s)
        nat.check(rc, 'diagonal')
        self._used_here(self._store)
        cur = cur[:n_take]
        dims = list(plan.take_dims)
        for q, a in enumerate(plan.axes):
            if a.transform is None:
                continue
            nxt_dims = list(dims)
            nxt_dims[q] = a.dim_hilbert
            nxt = torch.empty(max(math.prod(nxt_dims), 1), dtype=torch.float64,
                              device=self._device)
            tr = a.transform
            ws_bytes = L.qsv_apply_ptm_workspace_bytes(
                nat.QSV_F64, 1, nat.i32([tr.shape[1]]), nat.i32([tr.shape[0]]))
            nat.check(ws_bytes, 'diagonal')
            ws = _Scratch.get(self._device, 'ptm', ws_bytes).data_ptr() if ws_bytes else None
            rc = L.qsv_apply_ptm(nat.QSV_F64, cur.data_ptr(), nxt.data_ptr(), rank,
                                 nat.i64(dims), 1, nat.i32([q]), nat.i32([tr.shape[0]]),
                                 tr.ctypes.data_as(nat._c_f64p), ws, ws_bytes, s)
            nat.check(rc, 'diagonal')
            cur, dims = nxt[:math.prod(nxt_dims)], nxt_dims
        return cur.view(tuple(plan.dims_hilbert))

    def diagonal(self, *, get_data=True):
        """Populations <j|rho|j>, flat, length prod(dim_hilbert), float64
        (quantumsim/pauli_vectors/numpy.py:66-80); levels missing from a
        subbasis read 0.  ``get_data=False`` returns the same flat vector
        as a device tensor (no host copy)."""
        d = self._diagonal_dev().reshape(-1)
        if not get_data:
            return d
        return d.cpu().numpy()

    def multitake(self, *indices, get_data=True):
        """``to_pv()[np.ix_(*indices)]``, one index list per qudit, gathered
        on the device bit-exactly (the reference's multitake kernel,
        quantumsim/pauli_vectors/primitives.cu:92-130).  Returns float64 (a
        device tensor with ``get_data=False``)."""
        if len(indices) != len(self._shape):
            raise ValueError('multitake needs one index list per qudit ({}), got {}'
                             .format(len(self._shape), len(indices)))
        lists = []
        for q, idx in enumerate(indices):
            idx = [int(i) for i in np.asarray(idx).reshape(-1)]
            for i in idx:
                if i < 0 or i >= self._shape[q]:
                    raise ValueError('index {} out of range for qudit {} of dimension {}'
                                     .format(i, q, self._shape[q]))
            lists.append(idx)
        flat, offsets = [], []
        for idx in lists:
            offsets.append(len(flat))
            flat.extend(idx)
        out_dims = [len(idx) for idx in lists]
        n_out = math.prod(out_dims)
        out = torch.empty(max(n_out, 1), dtype=torch.float64, device=self._device)
        table = torch.tensor(flat if flat else [0], dtype=torch.int32, device=self._device)
        rc = nat.lib().qsv_multitake_strided(
            self._code(), nat.QSV_F64, self._store.data_ptr(), out.data_ptr(),
            len(self._shape), nat.i64(self._shape), nat.i64(self._pstr), nat.i64(out_dims),
            table.data_ptr(), nat.i32(offsets), _stream(self._device))
        nat.check(rc, 'multitake')
        self._used_here(self._store)
        res = out[:n_out].view(tuple(out_dims))
        return res.cpu().numpy() if get_data else res

    def _axis_sum(self, t, n_outer, d, n_inner):
        L = nat.lib()
        ws_bytes = nat.check(L.qsv_axis_sum_workspace_bytes(n_outer, d, n_inner), 'sum')
        ws = _Scratch.get(self._device, 'sum', ws_bytes)
        out = torch.empty(d, dtype=torch.float64, device=self._device)
        rc = L.qsv_axis_sum(nat.dtype_code(t.dtype), t.data_ptr(), n_outer, d, n_inner,
                            out.data_ptr(), ws.data_ptr(), ws_bytes, _stream(self._device))
        nat.check(rc, 'sum')
        return out

    def trace(self):
        """sum of the diagonal, float64 (numpy.py:82-84)."""
        d = self._diagonal_dev()
        return np.float64(self._axis_sum(d, 1, 1, d.numel()).item())

    def meas_prob(self, qubit):
        """Populations of the levels of one qudit, float64 ndarray of length
        d_hilbert (numpy.py:97-107)."""
        self._validate_qubit(qubit, 'qubit')
        d = self._diagonal_dev()
        dh = self.dim_hilbert
        n_outer = math.prod(dh[:qubit])
        n_inner = math.prod(dh[qubit + 1:])
        return self._axis_sum(d, n_outer, dh[qubit], n_inner).cpu().numpy()

    def partial_trace(self, *qubits):
        """Trace out every qudit not listed; a new PauliVectorHIP whose data
        axes keep ascending qudit order while ``bases`` follow the listed
        order (numpy.py:86-95)."""
        for q in qubits:
            self._validate_qubit(q, 'qubit')
        self._check_bases_match()
        traced = [i for i in range(self.n_qubits) if i not in qubits]
        cur, shape, strides = self._store, list(self._shape), self._pstr
        L = nat.lib()
        s = _stream(self._device)
        code = self._code()
        for q in reversed(traced):
            t = _basis_plan(self.bases[q]).tvec.reshape(1, -1)
            new_shape = list(shape)
            new_shape[q] = 1
            nxt = torch.empty(max(math.prod(new_shape), 1), dtype=self._dtype,
                              device=self._device)
            ws_bytes = nat.check(L.qsv_apply_ptm_workspace_bytes(
                code, 1, nat.i32([shape[q]]), nat.i32([1])), 'partial_trace')
            ws = _Scratch.get(self._device, 'ptm', ws_bytes).data_ptr() if ws_bytes else None
            rc = L.qsv_apply_ptm_strided(
                code, cur.data_ptr(), nxt.data_ptr(), len(shape), nat.i64(shape),
                nat.i64(strides) if strides is not None else None, None, 1, nat.i32([q]),
                nat.i32([1]), np.ascontiguousarray(t).ctypes.data_as(nat._c_f64p), ws,
                ws_bytes, s)
            nat.check(rc, 'partial_trace')
            self._used_here(cur)
            cur, shape, strides = nxt[:math.prod(new_shape)], new_shape, None
        kept_shape = tuple(self._shape[i] for i in range(self.n_qubits) if i in qubits)
        data = cur.reshape(kept_shape) if traced else self.to_torch()
        return self.__class__([self.bases[q] for q in qubits], data, force=True,
                              dtype=self._dtype, device=self._device)

    def renormalize(self):
        """Scale to unit trace when the trace exceeds 1e-8, else warn; returns
        the trace before scaling (numpy.py:109-117)."""
        tr = self.trace()
        if tr > 1e-8:
            rc = nat.lib().qsv_scale(self._code(), self._data.data_ptr(), self._data.numel(),
                                     float(tr ** -1), _stream(self._device))
            nat.check(rc, 'renormalize')
            self._used_here(self._store)
        else:
            warnings.warn(
                "Density matrix trace is 0; likely your further computation "
                "will fail. Have you projected DM on a state with zero weight?")
        return tr

    @classmethod
    def from_dm(cls, dm, bases, *, force=False, dtype=None, device=None):
        """From a density matrix, converted on the device
        (quantumsim/pauli_vectors/pauli_vector.py:48-53 with
        quantumsim/algebra/algebra.py:44-53; the qudits may have different
        Hilbert dimensions)."""
        from ..algebra import device as dalg
        if not hasattr(bases, '__iter__'):
            n_qubits = len(dm) // bases.dim_hilbert      # the reference's own rule
            bases = [bases] * n_qubits
        bases = list(bases)
        dev = _resolve_device(device)
        return cls(bases, dalg.dm_to_pv(dm, bases, device=dev), force=force, dtype=dtype,
                   device=dev)

    def to_dm(self):
        """Density matrix (numpy complex), converted on the device
        (quantumsim/pauli_vectors/pauli_vector.py:55-56, algebra.py:56-62)."""
        from ..algebra import device as dalg
        self._check_bases_match()
        return dalg.pv_to_dm(self.to_torch(), self.bases, device=self._device).cpu().numpy()

    def exp_value(self, operator):
        """tr(O rho) on the device, as quantumsim/states/state.py:43-83
        computes it for a State: ``operator`` a string of I, X, Y, Z or an
        operator matrix; returns a complex number."""
        from ..algebra import device as dalg
        self._check_bases_match()
        return dalg.exp_value(self.to_torch(), self.bases, operator, device=self._device)

    def copy(self):
        """Deep copy on the same device (the bases list is new, the basis
        objects are shared, as in numpy.py:119-120)."""
        new = self.__class__(self.bases, self.to_torch(), force=True,
                             dtype=self._dtype, device=self._device)
        self._used_here(self._store)
        return new
