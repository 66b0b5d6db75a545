"""Autotuner for the launch shapes of the batched apply and the PTM gradient.

The batched paths (``autograd.apply``, ``batch.CircuitBatch``) run the same
few geometries thousands of times: per element type, batch size, state shape,
targets and output dimensions.  For each such geometry the tuner times the
candidate launch shapes (``qsv_tile``, include/qsv.h) on scratch tensors of
the real size with HIP events, keeps the fastest and caches it:

* batched apply: the orbit-streaming kernel, or the LDS-tiled kernel with a
  tile of 32 ... 1024 orbits (infeasible tiles are skipped: the library
  answers QSV_ENOTSUP);
* PTM gradient: 4 ... 256 blocks per state.

Tuning happens only inside ``with tuning():`` (or ``Autotuner.tune_*``), so
a HIP-graph capture never times anything; outside, a geometry without an
entry uses the library's own planning.  The cache lives in memory and in a
JSON file (``QSV_TUNE_CACHE``, default ``~/.cache/quantumsim_amd/tiles.json``;
set it to an empty string to keep the cache in memory only).  Every
candidate computes the same result (the kernels differ only in how orbits
are grouped; the gradient only in the order of float64 partial sums), which
``tests/test_kernel_indexing.py`` checks by replaying each candidate on the
CPU and ``tests/test_autotune_gpu.py`` on the GPU.
"""
import contextlib
import json
import math
import os
import threading

import torch

from . import _native as nat

__all__ = ['Autotuner', 'get_tuner', 'set_tuner', 'tuning', 'APPLY_CANDIDATES',
           'GRAD_CANDIDATES']

APPLY_CANDIDATES = [(nat.KERNEL_DIRECT, 0)] + [(nat.KERNEL_LDS, o)
                                                for o in (32, 64, 128, 256, 512, 1024)]
GRAD_CANDIDATES = [4, 8, 16, 32, 64, 128, 256]


def _bucket(b):
    """Batch sizes are tuned per power-of-two bucket (a batcher's prefix
    shrinks as circuits finish)."""
    return 1 << max(0, int(b - 1).bit_length())


def _default_path():
    p = os.environ.get('QSV_TUNE_CACHE')
    if p is not None:
        return p or None
    return os.path.join(os.path.expanduser('~'), '.cache', 'quantumsim_amd', 'tiles.json')


class Autotuner:
    def __init__(self, path='default', reps=3):
        self.path = _default_path() if path == 'default' else path
        self.reps = reps
        self.entries = {}
        self.active = False
        self._lock = threading.Lock()
        if self.path and os.path.exists(self.path):
            try:
                with open(self.path) as fh:
                    self.entries = json.load(fh)
            except (OSError, ValueError):
                self.entries = {}

    # ------------------------------------------------------------------ keys
    _device_names = {}

    @classmethod
    def _key(cls, op, dtype, batch, shape, targets, dims_out):
        i = torch.cuda.current_device()
        dev = cls._device_names.get(i)
        if dev is None:
            dev = cls._device_names[i] = torch.cuda.get_device_name(i)
        return '|'.join([op, str(dtype).replace('torch.', ''), str(_bucket(batch)),
                         'x'.join(map(str, shape)), ','.join(map(str, targets)),
                         ','.join(map(str, dims_out)), dev])

    # --------------------------------------------------------------- lookups
    def apply_tile(self, dtype, batch, shape, targets, dims_out):
        """qsv_tile for a batched apply, tuning it first inside tuning()."""
        if not self.entries and not self.active:
            return None
        key = self._key('apply', dtype, batch, shape, targets, dims_out)
        e = self.entries.get(key)
        if e is None and self.active:
            e = self.tune_apply(dtype, batch, shape, targets, dims_out)
        if e is None or e['kernel'] is None:
            return None
        return nat.Tile(e['kernel'], e['tile_orbits'], 0, 0)

    def grad_tile(self, dtype, batch, shape, targets, dims_out):
        if not self.entries and not self.active:
            return None
        key = self._key('grad', dtype, batch, shape, targets, dims_out)
        e = self.entries.get(key)
        if e is None and self.active:
            e = self.tune_grad(dtype, batch, shape, targets, dims_out)
        return None if e is None else nat.Tile(0, 0, e['grad_blocks'], 0)

    # ---------------------------------------------------------------- timing
    def _time(self, fn):
        stream = torch.cuda.current_stream()
        fn()                                   # warm-up (and first-launch costs)
        ts = []
        for _ in range(self.reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ts.append(a.elapsed_time(b) * 1e3)
        ts.sort()
        return ts[len(ts) // 2]                # median, microseconds

    def tune_apply(self, dtype, batch, shape, targets, dims_out):
        L = nat.lib()
        B = _bucket(batch)
        n_in = math.prod(shape)
        oshape = list(shape)
        for t, d in zip(targets, dims_out):
            oshape[t] = d
        n_out = math.prod(oshape)
        din = math.prod(shape[t] for t in targets)
        dout = math.prod(dims_out)
        dev = torch.device('cuda', torch.cuda.current_device())
        x = torch.randn(B, n_in, dtype=dtype, device=dev)
        y = torch.empty(B, n_out, dtype=dtype, device=dev)
        P = torch.randn(B, dout, din, dtype=dtype, device=dev)
        stream = torch.cuda.current_stream().cuda_stream
        dims, tg, do = nat.i64(shape), nat.i32(targets), nat.i32(dims_out)
        code = nat.dtype_code(dtype)

        def run(tile):
            return L.qsv_apply_ptm_batched_tiled(code, x.data_ptr(), y.data_ptr(), B, len(shape),
                                                 dims, len(targets), tg, do, P.data_ptr(),
                                                 dout * din, None if tile is None else
                                                 nat.ctypes.byref(tile), stream)
        t_default = self._time(lambda: nat.check(run(None), 'apply'))
        best = None
        for kern, orbits in APPLY_CANDIDATES:
            tile = nat.Tile(kern, orbits, 0, 0)
            if run(tile) != 0:          # infeasible for this geometry
                continue
            t = self._time(lambda: run(tile))
            if best is None or t < best[0]:
                best = (t, kern, orbits)
        if best is None:
            # no launch shape applies (dense 64 x 64 PTMs run on the MFMA
            # kernel, which has none): the planner's choice
            e = {'kernel': None, 'tile_orbits': 0, 'us': round(t_default, 2),
                 'default_us': round(t_default, 2)}
        else:
            e = {'kernel': best[1], 'tile_orbits': best[2], 'us': round(best[0], 2),
                 'default_us': round(t_default, 2)}
        self._store(self._key('apply', dtype, batch, shape, targets, dims_out), e)
        return e

    def tune_grad(self, dtype, batch, shape, targets, dims_out):
        L = nat.lib()
        B = _bucket(batch)
        n_in = math.prod(shape)
        oshape = list(shape)
        for t, d in zip(targets, dims_out):
            oshape[t] = d
        n_out = math.prod(oshape)
        din = math.prod(shape[t] for t in targets)
        dout = math.prod(dims_out)
        dev = torch.device('cuda', torch.cuda.current_device())
        x = torch.randn(B, n_in, dtype=dtype, device=dev)
        yb = torch.randn(B, n_out, dtype=dtype, device=dev)
        out = torch.empty(B, dout, din, dtype=dtype, device=dev)
        dims, tg, do = nat.i64(shape), nat.i32(targets), nat.i32(dims_out)
        ws_bytes = nat.check(L.qsv_ptm_grad_workspace_bytes(B, len(shape), dims, len(targets),
                                                            tg, do), 'ptm_grad')
        ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream().cuda_stream
        code = nat.dtype_code(dtype)

        def run(tile):
            return L.qsv_ptm_grad_tiled(code, x.data_ptr(), yb.data_ptr(), B, len(shape), dims,
                                        len(targets), tg, do, out.data_ptr(), 0, ws.data_ptr(),
                                        ws_bytes, None if tile is None else nat.ctypes.byref(tile),
                                        stream)
        t_default = self._time(lambda: nat.check(run(None), 'ptm_grad'))
        best = None
        for nb in GRAD_CANDIDATES:
            tile = nat.Tile(0, 0, nb, 0)
            if run(tile) != 0:          # above the kernel's block limit (wide PTMs)
                continue
            t = self._time(lambda: nat.check(run(tile), 'ptm_grad'))
            if best is None or t < best[0]:
                best = (t, nb)
        e = {'grad_blocks': best[1], 'us': round(best[0], 2), 'default_us': round(t_default, 2)}
        self._store(self._key('grad', dtype, batch, shape, targets, dims_out), e)
        return e

    # ----------------------------------------------------------------- cache
    def _store(self, key, e):
        with self._lock:
            self.entries[key] = e

    def save(self):
        if not self.path:
            return
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        tmp = self.path + '.tmp'
        with open(tmp, 'w') as fh:
            json.dump(self.entries, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def summary(self):
        """Sum over the cached geometries of the default and tuned times."""
        out = {}
        for k, e in self.entries.items():
            op = k.split('|', 1)[0]
            s = out.setdefault(op, {'geometries': 0, 'default_us': 0.0, 'tuned_us': 0.0})
            s['geometries'] += 1
            s['default_us'] += e['default_us']
            s['tuned_us'] += e['us']
        for s in out.values():
            s['default_us'] = round(s['default_us'], 1)
            s['tuned_us'] = round(s['tuned_us'], 1)
        return out


_TUNER = None


def get_tuner():
    global _TUNER
    if _TUNER is None:
        _TUNER = Autotuner()
    return _TUNER


def set_tuner(tuner):
    """Install ``tuner`` as the process-wide tuner (returns the previous)."""
    global _TUNER
    prev, _TUNER = _TUNER, tuner
    return prev


@contextlib.contextmanager
def tuning(tuner=None):
    """Inside this block, batched applies and PTM gradients of geometries
    without a cache entry are tuned (eagerly, before they run)."""
    t = tuner or get_tuner()
    prev = t.active
    t.active = True
    try:
        yield t
    finally:
        t.active = prev
        t.save()
