"""Algorithmic HBM bytes of the library's batched launches, by kernel family.

While a ``Ledger`` is active (``with Ledger() as led: ...``), every launch
the differentiable and batched paths make through libqsv adds the bytes its
algorithm has to move:

* apply (``qsv_apply_ptm_batched``): (N_in + N_out) * s per state plus the
  PTMs read;
* PTM gradient (``qsv_ptm_grad``): both operand states, 2 * N * s per
  state for a square PTM, plus the gradient written;
* fused backward (``qsv_apply_ptm_batched_backward``): x and ybar read,
  xbar written, 3 * N * s per state, plus the PTMs read and the gradient
  written;
* persistent circuit kernels: the trajectory (the state before every gate,
  written forward and read back by the adjoint), the PTM pool, the float64
  populations and their gradients; the states themselves live in LDS.

With ``timed=True`` each launch is also bracketed by HIP events on the
stream it is queued on, so one instrumented eager step gives every family's
GPU time and rate.  Work on two streams overlaps, and an event pair then
also spans the other stream's kernels; ``isolate=True`` waits for the whole
device before and after every counted launch, so each one is timed alone
(as a kernel trace times it).  Off (the default) the ledger costs one global
lookup per launch.  No counterpart in the reference.
"""
import torch

__all__ = ['Ledger', 'active', 'APPLY_DENSE64', 'APPLY_SMALL', 'GRAD_DENSE64', 'GRAD_SMALL',
           'BWD_DENSE64', 'PERSISTENT_FWD', 'PERSISTENT_ADJ', 'KERNELS']

APPLY_DENSE64 = 'apply_dense64'
APPLY_SMALL = 'apply_small'
GRAD_DENSE64 = 'grad_dense64'
GRAD_SMALL = 'grad_small'
BWD_DENSE64 = 'bwd_dense64'
PERSISTENT_FWD = 'persistent_forward'
PERSISTENT_ADJ = 'persistent_adjoint'

# the HIP kernels behind each family (rocprofv3 names)
KERNELS = {
    APPLY_DENSE64: 'k_fused64 / k_apply64q (64 x 64 PTMs on MFMA)',
    APPLY_SMALL: 'k_apply_small / k_apply_lds (PTMs up to 16 x 16)',
    GRAD_DENSE64: 'k_ptm_grad64 / k_ptm_grad64d + k_ptm_grad_final',
    GRAD_SMALL: 'k_ptm_grad + k_ptm_grad_final',
    BWD_DENSE64: 'k_bwd64d / k_bwd64q + k_ptm_grad_final (one-pass backward of 64 x 64 blocks)',
    PERSISTENT_FWD: 'k_circuits_fwd',
    PERSISTENT_ADJ: 'k_circuits_adj',
}

_ACTIVE = None
# GPU clock cycles of the spin queued before an isolated launch (~100 us)
_SPIN_CYCLES = 200000


def active():
    """The ledger in force, or None."""
    return _ACTIVE


class Ledger:
    def __init__(self, timed=True, isolate=False):
        self.timed = timed
        self.isolate = isolate
        self.rows = {}        # family -> [launches, bytes, [(start, end) events], flops]

    def __enter__(self):
        global _ACTIVE
        self._prev = _ACTIVE
        _ACTIVE = self
        return self

    def __exit__(self, *exc):
        global _ACTIVE
        _ACTIVE = self._prev
        return False

    def begin(self):
        """Call right before a launch; returns its start event (or None)."""
        if not self.timed:
            return None
        if self.isolate:
            torch.cuda.synchronize()
            # keep the stream busy while the host queues the event and the
            # launch, so the start event fires right before the kernel runs
            # (not when the host gets to the launch)
            torch.cuda._sleep(_SPIN_CYCLES)
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def end(self, start, family, nbytes, flops=0):
        """Call right after the launch that ``start`` belongs to; ``flops``:
        the MFMA contraction's floating-point operations (dense families)."""
        row = self.rows.setdefault(family, [0, 0, [], 0])
        row[0] += 1
        row[1] += int(nbytes)
        row[3] += int(flops)
        if start is not None:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            row[2].append((start, ev))
            if self.isolate:
                torch.cuda.synchronize()

    def total_bytes(self):
        return sum(r[1] for r in self.rows.values())

    def summary(self, peak_gbs=None, peak_tflops=None):
        """{family: {launches, bytes, ms, GBs[, frac][, TFLOPs, mfma_frac]}}
        (waits for the GPU); ``frac`` against ``peak_gbs`` (HBM), and for the
        families that count flops ``mfma_frac`` against ``peak_tflops``
        (the dense MFMA peak of the states' element type)."""
        torch.cuda.synchronize()
        out = {}
        for fam, (n, b, evs, fl) in sorted(self.rows.items()):
            row = {'kernel': KERNELS.get(fam, fam), 'launches': n, 'bytes': b}
            if fl:
                row['flops'] = fl
            if evs:
                ms = sum(a.elapsed_time(c) for a, c in evs)
                row['ms'] = round(ms, 3)
                row['GBs'] = round(b / ms / 1e6, 1) if ms > 0 else None
                if peak_gbs and row['GBs']:
                    row['frac'] = round(row['GBs'] / peak_gbs, 4)
                if fl and ms > 0:
                    row['TFLOPs'] = round(fl / ms / 1e9, 2)
                    if peak_tflops:
                        row['mfma_frac'] = round(row['TFLOPs'] / peak_tflops, 4)
            out[fam] = row
        return out
