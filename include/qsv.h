/*
 * libqsv — MI355X (gfx950) kernels for Pauli-vector density matrices.
 *
 * C ABI only: plain pointers, sizes and a hipStream_t passed as void*.
 * Every call is stream-ordered and asynchronous; none allocates device
 * memory (callers hand in workspaces sized by the *_workspace_bytes
 * queries).  Return value 0 = success, negative = error; the message of
 * the last error on the calling thread is returned by qsv_last_error().
 *
 * Layout convention (quantumsim/pauli_vectors/cuda.py:170 and
 * quantumsim/pauli_vectors/numpy.py:55-64): a Pauli vector over n qudits is
 * a dense C-order tensor of shape dim_pauli[0..n-1]; qudit 0 is the most
 * significant (slowest) axis.  A k-qudit PTM is a C-order tensor of shape
 * (out_0..out_{k-1}, in_0..in_{k-1}) (quantumsim/operations/operation.py:476).
 */
#ifndef QSV_H
#define QSV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types of a Pauli vector on the device */
#define QSV_F32 0
#define QSV_F64 1

/* error codes */
#define QSV_OK 0
#define QSV_EINVAL (-1)   /* bad argument: shape, index, dtype, size     */
#define QSV_EHIP (-2)     /* a HIP runtime call failed                   */
#define QSV_ENOTSUP (-3)  /* valid request outside what the kernels do   */

#define QSV_MAX_RANK 32    /* qudits per Pauli vector                    */
#define QSV_MAX_TARGETS 8  /* qudits per PTM                             */

/* Launch shape of the batched apply and of the PTM gradient, as chosen by
 * an autotuner (quantumsim_amd/autotune.py times the candidates on the
 * shapes at hand).  Passing NULL, or zero fields, keeps the library's own
 * planning.  Plain data: the library keeps no tuning state. */
#define QSV_KERNEL_AUTO 0    /* planner's choice                          */
#define QSV_KERNEL_DIRECT 1  /* orbit-streaming kernel                    */
#define QSV_KERNEL_LDS 2     /* LDS-tiled kernel (QSV_ENOTSUP if the      */
                             /* geometry has no tiling)                   */
typedef struct qsv_tile {
  int32_t kernel;       /* QSV_KERNEL_*                                    */
  int32_t tile_orbits;  /* LDS tile size in orbits (0: planner's)        */
  int32_t grad_blocks;  /* gradient blocks per state, <= 256 (0: auto)   */
  int32_t reserved;     /* must be 0                                     */
} qsv_tile;

/* Tables of the persistent circuit kernels (device memory, built by the
 * caller).  A circuit: `n` qubits in the general(2) basis, gates
 * gates[gate0 .. gate0 + n_gates), trajectory element offset `traj` (room
 * for n_gates states of 4^n entries: the state before each gate),
 * population offset `pop` (2^n float64 entries).  A gate: k = 1 or 2
 * targets q0 (, q1) in PTM axis order, PTM at element offset `ptm` of the
 * PTM pool (4^k x 4^k, row-major, layout (out..., in...)); every gate owns
 * its PTM slot (the adjoint writes dL/dP there). */
typedef struct qsv_circuit {
  int32_t n;
  int32_t n_gates;
  int64_t gate0;
  int64_t traj;
  int64_t pop;
} qsv_circuit;

typedef struct qsv_gate {
  int32_t k;
  int32_t q0;
  int32_t q1;
  int32_t reserved;
  int64_t ptm;
} qsv_gate;

/* One block position of a brickwork window (two consecutive layers of a
 * layered noisy circuit), for qsv_window_blocks: the block's qubits
 * (ascending; qubits[0] is the most significant PTM digit), its output
 * array (element offset of a [n_windows][batch][D][D] array, D = 4^nq) and
 * up to four factors applied in order.  Factor f: kind 1 = a lone qubit of
 * layer f_layer (PTM D_q R_q on slot f_slot0), kind 2 = a bond of layer
 * f_layer on slots (f_slot0, f_slot1) (PTM (D_a (x) D_b) CZ (R_a (x) R_b)). */
typedef struct qsv_wblock {
  int32_t nq;
  int32_t qubits[3];
  int32_t nf;
  int32_t reserved;
  int64_t out_off;
  int32_t f_kind[4];
  int32_t f_layer[4];
  int32_t f_slot0[4];
  int32_t f_slot1[4];
} qsv_wblock;

/* Message of the last failed call on this thread ("" if none). */
const char* qsv_last_error(void);

/* ABI version (major*100 + minor). */
int qsv_abi_version(void);

/* Bytes of device workspace qsv_apply_ptm needs for this PTM shape
 * (0 when the PTM and its offset tables travel inside the kernel
 * arguments).  dims_in/dims_out are the k target dims. */
int64_t qsv_apply_ptm_workspace_bytes(int dtype, int k, const int32_t* dims_in_t,
                                      const int32_t* dims_out_t);

/* out = PTM applied to `in` on the qudits `targets` (k of them, in PTM axis
 * order; any order, no duplicates).
 *   dims[rank]      : input tensor shape (Pauli dims per qudit)
 *   dims_out_t[k]   : PTM output dims for the k targets (the tensor shape
 *                     of `out` is dims with dims[targets[j]] := dims_out_t[j])
 *   ptm_f64         : host pointer, C-order (out..., in...) float64 PTM
 *   workspace       : device memory of qsv_apply_ptm_workspace_bytes()
 *                     bytes (may be NULL when that is 0)
 * `in` and `out` must not overlap.
 * Replaces PauliVectorNumpy.apply_ptm (quantumsim/pauli_vectors/numpy.py:50-64)
 * and PauliVectorCuda._apply_{single,two}_qubit_ptm
 * (quantumsim/pauli_vectors/cuda.py:151-277 -> primitives.cu:15-89). */
int qsv_apply_ptm(int dtype, const void* in, void* out, int rank,
                  const int64_t* dims, int k, const int32_t* targets,
                  const int32_t* dims_out_t, const double* ptm_f64,
                  void* workspace, int64_t workspace_bytes, void* stream);

/* qsv_apply_ptm on strided tensors: in_strides[rank] / out_strides[rank]
 * are element strides per axis (NULL = C-contiguous) in C order without
 * overlap (stride[q] >= stride[q+1] * dims[q+1]); gaps are allowed.  The
 * backend pads the strides of the outermost axes of large vectors so that
 * the operand streams of a gate on two outer qudits do not alias in the
 * HBM channel/bank mapping. */
int qsv_apply_ptm_strided(int dtype, const void* in, void* out, int rank,
                          const int64_t* dims, const int64_t* in_strides,
                          const int64_t* out_strides, int k, const int32_t* targets,
                          const int32_t* dims_out_t, const double* ptm_f64,
                          void* workspace, int64_t workspace_bytes, void* stream);

/* Same contraction with the PTM already on the device, C-order
 * (out..., in...), in the state's element type (used when PTMs are built
 * on the device, e.g. from trainable parameters, and under graph capture).
 * `workspace` is used only by dense 256 x 256 PTMs (D * D elements); PTMs
 * of any other size up to 1024 entries a side need none (the table kernel
 * decodes its offsets on the device). */
int qsv_apply_ptm_dev(int dtype, const void* in, void* out, int rank,
                      const int64_t* dims, int k, const int32_t* targets,
                      const int32_t* dims_out_t, const void* ptm_dev,
                      void* workspace, int64_t workspace_bytes, void* stream);

/* qsv_apply_ptm_dev on strided tensors (strides as in qsv_apply_ptm_strided).
 * Dense 64 x 64 PTMs are gathered into the MFMA kernel's fragment order by
 * the kernel itself, 256 x 256 ones packed in the workspace by a device
 * kernel on `stream`, so a PTM built on the device never travels to the
 * host.  This is the backend call of the
 * device-side operation pipeline: the PTM of quantumsim's
 * PTMOperation.__call__ (quantumsim/operations/operation.py:529-551),
 * built and basis-converted with torch, applied as
 * PauliVectorNumpy.apply_ptm does (quantumsim/pauli_vectors/numpy.py:50-64). */
int qsv_apply_ptm_dev_strided(int dtype, const void* in, void* out, int rank,
                              const int64_t* dims, const int64_t* in_strides,
                              const int64_t* out_strides, int k, const int32_t* targets,
                              const int32_t* dims_out_t, const void* ptm_dev,
                              void* workspace, int64_t workspace_bytes, void* stream);

/* Batched form of qsv_apply_ptm_dev for PTMs of any size up to 1024
 * entries a side: inline kernels up to 16 x 16, dense 64 x 64 PTMs
 * (three-qubit blocks) on the MFMA kernel with each state's row-major PTM
 * gathered into fragment order inside the kernel, any other size (qutrit
 * pairs 81 x 81, qutrit + qubit 36 x 36, ...) on the table kernel with the
 * offsets decoded on the device; no workspace either way: `batch` states
 * of identical shape, stored back to back (in: batch x N_in, out: batch x
 * N_out), each contracted with its own device PTM
 * ptm_dev + b * ptm_batch_stride (elements of the state's type; a stride of
 * 0 applies one PTM to every state).  One launch, no host-device copies:
 * safe under HIP graph capture.  No counterpart in the reference (its
 * backend holds one state per object); used by the circuit batcher and the
 * differentiable path, which run many circuits or parameter sets at once. */
int qsv_apply_ptm_batched(int dtype, const void* in, void* out, int batch, int rank,
                          const int64_t* dims, int k, const int32_t* targets,
                          const int32_t* dims_out_t, const void* ptm_dev,
                          int64_t ptm_batch_stride, void* stream);

/* qsv_apply_ptm_batched with an explicit launch shape (NULL = planner's). */
int qsv_apply_ptm_batched_tiled(int dtype, const void* in, void* out, int batch, int rank,
                                const int64_t* dims, int k, const int32_t* targets,
                                const int32_t* dims_out_t, const void* ptm_dev,
                                int64_t ptm_batch_stride, const qsv_tile* tile, void* stream);

/* Largest qubit count whose state (and the adjoint's two states) fit the
 * LDS of one workgroup on the current device for the persistent circuit
 * kernels below. */
int qsv_circuits_max_qubits(int dtype);

/* LDS bytes per workgroup of the persistent forward (adjoint = 0) or
 * adjoint (1) kernel for circuits of up to max_qubits qubits. */
int64_t qsv_circuits_lds_bytes(int dtype, int max_qubits, int adjoint);

/* One persistent launch runs every circuit of the table (widths up to
 * max_qubits <= qsv_circuits_max_qubits, any depths): each workgroup keeps a
 * circuit's Pauli vector in LDS, starts from |0..0><0..0|, applies the
 * gates with the pool's PTMs, writes the state before every gate to `traj`
 * and the float64 populations to `pops`.  Safe under HIP graph capture (one
 * kernel, tables and buffers in device memory).  No counterpart in the
 * reference (one circuit per PauliVector object there); used by the circuit
 * batcher for its narrow circuits. */
int qsv_circuits_forward(int dtype, int n_circuits, const qsv_circuit* circs_dev,
                         const qsv_gate* gates_dev, const void* ptms, void* traj,
                         double* pops, int max_qubits, void* stream);

/* Adjoint of qsv_circuits_forward: from dL/dpops (float64) and the forward
 * trajectory, writes dL/dP of every gate (float64, fixed summation order)
 * into dptms at the gate's PTM offset.  One persistent launch. */
int qsv_circuits_adjoint(int dtype, int n_circuits, const qsv_circuit* circs_dev,
                         const qsv_gate* gates_dev, const void* ptms, const void* traj,
                         const double* dpops, double* dptms, int max_qubits, void* stream);

/* Dense blocks of every window of `batch` layered circuits of n_qubits
 * qubits, one launch: for window w, circuit b and position p of the table,
 * out[p.out_off + ((w * batch + b) * D + o) * D + i] = the product of the
 * position's factors built from
 *   rot  [batch][layers][n_qubits][4][4]  one-qubit PTM of layer 2w + l,
 *   damp [batch][n_qubits][4][4]          damping PTM of each qubit,
 *   cz   [16][16]                          the CZ PTM,
 * with the layer-1 factors replaced by the identity where keep[b * n_windows
 * + w] == 0 (a circuit whose depth ends inside the window).  float64
 * arithmetic, output in the state's element type.  table_host (validated on
 * the host) and table_dev hold the same n_pos entries.  No counterpart in
 * the reference (which composes PTMs on the host per operation); it replaces
 * the block composition of the batched path. */
int qsv_window_blocks(int dtype, int n_pos, const qsv_wblock* table_host,
                      const qsv_wblock* table_dev, int batch, int n_windows, int n_qubits,
                      int layers, const double* rot, const double* damp, const uint8_t* keep,
                      const double* cz, void* out, int64_t out_elems, void* stream);

/* Backward of qsv_window_blocks: from dL/dout (same layout, state type)
 * writes every entry of drot [batch][layers][n_qubits][4][4] and of
 * ddamp_part [batch][n_windows][2][n_qubits][4][4] (the damping gradient of
 * each window and layer; the caller sums them); float64. */
int qsv_window_blocks_backward(int dtype, int n_pos, const qsv_wblock* table_host,
                               const qsv_wblock* table_dev, int batch, int n_windows,
                               int n_qubits, int layers, const double* rot, const double* damp,
                               const uint8_t* keep, const double* cz, const void* dout,
                               int64_t out_elems, double* drot, double* ddamp_part,
                               void* stream);

/* CPU replay of both window-block kernels on host float64 buffers (the same
 * per-thread phases, every index checked); dout == NULL replays the forward
 * kernel only.  No counterpart in the reference. */
int qsv_emulate_window_blocks(int n_pos, const qsv_wblock* table, int batch, int n_windows,
                              int n_qubits, int layers, const double* rot, const double* damp,
                              const uint8_t* keep, const double* cz, double* out,
                              int64_t out_elems, const double* dout, double* drot,
                              double* ddamp_part);

/* Bytes of device workspace qsv_ptm_grad / qsv_ptm_grad_tiled need (enough
 * for any launch shape). */
int64_t qsv_ptm_grad_workspace_bytes(int batch, int rank, const int64_t* dims, int k,
                                     const int32_t* targets, const int32_t* dims_out_t);

/* Gradient of a PTM application with respect to the PTM, for `batch` states
 * back to back (the layout of qsv_apply_ptm_batched), PTMs up to 256
 * entries a side and 8192 entries (MFMA kernels up to 64 x 64, a float64
 * LDS kernel above: qutrit pairs 81 x 81, ...):
 *   dptm[b, o, i] (+)= sum over orbits of ybar[b, .., o, ..] * x[b, .., i, ..]
 * x: the states the PTM was applied to; ybar: the loss gradient with
 * respect to the results.  Accumulated in float64 in a fixed order
 * (deterministic; PTMs above 16 x 16 on float32 states sum in float32
 * within a wave, float64 beyond), written (accumulate = 0) or added
 * (accumulate = 1) to dptm in the state's element type.  The gradient with respect to x is
 * qsv_apply_ptm_batched with the transposed PTMs.  No counterpart in the
 * reference (which has no differentiation); it is the adjoint used by the
 * differentiable path. */
int qsv_ptm_grad(int dtype, const void* x, const void* ybar, int batch, int rank,
                 const int64_t* dims, int k, const int32_t* targets, const int32_t* dims_out_t,
                 void* dptm, int accumulate, void* workspace, int64_t workspace_bytes,
                 void* stream);

/* Size of the device scratch array of qsv_ptm_grad_noalloc (one per GPU,
 * held by the library until the process ends). */
#define QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES (16LL << 20)

/* qsv_ptm_grad without a workspace argument and without any allocation:
 * the partial tiles go to a file-scope device array of
 * QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES bytes (fewer blocks per state than
 * qsv_ptm_grad chooses when its tiles would not fit, which changes only
 * the order of the float64 partial sums).  Calls no hipMalloc/hipFree.
 * Requests whose smallest tiles (batch x dout x din x 8 bytes) exceed the
 * array return QSV_EINVAL.  The stream must belong to the current device.
 * Every call on one GPU shares that GPU's array: calls that overlap in
 * time on one GPU (different streams) race; calls ordered on one stream
 * are safe.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). */
int qsv_ptm_grad_noalloc(int dtype, const void* x, const void* ybar, int batch, int rank,
                         const int64_t* dims, int k, const int32_t* targets,
                         const int32_t* dims_out_t, void* dptm, int accumulate, void* stream);

/* QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES, as the library was built with it. */
int64_t qsv_ptm_grad_noalloc_scratch_bytes(void);

/* Backward of qsv_apply_ptm_batched for a square dense 64 x 64 PTM in one
 * pass: xbar = P^T ybar (per state, P = ptm_dev + b * ptm_batch_stride,
 * row-major in the state's element type) and dptm (+)= sum over orbits of
 * ybar x^T as qsv_ptm_grad computes it; reads x and ybar once and writes
 * xbar (three streams instead of four).  xbar is bitwise the transposed
 * apply's (the same k-steps of four entries in the same order).  workspace:
 * qsv_ptm_grad_workspace_bytes of the same geometry.  QSV_ENOTSUP unless the
 * states are 16-byte aligned and, for float64, the innermost axis is a
 * unit-stride bystander; for float32, the tiles of 32 orbits are whole and
 * either the orbits come in 4-orbit vectors (innermost axis a bystander)
 * or the PTM's lowest index is the innermost axis (entries in runs of
 * four): the two direct-to-LDS kernels.  Otherwise callers use
 * qsv_apply_ptm_batched with the transposed PTM and qsv_ptm_grad.  xbar
 * must not alias x or ybar.  No counterpart in the reference; the adjoint
 * of the differentiable and batched paths. */
int qsv_apply_ptm_batched_backward(int dtype, const void* x, const void* ybar, void* xbar,
                                   int batch, int rank, const int64_t* dims, int k,
                                   const int32_t* targets, const void* ptm_dev,
                                   int64_t ptm_batch_stride, void* dptm, int accumulate,
                                   void* workspace, int64_t workspace_bytes, void* stream);

/* qsv_ptm_grad with an explicit launch shape (tile->grad_blocks, at most 64
 * for PTMs above 16 x 16; NULL = automatic).  The result does not depend on
 * the shape beyond the order of the partial sums. */
int qsv_ptm_grad_tiled(int dtype, const void* x, const void* ybar, int batch, int rank,
                       const int64_t* dims, int k, const int32_t* targets,
                       const int32_t* dims_out_t, void* dptm, int accumulate, void* workspace,
                       int64_t workspace_bytes, const qsv_tile* tile, void* stream);

/* CPU replay of qsv_ptm_grad for one state of at most 2^24 entries: the
 * same plan, every lane's loads checked against the buffers (and vector
 * alignment), the MFMA contraction in hardware index order, the block
 * reduction and the fixed-order final sum, each orbit checked to be used
 * exactly once.  x, ybar: host float64 arrays (contiguous); dptm: host
 * float64 [dout * din].  grad_blocks: blocks per state, as the qsv_tile field (0:
 * automatic for a batch of one).  *vn_out (may be NULL) = orbits per lane
 * vector the GPU call would use for `dtype`, plus 256 when 64 x 64 PTMs
 * load runs of four entries as 16-byte vectors (innermost axis a target),
 * plus 512 when the LDS-staged 64 x 64 kernel runs (float64 states with
 * orbit vectors).  No counterpart in the reference; it lets CPU tests pin
 * the gradient kernel's indexing before GPU use. */
int qsv_emulate_ptm_grad(int dtype, const double* x, const double* ybar, int rank,
                         const int64_t* dims, int k, const int32_t* targets,
                         const int32_t* dims_out_t, double* dptm, int32_t grad_blocks,
                         int32_t* vn_out);

/* out = in[np.ix_(idx_0, ..., idx_{rank-1})] with an index of -1 producing
 * 0.  The index lists live in device memory as int32: idx_dev[offsets[r] +
 * j] is the j-th index of axis r, for j < out_dims[r]; offsets (host) has
 * rank entries.  in_dtype -> out_dtype conversion is exact (f32 -> f64) or
 * a plain copy.  Bit-exact gather.
 * Replaces the multitake kernel (quantumsim/pauli_vectors/primitives.cu:92-130,
 * host side cuda.py:307-336). */
int qsv_multitake(int in_dtype, int out_dtype, const void* in, void* out,
                  int rank, const int64_t* in_dims, const int64_t* out_dims,
                  const int32_t* idx_dev, const int32_t* offsets,
                  void* stream);

/* qsv_multitake on a strided input (in_strides[rank] element strides, NULL
 * = C-contiguous; C order without overlap, gaps allowed). */
int qsv_multitake_strided(int in_dtype, int out_dtype, const void* in, void* out, int rank,
                          const int64_t* in_dims, const int64_t* in_strides,
                          const int64_t* out_dims, const int32_t* idx_dev,
                          const int32_t* offsets, void* stream);

/* Bytes of device workspace qsv_axis_sum needs. */
int64_t qsv_axis_sum_workspace_bytes(int64_t n_outer, int d, int64_t n_inner);

/* out_f64[j] = sum_{o,i} in[o, j, i] for `in` viewed as [n_outer, d, n_inner],
 * accumulated in float64, deterministic order; out_f64 is a device pointer
 * of d doubles.  d <= 64.  With d = 1 this is the trace of a diagonal;
 * with the diagonal viewed around one qudit it is meas_prob.
 * Replaces pycuda's ReductionKernel sum_along_axis
 * (quantumsim/pauli_vectors/cuda.py:46-52, 366-375) and the host sums of
 * numpy.py:82-84, 97-107. */
int qsv_axis_sum(int dtype, const void* in, int64_t n_outer, int d,
                 int64_t n_inner, double* out_f64, void* workspace,
                 int64_t workspace_bytes, void* stream);

/* data[i] = data[i] * s for i < n, the product taken in float64 and rounded
 * once to the element type.  Replaces the scaling step of renormalize
 * (quantumsim/pauli_vectors/numpy.py:109-117, cuda.py:386-395). */
int qsv_scale(int dtype, void* data, int64_t n, double s, void* stream);

/* ---- diagnostics (host only, no GPU work) ------------------------------
 * Checks the library's 32-bit magic-number division against hardware
 * division for `n_divisors` divisors (1, 2, ..., plus all powers of 2, 3, 4,
 * 9 and random values) times a fixed set of numerators including 0 and
 * 2^32-1.  Returns the number of mismatches (0 = pass).  No counterpart in
 * the reference; it lets CPU tests pin the index arithmetic every kernel
 * uses. */
int64_t qsv_selftest_fastdiv(int32_t n_divisors, uint64_t seed);

/* CPU replay of qsv_apply_ptm for a vector of at most 2^24 entries: the
 * same planner picks the kernel path the GPU call would take for `dtype`
 * (reported in *path_out: 1 direct, 2 LDS-tiled, 3 table, 4 MFMA dense; for
 * the MFMA path bits 4-7 and 8-11 hold the elements per load and per store
 * access along the PTM index, bits 12-15 the orbits per lane vector), then every
 * thread's address arithmetic is executed on the host in float64 with each
 * global and LDS access checked against its buffer and every output entry
 * checked to be written exactly once (padding never).  in/out/ptm are host
 * float64 arrays of the strided extents (NULL strides = contiguous).
 * Returns QSV_EINVAL with a message on any violation.  No counterpart in
 * the reference; it lets CPU tests pin the kernels' indexing before GPU use. */
int qsv_emulate_apply_ptm(int dtype, const double* in, double* out, int rank,
                          const int64_t* dims, const int64_t* in_strides,
                          const int64_t* out_strides, int k, const int32_t* targets,
                          const int32_t* dims_out_t, const double* ptm_f64,
                          int32_t* path_out);

/* CPU replay of qsv_circuits_forward (and, with dpops != NULL, of
 * qsv_circuits_adjoint) on host tables and float64 host arrays, checking
 * every table entry against the array sizes, every orbit's LDS indices and
 * that each gate touches every state entry exactly once.  No counterpart in
 * the reference; it pins the table builder and the kernels' indexing. */
int qsv_emulate_circuits(int n_circuits, const qsv_circuit* circs, const qsv_gate* gates,
                         int64_t n_gates_total, const double* ptms, int64_t ptm_elems,
                         double* traj, int64_t traj_elems, double* pops, int64_t pop_elems,
                         const double* dpops, double* dptms, int max_qubits);

/* CPU replay of qsv_apply_ptm_batched_backward for one state (host float64
 * arrays; ptm row-major 64 x 64): every access checked, xbar written once
 * per entry, the MFMA maps of both parts; grad_blocks as in
 * qsv_emulate_ptm_grad.  No counterpart in the reference. */
int qsv_emulate_apply_ptm_batched_backward(int dtype, const double* x, const double* ybar,
                                           double* xbar, int rank, const int64_t* dims, int k,
                                           const int32_t* targets, const double* ptm,
                                           double* dptm, int32_t grad_blocks);

/* CPU replay of the direct-to-LDS float32 64 x 64 apply of the batched and
 * device-PTM paths (k_apply64q: the PTM's lowest index on the innermost
 * axis, whole 32-orbit tiles) for one state (host float64 arrays; ptm
 * row-major 64 x 64; blocks 0 = one block per four tiles): every load,
 * LDS read and store checked, y written once per entry.  QSV_ENOTSUP for
 * other geometries.  No counterpart in the reference. */
int qsv_emulate_apply_ptm_direct(int dtype, const double* x, double* y, int rank,
                                 const int64_t* dims, int k, const int32_t* targets,
                                 const double* ptm, int32_t blocks);

/* qsv_emulate_apply_ptm of the batched path's kernel choice for one state,
 * with an explicit launch shape (the autotuner's candidates are replayed
 * with it).  QSV_ENOTSUP when the tile is infeasible for the geometry. */
int qsv_emulate_apply_ptm_tiled(int dtype, const double* in, double* out, int rank,
                                const int64_t* dims, int k, const int32_t* targets,
                                const int32_t* dims_out_t, const double* ptm_f64,
                                const qsv_tile* tile, int32_t* path_out);

#ifdef __cplusplus
}
#endif

#endif /* QSV_H */
