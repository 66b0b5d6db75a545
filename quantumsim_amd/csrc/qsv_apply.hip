// Streaming PTM application for Pauli vectors on gfx950.
//
// Contraction (quantumsim/pauli_vectors/numpy.py:50-64, primitives.cu:15-89):
//   out[.., o_0, .., o_{k-1}, ..] = sum_i P[o, i] * in[.., i_0, .., i_{k-1}, ..]
// over the k target axes of a C-order tensor.
//
// Work decomposition: the non-target axes are merged into at most k+1
// "bystander groups" (runs of adjacent non-target axes).  One thread owns VEC
// consecutive bystander positions of the innermost group (an "orbit" is the
// D_in input entries that share one bystander index), loads its D_in operands
// with VEC-wide vector loads, and writes D_out results.  Lanes of a wavefront
// walk the innermost group, so every load and store instruction of a wave
// touches one contiguous 64*VEC*sizeof(T)-byte run whenever the innermost
// axis is not a target.  The PTM and the target offset tables are uniform
// across the grid and are read with scalar loads (kernel arguments or a small
// device table); nothing is staged per block.
#include "qsv_common.h"

#include <algorithm>
#include <string.h>
#include <type_traits>
#include <vector>

namespace qsv {

constexpr int kMaxGroups = QSV_MAX_TARGETS + 1;
constexpr int kBlock = 256;
constexpr int kInline = 16;  // PTMs up to 16 x 16 travel in kernel arguments

struct Geom {
  uint32_t n_threads;  // orbits / VEC
  int32_t ngroups;     // bystander groups, innermost first
  int32_t din, dout;   // products of the target input / output dims
  FastDiv gdiv[kMaxGroups];
  int64_t gs_in[kMaxGroups];
  int64_t gs_out[kMaxGroups];
};

template <typename T, int VEC>
struct VT {
  typedef T t __attribute__((ext_vector_type(VEC)));
};
template <typename T>
struct VT<T, 1> {
  typedef T t;
};

// --- operand sources.  p(o, i) is only called with o < dout, i < din.
// PTM padded to CH x CH (zeros) and both offset tables, by value.
template <typename T, int CH>
struct InlineOp {
  T ptm[CH * CH];
  int64_t off_in[CH];
  int64_t off_out[CH];
  static constexpr bool kPadded = true;
  __device__ __forceinline__ T p(int o, int i) const { return ptm[o * CH + i]; }
  __device__ __forceinline__ int64_t oin(int i) const { return off_in[i]; }
  __device__ __forceinline__ int64_t oout(int o) const { return off_out[o]; }
};

// Offsets by value, PTM (dout x din, unpadded) in device memory.
template <typename T, int CH>
struct PtrOp {
  int64_t off_in[CH];
  int64_t off_out[CH];
  const T* __restrict__ ptm;
  int32_t ld;
  static constexpr bool kPadded = false;
  __device__ __forceinline__ T p(int o, int i) const { return ptm[o * ld + i]; }
  __device__ __forceinline__ int64_t oin(int i) const { return off_in[i]; }
  __device__ __forceinline__ int64_t oout(int o) const { return off_out[o]; }
};

// Offsets and PTM in device memory (any size).  Table layout, built by the
// host: int64 off_in[din_pad]; int64 off_out[dout_pad]; and either the PTM
// zero padded to dout_pad x din_pad right after (padded = 1) or a separate
// caller-owned dout x din PTM (padded = 0).
template <typename T>
struct TableOp {
  const int64_t* __restrict__ off_in;
  const int64_t* __restrict__ off_out;
  const T* __restrict__ ptm;
  int32_t ld;
  int32_t padded;
};

__device__ __forceinline__ void orbit_base(const Geom& g, uint32_t m, int64_t& bin,
                                           int64_t& bout) {
  bin = 0;
  bout = 0;
#pragma unroll
  for (int j = 0; j < kMaxGroups; ++j) {
    if (j >= g.ngroups) break;
    if (j == g.ngroups - 1) {
      bin += int64_t(m) * g.gs_in[j];
      bout += int64_t(m) * g.gs_out[j];
    } else {
      const uint32_t q = fdiv(m, g.gdiv[j]);
      const uint32_t r = m - q * g.gdiv[j].d;
      bin += int64_t(r) * g.gs_in[j];
      bout += int64_t(r) * g.gs_out[j];
      m = q;
    }
  }
}

// ---------------------------------------------------------------- small
// din, dout <= CH: all D_in operands live in registers, outputs stream out.
template <typename T, int CH, int VEC, class Op>
__global__ void __launch_bounds__(kBlock)
    k_apply_small(const T* __restrict__ in, T* __restrict__ out, const Geom g, const Op op) {
  typedef typename VT<T, VEC>::t V;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= g.n_threads) return;
  int64_t bin, bout;
  orbit_base(g, t * VEC, bin, bout);

  V v[CH];
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    if (i < g.din)
      v[i] = *reinterpret_cast<const V*>(in + bin + op.oin(i));
    else
      v[i] = V(0);
  }
#pragma unroll
  for (int o = 0; o < CH; ++o) {
    if (o < g.dout) {
      V acc = V(0);
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (Op::kPadded || i < g.din) acc += op.p(o, i) * v[i];
      }
      *reinterpret_cast<V*>(out + bout + op.oout(o)) = acc;
    }
  }
}

// ---------------------------------------------------------------- table
// Any din/dout: operands CH at a time, outputs CH at a time.
template <typename T, int CH, int VEC>
__global__ void __launch_bounds__(kBlock)
    k_apply_table(const T* __restrict__ in, T* __restrict__ out, const Geom g,
                  const TableOp<T> op) {
  typedef typename VT<T, VEC>::t V;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= g.n_threads) return;
  int64_t bin, bout;
  orbit_base(g, t * VEC, bin, bout);

  for (int oc = 0; oc < g.dout; oc += CH) {
    V acc[CH];
#pragma unroll
    for (int o = 0; o < CH; ++o) acc[o] = V(0);
    for (int ic = 0; ic < g.din; ic += CH) {
      V v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (ic + i < g.din)
          v[i] = *reinterpret_cast<const V*>(in + bin + op.off_in[ic + i]);
        else
          v[i] = V(0);
      }
#pragma unroll
      for (int o = 0; o < CH; ++o) {
        if (oc + o < g.dout) {
          const T* prow = op.ptm + int64_t(oc + o) * op.ld + ic;
#pragma unroll
          for (int i = 0; i < CH; ++i) {
            if (op.padded || ic + i < g.din) acc[o] += prow[i] * v[i];
          }
        }
      }
    }
#pragma unroll
    for (int o = 0; o < CH; ++o) {
      if (oc + o < g.dout) *reinterpret_cast<V*>(out + bout + op.off_out[oc + o]) = acc[o];
    }
  }
}

// ------------------------------------------------------------ host side
struct Plan {
  Geom g;
  int64_t n_in, n_out;
  int vec;
  int din, dout;
  std::vector<int64_t> off_in, off_out;
};

static int make_plan(int dtype, int rank, const int64_t* dims, int k, const int32_t* targets,
                     const int32_t* dims_out_t, Plan& p) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(rank >= 0 && rank <= QSV_MAX_RANK, "rank %d outside [0, %d]", rank,
                QSV_MAX_RANK);
  QSV_CHECK_ARG(k >= 1 && k <= QSV_MAX_TARGETS, "number of targets %d outside [1, %d]", k,
                QSV_MAX_TARGETS);
  QSV_CHECK_ARG(k <= rank, "%d targets on a %d-qudit Pauli vector", k, rank);
  QSV_CHECK_ARG(dims && targets && dims_out_t, "null shape argument");
  bool is_target[QSV_MAX_RANK] = {false};
  int64_t dout_full[QSV_MAX_RANK];
  for (int q = 0; q < rank; ++q) {
    QSV_CHECK_ARG(dims[q] >= 1 && dims[q] <= 4096, "dims[%d] = %lld is invalid", q,
                  (long long)dims[q]);
    dout_full[q] = dims[q];
  }
  for (int j = 0; j < k; ++j) {
    const int q = targets[j];
    QSV_CHECK_ARG(q >= 0 && q < rank, "qubit %d does not exist in a %d-qudit Pauli vector",
                  q, rank);
    QSV_CHECK_ARG(!is_target[q], "qubit %d listed twice", q);
    QSV_CHECK_ARG(dims_out_t[j] >= 1 && dims_out_t[j] <= 4096, "PTM output dim %d invalid",
                  dims_out_t[j]);
    is_target[q] = true;
    dout_full[q] = dims_out_t[j];
  }
  int64_t sin[QSV_MAX_RANK], sout[QSV_MAX_RANK];
  int64_t acc_in = 1, acc_out = 1;
  for (int q = rank - 1; q >= 0; --q) {
    sin[q] = acc_in;
    sout[q] = acc_out;
    acc_in *= dims[q];
    acc_out *= dout_full[q];
    QSV_CHECK_ARG(acc_in < (int64_t(1) << 40) && acc_out < (int64_t(1) << 40),
                  "Pauli vector too large");
  }
  p.n_in = acc_in;
  p.n_out = acc_out;

  int64_t din = 1, dout = 1;
  for (int j = 0; j < k; ++j) {
    din *= dims[targets[j]];
    dout *= dims_out_t[j];
  }
  QSV_CHECK_ARG(din <= 65536 && dout <= 65536, "PTM of %lld x %lld is too large",
                (long long)dout, (long long)din);
  p.din = int(din);
  p.dout = int(dout);
  // offsets of the target sub-tensor, PTM index order (last target fastest)
  p.off_in.assign(din, 0);
  p.off_out.assign(dout, 0);
  for (int64_t i = 0; i < din; ++i) {
    int64_t r = i, off = 0;
    for (int j = k - 1; j >= 0; --j) {
      const int64_t d = dims[targets[j]];
      off += (r % d) * sin[targets[j]];
      r /= d;
    }
    p.off_in[i] = off;
  }
  for (int64_t o = 0; o < dout; ++o) {
    int64_t r = o, off = 0;
    for (int j = k - 1; j >= 0; --j) {
      const int64_t d = dims_out_t[j];
      off += (r % d) * sout[targets[j]];
      r /= d;
    }
    p.off_out[o] = off;
  }

  // bystander groups, innermost first; size-1 groups are dropped
  Geom& g = p.g;
  memset(&g, 0, sizeof(g));
  g.din = p.din;
  g.dout = p.dout;
  int ng = 0;
  int64_t M = 1;
  int q = rank - 1;
  while (q >= 0) {
    if (is_target[q]) {
      --q;
      continue;
    }
    int64_t size = 1;
    const int64_t s_in = sin[q], s_out = sout[q];
    while (q >= 0 && !is_target[q]) {
      size *= dims[q];
      --q;
    }
    if (size == 1) continue;
    QSV_CHECK_ARG(size < (int64_t(1) << 32), "bystander group too large");
    g.gs_in[ng] = s_in;
    g.gs_out[ng] = s_out;
    g.gdiv[ng] = make_fastdiv(uint32_t(size));
    M *= size;
    ++ng;
  }
  g.ngroups = ng;
  // vector width along the innermost group when it is unit-stride: 16-byte
  // accesses.  All other offsets are multiples of that group's size, so the
  // vector accesses stay aligned.
  int vec = 1;
  if (ng > 0 && g.gs_in[0] == 1 && g.gs_out[0] == 1) {
    const uint32_t g0 = g.gdiv[0].d;
    // f64: 2 x 8 B; f32: 4 x 4 B (a float2 variant needs 224 AGPRs on gfx950
    // and is not instantiated)
    const int want = dtype == QSV_F64 ? 2 : 4;
    vec = (g0 % want) == 0 ? want : 1;
  }
  p.vec = vec;
  QSV_CHECK_ARG(M / vec < (int64_t(1) << 32) - kBlock, "too many orbits for 32-bit indexing");
  g.n_threads = uint32_t(M / vec);
  return QSV_OK;
}

template <typename T, int CH, int VEC, class Op>
static int launch_small(const Plan& p, const Op& op, const void* in, void* out,
                        hipStream_t s) {
  if (p.g.n_threads == 0) return QSV_OK;
  const dim3 grid(ceil_div_u32(p.g.n_threads, kBlock));
  hipLaunchKernelGGL((k_apply_small<T, CH, VEC, Op>), grid, dim3(kBlock), 0, s,
                     static_cast<const T*>(in), static_cast<T*>(out), p.g, op);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename T, int CH, template <typename, int> class OpT>
static int dispatch_vec_small(const Plan& p, const OpT<T, CH>& op, const void* in, void* out,
                              hipStream_t s) {
  if constexpr (std::is_same<T, double>::value) {
    if (p.vec == 2) return launch_small<T, CH, 2>(p, op, in, out, s);
  } else {
    if (p.vec == 4) return launch_small<T, CH, 4>(p, op, in, out, s);
  }
  return launch_small<T, CH, 1>(p, op, in, out, s);
}

template <typename T, int CH>
static void fill_offsets(const Plan& p, int64_t* oin, int64_t* oout) {
  for (int i = 0; i < CH; ++i) oin[i] = i < p.din ? p.off_in[i] : 0;
  for (int o = 0; o < CH; ++o) oout[o] = o < p.dout ? p.off_out[o] : 0;
}

template <typename T, int CH>
static int run_inline(const Plan& p, const double* ptm, const void* in, void* out,
                      hipStream_t s) {
  InlineOp<T, CH> op;
  memset(&op, 0, sizeof(op));
  for (int o = 0; o < p.dout; ++o)
    for (int i = 0; i < p.din; ++i) op.ptm[o * CH + i] = T(ptm[int64_t(o) * p.din + i]);
  fill_offsets<T, CH>(p, op.off_in, op.off_out);
  return dispatch_vec_small<T, CH, InlineOp>(p, op, in, out, s);
}

template <typename T, int CH>
static int run_ptr(const Plan& p, const void* ptm_dev, const void* in, void* out,
                   hipStream_t s) {
  PtrOp<T, CH> op;
  memset(&op, 0, sizeof(op));
  fill_offsets<T, CH>(p, op.off_in, op.off_out);
  op.ptm = static_cast<const T*>(ptm_dev);
  op.ld = p.din;
  return dispatch_vec_small<T, CH, PtrOp>(p, op, in, out, s);
}

template <typename T, int CH, int VEC>
static int launch_table(const Plan& p, const TableOp<T>& op, const void* in, void* out,
                        hipStream_t s) {
  if (p.g.n_threads == 0) return QSV_OK;
  const dim3 grid(ceil_div_u32(p.g.n_threads, kBlock));
  hipLaunchKernelGGL((k_apply_table<T, CH, VEC>), grid, dim3(kBlock), 0, s,
                     static_cast<const T*>(in), static_cast<T*>(out), p.g, op);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename T>
static int run_table(const Plan& p, const TableOp<T>& op, const void* in, void* out,
                     hipStream_t s) {
  if constexpr (std::is_same<T, double>::value) {
    if (p.vec == 2) return launch_table<T, 16, 2>(p, op, in, out, s);
  } else {
    if (p.vec == 4) return launch_table<T, 16, 4>(p, op, in, out, s);
  }
  return launch_table<T, 16, 1>(p, op, in, out, s);
}

static inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

static bool fits_inline(int64_t din, int64_t dout) {
  return din <= kInline && dout <= kInline;
}

// bytes of the device table for a PTM of dout x din (0 when inline)
static int64_t table_bytes(int dtype, int64_t din, int64_t dout, bool with_ptm) {
  if (fits_inline(din, dout)) return 0;
  const int64_t dinp = round_up(din, 16), doutp = round_up(dout, 16);
  int64_t b = (dinp + doutp) * 8;
  if (with_ptm) b += dinp * doutp * dtype_size(dtype);
  return round_up(b, 256);
}

template <typename T>
static int run_host_table(const Plan& p, const double* ptm_f64, void* workspace,
                          const void* in, void* out, hipStream_t s) {
  const int64_t need = table_bytes(sizeof(T) == 8 ? QSV_F64 : QSV_F32, p.din, p.dout, true);
  const int64_t dinp = round_up(p.din, 16), doutp = round_up(p.dout, 16);
  std::vector<unsigned char> host(need, 0);
  int64_t* oin = reinterpret_cast<int64_t*>(host.data());
  int64_t* oout = oin + dinp;
  for (int i = 0; i < p.din; ++i) oin[i] = p.off_in[i];
  for (int o = 0; o < p.dout; ++o) oout[o] = p.off_out[o];
  T* P = reinterpret_cast<T*>(host.data() + (dinp + doutp) * 8);
  for (int o = 0; o < p.dout; ++o)
    for (int i = 0; i < p.din; ++i) P[o * dinp + i] = T(ptm_f64[int64_t(o) * p.din + i]);
  // pageable source: the runtime copies it to staging before returning
  QSV_CHECK_HIP(hipMemcpyAsync(workspace, host.data(), need, hipMemcpyHostToDevice, s));
  const unsigned char* wb = static_cast<const unsigned char*>(workspace);
  TableOp<T> op{reinterpret_cast<const int64_t*>(wb),
                reinterpret_cast<const int64_t*>(wb) + dinp,
                reinterpret_cast<const T*>(wb + (dinp + doutp) * 8), int32_t(dinp), 1};
  return run_table<T>(p, op, in, out, s);
}

template <typename T>
static int run_dev_table(const Plan& p, const void* ptm_dev, void* workspace, const void* in,
                         void* out, hipStream_t s) {
  const int64_t dinp = round_up(p.din, 16), doutp = round_up(p.dout, 16);
  std::vector<int64_t> host(dinp + doutp, 0);
  for (int i = 0; i < p.din; ++i) host[i] = p.off_in[i];
  for (int o = 0; o < p.dout; ++o) host[dinp + o] = p.off_out[o];
  QSV_CHECK_HIP(
      hipMemcpyAsync(workspace, host.data(), host.size() * 8, hipMemcpyHostToDevice, s));
  const int64_t* wb = static_cast<const int64_t*>(workspace);
  TableOp<T> op{wb, wb + dinp, static_cast<const T*>(ptm_dev), p.din, 0};
  return run_table<T>(p, op, in, out, s);
}

}  // namespace qsv

using namespace qsv;

extern "C" int64_t qsv_apply_ptm_workspace_bytes(int dtype, int k, const int32_t* dims_in_t,
                                                 const int32_t* dims_out_t) {
  if (dtype_size(dtype) == 0 || k < 1 || k > QSV_MAX_TARGETS || !dims_in_t || !dims_out_t) {
    set_error("qsv_apply_ptm_workspace_bytes: bad dtype, k or dims");
    return QSV_EINVAL;
  }
  int64_t din = 1, dout = 1;
  for (int j = 0; j < k; ++j) {
    din *= dims_in_t[j];
    dout *= dims_out_t[j];
  }
  return table_bytes(dtype, din, dout, true);
}

extern "C" int qsv_apply_ptm(int dtype, const void* in, void* out, int rank,
                             const int64_t* dims, int k, const int32_t* targets,
                             const int32_t* dims_out_t, const double* ptm_f64,
                             void* workspace, int64_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_plan(dtype, rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  QSV_CHECK_ARG(in && out && ptm_f64, "null pointer argument");
  QSV_CHECK_ARG(in != out, "in and out must not alias");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fits_inline(p.din, p.dout)) {
    const bool small = std::max(p.din, p.dout) <= 4;
    if (dtype == QSV_F64)
      return small ? run_inline<double, 4>(p, ptm_f64, in, out, s)
                   : run_inline<double, 16>(p, ptm_f64, in, out, s);
    return small ? run_inline<float, 4>(p, ptm_f64, in, out, s)
                 : run_inline<float, 16>(p, ptm_f64, in, out, s);
  }
  const int64_t need = table_bytes(dtype, p.din, p.dout, true);
  QSV_CHECK_ARG(workspace && workspace_bytes >= need,
                "a %d x %d PTM needs a workspace of %lld bytes", p.dout, p.din,
                (long long)need);
  return dtype == QSV_F64 ? run_host_table<double>(p, ptm_f64, workspace, in, out, s)
                          : run_host_table<float>(p, ptm_f64, workspace, in, out, s);
}

extern "C" int qsv_apply_ptm_dev(int dtype, const void* in, void* out, int rank,
                                 const int64_t* dims, int k, const int32_t* targets,
                                 const int32_t* dims_out_t, const void* ptm_dev,
                                 void* workspace, int64_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_plan(dtype, rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  QSV_CHECK_ARG(in && out && ptm_dev, "null pointer argument");
  QSV_CHECK_ARG(in != out, "in and out must not alias");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fits_inline(p.din, p.dout)) {
    const bool small = std::max(p.din, p.dout) <= 4;
    if (dtype == QSV_F64)
      return small ? run_ptr<double, 4>(p, ptm_dev, in, out, s)
                   : run_ptr<double, 16>(p, ptm_dev, in, out, s);
    return small ? run_ptr<float, 4>(p, ptm_dev, in, out, s)
                 : run_ptr<float, 16>(p, ptm_dev, in, out, s);
  }
  const int64_t need = table_bytes(dtype, p.din, p.dout, false);
  QSV_CHECK_ARG(workspace && workspace_bytes >= need,
                "offset tables of a %d x %d PTM need a workspace of %lld bytes", p.dout,
                p.din, (long long)need);
  return dtype == QSV_F64 ? run_dev_table<double>(p, ptm_dev, workspace, in, out, s)
                          : run_dev_table<float>(p, ptm_dev, workspace, in, out, s);
}
