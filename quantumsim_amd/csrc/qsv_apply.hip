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
#include "qsv_geom.h"
#include "qsv_fused.h"

#include <algorithm>
#include <stdlib.h>
#include <string.h>
#include <type_traits>
#include <vector>

namespace qsv {

constexpr int kInline = 16;  // PTMs up to 16 x 16 travel in kernel arguments
// --- operand sources.  p(o, i) is only called with o < dout, i < din.
// PTM padded to CH x CH (zeros) and both offset tables, by value.
template <typename T, int CH>
struct InlineOp {
  T ptm[CH * CH];
  int64_t off_in[CH];
  int64_t off_out[CH];
  static constexpr bool kPadded = true;
  __device__ __forceinline__ T p(int o, int i, uint32_t) const { return ptm[o * CH + i]; }
  __device__ __forceinline__ int64_t oin(int i) const { return off_in[i]; }
  __device__ __forceinline__ int64_t oout(int o) const { return off_out[o]; }
};

// Offsets by value, PTM (dout x din, unpadded) in device memory.
template <typename T, int CH>
struct PtrOp {
  int64_t off_in[CH];
  int64_t off_out[CH];
  const T* __restrict__ ptm;
  int64_t pstride;  // PTM elements between batch entries (0: one shared PTM)
  int32_t ld;
  static constexpr bool kPadded = false;
  __device__ __forceinline__ T p(int o, int i, uint32_t b) const {
    return ptm[int64_t(b) * pstride + o * ld + i];
  }
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

// ---------------------------------------------------------------- small
// din, dout <= CH: all D_in operands live in registers, outputs stream out.
template <typename T, int CH, int VEC, bool NT, bool FULL, class Op>
__global__ void __launch_bounds__(kBlock)
    k_apply_small(const T* __restrict__ in, T* __restrict__ out, const Geom g, const Op op) {
  typedef typename VT<T, VEC>::t V;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= g.n_threads) return;
  const uint32_t b = blockIdx.y;
  in += int64_t(b) * g.bs_in;
  out += int64_t(b) * g.bs_out;
  int64_t bin, bout;
  orbit_base(g, t * VEC, bin, bout);

  // FULL: din == dout == CH, no per-index guards (all CH loads issue back
  // to back in one basic block)
  V v[CH];
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    if (FULL || i < g.din)
      v[i] = gload<NT>(reinterpret_cast<const V*>(in + bin + op.oin(i)));
    else
      v[i] = V(0);
  }
#pragma unroll
  for (int o = 0; o < CH; ++o) {
    if (FULL || o < g.dout) {
      V acc = V(0);
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (FULL || Op::kPadded || i < g.din) acc += op.p(o, i, b) * v[i];
      }
      gstore<NT>(reinterpret_cast<V*>(out + bout + op.oout(o)), acc);
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

// ------------------------------------------------------- table, batched
// The table kernel for PTMs in device memory of any size (qutrit pairs 81 x
// 81, qubit-qutrit 36 x 36, ...), batched (grid.y = state, each state's PTM
// pstride elements apart, 0 = shared), with the entry offsets decoded by
// every block into LDS from the targets' dims and strides: no host table,
// no copy, so it is safe under graph capture.
struct TableGeo {
  int32_t k;
  int32_t din_t[QSV_MAX_TARGETS], dout_t[QSV_MAX_TARGETS];
  int64_t sin_t[QSV_MAX_TARGETS], sout_t[QSV_MAX_TARGETS];
};
constexpr int kTableMaxD = 1024;  // entries per side the LDS offsets hold

// Offset of PTM index e on one side: mixed radix over the targets in PTM
// axis order, the last target fastest.
__host__ __device__ inline int64_t table_offset(int k, const int32_t* d, const int64_t* st,
                                                int e) {
  int64_t off = 0;
  for (int j = k - 1; j >= 0; --j) {
    off += int64_t(e % d[j]) * st[j];
    e /= d[j];
  }
  return off;
}

template <typename T, int CH, int VEC>
__global__ void __launch_bounds__(kBlock)
    k_apply_table_b(const T* __restrict__ in, T* __restrict__ out, const Geom g,
                    const TableGeo tg, const T* __restrict__ ptm, const int64_t pstride) {
  typedef typename VT<T, VEC>::t V;
  __shared__ int64_t so_in[kTableMaxD], so_out[kTableMaxD];
  for (int e = threadIdx.x; e < g.din; e += kBlock)
    so_in[e] = table_offset(tg.k, tg.din_t, tg.sin_t, e);
  for (int e = threadIdx.x; e < g.dout; e += kBlock)
    so_out[e] = table_offset(tg.k, tg.dout_t, tg.sout_t, e);
  __syncthreads();
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= g.n_threads) return;
  const uint32_t b = blockIdx.y;
  in += int64_t(b) * g.bs_in;
  out += int64_t(b) * g.bs_out;
  ptm += int64_t(b) * pstride;
  int64_t bin, bout;
  orbit_base(g, t * VEC, bin, bout);
  for (int oc = 0; oc < g.dout; oc += CH) {
    V acc[CH];
#pragma unroll
    for (int o = 0; o < CH; ++o) acc[o] = V(0);
    for (int ic = 0; ic < g.din; ic += CH) {
      V v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i)
        v[i] = ic + i < g.din ? *reinterpret_cast<const V*>(in + bin + so_in[ic + i]) : V(0);
#pragma unroll
      for (int o = 0; o < CH; ++o) {
        if (oc + o < g.dout) {
          const T* prow = ptm + int64_t(oc + o) * g.din + ic;
#pragma unroll
          for (int i = 0; i < CH; ++i)
            if (ic + i < g.din) acc[o] += prow[i] * v[i];
        }
      }
    }
#pragma unroll
    for (int o = 0; o < CH; ++o)
      if (oc + o < g.dout) *reinterpret_cast<V*>(out + bout + so_out[oc + o]) = acc[o];
  }
}


// ------------------------------------------------------------ LDS tiled
// For targets near the innermost axis the orbit entries of neighbouring
// lanes are strided, so direct loads touch many cache lines per
// instruction.  Here a block owns a "tile": up to 256 orbits whose entries
// form a few long contiguous runs (one run per combination of the targets
// outside the tile's region).  The runs are copied into LDS with coalesced
// 16-byte loads, each thread contracts one orbit out of LDS, and the output
// runs are written back coalesced.  LDS rows are padded by one element per
// 16 so that orbits 16 elements apart fall into different banks.
constexpr int kMaxRuns = 16;

struct TileGeom {
  uint32_t n_tiles;
  int32_t ntg;  // tile-index groups (global), innermost first
  FastDiv tdiv[kMaxGroups];
  int64_t ts_in[kMaxGroups], ts_out[kMaxGroups];
  int32_t runs_in, runs_out;  // runs per tile (combos of targets outside the region)
  int32_t L_in, L_out;        // run lengths, elements
  int64_t run_in[kMaxRuns], run_out[kMaxRuns];  // run offsets from the tile base
  int32_t orbits;             // orbits per tile (<= 256)
  FastDiv pdiv;               // orbits per region block
  int32_t R_in, R_out;        // region block sizes
  int32_t nrg;                // bystander groups inside the region
  FastDiv rdiv[kMaxGroups];
  int32_t rs_in[kMaxGroups], rs_out[kMaxGroups];
  int32_t din, dout;
  int32_t slot_in[16], slot_out[16];  // LDS offsets of the PTM indices
  int32_t in_place;
  int64_t bs_in, bs_out;  // element stride between batch entries (grid.y)
};

__host__ __device__ __forceinline__ int lds_pad(int e) { return e + (e >> 4); }

// global bases of tile `m`
__host__ __device__ __forceinline__ void tile_base(const TileGeom& tg, uint32_t m, int64_t& gin,
                                                   int64_t& gout) {
  gin = 0;
  gout = 0;
#pragma unroll
  for (int j = 0; j < kMaxGroups; ++j) {
    if (j >= tg.ntg) break;
    if (j == tg.ntg - 1) {
      gin += int64_t(m) * tg.ts_in[j];
      gout += int64_t(m) * tg.ts_out[j];
    } else {
      const uint32_t q = fdiv(m, tg.tdiv[j]);
      const uint32_t r = m - q * tg.tdiv[j].d;
      gin += int64_t(r) * tg.ts_in[j];
      gout += int64_t(r) * tg.ts_out[j];
      m = q;
    }
  }
}

// LDS bases (before padding) of the orbit owned by thread `tid` of a tile
__host__ __device__ __forceinline__ void tile_orbit_base(const TileGeom& tg, int tid, int& bin,
                                                         int& bout) {
  const uint32_t xc = fdiv(uint32_t(tid), tg.pdiv);
  uint32_t rb = uint32_t(tid) - xc * tg.pdiv.d;
  bin = int(xc) * tg.R_in;
  bout = int(xc) * tg.R_out;
#pragma unroll
  for (int j = 0; j < kMaxGroups; ++j) {
    if (j >= tg.nrg) break;
    if (j == tg.nrg - 1) {
      bin += int(rb) * tg.rs_in[j];
      bout += int(rb) * tg.rs_out[j];
    } else {
      const uint32_t q = fdiv(rb, tg.rdiv[j]);
      const uint32_t r = rb - q * tg.rdiv[j].d;
      bin += int(r) * tg.rs_in[j];
      bout += int(r) * tg.rs_out[j];
      rb = q;
    }
  }
}

template <typename T, int VEC, bool NT, int CH, int ORB, bool FULL, class Op>
__global__ void __launch_bounds__(kBlock)
    k_apply_lds(const T* __restrict__ in, T* __restrict__ out, const TileGeom tg, const Op op) {
  typedef typename VT<T, VEC>::t V;
  extern __shared__ __align__(16) unsigned char qsv_smem[];
  T* lin = reinterpret_cast<T*>(qsv_smem);
  T* lout = tg.in_place ? lin : lin + lds_pad(tg.runs_in * tg.L_in) + 1;
  const int tid = threadIdx.x;
  const uint32_t b = blockIdx.y;
  in += int64_t(b) * tg.bs_in;
  out += int64_t(b) * tg.bs_out;

  int64_t gin, gout;
  tile_base(tg, blockIdx.x, gin, gout);

  // stage the input runs: kStage vectors per thread are loaded before any
  // is written to LDS, so that many loads are in flight per wave
  constexpr int kStage = 4;
  for (int a = 0; a < tg.runs_in; ++a) {
    const T* src = in + gin + tg.run_in[a];
    const int base = a * tg.L_in;
    for (int e0 = tid * VEC; e0 < tg.L_in; e0 += kStage * kBlock * VEC) {
      V val[kStage];
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int e = e0 + u * kBlock * VEC;
        if (e < tg.L_in) val[u] = gload<NT>(reinterpret_cast<const V*>(src + e));
      }
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int e = e0 + u * kBlock * VEC;
        if (e < tg.L_in) {
          if constexpr (VEC == 1) {
            lin[lds_pad(base + e)] = val[u];
          } else {
#pragma unroll
            for (int c = 0; c < VEC; ++c) lin[lds_pad(base + e + c)] = val[u][c];
          }
        }
      }
    }
  }
  __syncthreads();

  // contract ORB orbits per thread: orbits tid, tid + 256, ...
  int bout[ORB];
  T v[ORB][CH];
#pragma unroll
  for (int r = 0; r < ORB; ++r) {
    const int j = tid + r * kBlock;
    bout[r] = 0;
    if (j < tg.orbits) {
      int bin;
      tile_orbit_base(tg, j, bin, bout[r]);
#pragma unroll
      for (int i = 0; i < CH; ++i)
        v[r][i] = (FULL || i < tg.din) ? lin[lds_pad(bin + tg.slot_in[i])] : T(0);
    }
  }
  if (tg.in_place) __syncthreads();
#pragma unroll
  for (int r = 0; r < ORB; ++r) {
    if (tid + r * kBlock < tg.orbits) {
#pragma unroll
      for (int o = 0; o < CH; ++o) {
        if (FULL || o < tg.dout) {
          T acc = T(0);
#pragma unroll
          for (int i = 0; i < CH; ++i) {
            if (FULL || Op::kPadded || i < tg.din) acc += op.p(o, i, b) * v[r][i];
          }
          lout[lds_pad(bout[r] + tg.slot_out[o])] = acc;
        }
      }
    }
  }
  __syncthreads();

  // write the output runs
  for (int a = 0; a < tg.runs_out; ++a) {
    T* dst = out + gout + tg.run_out[a];
    const int base = a * tg.L_out;
    for (int e = tid * VEC; e < tg.L_out; e += kBlock * VEC) {
      V val;
      if constexpr (VEC == 1) {
        val = lout[lds_pad(base + e)];
      } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) val[c] = lout[lds_pad(base + e + c)];
      }
      gstore<NT>(reinterpret_cast<V*>(dst + e), val);
    }
  }
}

// ------------------------------------------------------------ host side
// Combined in+out bytes above which global accesses use the nontemporal
// hint (QSV_NT_BYTES overrides it).
int64_t nt_threshold_bytes() {
  static int64_t v = [] {
    const char* e = getenv("QSV_NT_BYTES");
    return e ? int64_t(atoll(e)) : (int64_t(256) << 20);
  }();
  return v;
}

struct Plan {
  Geom g;
  uint32_t batch;  // states processed by one launch (grid.y)
  int64_t n_in, n_out;
  int vec;
  bool nt;  // streaming (nontemporal) global accesses
  int din, dout;
  std::vector<int64_t> off_in, off_out;
  // per-axis description kept for the LDS-tiled planner
  int rank, k;
  int32_t targets[QSV_MAX_TARGETS];
  int64_t dims_in[QSV_MAX_RANK], dims_out[QSV_MAX_RANK];
  int64_t sin[QSV_MAX_RANK], sout[QSV_MAX_RANK];
  bool is_target[QSV_MAX_RANK];
};

// in_strides / out_strides: element strides per axis (NULL = C-contiguous).
// Strided layouts must keep C order without overlap: stride[q] >=
// stride[q+1] * dims[q+1]; gaps (padding) are allowed.
static int make_plan(int dtype, int rank, const int64_t* dims, const int64_t* in_strides,
                     const int64_t* out_strides, int k, const int32_t* targets,
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
    sin[q] = in_strides ? in_strides[q] : acc_in;
    sout[q] = out_strides ? out_strides[q] : acc_out;
    QSV_CHECK_ARG(sin[q] >= acc_in && sout[q] >= acc_out,
                  "strides of axis %d overlap the axes below it", q);
    acc_in = sin[q] * dims[q];
    acc_out = sout[q] * dout_full[q];
    QSV_CHECK_ARG(acc_in < (int64_t(1) << 40) && acc_out < (int64_t(1) << 40),
                  "Pauli vector too large");
  }
  // extents: one past the largest address
  p.n_in = acc_in;
  p.n_out = acc_out;
  // vectors that cannot stay resident in the last-level caches between
  // gates are streamed with nontemporal accesses
  p.nt = (acc_in + acc_out) * dtype_size(dtype) > nt_threshold_bytes();
  p.rank = rank;
  p.k = k;
  for (int j = 0; j < k; ++j) p.targets[j] = targets[j];
  for (int q = 0; q < rank; ++q) {
    p.dims_in[q] = dims[q];
    p.dims_out[q] = dout_full[q];
    p.sin[q] = sin[q];
    p.sout[q] = sout[q];
    p.is_target[q] = is_target[q];
  }

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
    if (is_target[q] || dims[q] == 1) {
      --q;
      continue;
    }
    // a group: adjacent bystander axes whose strides nest without gaps
    int64_t size = dims[q];
    const int64_t s_in = sin[q], s_out = sout[q];
    --q;
    while (q >= 0 && !is_target[q] && sin[q] == s_in * size && sout[q] == s_out * size) {
      size *= dims[q];
      --q;
    }
    QSV_CHECK_ARG(size < (int64_t(1) << 32), "bystander group too large");
    QSV_CHECK_ARG(ng < kMaxGroups, "too many bystander groups");
    g.gs_in[ng] = s_in;
    g.gs_out[ng] = s_out;
    g.gdiv[ng] = make_fastdiv(uint32_t(size));
    M *= size;
    ++ng;
  }
  g.ngroups = ng;
  g.bs_in = p.n_in;
  g.bs_out = p.n_out;
  p.batch = 1;
  // vector width along the innermost group when it is unit-stride: 16-byte
  // accesses.  All other offsets are multiples of that group's size, so the
  // vector accesses stay aligned.
  int vec = 1;
  if (ng > 0 && g.gs_in[0] == 1 && g.gs_out[0] == 1) {
    // f64: 2 x 8 B; f32: 4 x 4 B (a float2 variant needs 224 AGPRs on gfx950
    // and is not instantiated).  Every other stride and target offset must
    // be a multiple of the width for the vector accesses to stay aligned.
    const int want = dtype == QSV_F64 ? 2 : 4;
    bool aligned = g.gdiv[0].d % want == 0;
    for (int j = 1; j < ng; ++j)
      aligned = aligned && g.gs_in[j] % want == 0 && g.gs_out[j] % want == 0;
    for (int i = 0; i < p.din; ++i) aligned = aligned && p.off_in[i] % want == 0;
    for (int o = 0; o < p.dout; ++o) aligned = aligned && p.off_out[o] % want == 0;
    vec = aligned ? want : 1;
  }
  p.vec = vec;
  QSV_CHECK_ARG(M / vec < (int64_t(1) << 32) - kBlock, "too many orbits for 32-bit indexing");
  g.n_threads = uint32_t(M / vec);
  return QSV_OK;
}


// Byte threshold on the contiguous run below the innermost target under
// which the LDS-tiled kernel replaces the direct one (QSV_LDS_RUN_BYTES
// overrides it for tuning).  Measured on 14-qubit vectors: 512-byte runs
// (float64 bond (9,10)) stream faster staged (6.1 vs 5.65 TB/s); longer runs
// gain nothing from staging.
static int64_t lds_run_threshold() {
  static int64_t v = [] {
    const char* e = getenv("QSV_LDS_RUN_BYTES");
    return e ? int64_t(atoll(e)) : int64_t(1024);
  }();
  return v;
}

struct LdsPlan {
  TileGeom tg;
  uint32_t batch;
  int vec;
  bool nt;
  int ch;   // register tile width: 4 or 16
  int orb;  // orbits per thread: 1, 2 or 4
  size_t lds_bytes;
};

// Plans the LDS-tiled variant; returns false when the direct kernel is the
// better (or only) choice.
// Orbits per LDS tile: about 8 KB of staged input (measured best on MI355X
// for 14-qubit vectors against 4 and 16 KB), at most 4 orbits per thread.
// QSV_LDS_ORBITS overrides it for tuning.
constexpr int kLdsMaxOrbits = 4 * kBlock;
static int64_t lds_tile_orbits(int64_t dmax, int elem_bytes) {
  static const int64_t forced = [] {
    const char* e = getenv("QSV_LDS_ORBITS");
    return e ? int64_t(atoll(e)) : int64_t(0);
  }();
  constexpr int64_t tile_bytes = 8192;
  int64_t x = forced > 0 ? forced : tile_bytes / (dmax * elem_bytes);
  if (dmax > 4 && x > 2 * kBlock) x = 2 * kBlock;  // 16-wide register tiles: <= 2 orbits
  return x < 16 ? int64_t(16) : (x > kLdsMaxOrbits ? int64_t(kLdsMaxOrbits) : x);
}

// `tile` (may be NULL) forces the choice: QSV_KERNEL_DIRECT never tiles,
// QSV_KERNEL_LDS tiles whatever the run length, tile_orbits caps the tile.
static bool make_lds_plan(const Plan& p, int elem_bytes, LdsPlan& lp,
                          const qsv_tile* tile = nullptr) {
  if (p.din > 16 || p.dout > 16) return false;
  if (tile && tile->kernel == QSV_KERNEL_DIRECT) return false;
  int64_t cap = lds_tile_orbits(std::max(p.din, p.dout), elem_bytes);
  if (tile && tile->tile_orbits > 0) {
    cap = tile->tile_orbits;
    if (std::max(p.din, p.dout) > 4 && cap > 2 * kBlock) cap = 2 * kBlock;
    cap = std::min<int64_t>(std::max<int64_t>(cap, 16), kLdsMaxOrbits);
  }
  int t_in = -1;
  for (int j = 0; j < p.k; ++j) t_in = std::max(t_in, int(p.targets[j]));
  int64_t Z = 1;
  for (int q = t_in + 1; q < p.rank; ++q) Z *= p.dims_in[q];
  const bool forced = tile && tile->kernel == QSV_KERNEL_LDS;
  if (!forced && Z * elem_bytes >= lds_run_threshold()) return false;
  // axis c nests without a gap on the axes below it (input and output)
  auto contig = [&](int c) {
    if (c == p.rank - 1) return p.sin[c] == 1 && p.sout[c] == 1;
    return p.sin[c] == p.sin[c + 1] * p.dims_in[c + 1] &&
           p.sout[c] == p.sout[c + 1] * p.dims_out[c + 1];
  };
  for (int q = t_in; q < p.rank; ++q)
    if (!contig(q)) return false;

  TileGeom& tg = lp.tg;
  memset(&tg, 0, sizeof(tg));
  // grow the region [c0, rank) upward while a tile stays <= 256 orbits
  int c0 = t_in;
  int64_t P = Z;  // orbits per region block
  while (c0 > 0) {
    const int c = c0 - 1;
    if (!contig(c)) break;
    const int64_t np = p.is_target[c] ? P : P * p.dims_in[c];
    if (np > cap) break;
    P = np;
    c0 = c;
  }
  int64_t R_in = 1, R_out = 1;
  for (int q = c0; q < p.rank; ++q) {
    R_in *= p.dims_in[q];
    R_out *= p.dims_out[q];
  }
  // chunk of the bystander axis right above the region
  int64_t Tx = 1;
  if (c0 > 0 && !p.is_target[c0 - 1] && contig(c0 - 1)) {
    const int64_t d = p.dims_in[c0 - 1];  // a bystander next to the region
    for (int64_t t = d; t >= 1; --t)
      if (d % t == 0 && P * t <= cap) {
        Tx = t;
        break;
      }
  }
  const int64_t orbits = Tx * P;
  if (orbits < 1 || orbits > cap) return false;
  tg.orbits = int32_t(orbits);
  tg.pdiv = make_fastdiv(uint32_t(P));
  tg.R_in = int32_t(R_in);
  tg.R_out = int32_t(R_out);
  tg.L_in = int32_t(Tx * R_in);
  tg.L_out = int32_t(Tx * R_out);

  // runs: combinations of targets outside the region (ascending axis order)
  int outer[QSV_MAX_TARGETS];
  int nouter = 0;
  for (int q = 0; q < c0; ++q)
    if (p.is_target[q]) outer[nouter++] = q;
  int64_t runs_in = 1, runs_out = 1;
  for (int j = 0; j < nouter; ++j) {
    runs_in *= p.dims_in[outer[j]];
    runs_out *= p.dims_out[outer[j]];
  }
  if (runs_in > kMaxRuns || runs_out > kMaxRuns) return false;
  tg.runs_in = int32_t(runs_in);
  tg.runs_out = int32_t(runs_out);
  for (int64_t a = 0; a < runs_in; ++a) {
    int64_t r = a, off = 0;
    for (int j = nouter - 1; j >= 0; --j) {
      off += (r % p.dims_in[outer[j]]) * p.sin[outer[j]];
      r /= p.dims_in[outer[j]];
    }
    tg.run_in[a] = off;
  }
  for (int64_t a = 0; a < runs_out; ++a) {
    int64_t r = a, off = 0;
    for (int j = nouter - 1; j >= 0; --j) {
      off += (r % p.dims_out[outer[j]]) * p.sout[outer[j]];
      r /= p.dims_out[outer[j]];
    }
    tg.run_out[a] = off;
  }

  // LDS offset of every PTM index: run * L + offset inside the region
  auto slot = [&](int64_t i, bool is_out) {
    int64_t r = i, run = 0, run_mul = 1, off = 0;
    // digits of i, last target fastest
    int64_t digit[QSV_MAX_TARGETS];
    for (int j = p.k - 1; j >= 0; --j) {
      const int q = p.targets[j];
      const int64_t d = is_out ? p.dims_out[q] : p.dims_in[q];
      digit[j] = r % d;
      r /= d;
    }
    for (int jo = nouter - 1; jo >= 0; --jo) {
      const int q = outer[jo];
      int j = 0;
      while (p.targets[j] != q) ++j;
      run += digit[j] * run_mul;
      run_mul *= is_out ? p.dims_out[q] : p.dims_in[q];
    }
    for (int j = 0; j < p.k; ++j) {
      const int q = p.targets[j];
      if (q >= c0) off += digit[j] * (is_out ? p.sout[q] : p.sin[q]);
    }
    return run * (is_out ? tg.L_out : tg.L_in) + off;
  };
  tg.din = p.din;
  tg.dout = p.dout;
  for (int i = 0; i < p.din; ++i) tg.slot_in[i] = int32_t(slot(i, false));
  for (int o = 0; o < p.dout; ++o) tg.slot_out[o] = int32_t(slot(o, true));

  // bystander groups inside the region (innermost first)
  {
    int ng = 0;
    int q = p.rank - 1;
    while (q >= c0) {
      if (p.is_target[q]) {
        --q;
        continue;
      }
      int64_t size = 1;
      const int64_t s_in = p.sin[q], s_out = p.sout[q];
      while (q >= c0 && !p.is_target[q]) {
        size *= p.dims_in[q];
        --q;
      }
      if (size == 1) continue;
      tg.rdiv[ng] = make_fastdiv(uint32_t(size));
      tg.rs_in[ng] = int32_t(s_in);
      tg.rs_out[ng] = int32_t(s_out);
      ++ng;
    }
    tg.nrg = ng;
  }

  // tile-index groups: bystander axes above the region; the chunked axis
  // contributes dims / Tx at stride Tx * s
  {
    int ng = 0;
    int64_t n_tiles = 1;
    int q = c0 - 1;
    while (q >= 0) {
      if (p.is_target[q] || (p.dims_in[q] == 1 && q != c0 - 1)) {
        --q;
        continue;
      }
      // first axis of the group (the chunked axis contributes dims / Tx at
      // stride Tx * s); more axes join while their strides nest
      int64_t d0 = p.dims_in[q];
      int64_t s_in = p.sin[q], s_out = p.sout[q];
      if (q == c0 - 1) {
        d0 /= Tx;
        s_in *= Tx;
        s_out *= Tx;
      }
      int64_t size = d0;
      --q;
      while (q >= 0 && !p.is_target[q] && p.sin[q] == s_in * size &&
             p.sout[q] == s_out * size) {
        size *= p.dims_in[q];
        --q;
      }
      if (size == 1) continue;
      if (ng >= kMaxGroups) return false;
      if (size >= (int64_t(1) << 32)) return false;
      tg.tdiv[ng] = make_fastdiv(uint32_t(size));
      tg.ts_in[ng] = s_in;
      tg.ts_out[ng] = s_out;
      n_tiles *= size;
      ++ng;
    }
    tg.ntg = ng;
    if (n_tiles >= (int64_t(1) << 31)) return false;
    tg.n_tiles = uint32_t(n_tiles);
  }

  bool square = true;
  for (int j = 0; j < p.k; ++j)
    square = square && p.dims_in[p.targets[j]] == p.dims_out[p.targets[j]];
  tg.in_place = square ? 1 : 0;
  // vector width: the run lengths, run offsets and tile strides must all be
  // multiples of it (padded layouts may break this)
  const int want = elem_bytes == 8 ? 2 : 4;
  bool aligned = R_in % want == 0 && R_out % want == 0;
  for (int a = 0; a < tg.runs_in; ++a) aligned = aligned && tg.run_in[a] % want == 0;
  for (int a = 0; a < tg.runs_out; ++a) aligned = aligned && tg.run_out[a] % want == 0;
  for (int j = 0; j < tg.ntg; ++j)
    aligned = aligned && tg.ts_in[j] % want == 0 && tg.ts_out[j] % want == 0;
  lp.vec = aligned ? want : 1;
  const int64_t n_in = int64_t(tg.runs_in) * tg.L_in, n_out = int64_t(tg.runs_out) * tg.L_out;
  const int64_t elems = (n_in + n_in / 16 + 1) + (square ? 0 : n_out + n_out / 16 + 1);
  lp.lds_bytes = size_t(elems * elem_bytes);
  // the direct kernel's rule: nontemporal loads and stores for vectors that
  // do not stay in the caches between gates (the staged runs are whole lines)
  lp.nt = p.nt;
  lp.batch = p.batch;
  tg.bs_in = p.n_in;
  tg.bs_out = p.n_out;
  lp.ch = std::max(p.din, p.dout) <= 4 ? 4 : 16;
  lp.orb = int((orbits + kBlock - 1) / kBlock);
  if (lp.orb == 3) lp.orb = 4;
  if (lp.ch == 16 && lp.orb > 2) return false;
  if (lp.lds_bytes > 64 * 1024) return false;
  return true;
}

template <typename T, int VEC, int CH, int ORB, class Op>
static int launch_lds(const LdsPlan& lp, const Op& op, const void* in, void* out,
                      hipStream_t s) {
  if (lp.tg.n_tiles == 0 || lp.batch == 0) return QSV_OK;
  const dim3 grid(lp.tg.n_tiles, lp.batch);
  const bool full = lp.tg.din == CH && lp.tg.dout == CH;
  if (lp.nt && full)
    hipLaunchKernelGGL((k_apply_lds<T, VEC, true, CH, ORB, true, Op>), grid, dim3(kBlock),
                       lp.lds_bytes, s, static_cast<const T*>(in), static_cast<T*>(out), lp.tg,
                       op);
  else if (lp.nt)
    hipLaunchKernelGGL((k_apply_lds<T, VEC, true, CH, ORB, false, Op>), grid, dim3(kBlock),
                       lp.lds_bytes, s, static_cast<const T*>(in), static_cast<T*>(out), lp.tg,
                       op);
  else if (full)
    hipLaunchKernelGGL((k_apply_lds<T, VEC, false, CH, ORB, true, Op>), grid, dim3(kBlock),
                       lp.lds_bytes, s, static_cast<const T*>(in), static_cast<T*>(out), lp.tg,
                       op);
  else
    hipLaunchKernelGGL((k_apply_lds<T, VEC, false, CH, ORB, false, Op>), grid, dim3(kBlock),
                       lp.lds_bytes, s, static_cast<const T*>(in), static_cast<T*>(out), lp.tg,
                       op);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename T, int VEC, int CH, class Op>
static int dispatch_lds_orb(const LdsPlan& lp, const Op& op, const void* in, void* out,
                            hipStream_t s) {
  if constexpr (CH == 4) {
    if (lp.orb == 4) return launch_lds<T, VEC, CH, 4>(lp, op, in, out, s);
  }
  if (lp.orb == 2) return launch_lds<T, VEC, CH, 2>(lp, op, in, out, s);
  return launch_lds<T, VEC, CH, 1>(lp, op, in, out, s);
}

template <typename T, int CH, class Op>
static int dispatch_lds(const LdsPlan& lp, const Op& op, const void* in, void* out,
                        hipStream_t s) {
  if constexpr (std::is_same<T, double>::value) {
    if (lp.vec == 2) return dispatch_lds_orb<T, 2, CH>(lp, op, in, out, s);
  } else {
    if (lp.vec == 4) return dispatch_lds_orb<T, 4, CH>(lp, op, in, out, s);
  }
  return dispatch_lds_orb<T, 1, CH>(lp, op, in, out, s);
}

template <typename T, int CH, int VEC, class Op>
static int launch_small(const Plan& p, const Op& op, const void* in, void* out,
                        hipStream_t s) {
  if (p.g.n_threads == 0 || p.batch == 0) return QSV_OK;
  const dim3 grid(ceil_div_u32(p.g.n_threads, kBlock), p.batch);
  const bool full = p.din == CH && p.dout == CH;
  if (p.nt && full)
    hipLaunchKernelGGL((k_apply_small<T, CH, VEC, true, true, Op>), grid, dim3(kBlock), 0, s,
                       static_cast<const T*>(in), static_cast<T*>(out), p.g, op);
  else if (p.nt)
    hipLaunchKernelGGL((k_apply_small<T, CH, VEC, true, false, Op>), grid, dim3(kBlock), 0, s,
                       static_cast<const T*>(in), static_cast<T*>(out), p.g, op);
  else if (full)
    hipLaunchKernelGGL((k_apply_small<T, CH, VEC, false, true, Op>), grid, dim3(kBlock), 0, s,
                       static_cast<const T*>(in), static_cast<T*>(out), p.g, op);
  else
    hipLaunchKernelGGL((k_apply_small<T, CH, VEC, false, false, Op>), grid, dim3(kBlock), 0,
                       s, static_cast<const T*>(in), static_cast<T*>(out), p.g, op);
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
  LdsPlan lp;
  if (make_lds_plan(p, sizeof(T), lp)) return dispatch_lds<T, CH>(lp, op, in, out, s);
  fill_offsets<T, CH>(p, op.off_in, op.off_out);
  return dispatch_vec_small<T, CH, InlineOp>(p, op, in, out, s);
}

template <typename T, int CH>
static int run_ptr(const Plan& p, const void* ptm_dev, const void* in, void* out,
                   hipStream_t s, int64_t pstride = 0, const qsv_tile* tile = nullptr) {
  PtrOp<T, CH> op;
  memset(&op, 0, sizeof(op));
  op.ptm = static_cast<const T*>(ptm_dev);
  op.pstride = pstride;
  op.ld = p.din;
  LdsPlan lp;
  if (make_lds_plan(p, sizeof(T), lp, tile)) return dispatch_lds<T, CH>(lp, op, in, out, s);
  if (tile && tile->kernel == QSV_KERNEL_LDS) {
    set_error("no LDS tiling for this geometry");
    return QSV_ENOTSUP;
  }
  fill_offsets<T, CH>(p, op.off_in, op.off_out);
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

static bool use_fused(const Plan& p) {
  return p.din == p.dout && (p.din == 64 || p.din == 256);
}

// Plan of the MFMA kernels (index maps in qsv_fused.h).  False when the
// target offsets are not additive over the bits of the PTM index, which
// cannot happen for Pauli dimensions (perfect squares whose product is 64 or
// 256 are powers of 4) but is checked rather than assumed.
// batch > 1: states n_in / n_out elements apart, so vector accesses also
// need those strides to keep 16-byte alignment.
static bool make_fused_plan(const Plan& p, int elem_bytes, const void* in, const void* out,
                            FusedPlan& fp, uint32_t batch = 1) {
  if (!use_fused(p)) return false;
  const bool in_al = reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                     (batch <= 1 || (p.n_in * elem_bytes) % 16 == 0);
  const bool out_al = reinterpret_cast<uintptr_t>(out) % 16 == 0 &&
                      (batch <= 1 || (p.n_out * elem_bytes) % 16 == 0);
  const int D = p.din;
  for (int k = 0; k < D; ++k) {
    int64_t si = 0, so = 0;
    for (int b = 1; b < D; b <<= 1)
      if (k & b) {
        si += p.off_in[b];
        so += p.off_out[b];
      }
    if (si != p.off_in[k] || so != p.off_out[k]) return false;
  }
  const int VL = 16 / elem_bytes;
  auto width = [&](const std::vector<int64_t>& off, const int64_t* gs, bool aligned) {
    if (!aligned) return 1;
    for (int e = 1; e < VL; ++e)
      if (off[e] != e) return 1;
    for (int b = VL; b < D; b <<= 1)
      if (off[b] % VL) return 1;
    for (int q = 0; q < p.g.ngroups; ++q)
      if (gs[q] % VL) return 1;
    return VL;
  };
  fp.D = D;
  fp.vi = width(p.off_in, p.g.gs_in, in_al);
  fp.vo = width(p.off_out, p.g.gs_out, out_al);
  // float64 256 x 256: gathering vector stores out of the 64 accumulators
  // costs an extra 128 VGPRs (one wave per SIMD); scalar stores keep two
  if (D == 256 && elem_bytes == 8) fp.vo = 1;
  fp.n_orbits = uint32_t(int64_t(p.g.n_threads) * p.vec);
  // 64 x 64 with a unit-stride innermost bystander axis: lanes own p.vec
  // consecutive orbits (the planner's vector width already checks that every
  // stride and offset keeps such vectors aligned)
  fp.vn = 1;
  if (D == 64 && fp.vi == 1 && fp.vo == 1 && p.vec > 1 && in_al && out_al) fp.vn = p.vec;
  memset(&fp.offs, 0, sizeof(fp.offs));
  const int KS = D / 4;
  for (int u = 0; u < KS / fp.vi; ++u) fp.offs.in_u[u] = p.off_in[fused_k(fp.vi, fp.vi * u, 0)];
  for (int v = 0; v < KS / fp.vo; ++v) fp.offs.out_u[v] = p.off_out[fused_o(fp.vo, fp.vo * v, 0)];
  for (int j = 0; j < 4; ++j) {
    fp.offs.in_l[j] = p.off_in[fused_k(fp.vi, 0, j)];
    fp.offs.out_l[j] = p.off_out[fused_o(fp.vo, 0, j)];
  }
  return true;
}

// Dense 64 x 64 / 256 x 256 PTMs on MFMA: the workspace holds the PTM in
// fragment order.
template <typename T>
static int run_fused(const Plan& p, const FusedPlan& fp, const double* ptm_f64, void* workspace,
                     const void* in, void* out, hipStream_t s) {
  const int D = fp.D;
  std::vector<T> host(size_t(D) * D);
  fused_fragments<T>(D, int(sizeof(T)), fp.vi, fp.vo, ptm_f64, host.data());
  QSV_CHECK_HIP(hipMemcpyAsync(workspace, host.data(), host.size() * sizeof(T),
                               hipMemcpyHostToDevice, s));
  return launch_fused(sizeof(T) == 8 ? QSV_F64 : QSV_F32, fp, p.g, p.nt, in, out, workspace, s);
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


// TableGeo of a plan, checked against the plan's own offsets.
static int make_table_geo(const Plan& p, TableGeo& tg) {
  QSV_CHECK_ARG(p.din <= kTableMaxD && p.dout <= kTableMaxD,
                "PTM of %d x %d: the table kernel takes up to %d entries per side", p.dout,
                p.din, kTableMaxD);
  memset(&tg, 0, sizeof(tg));
  tg.k = p.k;
  for (int j = 0; j < p.k; ++j) {
    const int q = p.targets[j];
    tg.din_t[j] = int32_t(p.dims_in[q]);
    tg.dout_t[j] = int32_t(p.dims_out[q]);
    tg.sin_t[j] = p.sin[q];
    tg.sout_t[j] = p.sout[q];
  }
  for (int e = 0; e < p.din; ++e)
    QSV_CHECK_ARG(table_offset(tg.k, tg.din_t, tg.sin_t, e) == p.off_in[e],
                  "internal: table offsets disagree (input %d)", e);
  for (int e = 0; e < p.dout; ++e)
    QSV_CHECK_ARG(table_offset(tg.k, tg.dout_t, tg.sout_t, e) == p.off_out[e],
                  "internal: table offsets disagree (output %d)", e);
  return QSV_OK;
}

template <typename T>
static int run_table_b(const Plan& p, const void* ptm_dev, int64_t pstride, uint32_t batch,
                       const void* in, void* out, hipStream_t s) {
  TableGeo tg;
  const int rc = make_table_geo(p, tg);
  if (rc) return rc;
  if (p.g.n_threads == 0 || batch == 0) return QSV_OK;
  const dim3 grid(ceil_div_u32(p.g.n_threads, kBlock), batch);
  const T* pt = static_cast<const T*>(ptm_dev);
  if (p.vec == 16 / int(sizeof(T)))
    hipLaunchKernelGGL((k_apply_table_b<T, 16, 16 / sizeof(T)>), grid, dim3(kBlock), 0, s,
                       static_cast<const T*>(in), static_cast<T*>(out), p.g, tg, pt, pstride);
  else
    hipLaunchKernelGGL((k_apply_table_b<T, 16, 1>), grid, dim3(kBlock), 0, s,
                       static_cast<const T*>(in), static_cast<T*>(out), p.g, tg, pt, pstride);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
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
  return qsv_apply_ptm_strided(dtype, in, out, rank, dims, nullptr, nullptr, k, targets,
                               dims_out_t, ptm_f64, workspace, workspace_bytes, stream);
}

extern "C" int qsv_apply_ptm_strided(int dtype, const void* in, void* out, int rank,
                                     const int64_t* dims, const int64_t* in_strides,
                                     const int64_t* out_strides, int k, const int32_t* targets,
                                     const int32_t* dims_out_t, const double* ptm_f64,
                                     void* workspace, int64_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_plan(dtype, rank, dims, in_strides, out_strides, k, targets, dims_out_t, p);
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
  FusedPlan fp;
  if (make_fused_plan(p, dtype_size(dtype), in, out, fp))
    return dtype == QSV_F64 ? run_fused<double>(p, fp, ptm_f64, workspace, in, out, s)
                            : run_fused<float>(p, fp, ptm_f64, workspace, in, out, s);
  return dtype == QSV_F64 ? run_host_table<double>(p, ptm_f64, workspace, in, out, s)
                          : run_host_table<float>(p, ptm_f64, workspace, in, out, s);
}

extern "C" int qsv_apply_ptm_dev_strided(int dtype, const void* in, void* out, int rank,
                                         const int64_t* dims, const int64_t* in_strides,
                                         const int64_t* out_strides, int k,
                                         const int32_t* targets, const int32_t* dims_out_t,
                                         const void* ptm_dev, void* workspace,
                                         int64_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_plan(dtype, rank, dims, in_strides, out_strides, k, targets, dims_out_t, p);
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
  FusedPlan fp;
  if (make_fused_plan(p, dtype_size(dtype), in, out, fp)) {
    // dense 64 x 64: the kernel gathers the row-major PTM into fragment
    // order itself; 256 x 256: packed into fragment order in the workspace
    // by a kernel on the same stream.  No host copy either way.
    if (fp.D == 64)
      return launch_fused64_batched(dtype, fp, p.g, p.nt, in, out, ptm_dev, 0, 1, s);
    const int64_t need = int64_t(fp.D) * fp.D * dtype_size(dtype);
    QSV_CHECK_ARG(workspace && workspace_bytes >= need,
                  "a %d x %d PTM needs a workspace of %lld bytes", fp.D, fp.D, (long long)need);
    const int rp = launch_fused_pack(dtype, fp, ptm_dev, workspace, s);
    if (rp) return rp;
    return launch_fused(dtype, fp, p.g, p.nt, in, out, workspace, s);
  }
  // any other size: the table kernel with its offsets decoded on the device
  // (no host table, no copy)
  (void)workspace;
  (void)workspace_bytes;
  return dtype == QSV_F64 ? run_table_b<double>(p, ptm_dev, 0, 1, in, out, s)
                          : run_table_b<float>(p, ptm_dev, 0, 1, in, out, s);
}

extern "C" int qsv_apply_ptm_dev(int dtype, const void* in, void* out, int rank,
                                 const int64_t* dims, int k, const int32_t* targets,
                                 const int32_t* dims_out_t, const void* ptm_dev,
                                 void* workspace, int64_t workspace_bytes, void* stream) {
  return qsv_apply_ptm_dev_strided(dtype, in, out, rank, dims, nullptr, nullptr, k, targets,
                                   dims_out_t, ptm_dev, workspace, workspace_bytes, stream);
}

// ------------------------------------------------------ host emulation
// Replays the planner and the per-thread index arithmetic of the kernel
// path the GPU call would take, on the CPU in float64, checking every
// global and LDS access against its buffer and that every output element is
// written exactly once.  Diagnostic only: no GPU work.
namespace qsv {

// Every logical output entry written exactly once, padding never.
static int check_coverage(const Plan& p, const std::vector<int>& written, const char* who) {
  int64_t n_logical = 1;
  for (int q = 0; q < p.rank; ++q) n_logical *= p.dims_out[q];
  int64_t total = 0;
  for (int64_t a = 0; a < p.n_out; ++a) total += written[a];
  QSV_CHECK_ARG(total == n_logical, "%s: %lld writes for %lld entries", who, (long long)total,
                (long long)n_logical);
  std::vector<int64_t> idx(p.rank, 0);
  for (int64_t e = 0; e < n_logical; ++e) {
    int64_t r = e, a = 0;
    for (int q = p.rank - 1; q >= 0; --q) {
      a += (r % p.dims_out[q]) * p.sout[q];
      r /= p.dims_out[q];
    }
    QSV_CHECK_ARG(written[a] == 1, "%s: output %lld written %d times", who, (long long)a,
                  written[a]);
  }
  return QSV_OK;
}

static int emulate_direct(const Plan& p, const double* ptm, const double* in, double* out) {
  std::vector<int> written(p.n_out, 0);
  const int vec = p.vec;
  for (uint32_t t = 0; t < p.g.n_threads; ++t) {
    int64_t bin, bout;
    orbit_base(p.g, t * vec, bin, bout);
    for (int c = 0; c < vec; ++c) {
      std::vector<double> v(p.din);
      for (int i = 0; i < p.din; ++i) {
        const int64_t a = bin + p.off_in[i] + c;
        QSV_CHECK_ARG(a >= 0 && a < p.n_in, "direct: load %lld out of bounds", (long long)a);
        v[i] = in[a];
      }
      for (int o = 0; o < p.dout; ++o) {
        double acc = 0;
        for (int i = 0; i < p.din; ++i) acc += ptm[int64_t(o) * p.din + i] * v[i];
        const int64_t a = bout + p.off_out[o] + c;
        QSV_CHECK_ARG(a >= 0 && a < p.n_out, "direct: store %lld out of bounds", (long long)a);
        out[a] = acc;
        ++written[a];
      }
    }
  }
  return check_coverage(p, written, "direct");
}

static int emulate_lds(const Plan& p, const LdsPlan& lp, int elem_bytes, const double* ptm,
                       const double* in, double* out) {
  const TileGeom& tg = lp.tg;
  const int vec = lp.vec;
  const int64_t lds_elems = int64_t(lp.lds_bytes / elem_bytes);
  const int out_base = tg.in_place ? 0 : lds_pad(tg.runs_in * tg.L_in) + 1;
  std::vector<int> written(p.n_out, 0);
  QSV_CHECK_ARG(tg.L_in % vec == 0 && tg.L_out % vec == 0, "lds: run length not a multiple "
                "of the vector width");
  for (uint32_t tile = 0; tile < tg.n_tiles; ++tile) {
    std::vector<double> lds(lds_elems, 0.0);
    int64_t gin, gout;
    tile_base(tg, tile, gin, gout);
    for (int a = 0; a < tg.runs_in; ++a)
      for (int e = 0; e < tg.L_in; ++e) {
        const int64_t g = gin + tg.run_in[a] + e;
        QSV_CHECK_ARG(g >= 0 && g < p.n_in, "lds: load %lld out of bounds", (long long)g);
        QSV_CHECK_ARG(((gin + tg.run_in[a]) % vec) == 0, "lds: misaligned run");
        const int l = lds_pad(a * tg.L_in + e);
        QSV_CHECK_ARG(l >= 0 && l < lds_elems, "lds: LDS write %d out of bounds", l);
        lds[l] = in[g];
      }
    QSV_CHECK_ARG(tg.orbits <= kBlock * lp.orb, "lds: %d orbits exceed %d threads x %d",
                  tg.orbits, kBlock, lp.orb);
    QSV_CHECK_ARG(tg.din <= lp.ch && tg.dout <= lp.ch, "lds: PTM wider than the register tile");
    std::vector<std::vector<double>> vals(tg.orbits);
    std::vector<int> bouts(tg.orbits, 0);
    for (int tid = 0; tid < tg.orbits; ++tid) {
      int bin, bout;
      tile_orbit_base(tg, tid, bin, bout);
      bouts[tid] = bout;
      vals[tid].resize(tg.din);
      for (int i = 0; i < tg.din; ++i) {
        const int l = lds_pad(bin + tg.slot_in[i]);
        QSV_CHECK_ARG(l >= 0 && l < lds_pad(tg.runs_in * tg.L_in) + 1,
                      "lds: LDS read %d out of bounds", l);
        vals[tid][i] = lds[l];
      }
    }
    std::vector<int> lds_written(lds_elems, 0);
    for (int tid = 0; tid < tg.orbits; ++tid) {
      for (int o = 0; o < tg.dout; ++o) {
        double acc = 0;
        for (int i = 0; i < tg.din; ++i) acc += ptm[int64_t(o) * p.din + i] * vals[tid][i];
        const int l = out_base + lds_pad(bouts[tid] + tg.slot_out[o]);
        QSV_CHECK_ARG(l >= 0 && l < lds_elems, "lds: LDS result write %d out of bounds", l);
        lds[l] = acc;
        ++lds_written[l];
      }
    }
    for (int a = 0; a < tg.runs_out; ++a)
      for (int e = 0; e < tg.L_out; ++e) {
        const int64_t g = gout + tg.run_out[a] + e;
        QSV_CHECK_ARG(g >= 0 && g < p.n_out, "lds: store %lld out of bounds", (long long)g);
        const int l = out_base + lds_pad(a * tg.L_out + e);
        QSV_CHECK_ARG(l >= 0 && l < lds_elems, "lds: LDS read-out %d out of bounds", l);
        QSV_CHECK_ARG(lds_written[l] == 1, "lds: LDS element %d written %d times before "
                      "read-out", l, lds_written[l]);
        out[g] = lds[l];
        ++written[g];
      }
  }
  return check_coverage(p, written, "lds");
}

// Replays k_fused64 / k_fused256 lane by lane: the B-fragment loads (with
// their vector alignment and, for orbit vectors, the consecutive-orbit
// assumption), the fragment-ordered PTM reads, the MFMA contraction in
// hardware index order and the accumulator stores.
static int emulate_fused(const Plan& p, int elem_bytes, const double* ptm, const double* in,
                         double* out, uint32_t batch = 1) {
  FusedPlan fp;
  QSV_CHECK_ARG(make_fused_plan(p, elem_bytes, nullptr, nullptr, fp, batch),
                "fused: offsets not additive");
  const int D = fp.D, MB = D / 16, KS = D / 4;
  const int vi = fp.vi, vo = fp.vo, vn = fp.vn;
  std::vector<double> pf(int64_t(D) * D);
  fused_fragments<double>(D, elem_bytes, vi, vo, ptm, pf.data());
  {
    // the fragment table must hold every PTM entry exactly once
    std::vector<int> used(int64_t(D) * D, 0);
    for (int s = 0; s < KS; ++s)
      for (int mb = 0; mb < MB; ++mb)
        for (int lane = 0; lane < 64; ++lane) {
          int j, r;
          mfma_row_owner(elem_bytes, lane & 15, j, r);
          const int o = fused_o(vo, 4 * mb + r, j), k = fused_k(vi, s, lane >> 4);
          QSV_CHECK_ARG(o >= 0 && o < D && k >= 0 && k < D, "fused: map out of range");
          ++used[int64_t(o) * D + k];
        }
    for (int64_t e = 0; e < int64_t(D) * D; ++e)
      QSV_CHECK_ARG(used[e] == 1, "fused: PTM entry %lld used %d times", (long long)e, used[e]);
  }
  const int NB = fused_nb(elem_bytes, D, vn);
  const uint32_t cols = 16 * NB, n_orbits = fp.n_orbits;
  QSV_CHECK_ARG(n_orbits % vn == 0, "fused: orbit count not a multiple of the vector");
  const uint32_t ntiles = (n_orbits + cols - 1) / cols;
  std::vector<int> written(p.n_out, 0);
  std::vector<double> B(size_t(64) * NB * KS);
  // orbit of (tile, nb, n) and its base addresses; vn > 1 derives the bases
  // of a lane's vn orbits from the first one, as the kernel does
  auto bases = [&](uint32_t t, int nb, int n, bool clamp, int64_t& bin, int64_t& bout,
                   uint32_t& m) -> int {
    const uint32_t m0 = fused64_orbit_of(NB, vn, t, 0, n);
    m = fused64_orbit_of(NB, vn, t, nb, n);
    if (vn > 1) {
      const uint32_t mb0 = clamp && m0 >= n_orbits ? n_orbits - vn : m0;
      int64_t b0i, b0o, bi, bo;
      orbit_base(p.g, mb0, b0i, b0o);
      orbit_base(p.g, mb0 + nb, bi, bo);
      QSV_CHECK_ARG(bi == b0i + nb && bo == b0o + nb, "fused: orbits of a vector not adjacent");
      QSV_CHECK_ARG(b0i % vn == 0 && b0o % vn == 0, "fused: orbit vector misaligned");
      bin = bi;
      bout = bo;
    } else {
      orbit_base(p.g, clamp && m >= n_orbits ? n_orbits - 1 : m, bin, bout);
    }
    return QSV_OK;
  };
  for (uint32_t t = 0; t < ntiles; ++t) {
    for (int lane = 0; lane < 64; ++lane) {
      const int n = lane & 15, j = lane >> 4;
      for (int nb = 0; nb < NB; ++nb) {
        int64_t bin, bout;
        uint32_t m;
        if (bases(t, nb, n, true, bin, bout, m)) return QSV_EINVAL;
        for (int u = 0; u < KS / vi; ++u) {
          const int64_t a0 = bin + fp.offs.in_l[j] + fp.offs.in_u[u];
          QSV_CHECK_ARG((a0 - (vn > 1 ? nb : 0)) % std::max(vi, vn) == 0,
                        "fused: misaligned vector load");
          for (int e = 0; e < vi; ++e) {
            QSV_CHECK_ARG(a0 + e >= 0 && a0 + e < p.n_in, "fused: load %lld out of bounds",
                          (long long)(a0 + e));
            B[(size_t(lane) * NB + nb) * KS + u * vi + e] = in[a0 + e];
          }
        }
      }
    }
    for (int lane = 0; lane < 64; ++lane) {
      const int n = lane & 15, j = lane >> 4;
      for (int nb = 0; nb < NB; ++nb) {
        int64_t bin, bout;
        uint32_t m;
        if (bases(t, nb, n, false, bin, bout, m)) return QSV_EINVAL;
        if (m >= n_orbits) continue;
        for (int v = 0; v < KS / vo; ++v) {
          const int64_t a0 = bout + fp.offs.out_l[j] + fp.offs.out_u[v];
          QSV_CHECK_ARG((a0 - (vn > 1 ? nb : 0)) % std::max(vo, vn) == 0,
                        "fused: misaligned vector store");
          for (int e = 0; e < vo; ++e) {
            const int q = v * vo + e, mb = q >> 2, r = q & 3;
            const int row16 = elem_bytes == 8 ? j + 4 * r : 4 * j + r;
            // D[row][n] = sum over k-steps s and k-lanes jk of A[row][4 s + jk] B[4 s + jk][n]
            double acc = 0;
            for (int s = 0; s < KS; ++s)
              for (int jk = 0; jk < 4; ++jk) {
                const int64_t f = fused_frag_slot(D, s, mb, row16 + 16 * jk);
                QSV_CHECK_ARG(f >= 0 && f < int64_t(D) * D, "fused: fragment slot out of range");
                acc += pf[f] * B[(size_t(n + 16 * jk) * NB + nb) * KS + s];
              }
            const int64_t a = a0 + e;
            QSV_CHECK_ARG(a >= 0 && a < p.n_out, "fused: store %lld out of bounds", (long long)a);
            out[a] = acc;
            ++written[a];
          }
        }
      }
    }
  }
  return check_coverage(p, written, "fused");
}

}  // namespace qsv

extern "C" int qsv_emulate_apply_ptm(int dtype, const double* in, double* out, int rank,
                                     const int64_t* dims, const int64_t* in_strides,
                                     const int64_t* out_strides, int k, const int32_t* targets,
                                     const int32_t* dims_out_t, const double* ptm_f64,
                                     int32_t* path_out) {
  Plan p;
  const int rc = make_plan(dtype, rank, dims, in_strides, out_strides, k, targets, dims_out_t, p);
  if (rc) return rc;
  QSV_CHECK_ARG(in && out && ptm_f64, "null pointer argument");
  QSV_CHECK_ARG(p.n_in <= (int64_t(1) << 24) && p.n_out <= (int64_t(1) << 24),
                "emulation is for small vectors");
  const int es = dtype_size(dtype);
  LdsPlan lp;
  if (fits_inline(p.din, p.dout) && make_lds_plan(p, es, lp)) {
    if (path_out) *path_out = 2;
    return emulate_lds(p, lp, es, ptm_f64, in, out);
  }
  FusedPlan fp;
  if (!fits_inline(p.din, p.dout) && make_fused_plan(p, es, nullptr, nullptr, fp)) {
    if (path_out) *path_out = 4 | (fp.vi << 4) | (fp.vo << 8) | (fp.vn << 12);
    return emulate_fused(p, es, ptm_f64, in, out);
  }
  if (path_out) *path_out = fits_inline(p.din, p.dout) ? 1 : 3;
  return emulate_direct(p, ptm_f64, in, out);
}

static int check_tile(const qsv_tile* t) {
  if (!t) return QSV_OK;
  QSV_CHECK_ARG(t->kernel >= QSV_KERNEL_AUTO && t->kernel <= QSV_KERNEL_LDS,
                "tile: unknown kernel %d", t->kernel);
  QSV_CHECK_ARG(t->tile_orbits >= 0 && t->tile_orbits <= kLdsMaxOrbits,
                "tile: tile_orbits %d outside [0, %d]", t->tile_orbits, kLdsMaxOrbits);
  QSV_CHECK_ARG(t->grad_blocks >= 0 && t->grad_blocks <= 256,
                "tile: grad_blocks %d outside [0, 256]", t->grad_blocks);
  QSV_CHECK_ARG(t->reserved == 0, "tile: reserved field must be 0");
  return QSV_OK;
}

extern "C" int qsv_apply_ptm_batched(int dtype, const void* in, void* out, int batch, int rank,
                                     const int64_t* dims, int k, const int32_t* targets,
                                     const int32_t* dims_out_t, const void* ptm_dev,
                                     int64_t ptm_batch_stride, void* stream) {
  return qsv_apply_ptm_batched_tiled(dtype, in, out, batch, rank, dims, k, targets, dims_out_t,
                                     ptm_dev, ptm_batch_stride, nullptr, stream);
}

extern "C" int qsv_apply_ptm_batched_tiled(int dtype, const void* in, void* out, int batch,
                                           int rank, const int64_t* dims, int k,
                                           const int32_t* targets, const int32_t* dims_out_t,
                                           const void* ptm_dev, int64_t ptm_batch_stride,
                                           const qsv_tile* tile, void* stream) {
  if (const int rt = check_tile(tile)) return rt;
  Plan p;
  const int rc = make_plan(dtype, rank, dims, nullptr, nullptr, k, targets, dims_out_t, p);
  if (rc) return rc;
  QSV_CHECK_ARG(batch >= 0 && batch <= 65535, "batch %d outside [0, 65535]", batch);
  QSV_CHECK_ARG(in && out && ptm_dev, "null pointer argument");
  QSV_CHECK_ARG(ptm_batch_stride >= 0, "negative PTM batch stride");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the whole batch streams through the caches, so the nontemporal rule
  // applies to its total size (measured: 64 twelve-qubit float32 states,
  // 128 MB each, forward + adjoint 3-5 % faster than with the per-state rule)
  p.nt = (p.n_in + p.n_out) * int64_t(dtype_size(dtype)) * int64_t(batch > 0 ? batch : 1) >
         nt_threshold_bytes();
  if (!fits_inline(p.din, p.dout)) {
    // dense 64 x 64 (three qubits): MFMA kernel, each state's row-major PTM
    // gathered into fragment order by the kernel itself
    FusedPlan fp;
    if (p.din == 64 && p.dout == 64 &&
        make_fused_plan(p, dtype_size(dtype), in, out, fp, uint32_t(batch))) {
      QSV_CHECK_ARG(!tile || tile->kernel == QSV_KERNEL_AUTO,
                    "dense 64 x 64 PTMs take no launch shape");
      return launch_fused64_batched(dtype, fp, p.g, p.nt, in, out, ptm_dev, ptm_batch_stride,
                                    uint32_t(batch), s);
    }
    // any other size (qutrit pairs, qubit-qutrit pairs, ...): the table
    // kernel with the offsets decoded on the device
    QSV_CHECK_ARG(!tile || tile->kernel == QSV_KERNEL_AUTO,
                  "%d x %d PTMs take no launch shape", p.dout, p.din);
    return dtype == QSV_F64
               ? run_table_b<double>(p, ptm_dev, ptm_batch_stride, uint32_t(batch), in, out, s)
               : run_table_b<float>(p, ptm_dev, ptm_batch_stride, uint32_t(batch), in, out, s);
  }
  p.batch = uint32_t(batch);
  const bool small = std::max(p.din, p.dout) <= 4;
  if (dtype == QSV_F64)
    return small ? run_ptr<double, 4>(p, ptm_dev, in, out, s, ptm_batch_stride, tile)
                 : run_ptr<double, 16>(p, ptm_dev, in, out, s, ptm_batch_stride, tile);
  return small ? run_ptr<float, 4>(p, ptm_dev, in, out, s, ptm_batch_stride, tile)
               : run_ptr<float, 16>(p, ptm_dev, in, out, s, ptm_batch_stride, tile);
}

extern "C" int qsv_emulate_apply_ptm_tiled(int dtype, const double* in, double* out, int rank,
                                           const int64_t* dims, int k, const int32_t* targets,
                                           const int32_t* dims_out_t, const double* ptm_f64,
                                           const qsv_tile* tile, int32_t* path_out) {
  if (const int rt = check_tile(tile)) return rt;
  Plan p;
  const int rc = make_plan(dtype, rank, dims, nullptr, nullptr, k, targets, dims_out_t, p);
  if (rc) return rc;
  QSV_CHECK_ARG(in && out && ptm_f64, "null pointer argument");
  QSV_CHECK_ARG(p.n_in <= (int64_t(1) << 24) && p.n_out <= (int64_t(1) << 24),
                "emulation is for small vectors");
  const int es = dtype_size(dtype);
  if (!fits_inline(p.din, p.dout)) {
    // the batched 64 x 64 MFMA path (16-byte aligned states assumed, as
    // torch allocations are; batch > 1 so the state strides count too)
    FusedPlan fp;
    if (p.din == 64 && p.dout == 64 && make_fused_plan(p, es, nullptr, nullptr, fp, 2)) {
      QSV_CHECK_ARG(!tile || tile->kernel == QSV_KERNEL_AUTO,
                    "dense 64 x 64 PTMs take no launch shape");
      if (path_out) *path_out = 4 | (fp.vi << 4) | (fp.vo << 8) | (fp.vn << 12);
      return emulate_fused(p, es, ptm_f64, in, out, 2);
    }
    // the batched table kernel: offsets as the kernel decodes them
    QSV_CHECK_ARG(!tile || tile->kernel == QSV_KERNEL_AUTO,
                  "%d x %d PTMs take no launch shape", p.dout, p.din);
    TableGeo tg;
    const int rt = make_table_geo(p, tg);
    if (rt) return rt;
    if (path_out) *path_out = 3;
    return emulate_direct(p, ptm_f64, in, out);
  }
  LdsPlan lp;
  if (make_lds_plan(p, es, lp, tile)) {
    if (path_out) *path_out = 2;
    return emulate_lds(p, lp, es, ptm_f64, in, out);
  }
  if (tile && tile->kernel == QSV_KERNEL_LDS) {
    set_error("no LDS tiling for this geometry");
    return QSV_ENOTSUP;
  }
  if (path_out) *path_out = 1;
  return emulate_direct(p, ptm_f64, in, out);
}
