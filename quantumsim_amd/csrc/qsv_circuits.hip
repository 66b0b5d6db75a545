// Persistent kernels that run whole circuits: many qubit circuits of
// different widths and depths in one launch, the Pauli vector of each
// circuit resident in LDS (4^n entries: up to 7 qubits in float32 on
// gfx950's 160 KB of LDS per workgroup for the forward pass).
//
// A workgroup takes circuits from the circuit table (round robin over the
// grid) and, for each, runs its gates from the gate table:
//
//   forward:  S = |0..0><0..0|; per gate: save S to the trajectory (the
//             state before the gate), stage the PTM in LDS, contract every
//             orbit in place (an orbit's D entries are read into registers,
//             then overwritten; orbits are disjoint, so no other thread
//             touches them); at the end write the float64 populations
//             (the entries whose qubit digits are all |0><0| or |1><1|).
//   adjoint:  Y = dL/dS (the population gradient scattered onto those
//             entries); per gate, last first: load the state before the gate
//             from the trajectory, dL/dP[o][i] = sum over orbits of
//             Y[o] X[i] on v_mfma_f64_16x16x4 (lane (c, k) supplies entry c
//             of orbit k: four orbits per k-step for 16 x 16 PTMs, sixteen
//             for 4 x 4 ones packed as four 4 x 4 diagonal blocks; float64
//             products and sums, waves summed in a fixed order), then
//             Y <- P^T Y per orbit in place.
//
// LDS holds the states XOR-swizzled (circ_swz): the entries of an orbit sit
// 4^q apart, and without the swizzle the sixteen entries the MFMA operand
// loads of one orbit touch would share an LDS bank.
//
// All qubits are in the general(2) basis (Pauli dimension 4: two bits per
// digit), qubit 0 most significant.  Gates act on 1 or 2 qubits with PTMs
// of 4 x 4 or 16 x 16, layout (out..., in...) in target order, as
// qsv_apply_ptm.  No counterpart in the reference (one circuit per
// PauliVector object there); the per-gate arithmetic is its apply_ptm
// (quantumsim/pauli_vectors/numpy.py:50-64).
#include "qsv_common.h"
#include "qsv_mfma.h"

#include <algorithm>
#include <vector>

namespace qsv {

// 512 threads (8 waves) per circuit: one workgroup per CU (the adjoint's two
// 7-qubit float32 states and its wave partials take 148 KB of LDS), so the
// waves of that one circuit are all the latency hiding there is; 256
// threads measured 1.5 % slower on the 512-circuit step.
constexpr int kCircBlock = 512;
constexpr int kCircMaxQubits = 8;
// Largest state the adjoint's register prefetch covers: 7 qubits in float32
// (160 KB of LDS holds the adjoint's two states only up to there), 6 in
// float64.
template <typename T>
constexpr uint32_t kCircMaxEntries = sizeof(T) == 4 ? (1u << 14) : (1u << 12);

// Index of orbit m with zero 2-bit digits inserted at bit positions lo < hi
// (hi < 0: one target).
__host__ __device__ inline uint32_t circ_orbit_base(uint32_t m, int lo, int hi) {
  uint32_t t = ((m >> lo) << (lo + 2)) | (m & ((1u << lo) - 1u));
  if (hi >= 0) t = ((t >> hi) << (hi + 2)) | (t & ((1u << hi) - 1u));
  return t;
}

// Offset of PTM index i (target order: first target slowest) inside an orbit.
__host__ __device__ inline uint32_t circ_offset(int k, uint32_t i, int s0, int s1) {
  return k == 1 ? (i << s0) : (((i >> 2) << s0) | ((i & 3u) << s1));
}

// Pauli index of population j (qubit 0 = most significant bit of j): bit t
// of j becomes bit 2 t.
__host__ __device__ inline uint32_t circ_spread(uint32_t j) {
  uint32_t r = 0;
  for (int t = 0; t < 16; ++t) r |= ((j >> t) & 1u) << (2 * t);
  return r;
}

// LDS position of state entry a: the low 6 bits XORed with a hash of the
// higher ones (a bijection on [0, 4^n); entries below 64 stay in place).
__host__ __device__ inline uint32_t circ_swz(uint32_t a) {
  const uint32_t h = a >> 6;
  return a ^ ((h ^ (h >> 6) ^ (h >> 12)) & 63u);
}

struct GateGeom {
  int k, D, s0, s1, lo, hi;
  uint32_t n_orbits;
};

__host__ __device__ inline GateGeom circ_gate_geom(const qsv_gate& gd, int n) {
  GateGeom G;
  G.k = gd.k;
  G.D = gd.k == 1 ? 4 : 16;
  G.s0 = 2 * (n - 1 - gd.q0);
  G.s1 = gd.k == 2 ? 2 * (n - 1 - gd.q1) : 0;
  if (gd.k == 1) {
    G.lo = G.s0;
    G.hi = -1;
  } else {
    G.lo = G.s0 < G.s1 ? G.s0 : G.s1;
    G.hi = G.s0 < G.s1 ? G.s1 : G.s0;
  }
  G.n_orbits = (1u << (2 * n)) / uint32_t(G.D);
  return G;
}

// In-place contraction of every orbit of one gate (D fixed at compile time
// so the orbit's operands stay in registers).  TRANS: Y <- P^T Y.
template <typename T, int D, bool TRANS>
__device__ inline void circ_orbits(T* S, const T* P, const GateGeom& G, int tid) {
  for (uint32_t m = tid; m < G.n_orbits; m += kCircBlock) {
    const uint32_t base = circ_orbit_base(m, G.lo, G.hi);
    T xin[D];
#pragma unroll
    for (int i = 0; i < D; ++i) xin[i] = S[circ_swz(base + circ_offset(G.k, i, G.s0, G.s1))];
#pragma unroll
    for (int o = 0; o < D; ++o) {
      T acc = T(0);
#pragma unroll
      for (int i = 0; i < D; ++i) acc += (TRANS ? P[i * D + o] : P[o * D + i]) * xin[i];
      S[circ_swz(base + circ_offset(G.k, o, G.s0, G.s1))] = acc;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kCircBlock)
    k_circuits_fwd(const qsv_circuit* __restrict__ circs, int n_circ,
                   const qsv_gate* __restrict__ gates, const T* __restrict__ ptms,
                   T* __restrict__ traj, double* __restrict__ pops, int max_qubits) {
  extern __shared__ __align__(16) unsigned char qsv_circ_smem[];
  T* S = reinterpret_cast<T*>(qsv_circ_smem);
  T* P = S + (1u << (2 * max_qubits));
  const int tid = threadIdx.x;
  for (int c = blockIdx.x; c < n_circ; c += gridDim.x) {
    const qsv_circuit cd = circs[c];
    const int n = cd.n;
    const uint32_t N = 1u << (2 * n);
    for (uint32_t i = tid; i < N; i += kCircBlock) S[circ_swz(i)] = i == 0 ? T(1) : T(0);
    // one gate of lookahead: its descriptor and this thread's PTM entry
    qsv_gate nxt = cd.n_gates > 0 ? gates[cd.gate0] : qsv_gate{};
    T pr = cd.n_gates > 0 && tid < (nxt.k == 1 ? 16 : 256) ? ptms[nxt.ptm + tid] : T(0);
    __syncthreads();
    for (int g = 0; g < cd.n_gates; ++g) {
      const qsv_gate gd = nxt;
      const GateGeom G = circ_gate_geom(gd, n);
      T* tr = traj + cd.traj + int64_t(g) * N;
      for (uint32_t i = tid; i < N; i += kCircBlock) tr[i] = S[circ_swz(i)];
      if (tid < G.D * G.D) P[tid] = pr;
      __syncthreads();
      if (g + 1 < cd.n_gates) {
        nxt = gates[cd.gate0 + g + 1];
        if (tid < (nxt.k == 1 ? 16 : 256)) pr = ptms[nxt.ptm + tid];
      }
      if (G.k == 1)
        circ_orbits<T, 4, false>(S, P, G, tid);
      else
        circ_orbits<T, 16, false>(S, P, G, tid);
      __syncthreads();
    }
    for (uint32_t j = tid; j < (1u << n); j += kCircBlock)
      pops[cd.pop + j] = double(S[circ_swz(circ_spread(j))]);
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kCircBlock)
    k_circuits_adj(const qsv_circuit* __restrict__ circs, int n_circ,
                   const qsv_gate* __restrict__ gates, const T* __restrict__ ptms,
                   const T* __restrict__ traj, const double* __restrict__ dpops,
                   double* __restrict__ dptms, int max_qubits) {
  extern __shared__ __align__(16) unsigned char qsv_circ_smem[];
  double* red = reinterpret_cast<double*>(qsv_circ_smem);  // [waves][64 lanes][4]
  T* Y = reinterpret_cast<T*>(red + kCircBlock * 4);
  T* X = Y + (1u << (2 * max_qubits));
  T* P = X + (1u << (2 * max_qubits));
  typedef typename Mfma<double>::acc Acc;
  constexpr int W = kCircBlock / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 15, kk = lane >> 4;
  // the state before the next gate (backwards) is fetched into registers
  // while the current gate computes: entries j * 256 + tid
  constexpr int PER = kCircMaxEntries<T> / kCircBlock;
  T xr[PER];
  auto fetch = [&](const T* tr, uint32_t N) {
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (uint32_t(j * kCircBlock + tid) < N) xr[j] = tr[j * kCircBlock + tid];
  };
  for (int ci = blockIdx.x; ci < n_circ; ci += gridDim.x) {
    const qsv_circuit cd = circs[ci];
    const int n = cd.n;
    const uint32_t N = 1u << (2 * n);
    if (cd.n_gates > 0) fetch(traj + cd.traj + int64_t(cd.n_gates - 1) * N, N);
    for (uint32_t i = tid; i < N; i += kCircBlock) Y[i] = T(0);
    __syncthreads();
    for (uint32_t j = tid; j < (1u << n); j += kCircBlock)
      Y[circ_swz(circ_spread(j))] = T(dpops[cd.pop + j]);
    __syncthreads();
    qsv_gate nxt = cd.n_gates > 0 ? gates[cd.gate0 + cd.n_gates - 1] : qsv_gate{};
    T pr = cd.n_gates > 0 && tid < (nxt.k == 1 ? 16 : 256) ? ptms[nxt.ptm + tid] : T(0);
    for (int g = cd.n_gates - 1; g >= 0; --g) {
      const qsv_gate gd = nxt;
      const GateGeom G = circ_gate_geom(gd, n);
#pragma unroll
      for (int j = 0; j < PER; ++j)
        if (uint32_t(j * kCircBlock + tid) < N) X[circ_swz(j * kCircBlock + tid)] = xr[j];
      if (tid < G.D * G.D) P[tid] = pr;
      __syncthreads();
      if (g > 0) {
        fetch(traj + cd.traj + int64_t(g - 1) * N, N);
        nxt = gates[cd.gate0 + g - 1];
        if (tid < (nxt.k == 1 ? 16 : 256)) pr = ptms[nxt.ptm + tid];
      }
      // dL/dP on MFMA: A[row][k] = Y entry, B[k][col] = X entry, k = orbit
      Acc acc = Acc(0);
      if (G.k == 2) {
        const uint32_t oc = circ_offset(2, c, G.s0, G.s1);
        for (uint32_t m0 = 4 * w; m0 < G.n_orbits; m0 += 4 * W) {
          const uint32_t m = m0 + kk;
          double a = 0.0, b = 0.0;
          if (m < G.n_orbits) {
            const uint32_t e = circ_swz(circ_orbit_base(m, G.lo, G.hi) + oc);
            a = double(Y[e]);
            b = double(X[e]);
          }
          acc = Mfma<double>::op(a, b, acc);
        }
      } else {
        // sixteen orbits per k-step: lane column c = 4 u + (entry), orbit
        // m0 + 4 kk + u; dL/dP = sum of the four 4 x 4 diagonal blocks
        const uint32_t oc = circ_offset(1, c & 3, G.s0, 0);
        for (uint32_t m0 = 16 * w; m0 < G.n_orbits; m0 += 16 * W) {
          const uint32_t m = m0 + 4 * kk + (c >> 2);
          double a = 0.0, b = 0.0;
          if (m < G.n_orbits) {
            const uint32_t e = circ_swz(circ_orbit_base(m, G.lo, G.hi) + oc);
            a = double(Y[e]);
            b = double(X[e]);
          }
          acc = Mfma<double>::op(a, b, acc);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) red[(w * 64 + lane) * 4 + r] = acc[r];
      __syncthreads();
      // entry (o, i): float64 accumulator layout, row = lane / 16 + 4 r,
      // column = lane % 16
      if (tid < G.D * G.D) {
        double sum = 0.0;
        if (G.k == 2) {
          const int o = tid >> 4, i = tid & 15;
          const int src = (i + 16 * (o & 3)) * 4 + (o >> 2);
          for (int ww = 0; ww < W; ++ww) sum += red[ww * 256 + src];
        } else {
          const int o = tid >> 2, i = tid & 3;
          for (int u = 0; u < 4; ++u) {
            const int src = (4 * u + i + 16 * o) * 4 + u;  // row 4 u + o, column 4 u + i
            for (int ww = 0; ww < W; ++ww) sum += red[ww * 256 + src];
          }
        }
        dptms[gd.ptm + tid] = sum;
      }
      // Y <- P^T Y, orbit by orbit in place
      if (G.k == 1)
        circ_orbits<T, 4, true>(Y, P, G, tid);
      else
        circ_orbits<T, 16, true>(Y, P, G, tid);
      __syncthreads();
    }
  }
}

static size_t circ_lds_bytes(int elem_bytes, int max_qubits, bool adjoint) {
  const size_t N = size_t(1) << (2 * max_qubits);
  if (!adjoint) return (N + 256) * elem_bytes;
  return kCircBlock * 4 * 8 + (2 * N + 256) * elem_bytes;
}

static int lds_limit() {
  static int lim = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    (void)hipGetLastError();
    return v;
  }();
  return lim;
}

template <typename T>
static uint32_t circ_grid(const void* kern, size_t lds, int n_circ) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kCircBlock, lds) != hipSuccess ||
      per <= 0)
    per = 1;
  (void)hipGetLastError();
  const uint32_t cap = uint32_t(per) * uint32_t(cus);
  return std::min<uint32_t>(cap, uint32_t(n_circ));
}

}  // namespace qsv

using namespace qsv;

extern "C" int qsv_circuits_max_qubits(int dtype) {
  const int es = dtype_size(dtype);
  QSV_CHECK_ARG(es > 0, "unknown dtype %d", dtype);
  const uint32_t max_entries = es == 4 ? kCircMaxEntries<float> : kCircMaxEntries<double>;
  int best = 0;
  for (int n = 1; n <= kCircMaxQubits; ++n)
    if (circ_lds_bytes(es, n, true) <= size_t(lds_limit()) && (1u << (2 * n)) <= max_entries)
      best = n;
  return best;
}

extern "C" int64_t qsv_circuits_lds_bytes(int dtype, int max_qubits, int adjoint) {
  const int es = dtype_size(dtype);
  QSV_CHECK_ARG(es > 0, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(max_qubits >= 1 && max_qubits <= kCircMaxQubits, "max_qubits %d outside [1, %d]",
                max_qubits, kCircMaxQubits);
  return int64_t(circ_lds_bytes(es, max_qubits, adjoint != 0));
}

template <typename T, typename K, typename... Args>
static int launch_circ(K kern, size_t lds, int n_circ, hipStream_t s, Args... args) {
  static size_t attr = 0;
  if (lds > 65536 && lds > attr) {
    QSV_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    attr = lds;
  }
  const uint32_t grid = circ_grid<T>(reinterpret_cast<const void*>(kern), lds, n_circ);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kCircBlock), lds, s, args...);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

extern "C" int qsv_circuits_forward(int dtype, int n_circuits, const qsv_circuit* circs_dev,
                                    const qsv_gate* gates_dev, const void* ptms, void* traj,
                                    double* pops, int max_qubits, void* stream) {
  const int es = dtype_size(dtype);
  QSV_CHECK_ARG(es > 0, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(n_circuits >= 0, "negative circuit count");
  if (n_circuits == 0) return QSV_OK;
  QSV_CHECK_ARG(circs_dev && gates_dev && ptms && traj && pops, "null pointer argument");
  QSV_CHECK_ARG(max_qubits >= 1 && max_qubits <= qsv_circuits_max_qubits(dtype),
                "max_qubits %d outside [1, %d] (LDS per workgroup)", max_qubits,
                qsv_circuits_max_qubits(dtype));
  const size_t lds = circ_lds_bytes(es, max_qubits, false);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == QSV_F64)
    return launch_circ<double>(k_circuits_fwd<double>, lds, n_circuits, s, circs_dev, n_circuits,
                               gates_dev, static_cast<const double*>(ptms),
                               static_cast<double*>(traj), pops, max_qubits);
  return launch_circ<float>(k_circuits_fwd<float>, lds, n_circuits, s, circs_dev, n_circuits,
                            gates_dev, static_cast<const float*>(ptms), static_cast<float*>(traj),
                            pops, max_qubits);
}

extern "C" int qsv_circuits_adjoint(int dtype, int n_circuits, const qsv_circuit* circs_dev,
                                    const qsv_gate* gates_dev, const void* ptms,
                                    const void* traj, const double* dpops, double* dptms,
                                    int max_qubits, void* stream) {
  const int es = dtype_size(dtype);
  QSV_CHECK_ARG(es > 0, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(n_circuits >= 0, "negative circuit count");
  if (n_circuits == 0) return QSV_OK;
  QSV_CHECK_ARG(circs_dev && gates_dev && ptms && traj && dpops && dptms, "null pointer argument");
  QSV_CHECK_ARG(max_qubits >= 1 && max_qubits <= qsv_circuits_max_qubits(dtype),
                "max_qubits %d outside [1, %d] (LDS per workgroup)", max_qubits,
                qsv_circuits_max_qubits(dtype));
  const size_t lds = circ_lds_bytes(es, max_qubits, true);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == QSV_F64)
    return launch_circ<double>(k_circuits_adj<double>, lds, n_circuits, s, circs_dev, n_circuits,
                               gates_dev, static_cast<const double*>(ptms),
                               static_cast<const double*>(traj), dpops, dptms, max_qubits);
  return launch_circ<float>(k_circuits_adj<float>, lds, n_circuits, s, circs_dev, n_circuits,
                            gates_dev, static_cast<const float*>(ptms),
                            static_cast<const float*>(traj), dpops, dptms, max_qubits);
}

// Host replay of both kernels on host tables and arrays (float64), with
// every table entry, trajectory, population, PTM and LDS access checked.
extern "C" int qsv_emulate_circuits(int n_circuits, const qsv_circuit* circs,
                                    const qsv_gate* gates, int64_t n_gates_total,
                                    const double* ptms, int64_t ptm_elems, double* traj,
                                    int64_t traj_elems, double* pops, int64_t pop_elems,
                                    const double* dpops, double* dptms, int max_qubits) {
  QSV_CHECK_ARG(n_circuits >= 0 && circs && gates && ptms && traj && pops, "null argument");
  QSV_CHECK_ARG(max_qubits >= 1 && max_qubits <= kCircMaxQubits, "bad max_qubits");
  const uint32_t Nmax = 1u << (2 * max_qubits);
  std::vector<double> S(Nmax), Y(Nmax), X(Nmax);
  {
    // every gate owns its PTM slot (the adjoint writes dL/dP there)
    std::vector<std::pair<int64_t, int64_t>> slots;
    for (int c = 0; c < n_circuits; ++c)
      for (int g = 0; g < circs[c].n_gates && circs[c].gate0 + g < n_gates_total; ++g) {
        const qsv_gate gd = gates[circs[c].gate0 + g];
        slots.push_back({gd.ptm, gd.ptm + (gd.k == 1 ? 16 : 256)});
      }
    std::sort(slots.begin(), slots.end());
    for (size_t i = 1; i < slots.size(); ++i)
      QSV_CHECK_ARG(slots[i].first >= slots[i - 1].second, "two gates share PTM elements");
  }
  for (int c = 0; c < n_circuits; ++c) {
    const qsv_circuit cd = circs[c];
    QSV_CHECK_ARG(cd.n >= 1 && cd.n <= max_qubits, "circuit %d: %d qubits", c, cd.n);
    QSV_CHECK_ARG(cd.n_gates >= 0 && cd.gate0 >= 0 && cd.gate0 + cd.n_gates <= n_gates_total,
                  "circuit %d: gate range out of the table", c);
    const uint32_t N = 1u << (2 * cd.n);
    {
      std::vector<char> seen(N, 0);
      for (uint32_t i = 0; i < N; ++i) {
        const uint32_t p = circ_swz(i);
        QSV_CHECK_ARG(p < N && !seen[p], "LDS swizzle not a bijection for %d qubits", cd.n);
        seen[p] = 1;
      }
    }
    QSV_CHECK_ARG(cd.traj >= 0 && cd.traj + int64_t(cd.n_gates) * N <= traj_elems,
                  "circuit %d: trajectory out of range", c);
    QSV_CHECK_ARG(cd.pop >= 0 && cd.pop + (int64_t(1) << cd.n) <= pop_elems,
                  "circuit %d: populations out of range", c);
    for (int g = 0; g < cd.n_gates; ++g) {
      const qsv_gate gd = gates[cd.gate0 + g];
      QSV_CHECK_ARG(gd.k == 1 || gd.k == 2, "circuit %d gate %d: k = %d", c, g, gd.k);
      QSV_CHECK_ARG(gd.q0 >= 0 && gd.q0 < cd.n && (gd.k == 1 || (gd.q1 >= 0 && gd.q1 < cd.n &&
                                                                  gd.q1 != gd.q0)),
                    "circuit %d gate %d: bad targets", c, g);
      const int D = gd.k == 1 ? 4 : 16;
      QSV_CHECK_ARG(gd.ptm >= 0 && gd.ptm + D * D <= ptm_elems, "circuit %d gate %d: PTM offset",
                    c, g);
    }
    // forward
    std::fill(S.begin(), S.end(), 0.0);
    S[0] = 1.0;
    for (int g = 0; g < cd.n_gates; ++g) {
      const qsv_gate gd = gates[cd.gate0 + g];
      const GateGeom G = circ_gate_geom(gd, cd.n);
      for (uint32_t i = 0; i < N; ++i) traj[cd.traj + int64_t(g) * N + i] = S[i];
      std::vector<int> touched(N, 0);
      for (uint32_t m = 0; m < G.n_orbits; ++m) {
        const uint32_t base = circ_orbit_base(m, G.lo, G.hi);
        double xin[16];
        for (int i = 0; i < G.D; ++i) {
          const uint32_t a = base + circ_offset(G.k, i, G.s0, G.s1);
          QSV_CHECK_ARG(a < N, "circuit %d gate %d: LDS index %u out of range", c, g, a);
          xin[i] = S[a];
          ++touched[a];
        }
        for (int o = 0; o < G.D; ++o) {
          double acc = 0.0;
          for (int i = 0; i < G.D; ++i) acc += ptms[gd.ptm + o * G.D + i] * xin[i];
          S[base + circ_offset(G.k, o, G.s0, G.s1)] = acc;
        }
      }
      for (uint32_t a = 0; a < N; ++a)
        QSV_CHECK_ARG(touched[a] == 1, "circuit %d gate %d: entry %u in %d orbits", c, g, a,
                      touched[a]);
    }
    for (uint32_t j = 0; j < (1u << cd.n); ++j) {
      const uint32_t a = circ_spread(j);
      QSV_CHECK_ARG(a < N, "population index out of range");
      pops[cd.pop + j] = S[a];
    }
    if (!dpops) continue;
    QSV_CHECK_ARG(dptms != nullptr, "dptms missing");
    // adjoint
    std::fill(Y.begin(), Y.end(), 0.0);
    for (uint32_t j = 0; j < (1u << cd.n); ++j) Y[circ_spread(j)] = dpops[cd.pop + j];
    for (int g = cd.n_gates - 1; g >= 0; --g) {
      const qsv_gate gd = gates[cd.gate0 + g];
      const GateGeom G = circ_gate_geom(gd, cd.n);
      for (uint32_t i = 0; i < N; ++i) X[i] = traj[cd.traj + int64_t(g) * N + i];
      // MFMA replay: waves, k-steps and lanes as in k_circuits_adj
      constexpr int W = kCircBlock / 64;
      std::vector<double> red(size_t(W) * 64 * 4, 0.0);
      std::vector<int> used(G.n_orbits, 0);
      const uint32_t step = G.k == 2 ? 4 : 16;
      for (int w = 0; w < W; ++w)
        for (uint32_t m0 = step * w; m0 < G.n_orbits; m0 += step * W) {
          double A[64], B[64];
          for (int lane = 0; lane < 64; ++lane) {
            const int c = lane & 15, kk = lane >> 4;
            const uint32_t m = G.k == 2 ? m0 + kk : m0 + 4 * kk + (c >> 2);
            const uint32_t oc = G.k == 2 ? circ_offset(2, c, G.s0, G.s1)
                                         : circ_offset(1, c & 3, G.s0, 0);
            A[lane] = B[lane] = 0.0;
            if (m < G.n_orbits) {
              const uint32_t e = circ_orbit_base(m, G.lo, G.hi) + oc;
              QSV_CHECK_ARG(e < N, "adjoint: LDS index out of range");
              A[lane] = Y[e];
              B[lane] = X[e];
              if (G.k == 2 ? c == 0 : (c & 3) == 0) ++used[m];
            }
          }
          for (int lane = 0; lane < 64; ++lane)
            for (int r = 0; r < 4; ++r) {
              const int row = (lane >> 4) + 4 * r, col = lane & 15;
              double d = 0.0;
              for (int q = 0; q < 4; ++q) d += A[row + 16 * q] * B[col + 16 * q];
              red[(w * 64 + lane) * 4 + r] += d;
            }
        }
      for (uint32_t m = 0; m < G.n_orbits; ++m)
        QSV_CHECK_ARG(used[m] == 1, "adjoint: orbit %u used %d times", m, used[m]);
      for (int tid = 0; tid < G.D * G.D; ++tid) {
        double sum = 0.0;
        if (G.k == 2) {
          const int o = tid >> 4, i = tid & 15;
          const int src = (i + 16 * (o & 3)) * 4 + (o >> 2);
          for (int ww = 0; ww < W; ++ww) sum += red[ww * 256 + src];
        } else {
          const int o = tid >> 2, i = tid & 3;
          for (int u = 0; u < 4; ++u) {
            const int src = (4 * u + i + 16 * o) * 4 + u;
            for (int ww = 0; ww < W; ++ww) sum += red[ww * 256 + src];
          }
        }
        dptms[gd.ptm + tid] = sum;
      }
      for (uint32_t m = 0; m < G.n_orbits; ++m) {
        const uint32_t base = circ_orbit_base(m, G.lo, G.hi);
        double yin[16];
        for (int o = 0; o < G.D; ++o) yin[o] = Y[base + circ_offset(G.k, o, G.s0, G.s1)];
        for (int i = 0; i < G.D; ++i) {
          double acc = 0.0;
          for (int o = 0; o < G.D; ++o) acc += ptms[gd.ptm + o * G.D + i] * yin[o];
          Y[base + circ_offset(G.k, i, G.s0, G.s1)] = acc;
        }
      }
    }
  }
  return QSV_OK;
}
