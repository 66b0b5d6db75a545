// Orbit geometry shared by the apply kernels (streaming, LDS-tiled, MFMA).
#pragma once

#include "qsv_common.h"

namespace qsv {

constexpr int kMaxGroups = QSV_MAX_TARGETS + 1;
constexpr int kBlock = 256;

struct Geom {
  uint32_t n_threads;  // orbits / VEC
  int32_t ngroups;     // bystander groups, innermost first
  int32_t din, dout;   // products of the target input / output dims
  FastDiv gdiv[kMaxGroups];
  int64_t gs_in[kMaxGroups];
  int64_t gs_out[kMaxGroups];
  int64_t bs_in, bs_out;  // element stride between batch entries (grid.y)
};

template <typename T, int VEC>
struct VT {
  typedef T t __attribute__((ext_vector_type(VEC)));
};
template <typename T>
struct VT<T, 1> {
  typedef T t;
};

// Global accesses of the streaming kernels.  NT = nontemporal (streaming)
// hint, used when the vector is too large to stay cache-resident between
// consecutive gates.
// Bytes a launch moves (reads + writes over the batch) above which the
// streaming kernels use nontemporal accesses (256 MiB; QSV_NT_BYTES).
int64_t nt_threshold_bytes();

template <bool NT, typename V>
__device__ __forceinline__ V gload(const V* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void gstore(V* p, const V& v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

__host__ __device__ __forceinline__ void orbit_base(const Geom& g, uint32_t m, int64_t& bin,
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


}  // namespace qsv
