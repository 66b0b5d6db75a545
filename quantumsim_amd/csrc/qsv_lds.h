// Helpers of the kernels that move operand tiles HBM -> LDS with the direct
// loads of gfx950 (global_load_lds_dwordx4: lane l's 16 bytes land at the
// wave's LDS base + 16 l, the data never passing through VGPRs).
//
// The compiler counts a direct load as a vector-memory write to LDS and
// waits for every outstanding one (s_waitcnt vmcnt(0)) before any LDS read
// it can see, and a workgroup fence waits for them too; so the kernels that
// keep the next tile's loads in flight while the current tile is consumed
// (1) wait for the current tile themselves with an explicit s_waitcnt,
// (2) synchronise with a bare s_barrier (lds_barrier), and (3) read their
// LDS operands in asm blocks that end with their own lgkmcnt wait.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qsv {

// LDS byte offset of a pointer into a __shared__ array (the low 32 bits of
// its generic address: the shared aperture is 4 GB aligned).
__device__ __forceinline__ uint32_t lds_off(const void* p) {
  return uint32_t(reinterpret_cast<uintptr_t>(p));
}

// Workgroup barrier without a memory fence; the empty asm statements keep
// the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// s_waitcnt immediate (gfx9 encoding) waiting for at most `n` outstanding
// vector-memory operations, leaving the export and LDS/scalar counters
// alone.
constexpr int waitcnt_vm(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }

}  // namespace qsv
