// Dense k-qudit PTM application on MFMA (64 x 64 and 256 x 256 PTMs).
//
//   out[o, g] = sum_k P[o, k] in[k, g],   g over the N / D orbits
//
// Every wave owns columns (orbits) 16 at a time and computes all D output
// rows of them with v_mfma_{f64,f32}_16x16x4.  The wave loads its B operand
// (the orbit's D input entries) straight from HBM into the B-fragment
// registers and stores the accumulators straight back: no LDS round trip
// for the state.  The PTM (A operand) sits in LDS in fragment order.
//
// Within one MFMA k-step lane (n, j) = (lane % 16, lane / 16) supplies
// B[4 s + j][n] and receives accumulator rows that depend on j.  Which PTM
// column ("k") and row ("o") those hardware indices stand for is free: the
// contraction does not care about the order of k, and the rows only decide
// where a result is stored.  The maps
//
//   k = fused_k(VI, s, j) = 4 VI (s / VI) + VI j + s % VI
//   o = fused_o(VO, q, j) = 4 VO (q / VO) + VO j + q % VO   (q = 4 mb + r)
//
// give every lane VI consecutive k (VO consecutive o) per access, so a
// 16-byte vector load (store) covers them when the lowest PTM index has unit
// stride in memory (the innermost axis is a target).  Otherwise VI = VO = 1
// and neighbouring lanes read neighbouring orbits.
#pragma once

#include "qsv_geom.h"

namespace qsv {

struct FusedOffs {
  int64_t in_u[64];   // off_in[fused_k(VI, VI u, 0)],  u < (D / 4) / VI
  int64_t out_u[64];  // off_out[fused_o(VO, VO v, 0)], v < (D / 4) / VO
  int64_t in_l[4];    // off_in[fused_k(VI, 0, j)]
  int64_t out_l[4];   // off_out[fused_o(VO, 0, j)]
};

struct FusedPlan {
  int D;             // 64 or 256
  int vi, vo;        // elements per load / store access along k / o
  int vn;            // 64 x 64 only: consecutive orbits per lane (vi = vo = 1)
  uint32_t n_orbits;
  FusedOffs offs;
};

__host__ __device__ inline int fused_k(int vi, int s, int j) {
  return 4 * vi * (s / vi) + vi * j + s % vi;
}
__host__ __device__ inline int fused_o(int vo, int q, int j) {
  return 4 * vo * (q / vo) + vo * j + q % vo;
}

// Accumulator layout of v_mfma_{f64,f32}_16x16x4 (measured on gfx950): slot
// r of lane group j holds row j + 4 r (float64) or 4 j + r (float32) of the
// 16 x 16 block.  Inverse: the (j, r) that hold row m.
__host__ __device__ inline void mfma_row_owner(int elem_bytes, int m, int& j, int& r) {
  if (elem_bytes == 8) {
    j = m & 3;
    r = m >> 2;
  } else {
    j = m >> 2;
    r = m & 3;
  }
}

// Index of the PTM's A fragment for k-step s, 16-row block mb, lane
// (m + 16 j') in the fragment-ordered table: k-step major, so that a run of
// k-steps is one contiguous chunk.
__host__ __device__ inline int64_t fused_frag_slot(int D, int s, int mb, int lane) {
  return (int64_t(s) * (D / 16) + mb) * 64 + lane;
}

// PTM row o and column k whose value sits at fragment slot `slot` (the
// inverse of fused_frag_slot composed with the row/column maps).
__host__ __device__ inline void fused_frag_coords(int D, int elem_bytes, int vi, int vo,
                                                  int64_t slot, int& o, int& k) {
  const int MB = D / 16;
  const int lane = int(slot & 63);
  const int64_t sm = slot >> 6;
  const int mb = int(sm % MB);
  const int s = int(sm / MB);
  const int m = lane & 15, jk = lane >> 4;
  int j, r;
  mfma_row_owner(elem_bytes, m, j, r);
  o = fused_o(vo, 4 * mb + r, j);
  k = fused_k(vi, s, jk);
}

// Row-major PTM index (o * D + k) of fragment slot `slot`.
__host__ __device__ inline int64_t fused_frag_source(int D, int elem_bytes, int vi, int vo,
                                                     int64_t slot) {
  int o, k;
  fused_frag_coords(D, elem_bytes, vi, vo, slot, o, k);
  return int64_t(o) * D + k;
}

// Fills pf[D * D] (fragment order) from the row-major float64 PTM.
template <typename T>
inline void fused_fragments(int D, int elem_bytes, int vi, int vo, const double* ptm, T* pf) {
  for (int64_t slot = 0; slot < int64_t(D) * D; ++slot)
    pf[slot] = T(ptm[fused_frag_source(D, elem_bytes, vi, vo, slot)]);
}

// k-steps per LDS chunk of the 256 x 256 kernel: a chunk covers 16 (float64)
// or 32 (float32) consecutive PTM columns, i.e. one whole 128-byte line of an
// orbit's row when the targets are the innermost axes, so each line is read
// once (with 16 float32 columns half of every line was fetched twice).
constexpr int fused_chunk_steps(int elem_bytes) { return elem_bytes == 8 ? 4 : 8; }

int launch_fused(int dtype, const FusedPlan& fp, const Geom& g, bool nt, const void* in,
                 void* out, const void* ptm_frag, hipStream_t s);

// 64 x 64 over `batch` states (g.bs_in / g.bs_out elements apart), state b
// with the row-major device PTM ptm_rows + b * pstride (state's element
// type; pstride 0 = one PTM for all), gathered into fragment order in LDS by
// each block: no workspace, kernel launches only.
int launch_fused64_batched(int dtype, const FusedPlan& fp, const Geom& g, bool nt,
                           const void* in, void* out, const void* ptm_rows, int64_t pstride,
                           uint32_t batch, hipStream_t s);

// Float32 64 x 64 whose PTM's lowest index is the innermost axis (runs of
// four entries) on whole 32-orbit tiles: the direct-to-LDS apply.
bool apply64q_applies(int elem_bytes, const FusedPlan& fp);
int launch_apply64q(const FusedPlan& fp, const Geom& g, bool nt, const void* in, void* out,
                    const void* ptm_rows, int64_t pstride, uint32_t batch, hipStream_t s);

// Packs a device PTM (row-major D x D, state's element type) into fragment
// order at pf (D * D elements) on stream s; the device form of
// fused_fragments.
int launch_fused_pack(int dtype, const FusedPlan& fp, const void* ptm_dev, void* pf,
                      hipStream_t s);

// Column blocks per wave tile: the 64 x 64 kernel gives each wave tiles of
// 16 * nb orbits, the 256 x 256 kernel each block tiles of 4 x 16 orbits
// (one column block per wave).
int fused_nb(int elem_bytes, int D, int vn);

// Orbit handled by lane n in column block nb of tile `tile` (64 x 64 kernel).
__host__ __device__ inline uint32_t fused64_orbit_of(int nbs, int vn, uint32_t tile, int nb,
                                                     int n) {
  return vn > 1 ? tile * (16 * nbs) + vn * n + nb : tile * (16 * nbs) + 16 * nb + n;
}

}  // namespace qsv
