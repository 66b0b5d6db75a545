// Dense k-qudit PTM application on MFMA (64 x 64 and 256 x 256 PTMs).
#pragma once

#include "qsv_geom.h"

namespace qsv {

// Columns (orbits) per tile of the MFMA kernel for a D x D PTM.
constexpr int fused_cols(int D) { return 64; }

// Row of the 16 x 16 accumulator held in slot r by `lane` of
// v_mfma_{f64,f32}_16x16x4 (layouts measured on gfx950: float64 interleaves
// rows by 4, float32 keeps 4 consecutive rows per lane group).
__host__ __device__ inline int mfma_acc_row(int elem_bytes, int lane, int r) {
  return elem_bytes == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

// Index of P[row][k] in the fragment-ordered PTM: A operand of m-block
// row / 16, k-step k / 4, lane (row % 16) + 16 (k % 4).
__host__ __device__ inline int64_t fused_frag_index(int D, int row, int k) {
  return (int64_t(row >> 4) * (D >> 2) + (k >> 2)) * 64 + (row & 15) + 16 * (k & 3);
}

int launch_fused(int dtype, int D, const Geom& g, uint32_t n_orbits, const void* in, void* out,
                 const void* ptm_frag, const int64_t* off_in, const int64_t* off_out,
                 hipStream_t s);

}  // namespace qsv
