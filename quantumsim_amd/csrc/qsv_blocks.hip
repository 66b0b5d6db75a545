// Dense blocks of brickwork windows (two consecutive layers) for many
// circuits at once, built on the device from the per-qubit factors:
//
//   layer l of window w, qubit q:  R[b][2w + l][q]  (rotation PTM, 4 x 4)
//                                  D[b][q]          (damping PTM, 4 x 4)
//   a bond (qa, qb) of layer l:    (D_qa (x) D_qb) CZ (R_qa (x) R_qb)
//   a lone qubit q of layer l:     D_q R_q
//
// A block position p holds 2 or 3 qubits and up to 4 of those factors in
// application order (a table built on the host from the tiling of a window);
// the block is their product embedded in the block's 4^nq-dimensional PTM
// space.  Layer 1 of a window is the identity for a circuit whose depth
// ends inside the window (keep[b][w] = 0).  The backward kernel maps
// dL/dblock straight to dL/dR (one writer per slot) and to per-(window,
// layer) parts of dL/dD, summed by the caller.
//
// One workgroup per (circuit, window, position); all arithmetic in float64,
// the block written in the state's element type.  Every phase is a function
// of the thread index, so the host replay (qsv_emulate_window_blocks) runs
// the same index arithmetic serially with every access checked.
#include "qsv_common.h"

#include <string.h>
#include <vector>

namespace qsv {

constexpr int kWB = 256;  // threads per workgroup

struct WBGeom {
  int batch, nw, n, layers;  // circuits, windows, qubits, rotation layers
  const double* rot;         // [batch][layers][n][4][4]
  const double* damp;        // [batch][n][4][4]
  const uint8_t* keep;       // [batch][nw]
  const double* cz;          // [16][16]
};

// ------------------------------------------------------------ checking
// Device code trusts the host-validated table; the host replay checks every
// computed index against its buffer and records the first violation.
struct WBCheck {
  const char* what = nullptr;
  __host__ __device__ void operator()(bool ok, const char* w) {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (!ok && !what) what = w;
#else
    (void)ok;
    (void)w;
#endif
  }
};

__host__ __device__ __forceinline__ int wb_dim(int nq) { return nq == 3 ? 64 : (nq == 2 ? 16 : 4); }

// Row of the block index with the target slots' digits given and the other
// slots' digits taken from `rest` (most significant remaining slot first).
__host__ __device__ __forceinline__ int wb_row(int nq, int nt, int s0, int s1, int d0, int d1,
                                      int rest) {
  int row = 0;
  int rem = nq - nt;  // remaining slots, filled from rest
  for (int s = 0; s < nq; ++s) {
    int d;
    if (s == s0) {
      d = d0;
    } else if (nt == 2 && s == s1) {
      d = d1;
    } else {
      --rem;
      d = (rest >> (2 * rem)) & 3;
    }
    row = row * 4 + d;
  }
  return row;
}

// T = CZ (Ra (x) Rb) of bond factor f for (b, w), 16 x 16 into Tf.
__host__ __device__ __forceinline__ void wb_bond_T(int tid, int nthr, const qsv_wblock& blk,
                                                   int f, const WBGeom& g, int b, int w,
                                                   double* Tf, WBCheck& chk) {
  const int layer = 2 * w + blk.f_layer[f];
  const int qa = blk.qubits[blk.f_slot0[f]], qb = blk.qubits[blk.f_slot1[f]];
  chk(qa >= 0 && qa < g.n && qb >= 0 && qb < g.n, "bond factor qubit out of range");
  const double* Ra = g.rot + ((int64_t(b) * g.layers + layer) * g.n + qa) * 16;
  const double* Rb = g.rot + ((int64_t(b) * g.layers + layer) * g.n + qb) * 16;
  for (int e = tid; e < 256; e += nthr) {
    const int i = e >> 4, j = e & 15;
    double s = 0.0;
    for (int m = 0; m < 16; ++m)
      s += g.cz[i * 16 + m] * Ra[(m >> 2) * 4 + (j >> 2)] * Rb[(m & 3) * 4 + (j & 3)];
    Tf[e] = s;
  }
}

// Factor f of position `blk` for (b, w): F[f] (16 x 16 for a bond, 4 x 4 in
// the first 16 entries for a single) and the bond's T = CZ (Ra (x) Rb).
// Phase A: singles complete, bonds' T; phase B: bonds' F = (Da (x) Db) T.
__host__ __device__ __forceinline__ void wb_factors_a(int tid, int nthr, const qsv_wblock& blk,
                                             const WBGeom& g, int b, int w, double* F, double* T,
                                             WBCheck& chk) {
  const bool keep1 = g.keep[b * g.nw + w] != 0;
  for (int f = 0; f < blk.nf; ++f) {
    const int l = blk.f_layer[f];
    const bool ident = l == 1 && !keep1;
    const int layer = 2 * w + l;
    chk(layer < g.layers, "window layer beyond the rotation layers");
    if (blk.f_kind[f] == 1) {
      const int q = blk.qubits[blk.f_slot0[f]];
      chk(q >= 0 && q < g.n, "single factor qubit out of range");
      const double* R = g.rot + ((int64_t(b) * g.layers + layer) * g.n + q) * 16;
      const double* Dm = g.damp + (int64_t(b) * g.n + q) * 16;
      for (int e = tid; e < 16; e += nthr) {
        const int r = e >> 2, c = e & 3;
        double s = 0.0;
        for (int m = 0; m < 4; ++m) s += Dm[r * 4 + m] * R[m * 4 + c];
        F[f * 256 + e] = ident ? (r == c ? 1.0 : 0.0) : s;
      }
    } else {
      wb_bond_T(tid, nthr, blk, f, g, b, w, T + f * 256, chk);
    }
  }
}

__host__ __device__ __forceinline__ void wb_factors_b(int tid, int nthr, const qsv_wblock& blk,
                                             const WBGeom& g, int b, int w, double* F,
                                             const double* T) {
  const bool keep1 = g.keep[b * g.nw + w] != 0;
  for (int f = 0; f < blk.nf; ++f) {
    if (blk.f_kind[f] != 2) continue;
    const bool ident = blk.f_layer[f] == 1 && !keep1;
    const int qa = blk.qubits[blk.f_slot0[f]], qb = blk.qubits[blk.f_slot1[f]];
    const double* Da = g.damp + (int64_t(b) * g.n + qa) * 16;
    const double* Db = g.damp + (int64_t(b) * g.n + qb) * 16;
    for (int e = tid; e < 256; e += nthr) {
      const int i = e >> 4, j = e & 15;
      double s = 0.0;
      for (int m = 0; m < 16; ++m)
        s += Da[(i >> 2) * 4 + (m >> 2)] * Db[(i & 3) * 4 + (m & 3)] * T[f * 256 + m * 16 + j];
      F[f * 256 + e] = ident ? (i == j ? 1.0 : 0.0) : s;
    }
  }
}

// Row stride of the backward's D x D matrices in LDS: one double of padding
// per row, so that lanes reading the same column of different rows (the
// column apply, the partial trace) fall into different banks.
__host__ __device__ __forceinline__ int wb_ld(int D) { return D + 1; }

// M (D x D, row stride ld) <- E(F_f) M, or E(F_f)^T M with `transpose`: the
// factor (DT = 4: one slot, 16: two slots) acts on the row index's target
// digits.  Thread per (rest, column): lanes walk the columns of a row.
template <int DT>
__host__ __device__ __forceinline__ void wb_apply_t(int tid, int nthr, const qsv_wblock& blk, int f,
                                           const double* F, double* M, int ld, bool transpose,
                                           WBCheck& chk) {
  constexpr int NT = DT == 16 ? 2 : 1;
  const int nq = blk.nq, D = wb_dim(nq);
  const int s0 = blk.f_slot0[f], s1 = blk.f_slot1[f];
  const int nrest = D / DT;
  const double* Ff = F + f * 256;
  for (int gi = tid; gi < nrest * D; gi += nthr) {
    const int col = gi % D, rest = gi / D;
    double v[DT];
    int rows[DT];
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      rows[t] = wb_row(nq, NT, s0, s1, NT == 2 ? t >> 2 : t, t & 3, rest);
      chk(rows[t] >= 0 && rows[t] < D, "row out of range");
      v[t] = M[rows[t] * ld + col];
    }
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < DT; ++u) s += (transpose ? Ff[u * DT + t] : Ff[t * DT + u]) * v[u];
      M[rows[t] * ld + col] = s;
    }
  }
}

__host__ __device__ __forceinline__ void wb_apply(int tid, int nthr, const qsv_wblock& blk, int f,
                                         const double* F, double* M, int ld, bool transpose,
                                         WBCheck& chk) {
  if (blk.f_kind[f] == 2)
    wb_apply_t<16>(tid, nthr, blk, f, F, M, ld, transpose, chk);
  else
    wb_apply_t<4>(tid, nthr, blk, f, F, M, ld, transpose, chk);
}

// W (D x D, row stride ld) <- W E(F_f)^T: the factor acts on the column
// index's target digits, W'[r][col(t, c)] = sum_u F[t][u] W[r][col(u, c)].
// Thread per (row, rest): lanes walk the rows of a column (padded stride).
template <int DT>
__host__ __device__ __forceinline__ void wb_apply_col_t(int tid, int nthr, const qsv_wblock& blk, int f,
                                               const double* F, double* W, int ld,
                                               WBCheck& chk) {
  constexpr int NT = DT == 16 ? 2 : 1;
  const int nq = blk.nq, D = wb_dim(nq);
  const int s0 = blk.f_slot0[f], s1 = blk.f_slot1[f];
  const int nrest = D / DT;
  const double* Ff = F + f * 256;
  for (int gi = tid; gi < nrest * D; gi += nthr) {
    const int r = gi % D, rest = gi / D;
    double v[DT];
    int cols[DT];
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      cols[t] = wb_row(nq, NT, s0, s1, NT == 2 ? t >> 2 : t, t & 3, rest);
      chk(cols[t] >= 0 && cols[t] < D, "column out of range");
      v[t] = W[r * ld + cols[t]];
    }
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < DT; ++u) s += Ff[t * DT + u] * v[u];
      W[r * ld + cols[t]] = s;
    }
  }
}

__host__ __device__ __forceinline__ void wb_apply_col(int tid, int nthr, const qsv_wblock& blk, int f,
                                             const double* F, double* W, int ld, WBCheck& chk) {
  if (blk.f_kind[f] == 2)
    wb_apply_col_t<16>(tid, nthr, blk, f, F, W, ld, chk);
  else
    wb_apply_col_t<4>(tid, nthr, blk, f, F, W, ld, chk);
}

__host__ __device__ __forceinline__ void wb_identity(int tid, int nthr, int D, double* M) {
  for (int e = tid; e < D * D; e += nthr) M[e] = (e / D == e % D) ? 1.0 : 0.0;
}

__host__ __device__ __forceinline__ void wb_copy(int tid, int nthr, int D, int ld, const double* src,
                                        double* dst) {
  for (int e = tid; e < D * D; e += nthr) {
    const int r = e / D, c = e - r * D;
    dst[r * ld + c] = src[r * ld + c];
  }
}

// dF[ot][it] = sum over the other digits c of W[row(ot, c)][row(it, c)]:
// the gradient of the embedded factor E(F_f) = F_f (x) 1 when
// W = dL/dE(F_f) (row stride ld).
__host__ __device__ __forceinline__ void wb_trace(int tid, int nthr, const qsv_wblock& blk, int f,
                                         const double* W, int ld, double* dF, WBCheck& chk) {
  const int nq = blk.nq, D = wb_dim(nq);
  const int nt = blk.f_kind[f] == 2 ? 2 : 1, dt = nt == 2 ? 16 : 4;
  const int s0 = blk.f_slot0[f], s1 = blk.f_slot1[f];
  for (int e = tid; e < dt * dt; e += nthr) {
    const int ot = e / dt, it = e % dt;
    const int o0 = nt == 2 ? ot >> 2 : ot, o1 = ot & 3;
    const int i0 = nt == 2 ? it >> 2 : it, i1 = it & 3;
    double s = 0.0;
    for (int c = 0; c < D / dt; ++c) {
      const int ro = wb_row(nq, nt, s0, s1, o0, o1, c), ri = wb_row(nq, nt, s0, s1, i0, i1, c);
      chk(ro >= 0 && ro < D && ri >= 0 && ri < D, "trace entry out of range");
      s += W[ro * ld + ri];
    }
    dF[e] = s;
  }
}

// dF of factor f -> dR (written) and the (window, layer) part of dD.
// Bond: X = (Da (x) Db)^T dF, dK = CZ^T X, dL = dF T^T; then the four 4 x 4
// partial contractions.  Masked layer-1 factors write zeros.
__host__ __device__ __forceinline__ void wb_factor_grads_a(int tid, int nthr, const qsv_wblock& blk,
                                                  int f, const WBGeom& g, int b, int w,
                                                  const double* dF, double* X) {
  if (blk.f_kind[f] != 2) return;
  const int qa = blk.qubits[blk.f_slot0[f]], qb = blk.qubits[blk.f_slot1[f]];
  const double* Da = g.damp + (int64_t(b) * g.n + qa) * 16;
  const double* Db = g.damp + (int64_t(b) * g.n + qb) * 16;
  for (int e = tid; e < 256; e += nthr) {
    const int i = e >> 4, j = e & 15;
    double s = 0.0;
    for (int m = 0; m < 16; ++m)
      s += Da[(m >> 2) * 4 + (i >> 2)] * Db[(m & 3) * 4 + (i & 3)] * dF[m * 16 + j];
    X[e] = s;
  }
}

__host__ __device__ __forceinline__ void wb_factor_grads_b(int tid, int nthr, const qsv_wblock& blk,
                                                  int f, const WBGeom& g, const double* dF,
                                                  const double* X, const double* Tf, double* dK,
                                                  double* dL) {
  if (blk.f_kind[f] != 2) return;
  for (int e = tid; e < 256; e += nthr) {
    const int i = e >> 4, j = e & 15;
    double sk = 0.0, sl = 0.0;
    for (int m = 0; m < 16; ++m) {
      sk += g.cz[m * 16 + i] * X[m * 16 + j];
      sl += dF[i * 16 + m] * Tf[j * 16 + m];
    }
    dK[e] = sk;
    dL[e] = sl;
  }
}

__host__ __device__ __forceinline__ void wb_factor_grads_c(int tid, int nthr, const qsv_wblock& blk,
                                                  int f, const WBGeom& g, int b, int w,
                                                  const double* dF, const double* dK,
                                                  const double* dL, double* drot,
                                                  double* ddamp, WBCheck& chk) {
  const bool masked = blk.f_layer[f] == 1 && g.keep[b * g.nw + w] == 0;
  const int l = blk.f_layer[f], layer = 2 * w + l;
  auto rslot = [&](int q) { return drot + ((int64_t(b) * g.layers + layer) * g.n + q) * 16; };
  auto dslot = [&](int q) {
    return ddamp + (((int64_t(b) * g.nw + w) * 2 + l) * g.n + q) * 16;
  };
  if (blk.f_kind[f] == 1) {
    const int q = blk.qubits[blk.f_slot0[f]];
    const double* R = g.rot + ((int64_t(b) * g.layers + layer) * g.n + q) * 16;
    const double* Dm = g.damp + (int64_t(b) * g.n + q) * 16;
    for (int e = tid; e < 16; e += nthr) {
      const int r = e >> 2, c = e & 3;
      double sd = 0.0, sr = 0.0;
      for (int m = 0; m < 4; ++m) {
        sd += dF[r * 4 + m] * R[c * 4 + m];    // dS R^T
        sr += Dm[m * 4 + r] * dF[m * 4 + c];   // D^T dS
      }
      dslot(q)[e] = masked ? 0.0 : sd;
      rslot(q)[e] = masked ? 0.0 : sr;
    }
    return;
  }
  const int qa = blk.qubits[blk.f_slot0[f]], qb = blk.qubits[blk.f_slot1[f]];
  chk(qa != qb, "bond on one qubit");
  const double* Ra = g.rot + ((int64_t(b) * g.layers + layer) * g.n + qa) * 16;
  const double* Rb = g.rot + ((int64_t(b) * g.layers + layer) * g.n + qb) * 16;
  const double* Da = g.damp + (int64_t(b) * g.n + qa) * 16;
  const double* Db = g.damp + (int64_t(b) * g.n + qb) * 16;
  // 64 outputs: (which, r, c), which = dRa, dRb, dDa, dDb
  for (int e = tid; e < 64; e += nthr) {
    const int which = e >> 4, r = (e >> 2) & 3, c = e & 3;
    const double* Gm = which < 2 ? dK : dL;
    const double* other = which == 0 ? Rb : which == 1 ? Ra : which == 2 ? Db : Da;
    double s = 0.0;
    for (int x = 0; x < 4; ++x)
      for (int y = 0; y < 4; ++y) {
        // first-target factor: Gm[(r, x), (c, y)] other[x][y];
        // second-target factor: Gm[(x, r), (y, c)] other[x][y]
        const int i = (which & 1) ? x * 4 + r : r * 4 + x;
        const int j = (which & 1) ? y * 4 + c : c * 4 + y;
        s += Gm[i * 16 + j] * other[x * 4 + y];
      }
    double* dst = which == 0 ? rslot(qa) : which == 1 ? rslot(qb) : which == 2 ? dslot(qa)
                                                                               : dslot(qb);
    dst[r * 4 + c] = masked ? 0.0 : s;
  }
}

template <typename T>
__host__ __device__ __forceinline__ void wb_store(int tid, int nthr, const qsv_wblock& blk,
                                         const WBGeom& g, int b, int w, const double* M, T* out,
                                         int64_t out_elems, WBCheck& chk) {
  const int D = wb_dim(blk.nq);
  T* dst = out + blk.out_off + (int64_t(w) * g.batch + b) * D * D;
  chk(blk.out_off >= 0 && blk.out_off + (int64_t(w) * g.batch + b + 1) * D * D <= out_elems,
      "block output out of range");
  for (int e = tid; e < D * D; e += nthr) dst[e] = T(M[e]);
}

template <typename T>
__host__ __device__ __forceinline__ void wb_load(int tid, int nthr, const qsv_wblock& blk,
                                        const WBGeom& g, int b, int w, const T* dout, double* G,
                                        int64_t out_elems, WBCheck& chk) {
  const int D = wb_dim(blk.nq);
  const T* src = dout + blk.out_off + (int64_t(w) * g.batch + b) * D * D;
  chk(blk.out_off >= 0 && blk.out_off + (int64_t(w) * g.batch + b + 1) * D * D <= out_elems,
      "block gradient out of range");
  const int ld = wb_ld(D);
  for (int e = tid; e < D * D; e += nthr) G[(e / D) * ld + e % D] = double(src[e]);
}

// ------------------------------------------------------------ kernels
template <typename T>
__global__ void __launch_bounds__(kWB)
    k_window_blocks(const qsv_wblock* __restrict__ table, const WBGeom g, T* __restrict__ out,
                    int64_t out_elems) {
  __shared__ double F[4 * 256], Tm[4 * 256], M[64 * 64];
  const int b = blockIdx.x, w = blockIdx.y;
  const qsv_wblock& blk = table[blockIdx.z];  // uniform: scalar loads, no private copy
  const int tid = threadIdx.x;
  WBCheck chk;
  wb_factors_a(tid, kWB, blk, g, b, w, F, Tm, chk);
  wb_identity(tid, kWB, wb_dim(blk.nq), M);
  __syncthreads();
  wb_factors_b(tid, kWB, blk, g, b, w, F, Tm);
  __syncthreads();
  for (int f = 0; f < blk.nf; ++f) {
    wb_apply(tid, kWB, blk, f, F, M, wb_dim(blk.nq), false, chk);
    __syncthreads();
  }
  wb_store<T>(tid, kWB, blk, g, b, w, M, out, out_elems, chk);
}

// Backward of one block B = E_{nf-1} ... E_0 (E_f = F_f embedded).  With
// G_f = (E_{nf-1} ... E_{f+1})^T dL/dB (G updated in place, last factor
// first) and P_f = E_{f-1} ... E_0, dL/dE_f = G_f P_f^T: W = G_f is copied
// and multiplied from the right by E_0^T, ..., E_{f-1}^T (column applies),
// and dL/dF_f is W's partial trace over the digits F_f does not touch.
// LDS: the factors, G, W and dF (76.8 KB: two workgroups per CU).  W holds
// the bonds' T = CZ (Ra (x) Rb) while the factors are built, and after each
// partial trace the factor-gradient scratch: X, dK, dL and the factor's T
// (recomputed), at offsets 0, 256, 512 and 768.
template <typename T>
__global__ void __launch_bounds__(kWB)
    k_window_blocks_bwd(const qsv_wblock* __restrict__ table, const WBGeom g,
                        const T* __restrict__ dout, int64_t out_elems, double* __restrict__ drot,
                        double* __restrict__ ddamp) {
  __shared__ double F[4 * 256], G[64 * 65], W[64 * 65], dF[256];
  const int b = blockIdx.x, w = blockIdx.y;
  const qsv_wblock& blk = table[blockIdx.z];  // uniform: scalar loads, no private copy
  const int tid = threadIdx.x, D = wb_dim(blk.nq), ld = wb_ld(D);
  double* const X = W;
  double* const dK = W + 256;
  double* const dL = W + 512;
  double* const Tf = W + 768;
  WBCheck chk;
  wb_factors_a(tid, kWB, blk, g, b, w, F, W, chk);
  wb_load<T>(tid, kWB, blk, g, b, w, dout, G, out_elems, chk);
  __syncthreads();
  wb_factors_b(tid, kWB, blk, g, b, w, F, W);
  __syncthreads();
  for (int f = blk.nf - 1; f >= 0; --f) {
    const double* Wf = G;
    if (f > 0) {
      wb_copy(tid, kWB, D, ld, G, W);
      __syncthreads();
      for (int h = 0; h < f; ++h) {
        wb_apply_col(tid, kWB, blk, h, F, W, ld, chk);
        __syncthreads();
      }
      Wf = W;
    }
    wb_trace(tid, kWB, blk, f, Wf, ld, dF, chk);
    __syncthreads();
    if (f > 0) wb_apply(tid, kWB, blk, f, F, G, ld, true, chk);
    wb_factor_grads_a(tid, kWB, blk, f, g, b, w, dF, X);
    if (blk.f_kind[f] == 2) wb_bond_T(tid, kWB, blk, f, g, b, w, Tf, chk);
    __syncthreads();
    wb_factor_grads_b(tid, kWB, blk, f, g, dF, X, Tf, dK, dL);
    __syncthreads();
    wb_factor_grads_c(tid, kWB, blk, f, g, b, w, dF, dK, dL, drot, ddamp, chk);
    __syncthreads();
  }
}

// ------------------------------------------------------------ host side
static int check_table(const qsv_wblock* t, int npos, int n, int64_t out_elems, int batch,
                       int nw) {
  QSV_CHECK_ARG(t && npos >= 1 && npos <= 64, "window-block table: %d positions", npos);
  for (int p = 0; p < npos; ++p) {
    const qsv_wblock& b = t[p];
    QSV_CHECK_ARG(b.nq >= 1 && b.nq <= 3, "position %d: %d qubits", p, b.nq);
    QSV_CHECK_ARG(b.nf >= 1 && b.nf <= 4, "position %d: %d factors", p, b.nf);
    for (int s = 0; s < b.nq; ++s)
      QSV_CHECK_ARG(b.qubits[s] >= 0 && b.qubits[s] < n && (s == 0 || b.qubits[s] > b.qubits[s - 1]),
                    "position %d: qubits not ascending in [0, %d)", p, n);
    const int64_t D = wb_dim(b.nq);
    QSV_CHECK_ARG(b.out_off >= 0 && b.out_off + int64_t(nw) * batch * D * D <= out_elems,
                  "position %d: output range outside the buffer", p);
    for (int f = 0; f < b.nf; ++f) {
      QSV_CHECK_ARG(b.f_kind[f] == 1 || b.f_kind[f] == 2, "position %d: factor kind", p);
      QSV_CHECK_ARG(b.f_layer[f] == 0 || b.f_layer[f] == 1, "position %d: factor layer", p);
      QSV_CHECK_ARG(b.f_slot0[f] >= 0 && b.f_slot0[f] < b.nq, "position %d: slot", p);
      if (b.f_kind[f] == 2)
        QSV_CHECK_ARG(b.f_slot1[f] >= 0 && b.f_slot1[f] < b.nq && b.f_slot1[f] != b.f_slot0[f],
                      "position %d: bond slots", p);
    }
  }
  return QSV_OK;
}

}  // namespace qsv

using namespace qsv;

static int wb_geom(int batch, int n_windows, int n_qubits, int layers, const double* rot,
                   const double* damp, const uint8_t* keep, const double* cz, WBGeom& g) {
  QSV_CHECK_ARG(batch >= 1 && batch <= 65535 && n_windows >= 1 && n_windows <= 65535,
                "batch %d / windows %d out of range", batch, n_windows);
  QSV_CHECK_ARG(n_qubits >= 1 && n_qubits <= QSV_MAX_RANK, "qubits %d", n_qubits);
  QSV_CHECK_ARG(layers >= 2 * n_windows, "%d rotation layers for %d windows", layers, n_windows);
  QSV_CHECK_ARG(rot && damp && keep && cz, "null pointer argument");
  g.batch = batch;
  g.nw = n_windows;
  g.n = n_qubits;
  g.layers = layers;
  g.rot = rot;
  g.damp = damp;
  g.keep = keep;
  g.cz = cz;
  return QSV_OK;
}

extern "C" int qsv_window_blocks(int dtype, int n_pos, const qsv_wblock* table_host,
                                 const qsv_wblock* table_dev, int batch, int n_windows,
                                 int n_qubits, int layers, const double* rot, const double* damp,
                                 const uint8_t* keep, const double* cz, void* out,
                                 int64_t out_elems, void* stream) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  WBGeom g;
  int rc = wb_geom(batch, n_windows, n_qubits, layers, rot, damp, keep, cz, g);
  if (rc) return rc;
  rc = check_table(table_host, n_pos, n_qubits, out_elems, batch, n_windows);
  if (rc) return rc;
  QSV_CHECK_ARG(table_dev && out, "null pointer argument");
  const dim3 grid(batch, n_windows, n_pos);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_window_blocks<double>, grid, dim3(kWB), 0, s, table_dev, g,
                       static_cast<double*>(out), out_elems);
  else
    hipLaunchKernelGGL(k_window_blocks<float>, grid, dim3(kWB), 0, s, table_dev, g,
                       static_cast<float*>(out), out_elems);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

extern "C" int qsv_window_blocks_backward(int dtype, int n_pos, const qsv_wblock* table_host,
                                          const qsv_wblock* table_dev, int batch, int n_windows,
                                          int n_qubits, int layers, const double* rot,
                                          const double* damp, const uint8_t* keep,
                                          const double* cz, const void* dout, int64_t out_elems,
                                          double* drot, double* ddamp_part, void* stream) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  WBGeom g;
  int rc = wb_geom(batch, n_windows, n_qubits, layers, rot, damp, keep, cz, g);
  if (rc) return rc;
  rc = check_table(table_host, n_pos, n_qubits, out_elems, batch, n_windows);
  if (rc) return rc;
  QSV_CHECK_ARG(table_dev && dout && drot && ddamp_part, "null pointer argument");
  const dim3 grid(batch, n_windows, n_pos);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_window_blocks_bwd<double>, grid, dim3(kWB), 0, s, table_dev, g,
                       static_cast<const double*>(dout), out_elems, drot, ddamp_part);
  else
    hipLaunchKernelGGL(k_window_blocks_bwd<float>, grid, dim3(kWB), 0, s, table_dev, g,
                       static_cast<const float*>(dout), out_elems, drot, ddamp_part);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

// Host replay of both kernels (float64 buffers): the same phase functions
// run for every thread index in turn between the kernels' barriers, every
// computed index checked.  dout / drot / ddamp_part may be NULL to replay
// the forward kernel only.
extern "C" int qsv_emulate_window_blocks(int n_pos, const qsv_wblock* table, int batch,
                                         int n_windows, int n_qubits, int layers,
                                         const double* rot, const double* damp,
                                         const uint8_t* keep, const double* cz, double* out,
                                         int64_t out_elems, const double* dout, double* drot,
                                         double* ddamp_part) {
  WBGeom g;
  int rc = wb_geom(batch, n_windows, n_qubits, layers, rot, damp, keep, cz, g);
  if (rc) return rc;
  rc = check_table(table, n_pos, n_qubits, out_elems, batch, n_windows);
  if (rc) return rc;
  QSV_CHECK_ARG(out, "null pointer argument");
  std::vector<double> F(4 * 256), Tm(4 * 256), M(64 * 64), G(64 * 65), W(64 * 65);
  std::vector<double> dF(256);
  double* const X = W.data();
  double* const dK = W.data() + 256;
  double* const dL = W.data() + 512;
  double* const Tf = W.data() + 768;
  WBCheck chk;
  auto each = [&](auto fn) {
    for (int t = 0; t < kWB; ++t) fn(t);
  };
  for (int p = 0; p < n_pos; ++p)
    for (int w = 0; w < n_windows; ++w)
      for (int b = 0; b < batch; ++b) {
        const qsv_wblock& blk = table[p];
        const int D = wb_dim(blk.nq);
        each([&](int t) { wb_factors_a(t, kWB, blk, g, b, w, F.data(), Tm.data(), chk); });
        each([&](int t) { wb_identity(t, kWB, D, M.data()); });
        each([&](int t) { wb_factors_b(t, kWB, blk, g, b, w, F.data(), Tm.data()); });
        for (int f = 0; f < blk.nf; ++f)
          each([&](int t) { wb_apply(t, kWB, blk, f, F.data(), M.data(), D, false, chk); });
        each([&](int t) { wb_store<double>(t, kWB, blk, g, b, w, M.data(), out, out_elems, chk); });
        if (!dout) continue;
        QSV_CHECK_ARG(drot && ddamp_part, "null gradient buffers");
        // the backward kernel's phases: factors rebuilt with W holding T
        each([&](int t) {
          wb_factors_a(t, kWB, blk, g, b, w, F.data(), W.data(), chk);
          wb_load<double>(t, kWB, blk, g, b, w, dout, G.data(), out_elems, chk);
        });
        each([&](int t) { wb_factors_b(t, kWB, blk, g, b, w, F.data(), W.data()); });
        const int ld = wb_ld(D);
        for (int f = blk.nf - 1; f >= 0; --f) {
          const double* Wf = G.data();
          if (f > 0) {
            each([&](int t) { wb_copy(t, kWB, D, ld, G.data(), W.data()); });
            for (int h = 0; h < f; ++h)
              each([&](int t) { wb_apply_col(t, kWB, blk, h, F.data(), W.data(), ld, chk); });
            Wf = W.data();
          }
          each([&](int t) { wb_trace(t, kWB, blk, f, Wf, ld, dF.data(), chk); });
          each([&](int t) {
            if (f > 0) wb_apply(t, kWB, blk, f, F.data(), G.data(), ld, true, chk);
            wb_factor_grads_a(t, kWB, blk, f, g, b, w, dF.data(), X);
            if (blk.f_kind[f] == 2) wb_bond_T(t, kWB, blk, f, g, b, w, Tf, chk);
          });
          each([&](int t) { wb_factor_grads_b(t, kWB, blk, f, g, dF.data(), X, Tf, dK, dL); });
          each([&](int t) {
            wb_factor_grads_c(t, kWB, blk, f, g, b, w, dF.data(), dK, dL, drot, ddamp_part, chk);
          });
        }
        QSV_CHECK_ARG(!chk.what, "window blocks: %s", chk.what);
      }
  QSV_CHECK_ARG(!chk.what, "window blocks: %s", chk.what);
  return QSV_OK;
}
