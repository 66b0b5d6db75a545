// Adjoint of PTM application with respect to the PTM:
//
//   dP[b, o, i] = sum over orbits g of  ybar[b, g, o] * x[b, g, i]
//
// where x is the state the PTM was applied to and ybar the gradient of the
// loss with respect to the result.  (The gradient with respect to x is the
// same contraction with the transposed PTM and reuses the apply kernels.)
//
// The sum over orbits is a (D_out x G) (G x D_in) matrix product with a
// huge inner dimension, done on v_mfma_f64_16x16x4f64: each k-step takes
// four orbits, lane (c, k) = (l % 16, l / 16) supplying ybar[g_k][c] as the
// A operand and x[g_k][c] as the B operand, so one lane loads both operands
// of its own orbit straight from HBM (float32 states are widened to float64:
// products and sums are float64).  When the innermost axis is a bystander
// with unit stride a lane loads VN consecutive orbits as one 16-byte vector
// (VN k-steps); otherwise lanes c of one orbit read the targets' entries (a
// contiguous run when the targets are the innermost axes).  Blocks write
// float64 partial tiles, a second kernel sums them in a fixed order
// (deterministic).
#include "qsv_fused.h"
#include "qsv_geom.h"
#include "qsv_lds.h"
#include "qsv_mfma.h"

#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

namespace qsv {

constexpr int kGradBlocks = 256;
constexpr int kGradUnroll = 4;

struct GradOffs {
  int64_t off_in[64];
  int64_t off_out[64];
};

template <typename T, int VN>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad(const T* __restrict__ x, const T* __restrict__ ybar, const Geom g,
               const GradOffs offs, const uint32_t n_orbits, double* __restrict__ partial) {
  typedef typename Mfma<double>::acc Acc;
  typedef typename VT<T, VN>::t Vec;
  __shared__ double red[kBlock / 64][64 * 4];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = lane & 15, k = lane >> 4;
  const bool vi = c < g.din, vo = c < g.dout;
  const int64_t oi = vi ? offs.off_in[c] : 0, oo = vo ? offs.off_out[c] : 0;
  Acc acc = Acc(0);
  constexpr uint32_t STEP = 4 * VN;  // orbits per wave step
  constexpr int U = kGradUnroll;     // steps whose loads are issued together
  const uint32_t nsteps = (n_orbits + STEP - 1) / STEP;
  const uint32_t nchunks = (nsteps + U - 1) / U;
  for (uint32_t tc = blockIdx.x * (kBlock / 64) + w; tc < nchunks;
       tc += gridDim.x * (kBlock / 64)) {
    // loads of all U steps first (addresses clamped to entry 0 and values
    // zeroed for lanes past the end: no branches between the loads)
    Vec xv[U], yv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t m = (tc * U + u) * STEP + VN * k;
      const bool ok = m < n_orbits;  // VN divides n_orbits: whole vectors
      int64_t bin, bout;
      orbit_base(g, ok ? m : 0, bin, bout);
      const bool lx = ok && vi, ly = ok && vo;
      xv[u] = *reinterpret_cast<const Vec*>(x + (lx ? bin + oi : 0));
      yv[u] = *reinterpret_cast<const Vec*>(ybar + (ly ? bout + oo : 0));
      if (!lx) xv[u] = Vec(0);
      if (!ly) yv[u] = Vec(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (VN == 1) {
        acc = Mfma<double>::op(double(yv[u]), double(xv[u]), acc);
      } else {
#pragma unroll
        for (int e = 0; e < VN; ++e)
          acc = Mfma<double>::op(double(yv[u][e]), double(xv[u][e]), acc);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) red[w][lane * 4 + r] = acc[r];
  __syncthreads();
  // thread t sums entry (o, i) = (t / 16, t % 16) over the four waves; it
  // sits in column i, row o = k + 4 r of the float64 accumulator layout
  {
    const int o = tid >> 4, i = tid & 15;
    const int src = (i + 16 * (o & 3)) * 4 + (o >> 2);
    double sum = 0.0;
#pragma unroll
    for (int ww = 0; ww < kBlock / 64; ++ww) sum += red[ww][src];
    if (o < g.dout && i < g.din)
      partial[((int64_t(b) * gridDim.x + blockIdx.x) * g.dout + o) * g.din + i] = sum;
  }
}

// PTMs of 17..64 rows or columns (dense three-qubit blocks): the same
// contraction with a 64 x 64 output, i.e. 4 x 4 MFMA tiles per wave.  Lane
// (c, k) loads entries c, 16 + c, 32 + c, 48 + c of ybar (the A operands of
// the four row blocks) and of x (the B operands of the four column blocks)
// of its orbit; each k-step is 16 MFMAs.  Float64 states use
// v_mfma_f64_16x16x4f64; float32 states v_mfma_f32_16x16x4f32 (float32 sums
// within a wave, i.e. over at most a few thousand orbits; 64 float64 MFMAs
// per 16 orbits would make the kernel MFMA-bound below the HBM rate).  The
// four waves' tiles are summed in float64 in LDS in a fixed order (32 KB),
// one row block at a time, then the block writes its partial tile.
constexpr int kGrad64Blocks = 64;
constexpr int kGrad64Unroll = 1;
// QUAD loads move 16 bytes of x and of ybar per lane and k-step (a quarter
// of the orbit-vector mode's), so four k-steps are loaded before their
// MFMAs: 3.8 -> 4.5 TB/s on 64 twelve-qubit float32 states (unrolling the
// orbit-vector mode does not pay)
constexpr int kGrad64UnrollQuad = 4;

// QUAD (64 x 64 PTMs whose lowest index has unit stride, i.e. the
// innermost axis is a target): lane c holds the four consecutive entries
// 4c .. 4c+3 of its orbit, loaded as 16-byte vectors, and row (column) block
// q of the tiles is the entries {4 m + q}; the contraction does not depend
// on which entry sits in which MFMA row, only the epilogue's map does.
template <typename T, int VN, bool QUAD>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad64(const T* __restrict__ x, const T* __restrict__ ybar, const Geom g,
                 const GradOffs offs, const uint32_t n_orbits, double* __restrict__ partial) {
  typedef typename Mfma<T>::acc Acc;
  typedef typename VT<T, VN>::t Vec;
  constexpr int QV = 16 / sizeof(T);  // elements per 16-byte vector
  typedef typename VT<T, QV>::t QVec;
  __shared__ double red[64 * 64];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = lane & 15, k = lane >> 4;
  int64_t oi[4], oo[4];
  bool vi[4], vo[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ent = QUAD ? 4 * c : 16 * q + c;
    vi[q] = ent < g.din;
    vo[q] = ent < g.dout;
    oi[q] = vi[q] ? offs.off_in[ent] : 0;
    oo[q] = vo[q] ? offs.off_out[ent] : 0;
  }
  Acc acc[4][4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = Acc(0);
  constexpr uint32_t STEP = 4 * VN;
  constexpr int U = QUAD ? kGrad64UnrollQuad : kGrad64Unroll;
  const uint32_t nsteps = (n_orbits + STEP - 1) / STEP;
  const uint32_t nchunks = (nsteps + U - 1) / U;
  for (uint32_t tc = blockIdx.x * (kBlock / 64) + w; tc < nchunks;
       tc += gridDim.x * (kBlock / 64)) {
    Vec xv[U][4], yv[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t m = (tc * U + u) * STEP + VN * k;
      const bool ok = m < n_orbits;  // VN divides n_orbits: whole vectors
      int64_t bin, bout;
      orbit_base(g, ok ? m : 0, bin, bout);
      if constexpr (QUAD) {
        // entries 4c .. 4c+3 (all valid: QUAD needs a 64 x 64 PTM)
        const T* px = x + (ok ? bin + oi[0] : 0);
        const T* py = ybar + (ok ? bout + oo[0] : 0);
#pragma unroll
        for (int h = 0; h < 4 / QV; ++h) {
          const QVec a = *reinterpret_cast<const QVec*>(px + h * QV);
          const QVec bb = *reinterpret_cast<const QVec*>(py + h * QV);
#pragma unroll
          for (int e = 0; e < QV; ++e) {
            xv[u][h * QV + e] = ok ? a[e] : T(0);
            yv[u][h * QV + e] = ok ? bb[e] : T(0);
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bool lx = ok && vi[q], ly = ok && vo[q];
          xv[u][q] = *reinterpret_cast<const Vec*>(x + (lx ? bin + oi[q] : 0));
          yv[u][q] = *reinterpret_cast<const Vec*>(ybar + (ly ? bout + oo[q] : 0));
          if (!lx) xv[u][q] = Vec(0);
          if (!ly) yv[u][q] = Vec(0);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < VN; ++e)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb) {
            if constexpr (VN == 1)
              acc[rb][cb] = Mfma<T>::op(yv[u][rb], xv[u][cb], acc[rb][cb]);
            else
              acc[rb][cb] = Mfma<T>::op(yv[u][rb][e], xv[u][cb][e], acc[rb][cb]);
          }
  }
  // slot r of lane (c, k) in tile (rb, cb) holds MFMA row `row` = k + 4 r
  // (float64) or 4 k + r (float32) and column c, i.e. entry
  // (16 rb + row, 16 cb + c), or (4 row + rb, 4 c + cb) with QUAD
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
    for (int ww = 0; ww < kBlock / 64; ++ww) {
      if (w == ww) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = sizeof(T) == 8 ? k + 4 * r : 4 * k + r;
            const int idx = QUAD ? (4 * row + rb) * 64 + 4 * c + cb
                                 : (16 * rb + row) * 64 + 16 * cb + c;
            red[idx] = (ww == 0 ? 0.0 : red[idx]) + double(acc[rb][cb][r]);
          }
      }
      __syncthreads();
    }
  }
  const int dd = g.dout * g.din;
  for (int j = tid; j < dd; j += kBlock) {
    const int o = j / g.din, i = j - o * g.din;
    partial[(int64_t(b) * gridDim.x + blockIdx.x) * dd + j] = red[o * 64 + i];
  }
}

// LDS-staged 64 x 64 PTM gradient for the orbit-vector mode (the innermost
// axis is a unit-stride bystander: VN consecutive orbits per 16-byte
// vector), used for float64 states: k_ptm_grad64 holds all 16 float64
// accumulator tiles per wave (321 registers, one wave per SIMD), this one 4
// per wave (4 waves per SIMD); measured on 12-qubit states, 2.7 -> 4.2 TB/s.
// (For float32 the register-resident k_ptm_grad64 runs at 3 waves per SIMD
// and stays ahead: 4.1-5.1 vs 3.8-4.1 TB/s.)  A block walks tiles of TT orbits: the tile's x and ybar (64
// entries x TT orbits each) are loaded with coalesced 16-byte vectors (the
// lanes of a wave cover whole runs of one entry) into one of two LDS
// buffers, while the block's four waves run the MFMAs of the previous tile
// out of the other.  Wave w owns output rows 16 w .. 16 w + 15 (four 16 x 16
// MFMA tiles, all 64 columns), so the waves never sum into each other:
// each writes its rows of the block's partial tile at the end.
template <typename T, int VN, int TT>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad64s(const T* __restrict__ x, const T* __restrict__ ybar, const Geom g,
                  const GradOffs offs, const uint32_t n_orbits, double* __restrict__ partial) {
  typedef typename Mfma<T>::acc Acc;
  typedef typename VT<T, VN>::t Vec;
  constexpr int QV = 16 / sizeof(T);       // elements per 16-byte vector
  constexpr int RS = TT + QV;              // padded LDS row (elements)
  constexpr int TV = TT / VN;              // vectors per entry per tile
  constexpr int R = 64 * TV / kBlock;      // vectors per thread per array
  static_assert(VN == QV, "orbit vectors are 16 bytes");
  static_assert(TT % 16 == 0 && (64 * TV) % kBlock == 0 && kBlock % TV == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) T sx[2][64 * RS];
  __shared__ __attribute__((aligned(16))) T sy[2][64 * RS];
  __shared__ int64_t soff[2][64];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid < 64) {
    soff[0][tid] = offs.off_in[tid];
    soff[1][tid] = offs.off_out[tid];
  }
  __syncthreads();
  // loader: vector v = tid % TV of entries tid / TV + r kBlock / TV
  const int v = tid % TV, e0 = tid / TV;
  int64_t lin[R], lout[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    lin[r] = soff[0][e0 + r * (kBlock / TV)];
    lout[r] = soff[1][e0 + r * (kBlock / TV)];
  }
  const uint32_t ntiles = (n_orbits + TT - 1) / TT;
  Vec rx[R], ry[R];
  auto load = [&](uint32_t t) {
    const uint32_t m = t * TT + v * VN;
    const bool ok = m < n_orbits;
    int64_t bin, bout;
    orbit_base(g, ok ? m : 0, bin, bout);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      rx[r] = *reinterpret_cast<const Vec*>(x + bin + lin[r]);
      ry[r] = *reinterpret_cast<const Vec*>(ybar + bout + lout[r]);
      if (!ok) {
        rx[r] = Vec(0);
        ry[r] = Vec(0);
      }
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = e0 + r * (kBlock / TV);
      *reinterpret_cast<Vec*>(&sx[buf][e * RS + v * VN]) = rx[r];
      *reinterpret_cast<Vec*>(&sy[buf][e * RS + v * VN]) = ry[r];
    }
  };
  const int c = lane & 15, k = lane >> 4;
  Acc acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) acc[cb] = Acc(0);
  typedef typename VT<T, 4>::t V4;
  auto compute = [&](int buf) {
    const T* ya = &sy[buf][(16 * w + c) * RS + 4 * k];
    const T* xb = &sx[buf][c * RS + 4 * k];
#pragma unroll
    for (int u = 0; u < TT / 16; ++u) {
      const V4 a = *reinterpret_cast<const V4*>(ya + 16 * u);
      V4 bb[4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
        bb[cb] = *reinterpret_cast<const V4*>(xb + 16 * cb * RS + 16 * u);
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[cb] = Mfma<T>::op(a[e], bb[cb][e], acc[cb]);
    }
  };
  uint32_t t = blockIdx.x;
  if (t < ntiles) {
    load(t);
    stash(0);
  }
  __syncthreads();
  int buf = 0;
  for (; t < ntiles; t += gridDim.x) {
    const uint32_t tn = t + gridDim.x;
    const bool more = tn < ntiles;
    if (more) load(tn);
    compute(buf);
    if (more) stash(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // slot r of lane (c, k) in tile cb: row k + 4 r (float64) or 4 k + r
  // (float32) of the wave's row block, column c of column block cb
  double* out = partial + (int64_t(b) * gridDim.x + blockIdx.x) * 64 * 64;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = sizeof(T) == 8 ? k + 4 * r : 4 * k + r;
      out[(16 * w + row) * 64 + 16 * cb + c] = double(acc[cb][r]);
    }
}


// Float32 64 x 64 PTM gradient in the orbit-vector mode with the operand
// tiles moved HBM -> LDS by the direct loads (global_load_lds_dwordx4: lane
// l's 16 bytes land at the wave's LDS base + 16 l), so several tiles are in
// flight without holding registers.  A tile is 32 orbits x 64 entries of x
// and of ybar (8 KB each); wave-load j (0..7) covers entries 8j..8j+7, lane
// l = e + 8 c taking entry 8j + e, orbits 4c..4c+3 (one 16-byte vector of
// the innermost, unit-stride bystander).  Wave-load blocks are 1152 bytes
// apart (128 bytes of padding) so that an MFMA operand read -- 16 entries
// x 4 orbits, spread over two blocks -- hits 64 distinct banks.  Two
// stages, in two separate LDS arrays: while the MFMAs of one tile read one
// array, the loads of the next tile fill the other.  Wave w owns output
// rows 16w..16w+15 (4 accumulator tiles): no cross-wave sum; the float32
// accumulators are added into float64 ones every kDirFlush tiles.
constexpr int kDirTile = 32;            // orbits per tile
constexpr int kDirBlk = 1152;           // bytes per wave-load block (1024 + pad)
constexpr int kDirArr = 8 * kDirBlk;    // bytes per operand tile
constexpr int kDirFlush = 8;            // tiles per float32 partial sum

// Cache policy of the direct loads: bit 1 = nontemporal (streaming states
// that do not stay in the caches until their next use).
template <bool NT>
__device__ __forceinline__ void dir_issue(const float* x, const float* ybar, const Geom& g,
                                          uint32_t t, int w, int c, const int64_t (&oin)[2],
                                          const int64_t (&oout)[2], char* lds) {
  constexpr int kPol = NT ? 2 : 0;
  int64_t bin, bout;
  orbit_base(g, t * kDirTile + 4 * c, bin, bout);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int j = 2 * w + h;
    __builtin_amdgcn_global_load_lds(x + bin + oin[h], lds + j * kDirBlk, 16, 0, kPol);
    __builtin_amdgcn_global_load_lds(ybar + bout + oout[h], lds + kDirArr + j * kDirBlk, 16, 0,
                                     kPol);
  }
}

// Two k-steps (ks, ks + 1) of one tile: the A value (ybar) and the four B
// values (x, one per column tile) of each, read from LDS in one asm block
// that waits for its own reads (the compiler sees no LDS read here, so it
// does not wait for the next tile's direct loads before them).  The outputs
// are early-clobber: every read is issued before the block's wait, so no
// output may share a register with the two address inputs.
#define QSV_DIR_READ2(KS)                                                          \
  asm volatile(                                                                    \
      "ds_read_b32 %0, %10 offset:%c12\n"                                          \
      "ds_read_b32 %1, %11 offset:%c12\n"                                          \
      "ds_read_b32 %2, %11 offset:%c13\n"                                          \
      "ds_read_b32 %3, %11 offset:%c14\n"                                          \
      "ds_read_b32 %4, %11 offset:%c15\n"                                          \
      "ds_read_b32 %5, %10 offset:%c16\n"                                          \
      "ds_read_b32 %6, %11 offset:%c16\n"                                          \
      "ds_read_b32 %7, %11 offset:%c17\n"                                          \
      "ds_read_b32 %8, %11 offset:%c18\n"                                          \
      "ds_read_b32 %9, %11 offset:%c19\n"                                          \
      "s_waitcnt lgkmcnt(0)\n"                                                     \
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]),          \
        "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7]), "=&v"(v[8]), "=&v"(v[9])            \
      : "v"(ya), "v"(xb), "i"(128 * (KS)), "i"(128 * (KS) + 2 * kDirBlk),         \
        "i"(128 * (KS) + 4 * kDirBlk), "i"(128 * (KS) + 6 * kDirBlk),             \
        "i"(128 * (KS + 1)), "i"(128 * (KS + 1) + 2 * kDirBlk),                   \
        "i"(128 * (KS + 1) + 4 * kDirBlk), "i"(128 * (KS + 1) + 6 * kDirBlk)      \
      : "memory")

__device__ __forceinline__ void dir_compute(const char* lds, int w, int lane,
                                            Mfma<float>::acc (&acc)[4]) {
  const int m = lane & 15, kk = lane >> 4;
  const uint32_t ya = lds_off(lds + kDirArr + (2 * w + (m >> 3)) * kDirBlk + 16 * (m & 7) + 4 * kk);
  const uint32_t xb = lds_off(lds + (m >> 3) * kDirBlk + 16 * (m & 7) + 4 * kk);
  float v[10];
#define QSV_DIR_STEP2(KS)                                       \
  QSV_DIR_READ2(KS);                                            \
  _Pragma("unroll") for (int h = 0; h < 2; ++h)                 \
  _Pragma("unroll") for (int cb = 0; cb < 4; ++cb)              \
    acc[cb] = Mfma<float>::op(v[5 * h], v[5 * h + 1 + cb], acc[cb]);
  QSV_DIR_STEP2(0)
  QSV_DIR_STEP2(2)
  QSV_DIR_STEP2(4)
  QSV_DIR_STEP2(6)
#undef QSV_DIR_STEP2
}

template <bool NT>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad64d(const float* __restrict__ x, const float* __restrict__ ybar, const Geom g,
                  const GradOffs offs, const uint32_t n_orbits, double* __restrict__ partial) {
  __shared__ __align__(16) char st0[2 * kDirArr];
  __shared__ __align__(16) char st1[2 * kDirArr];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int e = lane & 7, c = lane >> 3;
  int64_t oin[2], oout[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    oin[h] = offs.off_in[8 * (2 * w + h) + e];
    oout[h] = offs.off_out[8 * (2 * w + h) + e];
  }
  typedef Mfma<float>::acc Acc;
  Acc acc[4];
  double acc64[4][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    acc[cb] = Acc(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc64[cb][r] = 0.0;
  }
  auto flush = [&]() {
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[cb][r] += double(acc[cb][r]);
      acc[cb] = Acc(0);
    }
  };
  const uint32_t ntiles = n_orbits / kDirTile, stride = gridDim.x;
  uint32_t t = blockIdx.x;
  int n_acc = 0;
  if (t < ntiles) dir_issue<NT>(x, ybar, g, t, w, c, oin, oout, st0);
  // two tiles per trip: st0 holds tile t, st1 tile t + stride
  while (t < ntiles) {
    const uint32_t t1 = t + stride;
    if (t1 < ntiles) {
      dir_issue<NT>(x, ybar, g, t1, w, c, oin, oout, st1);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(4));  // this tile's loads done
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    dir_compute(st0, w, lane, acc);
    if (++n_acc == kDirFlush) { flush(); n_acc = 0; }
    lds_barrier();
    if (t1 >= ntiles) break;
    const uint32_t t2 = t1 + stride;
    if (t2 < ntiles) {
      dir_issue<NT>(x, ybar, g, t2, w, c, oin, oout, st0);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(4));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    dir_compute(st1, w, lane, acc);
    if (++n_acc == kDirFlush) { flush(); n_acc = 0; }
    lds_barrier();
    t = t2;
  }
  flush();
  // slot r of lane (col, k) in tile cb: row 4k + r of the wave's 16 rows
  const int col = lane & 15, k = lane >> 4;
  double* out = partial + (int64_t(b) * gridDim.x + blockIdx.x) * 64 * 64;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(16 * w + 4 * k + r) * 64 + 16 * cb + col] = acc64[cb][r];
}

// LDS byte offset of k-step s of the adjoint apply below: entries 4s..4s+3
// (wave-load block s / 2, entries 4 (s % 2) + kk of it).
constexpr int bwd_koff(int s) { return (s >> 1) * kDirBlk + 64 * (s & 1); }

typedef float qsv_f2 __attribute__((ext_vector_type(2)));

// Eight k-steps of the adjoint apply's A operand (ybar^T: orbits 2m, 2m + 1
// of entry 4s + kk, one 8-byte read per step), read and waited for in one
// asm block as in QSV_DIR_READ2 (early-clobber outputs).
#define QSV_BWD_READ8(S0)                                                           \
  asm volatile(                                                                     \
      "ds_read_b64 %0, %8 offset:%c9\n"                                             \
      "ds_read_b64 %1, %8 offset:%c10\n"                                            \
      "ds_read_b64 %2, %8 offset:%c11\n"                                            \
      "ds_read_b64 %3, %8 offset:%c12\n"                                            \
      "ds_read_b64 %4, %8 offset:%c13\n"                                            \
      "ds_read_b64 %5, %8 offset:%c14\n"                                            \
      "ds_read_b64 %6, %8 offset:%c15\n"                                            \
      "ds_read_b64 %7, %8 offset:%c16\n"                                            \
      "s_waitcnt lgkmcnt(0)\n"                                                      \
      : "=&v"(av[0]), "=&v"(av[1]), "=&v"(av[2]), "=&v"(av[3]), "=&v"(av[4]),       \
        "=&v"(av[5]), "=&v"(av[6]), "=&v"(av[7])                                    \
      : "v"(ya), "i"(bwd_koff(S0)), "i"(bwd_koff(S0 + 1)), "i"(bwd_koff(S0 + 2)),   \
        "i"(bwd_koff(S0 + 3)), "i"(bwd_koff(S0 + 4)), "i"(bwd_koff(S0 + 5)),        \
        "i"(bwd_koff(S0 + 6)), "i"(bwd_koff(S0 + 7))                                \
      : "memory")

// xbar = P^T ybar on the staged tile, wave w's 16 entries 16w..16w+15:
// xbar^T (32 orbits x 16 entries) = ybar^T (32 x 64) P[:, 16w..16w+15] as two
// row tiles, row m of row tile rt being orbit 2m + rt; lane (n, kk) holds
// B = P[4s + kk][16w + n] for all 16 k-steps in registers (pb).  Slot r of
// lane l in row tile rt is orbit 8 (l / 16) + 2r + rt of entry 16w + l % 16,
// so a lane ends with 8 consecutive orbits of one entry: two 16-byte stores.
__device__ __forceinline__ void bwd_apply(const char* lds, const float (&pb)[16], int lane,
                                          const Geom& g, uint32_t t, int64_t xo,
                                          float* __restrict__ xbar) {
  const int m = lane & 15, kk = lane >> 4;
  const uint32_t ya = lds_off(lds + kDirArr + 16 * kk + 128 * (m >> 1) + 8 * (m & 1));
  typedef Mfma<float>::acc Acc;
  Acc a0 = Acc(0), a1 = Acc(0);
  qsv_f2 av[8];
  QSV_BWD_READ8(0);
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    a0 = Mfma<float>::op(av[s].x, pb[s], a0);
    a1 = Mfma<float>::op(av[s].y, pb[s], a1);
  }
  QSV_BWD_READ8(8);
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    a0 = Mfma<float>::op(av[s].x, pb[8 + s], a0);
    a1 = Mfma<float>::op(av[s].y, pb[8 + s], a1);
  }
  const uint32_t m0 = t * kDirTile + 8 * uint32_t(lane >> 4);
  int64_t bin0, bout0, bin1, bout1;
  orbit_base(g, m0, bin0, bout0);
  orbit_base(g, m0 + 4, bin1, bout1);
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 v0, v1;
  v0[0] = a0[0]; v0[1] = a1[0]; v0[2] = a0[1]; v0[3] = a1[1];
  v1[0] = a0[2]; v1[1] = a1[2]; v1[2] = a0[3]; v1[3] = a1[3];
  *reinterpret_cast<f4*>(xbar + bin0 + xo) = v0;
  *reinterpret_cast<f4*>(xbar + bin1 + xo) = v1;
}

// Fused float32 backward of a dense 64 x 64 block on k_ptm_grad64d's
// staged tiles: per tile of 32 orbits, dP += ybar x^T (dir_compute, as in
// k_ptm_grad64d) and xbar = P^T ybar from the same ybar tile in LDS
// (bwd_apply), stored from the accumulators: x and ybar are read once, xbar
// written once (three streams where the transposed apply plus the gradient
// move four).  The wait before each tile is vmcnt(4) although the previous
// tile's xbar stores may be in flight too: vector-memory loads retire in
// order among themselves, so a load of this tile still in flight would
// keep the next tile's four later loads in flight as well, five or more in
// all; at most four outstanding therefore means this tile's loads are done.
template <bool NT>
__global__ void __launch_bounds__(kBlock)
    k_bwd64d(const float* __restrict__ x, const float* __restrict__ ybar,
             float* __restrict__ xbar, const Geom g, const uint32_t n_orbits,
             const float* __restrict__ ptm, const int64_t pstride, const GradOffs offs,
             double* __restrict__ partial) {
  __shared__ __align__(16) char st0[2 * kDirArr];
  __shared__ __align__(16) char st1[2 * kDirArr];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  xbar += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  ptm += int64_t(b) * pstride;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int e = lane & 7, c = lane >> 3;
  int64_t oin[2], oout[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    oin[h] = offs.off_in[8 * (2 * w + h) + e];
    oout[h] = offs.off_out[8 * (2 * w + h) + e];
  }
  const int n = lane & 15, kk = lane >> 4;
  const int64_t xo = offs.off_in[16 * w + n];
  float pb[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) pb[s] = ptm[(4 * s + kk) * 64 + 16 * w + n];
  __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));  // the PTM, before any tile load
  typedef Mfma<float>::acc Acc;
  Acc acc[4];
  double acc64[4][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    acc[cb] = Acc(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc64[cb][r] = 0.0;
  }
  auto flush = [&]() {
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[cb][r] += double(acc[cb][r]);
      acc[cb] = Acc(0);
    }
  };
  const uint32_t ntiles = n_orbits / kDirTile, stride = gridDim.x;
  uint32_t t = blockIdx.x;
  int n_acc = 0;
  if (t < ntiles) dir_issue<NT>(x, ybar, g, t, w, c, oin, oout, st0);
  while (t < ntiles) {
    const uint32_t t1 = t + stride;
    if (t1 < ntiles) {
      dir_issue<NT>(x, ybar, g, t1, w, c, oin, oout, st1);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(4));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    dir_compute(st0, w, lane, acc);
    bwd_apply(st0, pb, lane, g, t, xo, xbar);
    if (++n_acc == kDirFlush) { flush(); n_acc = 0; }
    lds_barrier();
    if (t1 >= ntiles) break;
    const uint32_t t2 = t1 + stride;
    if (t2 < ntiles) {
      dir_issue<NT>(x, ybar, g, t2, w, c, oin, oout, st0);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(4));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    dir_compute(st1, w, lane, acc);
    bwd_apply(st1, pb, lane, g, t1, xo, xbar);
    if (++n_acc == kDirFlush) { flush(); n_acc = 0; }
    lds_barrier();
    t = t2;
  }
  flush();
  const int col = lane & 15, k = lane >> 4;
  double* out = partial + (int64_t(b) * gridDim.x + blockIdx.x) * 64 * 64;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(16 * w + 4 * k + r) * 64 + 16 * cb + col] = acc64[cb][r];
}

// One-pass float32 backward of a dense 64 x 64 block whose last target (the
// PTM's lowest index) is the innermost axis (entries in aligned,
// contiguous runs of four: "quad" geometry).  Tile = 32 orbits; wave-load block ks (0..7) holds orbits
// 4ks..4ks+3, its lane l = 4q + okk taking entries 4q..4q+3 of orbit
// 4ks + okk (16 bytes; when the targets are the last three axes an orbit's
// 64 entries are one 256-byte run, so a wave-load reads 1 KB contiguous).
// Blocks are 1088 bytes apart (64 bytes of padding), which puts every
// operand read below on distinct banks.  Per tile:
//   dP += ybar x^T: wave w owns rows 16w..16w+15, k = the tile's orbits
//     (8 k-steps of 4), one 4-byte read per operand per lane;
//   xbar = P^T ybar: wave w computes output entries 16w..16w+15 (one row
//     tile) for both 16-orbit column tiles; k-step s covers entries
//     16 (s / 4) + 4 kk + s % 4 (k_fused64's grouping for 16-byte runs, so
//     xbar is bitwise the transposed apply's): lane (n, kk) reads a whole
//     quad of ybar (16 bytes, four k-steps), P's 16 fragments per lane
//     stay in registers, and accumulator slots r of lane (n, j) are the
//     quad 16w + 4j + r of orbit n: one 16-byte store per column tile.
constexpr int kQBlk = 1088;          // bytes per wave-load block (1024 + pad)
constexpr int kQArr = 8 * kQBlk;     // bytes per operand tile

template <bool NT>
__device__ __forceinline__ void quad_issue(const float* x, const float* ybar, const Geom& g,
                                           uint32_t t, int w, int lane, int64_t oiq,
                                           int64_t ooq, char* lds) {
  constexpr int kPol = NT ? 2 : 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int ks = 2 * w + h;
    int64_t bin, bout;
    orbit_base(g, t * kDirTile + 4 * ks + (lane & 3), bin, bout);
    __builtin_amdgcn_global_load_lds(x + bin + oiq, lds + ks * kQBlk, 16, 0, kPol);
    __builtin_amdgcn_global_load_lds(ybar + bout + ooq, lds + kQArr + ks * kQBlk, 16, 0,
                                     kPol);
  }
}

// Two k-steps (KS, KS + 1) of the gradient: A (ybar) and the four B (x)
// operands of each, read and waited for in one asm block.
#define QSV_Q_READ2(KS)                                                            \
  asm volatile(                                                                    \
      "ds_read_b32 %0, %10 offset:%c12\n"                                          \
      "ds_read_b32 %1, %11 offset:%c12\n"                                          \
      "ds_read_b32 %2, %11 offset:%c13\n"                                          \
      "ds_read_b32 %3, %11 offset:%c14\n"                                          \
      "ds_read_b32 %4, %11 offset:%c15\n"                                          \
      "ds_read_b32 %5, %10 offset:%c16\n"                                          \
      "ds_read_b32 %6, %11 offset:%c16\n"                                          \
      "ds_read_b32 %7, %11 offset:%c17\n"                                          \
      "ds_read_b32 %8, %11 offset:%c18\n"                                          \
      "ds_read_b32 %9, %11 offset:%c19\n"                                          \
      "s_waitcnt lgkmcnt(0)\n"                                                     \
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]),          \
        "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7]), "=&v"(v[8]), "=&v"(v[9])            \
      : "v"(ya), "v"(xb), "i"(kQBlk * (KS)), "i"(kQBlk * (KS) + 256),             \
        "i"(kQBlk * (KS) + 512), "i"(kQBlk * (KS) + 768), "i"(kQBlk * (KS + 1)),   \
        "i"(kQBlk * (KS + 1) + 256), "i"(kQBlk * (KS + 1) + 512),                 \
        "i"(kQBlk * (KS + 1) + 768)                                                \
      : "memory")

__device__ __forceinline__ void quad_grad(const char* lds, int w, int lane,
                                          Mfma<float>::acc (&acc)[4]) {
  const int m = lane & 15, kk = lane >> 4;
  const int row = 64 * (m >> 2) + 16 * kk + 4 * (m & 3);
  const uint32_t ya = lds_off(lds + kQArr + 256 * w + row);
  const uint32_t xb = lds_off(lds + row);
  float v[10];
#define QSV_Q_STEP2(KS)                                         \
  QSV_Q_READ2(KS);                                              \
  _Pragma("unroll") for (int h = 0; h < 2; ++h)                 \
  _Pragma("unroll") for (int cb = 0; cb < 4; ++cb)              \
    acc[cb] = Mfma<float>::op(v[5 * h], v[5 * h + 1 + cb], acc[cb]);
  QSV_Q_STEP2(0)
  QSV_Q_STEP2(2)
  QSV_Q_STEP2(4)
  QSV_Q_STEP2(6)
#undef QSV_Q_STEP2
}

typedef float qsv_f4 __attribute__((ext_vector_type(4)));

// The ybar quads of one column tile (four groups of four k-steps).
#define QSV_Q_READQ(CT)                                                            \
  asm volatile(                                                                    \
      "ds_read_b128 %0, %4 offset:%c5\n"                                           \
      "ds_read_b128 %1, %4 offset:%c6\n"                                           \
      "ds_read_b128 %2, %4 offset:%c7\n"                                           \
      "ds_read_b128 %3, %4 offset:%c8\n"                                           \
      "s_waitcnt lgkmcnt(0)\n"                                                     \
      : "=&v"(bq[0]), "=&v"(bq[1]), "=&v"(bq[2]), "=&v"(bq[3])                     \
      : "v"(yb), "i"(4 * kQBlk * (CT)), "i"(4 * kQBlk * (CT) + 256),              \
        "i"(4 * kQBlk * (CT) + 512), "i"(4 * kQBlk * (CT) + 768)                   \
      : "memory")

__device__ __forceinline__ void quad_apply(const char* lds, const float (&pa)[16], int lane,
                                           const Geom& g, uint32_t t, int64_t xo,
                                           float* __restrict__ xbar) {
  const int n = lane & 15, kk = lane >> 4;
  const uint32_t yb = lds_off(lds + kQArr + (n >> 2) * kQBlk + 64 * kk + 16 * (n & 3));
  typedef Mfma<float>::acc Acc;
  qsv_f4 bq[4];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    if (ct == 0) {
      QSV_Q_READQ(0);
    } else {
      QSV_Q_READQ(1);
    }
    Acc a = Acc(0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e) a = Mfma<float>::op(pa[4 * u + e], bq[u][e], a);
    int64_t bin, bout;
    orbit_base(g, t * kDirTile + 16 * ct + n, bin, bout);
    *reinterpret_cast<qsv_f4*>(xbar + bin + xo) = a;
  }
}

template <bool NT>
__global__ void __launch_bounds__(kBlock)
    k_bwd64q(const float* __restrict__ x, const float* __restrict__ ybar,
             float* __restrict__ xbar, const Geom g, const uint32_t n_orbits,
             const float* __restrict__ ptm, const int64_t pstride, const GradOffs offs,
             double* __restrict__ partial) {
  __shared__ __align__(16) char st0[2 * kQArr];
  __shared__ __align__(16) char st1[2 * kQArr];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  xbar += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  ptm += int64_t(b) * pstride;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t oiq = offs.off_in[4 * (lane >> 2)], ooq = offs.off_out[4 * (lane >> 2)];
  const int m = lane & 15, kk = lane >> 4;
  // A of the adjoint apply: P^T[16w + m][k] = P[k][16w + m], k-step s at
  // entry 16 (s / 4) + 4 kk + s % 4
  float pa[16];
#pragma unroll
  for (int sk = 0; sk < 16; ++sk)
    pa[sk] = ptm[(16 * (sk >> 2) + 4 * kk + (sk & 3)) * 64 + 16 * w + m];
  const int64_t xo = offs.off_in[16 * w + 4 * kk];
  __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));  // the PTM, before any tile load
  typedef Mfma<float>::acc Acc;
  Acc acc[4];
  double acc64[4][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    acc[cb] = Acc(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc64[cb][r] = 0.0;
  }
  auto flush = [&]() {
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[cb][r] += double(acc[cb][r]);
      acc[cb] = Acc(0);
    }
  };
  const uint32_t ntiles = n_orbits / kDirTile, stride = gridDim.x;
  uint32_t t = blockIdx.x;
  int n_acc = 0;
  if (t < ntiles) quad_issue<NT>(x, ybar, g, t, w, lane, oiq, ooq, st0);
  while (t < ntiles) {
    const uint32_t t1 = t + stride;
    if (t1 < ntiles) {
      quad_issue<NT>(x, ybar, g, t1, w, lane, oiq, ooq, st1);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(4));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    quad_grad(st0, w, lane, acc);
    quad_apply(st0, pa, lane, g, t, xo, xbar);
    if (++n_acc == kDirFlush) { flush(); n_acc = 0; }
    lds_barrier();
    if (t1 >= ntiles) break;
    const uint32_t t2 = t1 + stride;
    if (t2 < ntiles) {
      quad_issue<NT>(x, ybar, g, t2, w, lane, oiq, ooq, st0);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(4));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    quad_grad(st1, w, lane, acc);
    quad_apply(st1, pa, lane, g, t1, xo, xbar);
    if (++n_acc == kDirFlush) { flush(); n_acc = 0; }
    lds_barrier();
    t = t2;
  }
  flush();
  const int col = lane & 15, k = lane >> 4;
  double* out = partial + (int64_t(b) * gridDim.x + blockIdx.x) * 64 * 64;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(16 * w + 4 * k + r) * 64 + 16 * cb + col] = acc64[cb][r];
}

// Float32 64 x 64 apply y = P x for the quad geometry on direct-loaded
// tiles: x staged as in k_bwd64q (one operand), wave w computes output
// entries 16w..16w+15 of both 16-orbit column tiles with the k-steps of
// k_fused64's runs-of-four grouping (y bitwise k_fused64's), P's 16
// fragments per lane in registers, one 16-byte store per column tile.
template <bool NT>
__device__ __forceinline__ void quad_issue_x(const float* x, const Geom& g, uint32_t t, int w,
                                             int lane, int64_t oiq, char* lds) {
  constexpr int kPol = NT ? 2 : 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int ks = 2 * w + h;
    int64_t bin, bout;
    orbit_base(g, t * kDirTile + 4 * ks + (lane & 3), bin, bout);
    __builtin_amdgcn_global_load_lds(x + bin + oiq, lds + ks * kQBlk, 16, 0, kPol);
  }
}

__device__ __forceinline__ void quad_apply_fwd(const char* lds, const float (&pa)[16], int lane,
                                               const Geom& g, uint32_t t, int64_t yo,
                                               float* __restrict__ y) {
  const int n = lane & 15, kk = lane >> 4;
  const uint32_t yb = lds_off(lds + (n >> 2) * kQBlk + 64 * kk + 16 * (n & 3));
  typedef Mfma<float>::acc Acc;
  qsv_f4 bq[4];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    if (ct == 0) {
      QSV_Q_READQ(0);
    } else {
      QSV_Q_READQ(1);
    }
    Acc a = Acc(0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e) a = Mfma<float>::op(pa[4 * u + e], bq[u][e], a);
    int64_t bin, bout;
    orbit_base(g, t * kDirTile + 16 * ct + n, bin, bout);
    *reinterpret_cast<qsv_f4*>(y + bout + yo) = a;
  }
}

template <bool NT>
__global__ void __launch_bounds__(kBlock)
    k_apply64q(const float* __restrict__ x, float* __restrict__ y, const Geom g,
               const uint32_t n_orbits, const float* __restrict__ ptm, const int64_t pstride,
               const GradOffs offs) {
  __shared__ __align__(16) char st0[kQArr];
  __shared__ __align__(16) char st1[kQArr];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  y += int64_t(b) * g.bs_out;
  ptm += int64_t(b) * pstride;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t oiq = offs.off_in[4 * (lane >> 2)];
  const int m = lane & 15, kk = lane >> 4;
  float pa[16];
#pragma unroll
  for (int sk = 0; sk < 16; ++sk)
    pa[sk] = ptm[(16 * w + m) * 64 + 16 * (sk >> 2) + 4 * kk + (sk & 3)];
  const int64_t yo = offs.off_out[16 * w + 4 * kk];
  __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));  // the PTM, before any tile load
  const uint32_t ntiles = n_orbits / kDirTile, stride = gridDim.x;
  uint32_t t = blockIdx.x;
  if (t < ntiles) quad_issue_x<NT>(x, g, t, w, lane, oiq, st0);
  while (t < ntiles) {
    const uint32_t t1 = t + stride;
    if (t1 < ntiles) {
      quad_issue_x<NT>(x, g, t1, w, lane, oiq, st1);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(2));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    quad_apply_fwd(st0, pa, lane, g, t, yo, y);
    lds_barrier();
    if (t1 >= ntiles) break;
    const uint32_t t2 = t1 + stride;
    if (t2 < ntiles) {
      quad_issue_x<NT>(x, g, t2, w, lane, oiq, st0);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(2));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    lds_barrier();
    quad_apply_fwd(st1, pa, lane, g, t1, yo, y);
    lds_barrier();
    t = t2;
  }
}

// Resident blocks of `kernel` over the device (occupancy x CUs).
static uint32_t resident_blocks(const void* kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  (void)hipGetLastError();
  return uint32_t(per_cu) * uint32_t(cus);
}

// Host side of k_apply64q: natural per-entry offsets from the fused plan's
// split for runs of four (entry 16u + 4j + r at in_u[u] + in_l[j] + r).
bool apply64q_applies(int elem_bytes, const FusedPlan& fp) {
  return elem_bytes == 4 && fp.D == 64 && fp.vi == 4 && fp.vo == 4 && fp.vn == 1 &&
         fp.n_orbits % kDirTile == 0;
}

int launch_apply64q(const FusedPlan& fp, const Geom& g, bool nt, const void* in, void* out,
                    const void* ptm_rows, int64_t pstride, uint32_t batch, hipStream_t s) {
  GradOffs o;
  for (int u = 0; u < 4; ++u)
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 4; ++r) {
        o.off_in[16 * u + 4 * j + r] = fp.offs.in_u[u] + fp.offs.in_l[j] + r;
        o.off_out[16 * u + 4 * j + r] = fp.offs.out_u[u] + fp.offs.out_l[j] + r;
      }
  static const uint32_t cap =
      resident_blocks(reinterpret_cast<const void*>(k_apply64q<true>));
  const uint32_t tiles = fp.n_orbits / kDirTile;
  uint32_t nb = cap / std::max<uint32_t>(batch, 1);
  nb = std::max<uint32_t>(1, std::min(nb, std::max<uint32_t>(tiles / 4, 1)));
  auto kern = nt ? k_apply64q<true> : k_apply64q<false>;
  hipLaunchKernelGGL(kern, dim3(nb, batch), dim3(kBlock), 0, s, static_cast<const float*>(in),
                     static_cast<float*>(out), g, fp.n_orbits,
                     static_cast<const float*>(ptm_rows), pstride, o);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

// PTMs above 64 x 64 (qutrit pairs 81 x 81, ...): target dims and strides,
// the entry offsets decoded by the kernel (mixed radix, last target fastest).
constexpr int kGradBigMaxSide = 256;
constexpr int kGradBigMaxEntries = 8192;  // dout x din
constexpr int kGradBigOrbits = 8;         // orbits staged in LDS per step
constexpr int kGradBigBlocks = 64;
struct BigGeo {
  int32_t k;
  int32_t din_t[QSV_MAX_TARGETS], dout_t[QSV_MAX_TARGETS];
  int64_t sin_t[QSV_MAX_TARGETS], sout_t[QSV_MAX_TARGETS];
};

__host__ __device__ inline int64_t big_offset(int k, const int32_t* d, const int64_t* st, int e) {
  int64_t off = 0;
  for (int j = k - 1; j >= 0; --j) {
    off += int64_t(e % d[j]) * st[j];
    e /= d[j];
  }
  return off;
}

// PTMs above 64 x 64 (qutrit pairs, ...): dP[o][i] = sum over orbits of
// ybar[o] x[i] on the VALU in float64.  A block takes chunks of
// kGradBigOrbits orbits (chunk c, c + gridDim.x, ...), stages their x and
// ybar entries in LDS (offsets decoded from the targets' dims and strides),
// and thread t accumulates entries (o, i) = t, t + kBlock, ... of the
// block's partial tile (same layout as the other gradient kernels, summed
// by k_ptm_grad_final).
constexpr int kGradBigPerThread = kGradBigMaxEntries / kBlock;
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad_big(const T* __restrict__ x, const T* __restrict__ ybar, const Geom g,
                   const BigGeo bg, const uint32_t n_orbits, double* __restrict__ partial) {
  __shared__ int64_t so_in[kGradBigMaxSide], so_out[kGradBigMaxSide];
  __shared__ double sx[kGradBigOrbits * kGradBigMaxSide], sy[kGradBigOrbits * kGradBigMaxSide];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  const int tid = threadIdx.x, din = g.din, dout = g.dout, dd = din * dout;
  for (int e = tid; e < din; e += kBlock) so_in[e] = big_offset(bg.k, bg.din_t, bg.sin_t, e);
  for (int e = tid; e < dout; e += kBlock) so_out[e] = big_offset(bg.k, bg.dout_t, bg.sout_t, e);
  double acc[kGradBigPerThread];
#pragma unroll
  for (int r = 0; r < kGradBigPerThread; ++r) acc[r] = 0.0;
  const uint32_t nchunks = (n_orbits + kGradBigOrbits - 1) / kGradBigOrbits;
  __syncthreads();
  for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    for (int f = tid; f < kGradBigOrbits * din; f += kBlock) {
      const int ol = f / din, i = f - ol * din;
      const uint32_t m = c * kGradBigOrbits + ol;
      double v = 0.0;
      if (m < n_orbits) {
        int64_t bin, bout;
        orbit_base(g, m, bin, bout);
        v = double(x[bin + so_in[i]]);
      }
      sx[ol * kGradBigMaxSide + i] = v;
    }
    for (int f = tid; f < kGradBigOrbits * dout; f += kBlock) {
      const int ol = f / dout, o = f - ol * dout;
      const uint32_t m = c * kGradBigOrbits + ol;
      double v = 0.0;
      if (m < n_orbits) {
        int64_t bin, bout;
        orbit_base(g, m, bin, bout);
        v = double(ybar[bout + so_out[o]]);
      }
      sy[ol * kGradBigMaxSide + o] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kGradBigPerThread; ++r) {
      const int j = tid + r * kBlock;
      if (j < dd) {
        const int o = j / din, i = j - o * din;
        double s = acc[r];
#pragma unroll
        for (int ol = 0; ol < kGradBigOrbits; ++ol)
          s += sy[ol * kGradBigMaxSide + o] * sx[ol * kGradBigMaxSide + i];
        acc[r] = s;
      }
    }
    __syncthreads();
  }
  double* out = partial + (int64_t(b) * gridDim.x + blockIdx.x) * dd;
#pragma unroll
  for (int r = 0; r < kGradBigPerThread; ++r) {
    const int j = tid + r * kBlock;
    if (j < dd) out[j] = acc[r];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad_final(const double* __restrict__ partial, int nblocks, int dd,
                     T* __restrict__ out, double scale_out_accumulate) {
  // one thread per (b, o, i): fixed-order sum over the block partials
  const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (j >= dd) return;
  double s = 0.0;
  for (int k = 0; k < nblocks; ++k) s += partial[(b * nblocks + k) * dd + j];
  if (scale_out_accumulate != 0.0)
    out[b * dd + j] = T(double(out[b * dd + j]) + s);
  else
    out[b * dd + j] = T(s);
}

struct GradPlan {
  Geom g;
  GradOffs offs;
  uint32_t n_orbits;
  int din, dout;
  int vn;    // orbits per 16-byte lane vector (1: scalar loads)
  int quad;  // 64 x 64 with a unit-stride lowest index: 16-byte loads of 4 entries
  int64_t n_in, n_out;
  int big;   // above 64 on a side: k_ptm_grad_big, offsets decoded on the device
  BigGeo bg;
};

// Tile of the LDS-staged gradient (float64 states only).
constexpr int kGrad64sTile = 16;

// The staged kernel applies: float64, square 64 x 64, orbit vectors.
// Float32 64 x 64 with 16-byte orbit vectors and whole 32-orbit tiles: the
// direct-to-LDS kernel k_ptm_grad64d.
static bool grad64_direct(const GradPlan& p, int elem_bytes, int vn) {
  return elem_bytes == 4 && !p.big && p.din == 64 && p.dout == 64 && vn == 4 &&
         p.n_orbits % kDirTile == 0;
}

// Float32 64 x 64 whose entries come in aligned runs of four (the innermost
// axis a target) and whole 32-orbit tiles: the one-pass backward k_bwd64q.
static bool bwd64q_applies(const GradPlan& p) {
  return !p.big && p.din == 64 && p.dout == 64 && p.quad && p.n_orbits % kDirTile == 0;
}

static bool grad64_staged(const GradPlan& p, int elem_bytes, int vn) {
  return elem_bytes == 8 && p.din == 64 && p.dout == 64 && vn == 16 / elem_bytes;
}

// Geometry of the contraction without vector widths (one orbit per step).
static int make_grad_plan(int elem_bytes, int rank, const int64_t* dims, int k,
                          const int32_t* targets, const int32_t* dims_out_t, GradPlan& p) {
  QSV_CHECK_ARG(rank >= 1 && rank <= QSV_MAX_RANK, "rank %d outside [1, %d]", rank,
                QSV_MAX_RANK);
  QSV_CHECK_ARG(k >= 1 && k <= QSV_MAX_TARGETS && k <= rank, "bad number of targets %d", k);
  QSV_CHECK_ARG(dims && targets && dims_out_t, "null shape argument");
  bool is_target[QSV_MAX_RANK] = {false};
  int64_t dout_full[QSV_MAX_RANK];
  for (int q = 0; q < rank; ++q) {
    QSV_CHECK_ARG(dims[q] >= 1 && dims[q] <= 4096, "dims[%d] invalid", q);
    dout_full[q] = dims[q];
  }
  for (int j = 0; j < k; ++j) {
    const int q = targets[j];
    QSV_CHECK_ARG(q >= 0 && q < rank && !is_target[q], "bad or repeated target %d", q);
    QSV_CHECK_ARG(dims_out_t[j] >= 1, "bad output dim");
    is_target[q] = true;
    dout_full[q] = dims_out_t[j];
  }
  int64_t sin[QSV_MAX_RANK], sout[QSV_MAX_RANK], ai = 1, ao = 1;
  for (int q = rank - 1; q >= 0; --q) {
    sin[q] = ai;
    sout[q] = ao;
    ai *= dims[q];
    ao *= dout_full[q];
    QSV_CHECK_ARG(ai < (int64_t(1) << 40) && ao < (int64_t(1) << 40), "vector too large");
  }
  p.n_in = ai;
  p.n_out = ao;
  int64_t din = 1, dout = 1;
  for (int j = 0; j < k; ++j) {
    din *= dims[targets[j]];
    dout *= dims_out_t[j];
  }
  QSV_CHECK_ARG(din <= kGradBigMaxSide && dout <= kGradBigMaxSide &&
                    din * dout <= kGradBigMaxEntries,
                "PTM gradient supports up to %d entries per side and %d entries, got %lld x %lld",
                kGradBigMaxSide, kGradBigMaxEntries, (long long)dout, (long long)din);
  p.din = int(din);
  p.dout = int(dout);
  memset(&p.offs, 0, sizeof(p.offs));
  memset(&p.bg, 0, sizeof(p.bg));
  p.big = din > 64 || dout > 64;
  p.bg.k = k;
  for (int j = 0; j < k; ++j) {
    p.bg.din_t[j] = int32_t(dims[targets[j]]);
    p.bg.dout_t[j] = int32_t(dims_out_t[j]);
    p.bg.sin_t[j] = sin[targets[j]];
    p.bg.sout_t[j] = sout[targets[j]];
  }
  for (int64_t i = 0; i < din && !p.big; ++i) {
    int64_t r = i, off = 0;
    for (int j = k - 1; j >= 0; --j) {
      off += (r % dims[targets[j]]) * sin[targets[j]];
      r /= dims[targets[j]];
    }
    p.offs.off_in[i] = off;
  }
  for (int64_t o = 0; o < dout && !p.big; ++o) {
    int64_t r = o, off = 0;
    for (int j = k - 1; j >= 0; --j) {
      off += (r % dims_out_t[j]) * sout[targets[j]];
      r /= dims_out_t[j];
    }
    p.offs.off_out[o] = off;
  }
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
  g.bs_in = p.n_in;
  g.bs_out = p.n_out;
  QSV_CHECK_ARG(M < (int64_t(1) << 32) - kBlock, "too many orbits");
  p.n_orbits = uint32_t(M);
  g.n_threads = p.n_orbits;
  // lanes load 16-byte vectors of consecutive orbits when the innermost axis
  // is a unit-stride bystander and every base, offset and batch stride keeps
  // them aligned
  p.vn = 1;
  const int want = 16 / elem_bytes;
  if (!p.big && ng > 0 && g.gs_in[0] == 1 && g.gs_out[0] == 1 && g.gdiv[0].d % want == 0 &&
      p.n_in % want == 0 && p.n_out % want == 0) {
    bool aligned = true;
    for (int j = 1; j < ng; ++j)
      aligned = aligned && g.gs_in[j] % want == 0 && g.gs_out[j] % want == 0;
    for (int i = 0; i < p.din; ++i) aligned = aligned && p.offs.off_in[i] % want == 0;
    for (int o = 0; o < p.dout; ++o) aligned = aligned && p.offs.off_out[o] % want == 0;
    if (aligned) p.vn = want;
  }
  // 64 x 64 PTMs whose entries come in aligned runs of four (the innermost
  // axis is a target of Pauli dimension 4)
  p.quad = 0;
  if (!p.big && p.vn == 1 && p.din == 64 && p.dout == 64 && p.n_in % want == 0 &&
      p.n_out % want == 0) {
    bool ok = true;
    for (int e = 0; e < 64; ++e)
      ok = ok && p.offs.off_in[e] == p.offs.off_in[e & ~3] + (e & 3) &&
           p.offs.off_out[e] == p.offs.off_out[e & ~3] + (e & 3);
    for (int e = 0; e < 64; e += 4)
      ok = ok && p.offs.off_in[e] % want == 0 && p.offs.off_out[e] % want == 0;
    for (int j = 0; j < ng; ++j) ok = ok && g.gs_in[j] % want == 0 && g.gs_out[j] % want == 0;
    if (ok) p.quad = 1;
  }
  return QSV_OK;
}

// blocks per state: at least 1024 orbits per block, at most kGradBlocks,
// and no more than needed for ~4096 blocks over the whole batch (fewer
// partial tiles for the final sum to read)
static int grad_blocks(uint32_t n_orbits, int batch) {
  uint32_t b = (n_orbits + 1023) / 1024;
  const uint32_t fill = uint32_t((4096 + std::max(batch, 1) - 1) / std::max(batch, 1));
  b = std::min(b, fill);
  return int(b < 1 ? 1 : (b > kGradBlocks ? kGradBlocks : b));
}

// Larger than 16 x 16: 16 times the MFMA work per orbit, 32 KB partial
// tiles; at least 512 orbits per block, ~2048 blocks over the batch, at
// most kGrad64Blocks per state.
static bool grad_wide(const GradPlan& p) { return p.din > 16 || p.dout > 16; }

static int grad64_blocks(uint32_t n_orbits, int batch) {
  uint32_t b = (n_orbits + 511) / 512;
  const uint32_t fill = uint32_t((2048 + std::max(batch, 1) - 1) / std::max(batch, 1));
  b = std::min(b, fill);
  return int(b < 1 ? 1 : (b > kGrad64Blocks ? kGrad64Blocks : b));
}

// Above 64 x 64: at least 256 orbits per block, ~1024 blocks over the
// batch, at most kGradBigBlocks per state (partial tiles up to 64 KB).
static int grad_big_blocks(uint32_t n_orbits, int batch) {
  uint32_t b = (n_orbits + 255) / 256;
  const uint32_t fill = uint32_t((1024 + std::max(batch, 1) - 1) / std::max(batch, 1));
  b = std::min(b, fill);
  return int(b < 1 ? 1 : (b > kGradBigBlocks ? kGradBigBlocks : b));
}

static int max_grad_blocks(const GradPlan& p) {
  return p.big ? kGradBigBlocks : (grad_wide(p) ? kGrad64Blocks : kGradBlocks);
}

static int default_grad_blocks(const GradPlan& p, int batch) {
  return p.big ? grad_big_blocks(p.n_orbits, batch)
               : (grad_wide(p) ? grad64_blocks(p.n_orbits, batch)
                               : grad_blocks(p.n_orbits, batch));
}

// Fused backward of a dense 64 x 64 block of float64 states (float32 runs
// k_bwd64d) when the innermost axis is a unit-stride bystander:
//   xbar = P^T ybar   and   dP += sum over orbits ybar x^T
// in one pass, reading x and ybar once from HBM and writing xbar (three
// streams instead of the four of a transposed apply plus a gradient).  A
// wave walks tiles of 16 VN orbits; per tile:
//   A. the k_fused64 contraction with the transposed PTM (gathered into
//      fragment order in LDS): lane (n, j) loads VN-orbit vectors of ybar
//      for entries 4 s + j, the accumulators are stored to xbar;
//   B. the k_ptm_grad64 contraction over the same orbits, four k-steps of
//      4 VN orbits: lane (c, k) loads entries 16 q + c of x and of ybar
//      (ybar's second read of the tile comes from the caches).
// The dP tiles stay in registers across tiles; at the end the four waves
// are summed in LDS (the PTM's space, after a barrier) in a fixed order and
// the block writes its partial, summed by k_ptm_grad_final.
template <typename T, int VN>
__global__ void __launch_bounds__(kBlock)
    k_bwd64(const T* __restrict__ x, const T* __restrict__ ybar, T* __restrict__ xbar,
            const Geom g, const uint32_t n_orbits, const T* __restrict__ ptm,
            const int64_t pstride, const GradOffs offs, const FusedOffs fo,
            double* __restrict__ partial) {
  constexpr int MB = 4, KS = 16, VH = 2;
  constexpr uint32_t COLS = 16 * VN, STEP = 4 * VN;
  typedef typename Mfma<T>::acc Acc;
  typedef typename VT<T, VN>::t Vec;
  typedef typename VT<T, VH>::t VHv;
  __shared__ double smem[64 * 64];
  T* As = reinterpret_cast<T*>(smem);
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  xbar += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  ptm += int64_t(b) * pstride;
  for (int i = threadIdx.x; i < KS * MB * 64; i += kBlock) {
    int o, kk;
    fused_frag_coords(64, int(sizeof(T)), 1, 1, i, o, kk);
    As[i] = ptm[int64_t(kk) * 64 + o];  // fragment (o, kk) of P^T
  }
  __syncthreads();
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = lane & 15, j = lane >> 4;  // part A: orbit column n, k-group j
  const int c = n, kq = j;                 // part B: entry c, orbit group kq
  const uint32_t nw = gridDim.x * (kBlock / 64);
  const uint32_t wid = blockIdx.x * (kBlock / 64) + w;
  const uint32_t full = n_orbits / COLS;
  const int64_t li = fo.in_l[j], lo = fo.out_l[j];
  int64_t oi[4], oo[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    oi[q] = offs.off_in[16 * q + c];
    oo[q] = offs.off_out[16 * q + c];
  }
  Acc dacc[4][4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) dacc[rb][cb] = Acc(0);

  auto tile = [&](uint32_t t, bool tail) {
    // A: xbar = P^T ybar on orbits t COLS + VN n + e, in halves of VH
    // orbits (float32: 8-byte accesses, still whole 128-byte lines over 16
    // lanes) to keep the registers of both parts within two waves per SIMD
    {
      const uint32_t m0 = t * COLS + VN * n;
      const uint32_t m = m0 < n_orbits ? m0 : n_orbits - VN;  // tail: loaded, not stored
      int64_t bin, bout;
      orbit_base(g, m, bin, bout);
      const T* py = ybar + bout + li;
      T* px = xbar + bin + lo;
      const bool st = !tail || m0 < n_orbits;
#pragma unroll
      for (int h = 0; h < VN / VH; ++h) {
        T bv[VH][KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const VHv v = *reinterpret_cast<const VHv*>(py + h * VH + fo.in_u[s]);
#pragma unroll
          for (int e = 0; e < VH; ++e) bv[e][s] = v[e];
        }
        Acc acc[MB][VH];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int nb = 0; nb < VH; ++nb) acc[mb][nb] = Acc(0);
        int abase = lane;
        asm volatile("" : "+v"(abase));
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) {
            const T a = As[(s * MB + mb) * 64 + abase];
#pragma unroll
            for (int nb = 0; nb < VH; ++nb) acc[mb][nb] = Mfma<T>::op(a, bv[nb][s], acc[mb][nb]);
          }
        if (st) {
#pragma unroll
          for (int q = 0; q < 4 * MB; ++q) {
            VHv v;
#pragma unroll
            for (int e = 0; e < VH; ++e) v[e] = acc[q >> 2][e][q & 3];
            *reinterpret_cast<VHv*>(px + h * VH + fo.out_u[q]) = v;
          }
        }
      }
    }
    // B: dP over the same orbits, four k-steps of 4 VN orbits
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      const uint32_t m = t * COLS + tt * STEP + VN * kq;
      const bool ok = m < n_orbits;
      int64_t bin, bout;
      orbit_base(g, ok ? m : 0, bin, bout);
      Vec xv[4], yv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xv[q] = *reinterpret_cast<const Vec*>(x + (ok ? bin + oi[q] : 0));
        yv[q] = *reinterpret_cast<const Vec*>(ybar + (ok ? bout + oo[q] : 0));
        if (!ok) {
          xv[q] = Vec(0);
          yv[q] = Vec(0);
        }
      }
#pragma unroll
      for (int e = 0; e < VN; ++e)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
            dacc[rb][cb] = Mfma<T>::op(yv[rb][e], xv[cb][e], dacc[rb][cb]);
    }
  };
  for (uint32_t t = wid; t < full; t += nw) tile(t, false);
  if (full * COLS < n_orbits && wid == full % nw) tile(full, true);

  __syncthreads();  // every wave is done with the PTM: its LDS now sums dP
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
    for (int ww = 0; ww < kBlock / 64; ++ww) {
      if (w == ww) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = sizeof(T) == 8 ? kq + 4 * r : 4 * kq + r;
            const int idx = (16 * rb + row) * 64 + 16 * cb + c;
            smem[idx] = (ww == 0 ? 0.0 : smem[idx]) + double(dacc[rb][cb][r]);
          }
      }
      __syncthreads();
    }
  }
  for (int jj = tid; jj < 64 * 64; jj += kBlock)
    partial[(int64_t(b) * gridDim.x + blockIdx.x) * 64 * 64 + jj] = smem[jj];
}

// Offsets of the transposed apply (input: ybar, laid out like the forward's
// output; output: xbar, like the forward's input), vector widths 1 along
// the PTM index.
static FusedOffs bwd64_offsets(const GradPlan& p) {
  FusedOffs fo;
  memset(&fo, 0, sizeof(fo));
  for (int u = 0; u < 16; ++u) fo.in_u[u] = p.offs.off_out[fused_k(1, u, 0)];
  for (int v = 0; v < 16; ++v) fo.out_u[v] = p.offs.off_in[fused_o(1, v, 0)];
  for (int j = 0; j < 4; ++j) {
    fo.in_l[j] = p.offs.off_out[fused_k(1, 0, j)];
    fo.out_l[j] = p.offs.off_in[fused_o(1, 0, j)];
  }
  return fo;
}

// The fused backward applies: square 64 x 64 PTM, offsets additive over the
// index bits (as the MFMA maps need), orbit vectors (p.vn > 1).
static bool bwd64_applies(const GradPlan& p) {
  if (p.din != 64 || p.dout != 64 || p.vn <= 1) return false;
  for (int k = 0; k < 64; ++k) {
    int64_t si = 0, so = 0;
    for (int bit = 1; bit < 64; bit <<= 1)
      if (k & bit) {
        si += p.offs.off_in[bit];
        so += p.offs.off_out[bit];
      }
    if (si != p.offs.off_in[k] || so != p.offs.off_out[k]) return false;
  }
  return true;
}

}  // namespace qsv

using namespace qsv;

extern "C" int64_t qsv_ptm_grad_workspace_bytes(int batch, int rank, const int64_t* dims, int k,
                                                const int32_t* targets,
                                                const int32_t* dims_out_t) {
  GradPlan p;
  const int rc = make_grad_plan(8, rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  return int64_t(batch) * max_grad_blocks(p) * p.din * p.dout * 8;
}

template <typename T>
static void launch_grad(const GradPlan& p, int vn, bool quad, dim3 grid, const void* x,
                        const void* ybar, double* partial, hipStream_t s) {
  const T* xx = static_cast<const T*>(x);
  const T* yy = static_cast<const T*>(ybar);
  if (p.big) {
    hipLaunchKernelGGL(k_ptm_grad_big<T>, grid, dim3(kBlock), 0, s, xx, yy, p.g, p.bg,
                       p.n_orbits, partial);
    return;
  }
  if (grad_wide(p)) {
    if constexpr (sizeof(T) == 4) {
      if (grad64_direct(p, 4, vn)) {
        const bool nt = (p.n_in + p.n_out) * 4 * int64_t(grid.y) > nt_threshold_bytes();
        if (nt)
          hipLaunchKernelGGL(k_ptm_grad64d<true>, grid, dim3(kBlock), 0, s, xx, yy, p.g, p.offs,
                             p.n_orbits, partial);
        else
          hipLaunchKernelGGL(k_ptm_grad64d<false>, grid, dim3(kBlock), 0, s, xx, yy, p.g,
                             p.offs, p.n_orbits, partial);
        return;
      }
    }
    if (grad64_staged(p, int(sizeof(T)), vn))
      hipLaunchKernelGGL((k_ptm_grad64s<T, 16 / sizeof(T), kGrad64sTile>), grid, dim3(kBlock), 0,
                         s, xx, yy, p.g, p.offs, p.n_orbits, partial);
    else if (vn == 16 / int(sizeof(T)))
      hipLaunchKernelGGL((k_ptm_grad64<T, 16 / sizeof(T), false>), grid, dim3(kBlock), 0, s, xx,
                         yy, p.g, p.offs, p.n_orbits, partial);
    else if (quad)
      hipLaunchKernelGGL((k_ptm_grad64<T, 1, true>), grid, dim3(kBlock), 0, s, xx, yy, p.g,
                         p.offs, p.n_orbits, partial);
    else
      hipLaunchKernelGGL((k_ptm_grad64<T, 1, false>), grid, dim3(kBlock), 0, s, xx, yy, p.g,
                         p.offs, p.n_orbits, partial);
    return;
  }
  if (vn == 16 / int(sizeof(T)))
    hipLaunchKernelGGL((k_ptm_grad<T, 16 / sizeof(T)>), grid, dim3(kBlock), 0, s, xx, yy, p.g,
                       p.offs, p.n_orbits, partial);
  else
    hipLaunchKernelGGL((k_ptm_grad<T, 1>), grid, dim3(kBlock), 0, s, xx, yy, p.g, p.offs,
                       p.n_orbits, partial);
}

extern "C" int qsv_ptm_grad(int dtype, const void* x, const void* ybar, int batch, int rank,
                            const int64_t* dims, int k, const int32_t* targets,
                            const int32_t* dims_out_t, void* dptm, int accumulate,
                            void* workspace, int64_t workspace_bytes, void* stream) {
  return qsv_ptm_grad_tiled(dtype, x, ybar, batch, rank, dims, k, targets, dims_out_t, dptm,
                            accumulate, workspace, workspace_bytes, nullptr, stream);
}

// Launches of qsv_ptm_grad_tiled and qsv_ptm_grad_noalloc: the partial
// tiles go to `partial` (capacity `partial_bytes`).  With `shrink` the
// planner's blocks per state are lowered until the tiles fit (the result
// changes only in the order of the float64 partial sums); without it a
// workspace that is too small is an error.
static int ptm_grad_run(int dtype, const void* x, const void* ybar, int batch, int rank,
                        const int64_t* dims, int k, const int32_t* targets,
                        const int32_t* dims_out_t, void* dptm, int accumulate, void* partial_buf,
                        int64_t partial_bytes, bool shrink, const qsv_tile* tile,
                        void* stream) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(!tile || (tile->grad_blocks >= 0 && tile->grad_blocks <= kGradBlocks),
                "tile: grad_blocks outside [0, %d]", kGradBlocks);
  QSV_CHECK_ARG(batch >= 0 && batch <= 65535, "batch %d outside [0, 65535]", batch);
  GradPlan p;
  const int rc = make_grad_plan(dtype_size(dtype), rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  if (batch == 0) return QSV_OK;
  QSV_CHECK_ARG(x && ybar && dptm, "null pointer argument");
  QSV_CHECK_ARG(!tile || tile->grad_blocks <= max_grad_blocks(p),
                "tile: grad_blocks %d above %d for a %d x %d PTM", tile ? tile->grad_blocks : 0,
                max_grad_blocks(p), p.dout, p.din);
  int nb = tile && tile->grad_blocks > 0 ? tile->grad_blocks : default_grad_blocks(p, batch);
  const int64_t per_block = int64_t(batch) * p.din * p.dout * 8;
  if (shrink) {
    QSV_CHECK_ARG(per_block <= partial_bytes,
                  "PTM gradient of %d states with a %d x %d PTM needs at least %lld bytes of "
                  "scratch, above the %lld this entry point holds",
                  batch, p.dout, p.din, (long long)per_block, (long long)partial_bytes);
    nb = int(std::min<int64_t>(nb, partial_bytes / per_block));
  }
  const int64_t need = per_block * nb;
  QSV_CHECK_ARG(partial_buf && partial_bytes >= need, "PTM gradient needs %lld workspace bytes",
                (long long)need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* partial = static_cast<double*>(partial_buf);
  const dim3 grid(nb, batch);
  const bool aligned = reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(ybar) % 16 == 0;
  const int vn = aligned ? p.vn : 1;
  const bool quad = aligned && p.quad;
  if (dtype == QSV_F64)
    launch_grad<double>(p, vn, quad, grid, x, ybar, partial, s);
  else
    launch_grad<float>(p, vn, quad, grid, x, ybar, partial, s);
  QSV_CHECK_HIP(hipGetLastError());
  const int dd = p.din * p.dout;
  const dim3 g2((dd + kBlock - 1) / kBlock, batch);
  const double acc_flag = accumulate ? 1.0 : 0.0;
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_ptm_grad_final<double>, g2, dim3(kBlock), 0, s, partial, nb, dd,
                       static_cast<double*>(dptm), acc_flag);
  else
    hipLaunchKernelGGL(k_ptm_grad_final<float>, g2, dim3(kBlock), 0, s, partial, nb, dd,
                       static_cast<float*>(dptm), acc_flag);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

extern "C" int qsv_ptm_grad_tiled(int dtype, const void* x, const void* ybar, int batch, int rank,
                                  const int64_t* dims, int k, const int32_t* targets,
                                  const int32_t* dims_out_t, void* dptm, int accumulate,
                                  void* workspace, int64_t workspace_bytes, const qsv_tile* tile,
                                  void* stream) {
  return ptm_grad_run(dtype, x, ybar, batch, rank, dims, k, targets, dims_out_t, dptm,
                      accumulate, workspace, workspace_bytes, false, tile, stream);
}

// ---- zero-allocation variant ---------------------------------------------
// The partial tiles of qsv_ptm_grad_noalloc live in this file-scope device
// array: one copy per GPU, held until the process ends.  In host code the
// array's name is only a stand-in for it; its device address on each GPU is
// looked up once with hipGetSymbolAddress (on the first call with that
// device current) and that address is what the kernels receive.  Every
// block writes its whole partial tile before the final sum reads it, so
// nothing carried over from an earlier call is ever read and nothing needs
// clearing; calls that overlap on one GPU (two streams at once) race on it.
namespace qsv {
static __device__ __align__(256) double scratch[QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES / 8];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, which must be the
// device of `stream`.
static int noalloc_scratch(void* stream, double** out) {
  int dev = -1;
  QSV_CHECK_HIP(hipGetDevice(&dev));
  QSV_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d outside [0, %d)", dev, kMaxDevices);
  if (stream) {
    hipDevice_t sdev = -1;
    QSV_CHECK_HIP(hipStreamGetDevice(static_cast<hipStream_t>(stream), &sdev));
    QSV_CHECK_ARG(int(sdev) == dev, "the stream belongs to device %d, device %d is current",
                  int(sdev), dev);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    QSV_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    QSV_CHECK_ARG(ptr != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return QSV_OK;
}
}  // namespace qsv

extern "C" int64_t qsv_ptm_grad_noalloc_scratch_bytes(void) {
  return QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES;
}

extern "C" int qsv_ptm_grad_noalloc(int dtype, const void* x, const void* ybar, int batch,
                                    int rank, const int64_t* dims, int k, const int32_t* targets,
                                    const int32_t* dims_out_t, void* dptm, int accumulate,
                                    void* stream) {
  // argument checks first: a rejected call touches no device state
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(batch >= 0 && batch <= 65535, "batch %d outside [0, 65535]", batch);
  GradPlan p;
  const int rc = make_grad_plan(dtype_size(dtype), rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  if (batch == 0) return QSV_OK;
  const int64_t least = int64_t(batch) * p.din * p.dout * 8;
  QSV_CHECK_ARG(least <= QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES,
                "PTM gradient of %d states with a %d x %d PTM needs at least %lld bytes of "
                "scratch, above the %lld this entry point holds",
                batch, p.dout, p.din, (long long)least,
                (long long)QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES);
  QSV_CHECK_ARG(x && ybar && dptm, "null pointer argument");
  double* ptr = nullptr;
  const int rs = noalloc_scratch(stream, &ptr);
  if (rs) return rs;
  return ptm_grad_run(dtype, x, ybar, batch, rank, dims, k, targets, dims_out_t, dptm,
                      accumulate, ptr, QSV_PTM_GRAD_NOALLOC_SCRATCH_BYTES, true, nullptr,
                      stream);
}

// Host replay of k_ptm_grad64 + k_ptm_grad_final for one state (same checks
// as the replay of the 16 x 16 kernel below).
static int emulate_grad64(const GradPlan& p, int es, const double* x, const double* ybar, int nb,
                          double* dptm, int U = kGrad64Unroll) {
  const int VN = p.vn, waves = kBlock / 64;
  const uint32_t STEP = 4 * VN, nsteps = (p.n_orbits + STEP - 1) / STEP;
  const uint32_t nchunks = (nsteps + U - 1) / U;
  const int dd = p.din * p.dout;
  std::vector<int> used(p.n_orbits, 0);
  std::vector<double> partial(size_t(nb) * dd, 0.0);
  // acc[w][rb][cb][lane][r]
  std::vector<double> acc(size_t(waves) * 16 * 64 * 4);
  std::vector<double> A(size_t(4) * 64 * 16), B(size_t(4) * 64 * 16);  // [q][lane][e]
  for (int blk = 0; blk < nb; ++blk) {
    std::fill(acc.begin(), acc.end(), 0.0);
    for (int w = 0; w < waves; ++w)
      for (uint32_t tc = blk * waves + w; tc < nchunks; tc += nb * waves)
        for (uint32_t t = tc * U; t < tc * U + U; ++t) {
          for (int lane = 0; lane < 64; ++lane) {
            const int c = lane & 15, kk = lane >> 4;
            const uint32_t m = t * STEP + VN * kk;
            const bool ok = m < p.n_orbits;
            int64_t bin, bout;
            orbit_base(p.g, ok ? m : 0, bin, bout);
            if (ok && c == 0) {
              for (int ee = 0; ee < VN; ++ee) {
                QSV_CHECK_ARG(m + ee < p.n_orbits, "grad64: vector past the last orbit");
                int64_t bi2, bo2;
                orbit_base(p.g, m + ee, bi2, bo2);
                QSV_CHECK_ARG(bi2 == bin + ee && bo2 == bout + ee,
                              "grad64: orbits of a vector not adjacent");
                ++used[m + ee];
              }
            }
            for (int q = 0; q < 4; ++q)
              for (int e = 0; e < VN; ++e) {
                double& a = A[(size_t(q) * 64 + lane) * 16 + e];
                double& bb = B[(size_t(q) * 64 + lane) * 16 + e];
                a = bb = 0.0;
                if (!ok) continue;
                if (p.quad) {
                  // entries 4c + q: one run of four from the (aligned) first
                  const int QV = 16 / es;
                  const int64_t ai = bin + p.offs.off_in[4 * c], ao = bout + p.offs.off_out[4 * c];
                  QSV_CHECK_ARG(ai % QV == 0 && ao % QV == 0, "grad64: misaligned quad load");
                  QSV_CHECK_ARG(p.offs.off_in[4 * c + q] == p.offs.off_in[4 * c] + q &&
                                    p.offs.off_out[4 * c + q] == p.offs.off_out[4 * c] + q,
                                "grad64: quad entries not contiguous");
                  QSV_CHECK_ARG(ai + 3 < p.n_in && ao + 3 < p.n_out && ai >= 0 && ao >= 0,
                                "grad64: quad load out of bounds");
                  bb = x[ai + q];
                  a = ybar[ao + q];
                  continue;
                }
                const int ent = 16 * q + c;
                if (ent < p.din) {
                  const int64_t ad = bin + p.offs.off_in[ent] + e;
                  QSV_CHECK_ARG((ad - e) % VN == 0, "grad64: misaligned vector load");
                  QSV_CHECK_ARG(ad >= 0 && ad < p.n_in, "grad64: x load %lld out of bounds",
                                (long long)ad);
                  bb = x[ad];
                }
                if (ent < p.dout) {
                  const int64_t ad = bout + p.offs.off_out[ent] + e;
                  QSV_CHECK_ARG((ad - e) % VN == 0, "grad64: misaligned vector load");
                  QSV_CHECK_ARG(ad >= 0 && ad < p.n_out, "grad64: ybar load %lld out of bounds",
                                (long long)ad);
                  a = ybar[ad];
                }
              }
          }
          // tile (rb, cb): D[row][col] += sum_kk A_rb[row + 16 kk] B_cb[col + 16 kk];
          // slot r of lane l holds column l % 16 and row (l / 16) + 4 r
          // (float64 MFMA) or 4 (l / 16) + r (float32 MFMA)
          for (int e = 0; e < VN; ++e)
            for (int rb = 0; rb < 4; ++rb)
              for (int cb = 0; cb < 4; ++cb)
                for (int lane = 0; lane < 64; ++lane)
                  for (int r = 0; r < 4; ++r) {
                    const int row = es == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                    const int col = lane & 15;
                    double d = 0.0;
                    for (int kk = 0; kk < 4; ++kk)
                      d += A[(size_t(rb) * 64 + row + 16 * kk) * 16 + e] *
                           B[(size_t(cb) * 64 + col + 16 * kk) * 16 + e];
                    acc[((((size_t(w) * 4 + rb) * 4 + cb) * 64) + lane) * 4 + r] += d;
                  }
        }
    std::vector<double> red(64 * 64, 0.0);
    for (int w = 0; w < waves; ++w)
      for (int rb = 0; rb < 4; ++rb)
        for (int cb = 0; cb < 4; ++cb)
          for (int lane = 0; lane < 64; ++lane)
            for (int r = 0; r < 4; ++r) {
              const int row = es == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
              const int idx = p.quad ? (4 * row + rb) * 64 + 4 * (lane & 15) + cb
                                     : (16 * rb + row) * 64 + 16 * cb + (lane & 15);
              const double v = acc[((((size_t(w) * 4 + rb) * 4 + cb) * 64) + lane) * 4 + r];
              red[idx] = (w == 0 ? 0.0 : red[idx]) + v;
            }
    for (int j = 0; j < dd; ++j) {
      const int o = j / p.din, i = j - o * p.din;
      partial[size_t(blk) * dd + j] = red[o * 64 + i];
    }
  }
  for (uint32_t m = 0; m < p.n_orbits; ++m)
    QSV_CHECK_ARG(used[m] == 1, "grad64: orbit %u consumed %d times", m, used[m]);
  for (int j = 0; j < dd; ++j) {
    double s = 0.0;
    for (int blk = 0; blk < nb; ++blk) s += partial[size_t(blk) * dd + j];
    dptm[j] = s;
  }
  return QSV_OK;
}

// Host replay of k_ptm_grad + k_ptm_grad_final for one state: every lane's
// loads (bounds and vector alignment checked), the MFMA contraction in
// hardware index order (A = ybar, B = x, k = the four orbits of a k-step),
// the block reduction's accumulator-to-entry map and the fixed-order final
// sum.  Every orbit must be consumed exactly once.
// Host replay of k_ptm_grad64d + k_ptm_grad_final for one state: every
// direct load (16 bytes = orbits m..m+3 of one entry, in bounds, aligned,
// the four orbits consecutive in memory) and the LDS slot it lands in;
// every MFMA operand read checked to find, in the stage of the current
// tile, the (entry, orbit) the product needs; every entry of the partial
// tiles written once.
static int emulate_grad64d(const GradPlan& p, const double* x, const double* ybar, int nb,
                           double* dptm, const double* ptm = nullptr, double* xbar = nullptr) {
  QSV_CHECK_ARG(grad64_direct(p, 4, p.vn), "grad64d: not a direct geometry");
  std::vector<int> xwritten(xbar ? size_t(p.n_in) : 0, 0);
  const uint32_t ntiles = p.n_orbits / kDirTile;
  const int nf = 2 * kDirArr / 4;  // floats per stage
  // per stage float: value, entry, orbit, array (0 x, 1 ybar), tile
  std::vector<double> val(2 * size_t(nf));
  std::vector<int64_t> tag(2 * size_t(nf) * 4);
  std::vector<int> loaded(size_t(64) * p.n_orbits * 2, 0);
  std::vector<double> partial(size_t(nb) * 64 * 64, 0.0);
  for (int blk = 0; blk < nb; ++blk) {
    std::vector<double> D(4 * 4 * 16 * 16, 0.0);  // [w][cb][row][col]
    int stage = 0;
    for (uint32_t t = blk; t < ntiles; t += nb, stage ^= 1) {
      double* sv = val.data() + size_t(stage) * nf;
      int64_t* stg = tag.data() + size_t(stage) * nf * 4;
      std::fill(stg, stg + size_t(nf) * 4, int64_t(-1));
      for (int w = 0; w < kBlock / 64; ++w)
        for (int lane = 0; lane < 64; ++lane) {
          const int e = lane & 7, c = lane >> 3;
          const uint32_t m = t * kDirTile + 4 * c;
          int64_t bin, bout;
          orbit_base(p.g, m, bin, bout);
          for (int q = 1; q < 4; ++q) {
            int64_t bi2, bo2;
            orbit_base(p.g, m + q, bi2, bo2);
            QSV_CHECK_ARG(bi2 == bin + q && bo2 == bout + q,
                          "grad64d: orbits of a vector not consecutive");
          }
          for (int h = 0; h < 2; ++h) {
            const int j = 2 * w + h, ent = 8 * j + e;
            for (int arr = 0; arr < 2; ++arr) {
              const int64_t a = arr == 0 ? bin + p.offs.off_in[ent] : bout + p.offs.off_out[ent];
              const int64_t n_tot = arr == 0 ? p.n_in : p.n_out;
              QSV_CHECK_ARG(a >= 0 && a + 4 <= n_tot && a % 4 == 0,
                            "grad64d: direct load out of bounds or misaligned");
              const int byte = arr * kDirArr + j * kDirBlk + 16 * lane;
              QSV_CHECK_ARG(byte >= 0 && byte + 16 <= 2 * kDirArr, "grad64d: LDS slot");
              for (int q = 0; q < 4; ++q) {
                const int f = byte / 4 + q;
                QSV_CHECK_ARG(stg[f * 4] == -1, "grad64d: LDS float written twice");
                stg[f * 4 + 0] = ent;
                stg[f * 4 + 1] = m + q;
                stg[f * 4 + 2] = arr;
                stg[f * 4 + 3] = t;
                sv[f] = (arr == 0 ? x : ybar)[a + q];
                ++loaded[(size_t(m + q) * 64 + ent) * 2 + arr];
              }
            }
          }
        }
      // MFMA operands: lane (m, kk) of wave w, k-step ks
      auto fetch = [&](int byte, int ent, uint32_t orb, int arr, double& out) -> int {
        QSV_CHECK_ARG(byte >= 0 && byte + 4 <= 2 * kDirArr && byte % 4 == 0,
                      "grad64d: operand read outside the stage");
        const int f = byte / 4;
        QSV_CHECK_ARG(stg[f * 4] == ent && stg[f * 4 + 1] == orb && stg[f * 4 + 2] == arr &&
                          stg[f * 4 + 3] == int64_t(t),
                      "grad64d: operand read finds the wrong entry or orbit");
        out = sv[f];
        return QSV_OK;
      };
      for (int w = 0; w < kBlock / 64; ++w)
        for (int ks = 0; ks < kDirTile / 4; ++ks)
          for (int kk = 0; kk < 4; ++kk) {
            const uint32_t orb = t * kDirTile + 4 * ks + kk;
            double A[16], B[4][16];
            for (int mm = 0; mm < 16; ++mm) {
              const int ya = kDirArr + (2 * w + (mm >> 3)) * kDirBlk + 16 * (mm & 7) + 4 * kk;
              const int xb = (mm >> 3) * kDirBlk + 16 * (mm & 7) + 4 * kk;
              int rc = fetch(ya + 128 * ks, 16 * w + mm, orb, 1, A[mm]);
              if (rc) return rc;
              for (int cb = 0; cb < 4; ++cb) {
                rc = fetch(xb + 2 * cb * kDirBlk + 128 * ks, 16 * cb + mm, orb, 0, B[cb][mm]);
                if (rc) return rc;
              }
            }
            for (int cb = 0; cb < 4; ++cb)
              for (int r = 0; r < 16; ++r)
                for (int cc = 0; cc < 16; ++cc)
                  D[((w * 4 + cb) * 16 + r) * 16 + cc] += A[r] * B[cb][cc];
          }
      if (!xbar) continue;
      // k_bwd64d's adjoint apply: lane (m, kk) of wave w reads orbits 2m,
      // 2m + 1 of entry 4s + kk (8 bytes); B = P[4s + kk][16w + n]
      for (int w = 0; w < kBlock / 64; ++w) {
        double X[2][16][16] = {};  // [row tile][row m][column n]
        for (int sk = 0; sk < 16; ++sk)
          for (int kk = 0; kk < 4; ++kk) {
            const int ent = 4 * sk + kk;
            double A[2][16];
            for (int mm = 0; mm < 16; ++mm) {
              const int byte =
                  kDirArr + 16 * kk + 128 * (mm >> 1) + 8 * (mm & 1) + bwd_koff(sk);
              QSV_CHECK_ARG(byte % 8 == 0, "bwd64d: 8-byte read misaligned");
              for (int rt = 0; rt < 2; ++rt) {
                const int rc = fetch(byte + 4 * rt, ent, t * kDirTile + 2 * mm + rt, 1,
                                     A[rt][mm]);
                if (rc) return rc;
              }
            }
            for (int rt = 0; rt < 2; ++rt)
              for (int mm = 0; mm < 16; ++mm)
                for (int nn = 0; nn < 16; ++nn)
                  X[rt][mm][nn] += A[rt][mm] * ptm[ent * 64 + 16 * w + nn];
          }
        // slot r of lane l in row tile rt: row 4 (l / 16) + r, column l % 16
        for (int lane = 0; lane < 64; ++lane) {
          const int q = lane >> 4, nn = lane & 15;
          for (int hv = 0; hv < 2; ++hv) {
            const uint32_t m0 = t * kDirTile + 8 * q + 4 * hv;
            int64_t bin, bout;
            orbit_base(p.g, m0, bin, bout);
            const int64_t a = bin + p.offs.off_in[16 * w + nn];
            QSV_CHECK_ARG(a >= 0 && a + 4 <= p.n_in && a % 4 == 0,
                          "bwd64d: xbar store out of bounds or misaligned");
            for (int i = 0; i < 4; ++i) {
              const int r = 2 * hv + (i >> 1), rt = i & 1;
              QSV_CHECK_ARG(2 * (4 * q + r) + rt == 8 * q + 4 * hv + i,
                            "bwd64d: store slot holds another orbit");
              xbar[a + i] = X[rt][4 * q + r][nn];
              ++xwritten[a + i];
            }
          }
        }
      }
    }
    // accumulator slot r of lane l: row 4 (l / 16) + r, column l % 16
    std::vector<int> written(64 * 64, 0);
    for (int w = 0; w < kBlock / 64; ++w)
      for (int lane = 0; lane < 64; ++lane)
        for (int cb = 0; cb < 4; ++cb)
          for (int r = 0; r < 4; ++r) {
            const int row = 4 * (lane >> 4) + r, col = lane & 15;
            const int idx = (16 * w + row) * 64 + 16 * cb + col;
            ++written[idx];
            partial[size_t(blk) * 4096 + idx] = D[((w * 4 + cb) * 16 + row) * 16 + col];
          }
    for (int i = 0; i < 4096; ++i)
      QSV_CHECK_ARG(written[i] == 1, "grad64d: partial entry written %d times", written[i]);
  }
  for (size_t i = 0; i < loaded.size(); ++i)
    QSV_CHECK_ARG(loaded[i] == 1, "grad64d: operand loaded %d times", loaded[i]);
  for (size_t i = 0; i < xwritten.size(); ++i)
    QSV_CHECK_ARG(xwritten[i] == 1, "bwd64d: xbar entry %lld written %d times", (long long)i,
                  xwritten[i]);
  for (int j = 0; j < 4096; ++j) {
    double sum = 0.0;
    for (int b = 0; b < nb; ++b) sum += partial[size_t(b) * 4096 + j];
    dptm[j] = sum;
  }
  return QSV_OK;
}

// Host replay of k_ptm_grad64s + k_ptm_grad_final for one state: the
// loader's vectors (in bounds, aligned, the orbits of a vector adjacent,
// every entry and orbit loaded exactly once, LDS indices inside the padded
// rows), the MFMA operand slots of each wave, and the epilogue map (every
// (o, i) of a block's partial written once).
static int emulate_grad64s(const GradPlan& p, int es, const double* x, const double* ybar,
                           int nb, double* dptm) {
  const int VN = 16 / es, TT = kGrad64sTile, RS = TT + VN, TV = TT / VN;
  const int R = 64 * TV / kBlock;
  QSV_CHECK_ARG(p.din == 64 && p.dout == 64 && p.vn == VN, "grad64s: not a staged geometry");
  const uint32_t ntiles = (p.n_orbits + TT - 1) / TT;
  std::vector<int> loaded(size_t(64) * p.n_orbits, 0);
  std::vector<double> partial(size_t(nb) * 64 * 64, 0.0);
  std::vector<double> sx(size_t(64) * RS), sy(size_t(64) * RS);
  for (int blk = 0; blk < nb; ++blk) {
    std::vector<double> acc(64 * 64, 0.0);  // [o][i] of this block
    for (uint32_t t = blk; t < ntiles; t += nb) {
      std::fill(sx.begin(), sx.end(), 0.0);
      std::fill(sy.begin(), sy.end(), 0.0);
      for (int tid = 0; tid < kBlock; ++tid) {
        const int v = tid % TV, e0 = tid / TV;
        const uint32_t m = t * TT + v * VN;
        const bool ok = m < p.n_orbits;
        int64_t bin, bout;
        orbit_base(p.g, ok ? m : 0, bin, bout);
        for (int r = 0; r < R; ++r) {
          const int e = e0 + r * (kBlock / TV);
          QSV_CHECK_ARG(e < 64, "grad64s: entry %d", e);
          const int64_t ai = bin + p.offs.off_in[e], ao = bout + p.offs.off_out[e];
          QSV_CHECK_ARG(ai % VN == 0 && ao % VN == 0, "grad64s: misaligned vector load");
          QSV_CHECK_ARG(ai >= 0 && ai + VN <= p.n_in && ao >= 0 && ao + VN <= p.n_out,
                        "grad64s: load out of bounds");
          for (int i = 0; i < VN; ++i) {
            const int slot = v * VN + i;
            QSV_CHECK_ARG(slot < TT && e * RS + slot < 64 * RS, "grad64s: LDS index");
            if (!ok) continue;
            int64_t bi2, bo2;
            orbit_base(p.g, m + i, bi2, bo2);
            QSV_CHECK_ARG(m + i < p.n_orbits && bi2 == bin + i && bo2 == bout + i,
                          "grad64s: orbits of a vector not adjacent");
            sx[size_t(e) * RS + slot] = x[ai + i];
            sy[size_t(e) * RS + slot] = ybar[ao + i];
            ++loaded[size_t(e) * p.n_orbits + m + i];
          }
        }
      }
      // wave w, lane (c, k): k-step slot 16 u + 4 k + e; one MFMA per (u, e, cb)
      for (int w = 0; w < 4; ++w)
        for (int u = 0; u < TT / 16; ++u)
          for (int e = 0; e < 4; ++e)
            for (int cb = 0; cb < 4; ++cb)
              for (int ro = 0; ro < 16; ++ro)
                for (int ci = 0; ci < 16; ++ci) {
                  double sum = 0.0;
                  for (int k = 0; k < 4; ++k) {
                    const int slot = 16 * u + 4 * k + e;
                    sum += sy[size_t(16 * w + ro) * RS + slot] * sx[size_t(16 * cb + ci) * RS + slot];
                  }
                  acc[(16 * w + ro) * 64 + 16 * cb + ci] += sum;
                }
    }
    // epilogue: slot r of lane (c, k) in tile cb -> row k + 4 r (float64) or
    // 4 k + r (float32) of the wave's rows, column c of block cb
    std::vector<int> seen(64 * 64, 0);
    for (int w = 0; w < 4; ++w)
      for (int lane = 0; lane < 64; ++lane)
        for (int cb = 0; cb < 4; ++cb)
          for (int r = 0; r < 4; ++r) {
            const int c = lane & 15, k = lane >> 4;
            const int row = es == 8 ? k + 4 * r : 4 * k + r;
            const int o = 16 * w + row, i = 16 * cb + c;
            ++seen[o * 64 + i];
            partial[size_t(blk) * 4096 + o * 64 + i] = acc[o * 64 + i];
          }
    for (int j = 0; j < 4096; ++j)
      QSV_CHECK_ARG(seen[j] == 1, "grad64s: partial entry %d written %d times", j, seen[j]);
  }
  for (size_t j = 0; j < loaded.size(); ++j)
    QSV_CHECK_ARG(loaded[j] == 1, "grad64s: entry/orbit %zu loaded %d times", j, loaded[j]);
  for (int j = 0; j < 4096; ++j) {
    double sum = 0.0;
    for (int b = 0; b < nb; ++b) sum += partial[size_t(b) * 4096 + j];
    dptm[j] = sum;
  }
  return QSV_OK;
}

// Host replay of k_ptm_grad_big + k_ptm_grad_final for one state: the
// decoded offsets, every staged load in bounds, every (entry, orbit) staged
// exactly once, the LDS slots inside their rows.
static int emulate_grad_big(const GradPlan& p, const double* x, const double* ybar, int nb,
                            double* dptm) {
  const int din = p.din, dout = p.dout, dd = din * dout;
  std::vector<int64_t> oin(din), oout(dout);
  for (int e = 0; e < din; ++e) oin[e] = big_offset(p.bg.k, p.bg.din_t, p.bg.sin_t, e);
  for (int e = 0; e < dout; ++e) oout[e] = big_offset(p.bg.k, p.bg.dout_t, p.bg.sout_t, e);
  const uint32_t nchunks = (p.n_orbits + kGradBigOrbits - 1) / kGradBigOrbits;
  std::vector<int> used_in(size_t(din) * p.n_orbits, 0), used_out(size_t(dout) * p.n_orbits, 0);
  std::vector<double> partial(size_t(nb) * dd, 0.0), sx(kGradBigOrbits * kGradBigMaxSide),
      sy(kGradBigOrbits * kGradBigMaxSide);
  QSV_CHECK_ARG(dd <= kGradBigPerThread * kBlock, "big gradient: too many entries");
  for (int blk = 0; blk < nb; ++blk) {
    std::vector<double> acc(dd, 0.0);
    for (uint32_t c = blk; c < nchunks; c += nb) {
      for (int f = 0; f < kGradBigOrbits * din; ++f) {
        const int ol = f / din, i = f - ol * din;
        const uint32_t m = c * kGradBigOrbits + ol;
        double v = 0.0;
        if (m < p.n_orbits) {
          int64_t bin, bout;
          orbit_base(p.g, m, bin, bout);
          const int64_t a = bin + oin[i];
          QSV_CHECK_ARG(a >= 0 && a < p.n_in, "big gradient: x load out of bounds");
          v = x[a];
          ++used_in[size_t(i) * p.n_orbits + m];
        }
        sx[ol * kGradBigMaxSide + i] = v;
      }
      for (int f = 0; f < kGradBigOrbits * dout; ++f) {
        const int ol = f / dout, o = f - ol * dout;
        const uint32_t m = c * kGradBigOrbits + ol;
        double v = 0.0;
        if (m < p.n_orbits) {
          int64_t bin, bout;
          orbit_base(p.g, m, bin, bout);
          const int64_t a = bout + oout[o];
          QSV_CHECK_ARG(a >= 0 && a < p.n_out, "big gradient: ybar load out of bounds");
          v = ybar[a];
          ++used_out[size_t(o) * p.n_orbits + m];
        }
        sy[ol * kGradBigMaxSide + o] = v;
      }
      for (int j = 0; j < dd; ++j) {
        const int o = j / din, i = j - o * din;
        for (int ol = 0; ol < kGradBigOrbits; ++ol)
          acc[j] += sy[ol * kGradBigMaxSide + o] * sx[ol * kGradBigMaxSide + i];
      }
    }
    for (int j = 0; j < dd; ++j) partial[size_t(blk) * dd + j] = acc[j];
  }
  for (size_t j = 0; j < used_in.size(); ++j)
    QSV_CHECK_ARG(used_in[j] == 1, "big gradient: x entry staged %d times", used_in[j]);
  for (size_t j = 0; j < used_out.size(); ++j)
    QSV_CHECK_ARG(used_out[j] == 1, "big gradient: ybar entry staged %d times", used_out[j]);
  for (int j = 0; j < dd; ++j) {
    double sum = 0.0;
    for (int b = 0; b < nb; ++b) sum += partial[size_t(b) * dd + j];
    dptm[j] = sum;
  }
  return QSV_OK;
}

extern "C" int qsv_emulate_ptm_grad(int dtype, const double* x, const double* ybar, int rank,
                                    const int64_t* dims, int k, const int32_t* targets,
                                    const int32_t* dims_out_t, double* dptm,
                                    int32_t grad_blocks_forced, int32_t* vn_out) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  GradPlan p;
  const int rc = make_grad_plan(dtype_size(dtype), rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  QSV_CHECK_ARG(x && ybar && dptm, "null pointer argument");
  QSV_CHECK_ARG(p.n_in <= (int64_t(1) << 24) && p.n_out <= (int64_t(1) << 24),
                "emulation is for small vectors");
  QSV_CHECK_ARG(grad_blocks_forced >= 0 && grad_blocks_forced <= max_grad_blocks(p),
                "grad_blocks outside [0, %d]", max_grad_blocks(p));
  const bool staged = grad64_staged(p, dtype_size(dtype), p.vn);
  const bool direct = grad64_direct(p, dtype_size(dtype), p.vn);
  if (vn_out)
    *vn_out = p.vn | (p.quad << 8) | (staged ? 512 : 0) | (p.big ? 1024 : 0) |
              (direct ? 2048 : 0);
  if (direct)
    return emulate_grad64d(p, x, ybar,
                           grad_blocks_forced > 0 ? grad_blocks_forced
                                                  : grad64_blocks(p.n_orbits, 1),
                           dptm);
  if (p.big)
    return emulate_grad_big(p, x, ybar,
                            grad_blocks_forced > 0 ? grad_blocks_forced
                                                   : grad_big_blocks(p.n_orbits, 1),
                            dptm);
  if (staged)
    return emulate_grad64s(p, dtype_size(dtype), x, ybar,
                           grad_blocks_forced > 0 ? grad_blocks_forced
                                                  : grad64_blocks(p.n_orbits, 1),
                           dptm);
  if (grad_wide(p))
    return emulate_grad64(p, dtype_size(dtype), x, ybar,
                          grad_blocks_forced > 0 ? grad_blocks_forced
                                                 : grad64_blocks(p.n_orbits, 1),
                          dptm, p.quad ? kGrad64UnrollQuad : kGrad64Unroll);
  const int VN = p.vn, waves = kBlock / 64;
  const int nb = grad_blocks_forced > 0 ? grad_blocks_forced : grad_blocks(p.n_orbits, 1);
  if (vn_out) *vn_out = VN;
  const uint32_t STEP = 4 * VN, nsteps = (p.n_orbits + STEP - 1) / STEP;
  std::vector<int> used(p.n_orbits, 0);
  std::vector<double> partial(size_t(nb) * p.din * p.dout, 0.0);
  for (int blk = 0; blk < nb; ++blk) {
    // acc[w][lane][r]
    std::vector<double> acc(size_t(waves) * 64 * 4, 0.0);
    const uint32_t nchunks = (nsteps + kGradUnroll - 1) / kGradUnroll;
    for (int w = 0; w < waves; ++w)
      for (uint32_t tc = blk * waves + w; tc < nchunks; tc += nb * waves)
      for (uint32_t t = tc * kGradUnroll; t < tc * kGradUnroll + kGradUnroll; ++t) {
        // operands of this wave iteration: lane (c, kk) -> VN k-steps
        double A[64][16], B[64][16];  // [lane][e]
        for (int lane = 0; lane < 64; ++lane) {
          const int c = lane & 15, kk = lane >> 4;
          const uint32_t m = t * STEP + VN * kk;
          const bool ok = m < p.n_orbits;
          int64_t bin, bout;
          orbit_base(p.g, ok ? m : 0, bin, bout);
          for (int e = 0; e < VN; ++e) {
            A[lane][e] = B[lane][e] = 0.0;
            if (!ok) continue;
            if (e == 0 && c == 0) {
              for (int ee = 0; ee < VN; ++ee) {
                QSV_CHECK_ARG(m + ee < p.n_orbits, "grad: vector past the last orbit");
                int64_t bi2, bo2;
                orbit_base(p.g, m + ee, bi2, bo2);
                QSV_CHECK_ARG(bi2 == bin + ee && bo2 == bout + ee,
                              "grad: orbits of a vector not adjacent");
                ++used[m + ee];
              }
            }
            if (c < p.din) {
              const int64_t a = bin + p.offs.off_in[c] + e;
              QSV_CHECK_ARG((a - e) % VN == 0, "grad: misaligned vector load");
              QSV_CHECK_ARG(a >= 0 && a < p.n_in, "grad: x load %lld out of bounds", (long long)a);
              B[lane][e] = x[a];
            }
            if (c < p.dout) {
              const int64_t a = bout + p.offs.off_out[c] + e;
              QSV_CHECK_ARG((a - e) % VN == 0, "grad: misaligned vector load");
              QSV_CHECK_ARG(a >= 0 && a < p.n_out, "grad: ybar load %lld out of bounds",
                            (long long)a);
              A[lane][e] = ybar[a];
            }
          }
        }
        // D[row][col] += sum_kk A[row + 16 kk] B[col + 16 kk]; slot r of lane l
        // holds row (l / 16) + 4 r, column l % 16
        for (int e = 0; e < VN; ++e)
          for (int lane = 0; lane < 64; ++lane)
            for (int r = 0; r < 4; ++r) {
              const int row = (lane >> 4) + 4 * r, col = lane & 15;
              double d = 0.0;
              for (int kk = 0; kk < 4; ++kk) d += A[row + 16 * kk][e] * B[col + 16 * kk][e];
              acc[(size_t(w) * 64 + lane) * 4 + r] += d;
            }
      }
    for (int tid = 0; tid < kBlock; ++tid) {
      const int o = tid >> 4, i = tid & 15;
      const int src = (i + 16 * (o & 3)) * 4 + (o >> 2);
      double sum = 0.0;
      for (int w = 0; w < waves; ++w) sum += acc[size_t(w) * 64 * 4 + src];
      if (o < p.dout && i < p.din) partial[(size_t(blk) * p.dout + o) * p.din + i] = sum;
    }
  }
  for (uint32_t m = 0; m < p.n_orbits; ++m)
    QSV_CHECK_ARG(used[m] == 1, "grad: orbit %u consumed %d times", m, used[m]);
  const int dd = p.din * p.dout;
  for (int j = 0; j < dd; ++j) {
    double s = 0.0;
    for (int blk = 0; blk < nb; ++blk) s += partial[size_t(blk) * dd + j];
    dptm[j] = s;
  }
  return QSV_OK;
}

// ------------------------------------------------------ fused backward
static int bwd64_plan(int dtype, int rank, const int64_t* dims, int k, const int32_t* targets,
                      GradPlan& p) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(dims && targets && k >= 1 && k <= QSV_MAX_TARGETS && rank >= 1 &&
                    rank <= QSV_MAX_RANK,
                "bad shape arguments");
  int32_t dout[QSV_MAX_TARGETS];
  for (int j = 0; j < k; ++j) {
    QSV_CHECK_ARG(targets[j] >= 0 && targets[j] < rank, "bad target %d", targets[j]);
    dout[j] = int32_t(dims[targets[j]]);
  }
  const int rc = make_grad_plan(dtype_size(dtype), rank, dims, k, targets, dout, p);
  if (rc) return rc;
  if (dtype == QSV_F32 && bwd64q_applies(p)) return QSV_OK;
  if (!bwd64_applies(p)) {
    set_error("fused backward needs a square 64 x 64 PTM and a unit-stride bystander "
              "innermost axis (float32: or a target innermost axis of dimension 4)");
    return QSV_ENOTSUP;
  }
  if (dtype == QSV_F32 && !grad64_direct(p, 4, p.vn)) {
    set_error("float32 fused backward needs 4-orbit vectors and whole %d-orbit tiles",
              kDirTile);
    return QSV_ENOTSUP;
  }
  return QSV_OK;
}

static int bwd64_blocks(const GradPlan& p, int batch, uint32_t cap) {
  const uint32_t tiles = (p.n_orbits + 16 * p.vn - 1) / (16 * p.vn);
  uint32_t nb = (tiles + kBlock / 64 - 1) / (kBlock / 64);
  nb = std::min<uint32_t>(nb, std::max<uint32_t>(1, cap / std::max(batch, 1)));
  nb = std::min<uint32_t>(nb, kGrad64Blocks);
  return int(std::max<uint32_t>(nb, 1));
}

// k_bwd64d: one round of resident blocks over the batch (occupancy x CUs /
// batch per state; measured on 8-12-qubit states: 5-12 % faster than the
// gradient's 512 orbits per block), at least 4 tiles per block.

static int bwd64d_blocks(const GradPlan& p, int batch) {
  static const uint32_t cap_d =
      resident_blocks(reinterpret_cast<const void*>(k_bwd64d<true>));
  static const uint32_t cap_q =
      resident_blocks(reinterpret_cast<const void*>(k_bwd64q<true>));
  const uint32_t cap = p.quad ? cap_q : cap_d;
  const uint32_t tiles = p.n_orbits / kDirTile;
  uint32_t nb = cap / uint32_t(std::max(batch, 1));
  nb = std::min(nb, tiles / 4);
  nb = std::min<uint32_t>(nb, kGrad64Blocks);
  return int(std::max<uint32_t>(nb, 1));
}

template <typename T, int VN>
static uint32_t bwd64_cap() {
  static const uint32_t cap = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(k_bwd64<T, VN>), kBlock, 0) != hipSuccess ||
        per_cu <= 0)
      per_cu = 1;
    (void)hipGetLastError();
    return uint32_t(per_cu) * uint32_t(cus);
  }();
  return cap;
}

extern "C" int qsv_apply_ptm_batched_backward(int dtype, const void* x, const void* ybar,
                                              void* xbar, int batch, int rank,
                                              const int64_t* dims, int k,
                                              const int32_t* targets, const void* ptm_dev,
                                              int64_t ptm_batch_stride, void* dptm,
                                              int accumulate, void* workspace,
                                              int64_t workspace_bytes, void* stream) {
  QSV_CHECK_ARG(batch >= 0 && batch <= 65535, "batch %d outside [0, 65535]", batch);
  QSV_CHECK_ARG(ptm_batch_stride >= 0, "negative PTM batch stride");
  GradPlan p;
  const int rc = bwd64_plan(dtype, rank, dims, k, targets, p);
  if (rc) return rc;
  if (batch == 0) return QSV_OK;
  QSV_CHECK_ARG(x && ybar && xbar && ptm_dev && dptm, "null pointer argument");
  QSV_CHECK_ARG(xbar != ybar && xbar != x, "xbar must not alias x or ybar");
  const int es = dtype_size(dtype);
  auto al = [](const void* q) { return reinterpret_cast<uintptr_t>(q) % 16 == 0; };
  if (!al(x) || !al(ybar) || !al(xbar) || (p.n_in * es) % 16) {
    set_error("fused backward needs 16-byte aligned states");
    return QSV_ENOTSUP;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const FusedOffs fo = bwd64_offsets(p);
  // bwd64_plan admits only k_bwd64d's and k_bwd64q's float32 geometries
  const bool direct = dtype == QSV_F32;
  const int nb = direct ? bwd64d_blocks(p, batch) : bwd64_blocks(p, batch, bwd64_cap<double, 2>());
  const int64_t need = int64_t(batch) * nb * 64 * 64 * 8;
  QSV_CHECK_ARG(workspace && workspace_bytes >= need, "fused backward needs %lld workspace bytes",
                (long long)need);
  double* partial = static_cast<double*>(workspace);
  const dim3 grid(nb, batch);
  if (direct) {
    // x and ybar read, xbar written: the apply's nontemporal rule
    const bool nt = (2 * p.n_in + p.n_out) * 4 * int64_t(batch) > nt_threshold_bytes();
    auto kern = p.quad ? (nt ? k_bwd64q<true> : k_bwd64q<false>)
                       : (nt ? k_bwd64d<true> : k_bwd64d<false>);
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, s, static_cast<const float*>(x),
                       static_cast<const float*>(ybar), static_cast<float*>(xbar), p.g,
                       p.n_orbits, static_cast<const float*>(ptm_dev), ptm_batch_stride, p.offs,
                       partial);
  } else
    hipLaunchKernelGGL((k_bwd64<double, 2>), grid, dim3(kBlock), 0, s,
                       static_cast<const double*>(x), static_cast<const double*>(ybar),
                       static_cast<double*>(xbar), p.g, p.n_orbits,
                       static_cast<const double*>(ptm_dev), ptm_batch_stride, p.offs, fo, partial);
  QSV_CHECK_HIP(hipGetLastError());
  const dim3 g2((64 * 64 + kBlock - 1) / kBlock, batch);
  const double acc_flag = accumulate ? 1.0 : 0.0;
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_ptm_grad_final<double>, g2, dim3(kBlock), 0, s, partial, nb, 64 * 64,
                       static_cast<double*>(dptm), acc_flag);
  else
    hipLaunchKernelGGL(k_ptm_grad_final<float>, g2, dim3(kBlock), 0, s, partial, nb, 64 * 64,
                       static_cast<float*>(dptm), acc_flag);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

// Host replay of k_bwd64q + k_ptm_grad_final for one state: every direct
// load (bounds, alignment, its four entries contiguous, its LDS slot),
// every operand read of both contractions finding its (entry, orbit) in
// the current stage, the MFMAs in hardware index order, every xbar store
// in bounds and each xbar entry written once, each operand loaded once.
static int emulate_bwd64q(const GradPlan& p, const double* x, const double* ybar, int nb,
                          double* dptm, const double* ptm, double* xbar) {
  QSV_CHECK_ARG(bwd64q_applies(p), "bwd64q: not a quad geometry");
  const uint32_t ntiles = p.n_orbits / kDirTile;
  const int nf = 2 * kQArr / 4;
  std::vector<double> sv(nf);
  std::vector<int64_t> stg(size_t(nf) * 4);
  std::vector<int> loaded(size_t(64) * p.n_orbits * 2, 0), xwritten(p.n_in, 0);
  std::vector<double> partial(size_t(nb) * 4096, 0.0);
  for (int blk = 0; blk < nb; ++blk) {
    std::vector<double> D(4 * 4 * 16 * 16, 0.0);  // [w][cb][row][col]
    for (uint32_t t = blk; t < ntiles; t += nb) {
      std::fill(stg.begin(), stg.end(), int64_t(-1));
      for (int w = 0; w < kBlock / 64; ++w)
        for (int lane = 0; lane < 64; ++lane) {
          const int q = lane >> 2;
          for (int h = 0; h < 2; ++h) {
            const int ks = 2 * w + h;
            const uint32_t orb = t * kDirTile + 4 * ks + (lane & 3);
            int64_t bin, bout;
            orbit_base(p.g, orb, bin, bout);
            for (int arr = 0; arr < 2; ++arr) {
              const int64_t* off = arr == 0 ? p.offs.off_in : p.offs.off_out;
              const int64_t a = (arr == 0 ? bin : bout) + off[4 * q];
              const int64_t n_tot = arr == 0 ? p.n_in : p.n_out;
              QSV_CHECK_ARG(a >= 0 && a + 4 <= n_tot && a % 4 == 0,
                            "bwd64q: direct load out of bounds or misaligned");
              for (int r = 1; r < 4; ++r)
                QSV_CHECK_ARG(off[4 * q + r] == off[4 * q] + r, "bwd64q: quad not contiguous");
              const int byte = arr * kQArr + ks * kQBlk + 16 * lane;
              QSV_CHECK_ARG(byte >= 0 && byte + 16 <= 2 * kQArr, "bwd64q: LDS slot");
              for (int r = 0; r < 4; ++r) {
                const int f = byte / 4 + r;
                QSV_CHECK_ARG(stg[size_t(f) * 4] == -1, "bwd64q: LDS float written twice");
                stg[size_t(f) * 4 + 0] = 4 * q + r;
                stg[size_t(f) * 4 + 1] = orb;
                stg[size_t(f) * 4 + 2] = arr;
                stg[size_t(f) * 4 + 3] = t;
                sv[f] = (arr == 0 ? x : ybar)[a + r];
                ++loaded[(size_t(orb) * 64 + 4 * q + r) * 2 + arr];
              }
            }
          }
        }
      auto fetch = [&](int byte, int ent, uint32_t orb, int arr, double& out) -> int {
        QSV_CHECK_ARG(byte >= 0 && byte + 4 <= 2 * kQArr && byte % 4 == 0,
                      "bwd64q: operand read outside the stage");
        const int f = byte / 4;
        QSV_CHECK_ARG(stg[size_t(f) * 4] == ent && stg[size_t(f) * 4 + 1] == orb &&
                          stg[size_t(f) * 4 + 2] == arr && stg[size_t(f) * 4 + 3] == int64_t(t),
                      "bwd64q: operand read finds the wrong entry or orbit");
        out = sv[f];
        return QSV_OK;
      };
      for (int w = 0; w < kBlock / 64; ++w) {
        // dP: k-step ks over orbits 4ks + kk
        for (int ks = 0; ks < 8; ++ks)
          for (int kk = 0; kk < 4; ++kk) {
            const uint32_t orb = t * kDirTile + 4 * ks + kk;
            double A[16], B[4][16];
            for (int mm = 0; mm < 16; ++mm) {
              const int row = 64 * (mm >> 2) + 16 * kk + 4 * (mm & 3) + ks * kQBlk;
              int rc = fetch(kQArr + 256 * w + row, 16 * w + mm, orb, 1, A[mm]);
              if (rc) return rc;
              for (int cb = 0; cb < 4; ++cb) {
                rc = fetch(row + 256 * cb, 16 * cb + mm, orb, 0, B[cb][mm]);
                if (rc) return rc;
              }
            }
            for (int cb = 0; cb < 4; ++cb)
              for (int r = 0; r < 16; ++r)
                for (int cc = 0; cc < 16; ++cc)
                  D[((w * 4 + cb) * 16 + r) * 16 + cc] += A[r] * B[cb][cc];
          }
        if (!xbar) continue;
        // xbar: column tile ct, k-step s = 4u + e at entry 16u + 4kk + e
        for (int ct = 0; ct < 2; ++ct) {
          double X[16][16] = {};  // [row m][column n]
          for (int u = 0; u < 4; ++u)
            for (int kk = 0; kk < 4; ++kk)
              for (int nn = 0; nn < 16; ++nn) {
                const int byte = kQArr + (nn >> 2) * kQBlk + 64 * kk + 16 * (nn & 3) +
                                 4 * kQBlk * ct + 256 * u;
                QSV_CHECK_ARG(byte % 16 == 0, "bwd64q: 16-byte read misaligned");
                for (int e = 0; e < 4; ++e) {
                  const int ent = 16 * u + 4 * kk + e;
                  double bv;
                  const int rc =
                      fetch(byte + 4 * e, ent, t * kDirTile + 16 * ct + nn, 1, bv);
                  if (rc) return rc;
                  for (int mm = 0; mm < 16; ++mm) X[mm][nn] += ptm[ent * 64 + 16 * w + mm] * bv;
                }
              }
          // slot r of lane (n, j): row 4j + r = entry 16w + 4j + r of orbit n
          for (int lane = 0; lane < 64; ++lane) {
            const int nn = lane & 15, j = lane >> 4;
            int64_t bin, bout;
            orbit_base(p.g, t * kDirTile + 16 * ct + nn, bin, bout);
            const int64_t a = bin + p.offs.off_in[16 * w + 4 * j];
            QSV_CHECK_ARG(a >= 0 && a + 4 <= p.n_in && a % 4 == 0,
                          "bwd64q: xbar store out of bounds or misaligned");
            for (int r = 0; r < 4; ++r) {
              QSV_CHECK_ARG(p.offs.off_in[16 * w + 4 * j + r] == p.offs.off_in[16 * w + 4 * j] + r,
                            "bwd64q: xbar quad not contiguous");
              xbar[a + r] = X[4 * j + r][nn];
              ++xwritten[a + r];
            }
          }
        }
      }
    }
    std::vector<int> written(4096, 0);
    for (int w = 0; w < kBlock / 64; ++w)
      for (int lane = 0; lane < 64; ++lane)
        for (int cb = 0; cb < 4; ++cb)
          for (int r = 0; r < 4; ++r) {
            const int row = 4 * (lane >> 4) + r, col = lane & 15;
            const int idx = (16 * w + row) * 64 + 16 * cb + col;
            ++written[idx];
            partial[size_t(blk) * 4096 + idx] = D[((w * 4 + cb) * 16 + row) * 16 + col];
          }
    for (int i = 0; i < 4096; ++i)
      QSV_CHECK_ARG(written[i] == 1, "bwd64q: partial entry written %d times", written[i]);
  }
  for (size_t i = 0; i < loaded.size(); ++i)
    QSV_CHECK_ARG(loaded[i] == 1, "bwd64q: operand loaded %d times", loaded[i]);
  for (size_t i = 0; i < xwritten.size(); ++i)
    QSV_CHECK_ARG(xwritten[i] == 1, "bwd64q: xbar entry %lld written %d times", (long long)i,
                  xwritten[i]);
  for (int j = 0; j < 4096; ++j) {
    double sum = 0.0;
    for (int b = 0; b < nb; ++b) sum += partial[size_t(b) * 4096 + j];
    dptm[j] = sum;
  }
  return QSV_OK;
}

// Host replay of k_apply64q for one state: every direct load of x (bounds,
// alignment, its four entries contiguous, its LDS slot), every 16-byte
// operand read finding its (entry, orbit) in the current stage, the MFMA
// contraction in hardware index order, every y store in bounds and each y
// entry written once, each x entry loaded once.
static int emulate_apply64q(const GradPlan& p, const double* x, const double* ptm, int nb,
                            double* y) {
  QSV_CHECK_ARG(bwd64q_applies(p), "apply64q: not a quad geometry");
  const uint32_t ntiles = p.n_orbits / kDirTile;
  const int nf = kQArr / 4;
  std::vector<double> sv(nf);
  std::vector<int64_t> stg(size_t(nf) * 3);
  std::vector<int> loaded(size_t(64) * p.n_orbits, 0), written(p.n_out, 0);
  for (int blk = 0; blk < nb; ++blk)
    for (uint32_t t = blk; t < ntiles; t += nb) {
      std::fill(stg.begin(), stg.end(), int64_t(-1));
      for (int w = 0; w < kBlock / 64; ++w)
        for (int lane = 0; lane < 64; ++lane) {
          const int q = lane >> 2;
          for (int h = 0; h < 2; ++h) {
            const int ks = 2 * w + h;
            const uint32_t orb = t * kDirTile + 4 * ks + (lane & 3);
            int64_t bin, bout;
            orbit_base(p.g, orb, bin, bout);
            const int64_t a = bin + p.offs.off_in[4 * q];
            QSV_CHECK_ARG(a >= 0 && a + 4 <= p.n_in && a % 4 == 0,
                          "apply64q: direct load out of bounds or misaligned");
            const int byte = ks * kQBlk + 16 * lane;
            QSV_CHECK_ARG(byte >= 0 && byte + 16 <= kQArr, "apply64q: LDS slot");
            for (int r = 0; r < 4; ++r) {
              QSV_CHECK_ARG(p.offs.off_in[4 * q + r] == p.offs.off_in[4 * q] + r,
                            "apply64q: quad not contiguous");
              const int f = byte / 4 + r;
              QSV_CHECK_ARG(stg[size_t(f) * 3] == -1, "apply64q: LDS float written twice");
              stg[size_t(f) * 3 + 0] = 4 * q + r;
              stg[size_t(f) * 3 + 1] = orb;
              stg[size_t(f) * 3 + 2] = t;
              sv[f] = x[a + r];
              ++loaded[size_t(orb) * 64 + 4 * q + r];
            }
          }
        }
      for (int w = 0; w < kBlock / 64; ++w)
        for (int ct = 0; ct < 2; ++ct) {
          double Y[16][16] = {};  // [row m][column n]
          for (int u = 0; u < 4; ++u)
            for (int kk = 0; kk < 4; ++kk)
              for (int nn = 0; nn < 16; ++nn) {
                const int byte =
                    (nn >> 2) * kQBlk + 64 * kk + 16 * (nn & 3) + 4 * kQBlk * ct + 256 * u;
                QSV_CHECK_ARG(byte % 16 == 0 && byte + 16 <= kQArr,
                              "apply64q: 16-byte read misaligned or outside the stage");
                for (int e = 0; e < 4; ++e) {
                  const int ent = 16 * u + 4 * kk + e, f = byte / 4 + e;
                  QSV_CHECK_ARG(stg[size_t(f) * 3] == ent &&
                                    stg[size_t(f) * 3 + 1] == t * kDirTile + 16 * ct + nn &&
                                    stg[size_t(f) * 3 + 2] == int64_t(t),
                                "apply64q: operand read finds the wrong entry or orbit");
                  for (int mm = 0; mm < 16; ++mm) Y[mm][nn] += ptm[(16 * w + mm) * 64 + ent] * sv[f];
                }
              }
          for (int lane = 0; lane < 64; ++lane) {
            const int nn = lane & 15, j = lane >> 4;
            int64_t bin, bout;
            orbit_base(p.g, t * kDirTile + 16 * ct + nn, bin, bout);
            const int64_t a = bout + p.offs.off_out[16 * w + 4 * j];
            QSV_CHECK_ARG(a >= 0 && a + 4 <= p.n_out && a % 4 == 0,
                          "apply64q: y store out of bounds or misaligned");
            for (int r = 0; r < 4; ++r) {
              QSV_CHECK_ARG(p.offs.off_out[16 * w + 4 * j + r] == p.offs.off_out[16 * w + 4 * j] + r,
                            "apply64q: y quad not contiguous");
              y[a + r] = Y[4 * j + r][nn];
              ++written[a + r];
            }
          }
        }
    }
  for (size_t i = 0; i < loaded.size(); ++i)
    QSV_CHECK_ARG(loaded[i] == 1, "apply64q: x entry loaded %d times", loaded[i]);
  for (size_t i = 0; i < written.size(); ++i)
    QSV_CHECK_ARG(written[i] == 1, "apply64q: y entry %lld written %d times", (long long)i,
                  written[i]);
  return QSV_OK;
}

extern "C" int qsv_emulate_apply_ptm_direct(int dtype, const double* x, double* y, int rank,
                                            const int64_t* dims, int k, const int32_t* targets,
                                            const double* ptm, int32_t blocks) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(dims && targets && k >= 1 && k <= QSV_MAX_TARGETS && rank >= 1 &&
                    rank <= QSV_MAX_RANK,
                "bad shape arguments");
  int32_t dout[QSV_MAX_TARGETS];
  for (int j = 0; j < k; ++j) {
    QSV_CHECK_ARG(targets[j] >= 0 && targets[j] < rank, "bad target %d", targets[j]);
    dout[j] = int32_t(dims[targets[j]]);
  }
  GradPlan p;
  const int rc = make_grad_plan(dtype_size(dtype), rank, dims, k, targets, dout, p);
  if (rc) return rc;
  if (dtype != QSV_F32 || !bwd64q_applies(p)) {
    set_error("the direct 64 x 64 apply takes float32 states whose PTM's lowest index is the "
              "innermost axis, on whole %d-orbit tiles", kDirTile);
    return QSV_ENOTSUP;
  }
  QSV_CHECK_ARG(x && y && ptm, "null pointer argument");
  QSV_CHECK_ARG(p.n_in <= (int64_t(1) << 24), "emulation is for small vectors");
  QSV_CHECK_ARG(blocks >= 0 && blocks <= 4096, "blocks outside [0, 4096]");
  const uint32_t tiles = p.n_orbits / kDirTile;
  const int nb = blocks > 0 ? blocks : int(std::max<uint32_t>(1, tiles / 4));
  return emulate_apply64q(p, x, ptm, nb, y);
}

// Host replay of k_bwd64 + k_ptm_grad_final for one state.  Part A: every
// lane's loads and stores (bounds, vector alignment, each xbar entry written
// once) and the MFMA contraction in hardware index order with the
// transposed PTM's fragment table; part B: the wide-gradient replay with
// the kernel's four k-steps per tile.  Float32 geometries replay k_bwd64d
// (emulate_grad64d with its apply half) or k_bwd64q instead.
extern "C" int qsv_emulate_apply_ptm_batched_backward(int dtype, const double* x,
                                                      const double* ybar, double* xbar,
                                                      int rank, const int64_t* dims, int k,
                                                      const int32_t* targets,
                                                      const double* ptm, double* dptm,
                                                      int32_t grad_blocks) {
  GradPlan p;
  const int rc = bwd64_plan(dtype, rank, dims, k, targets, p);
  if (rc) return rc;
  QSV_CHECK_ARG(x && ybar && xbar && ptm && dptm, "null pointer argument");
  QSV_CHECK_ARG(p.n_in <= (int64_t(1) << 24), "emulation is for small vectors");
  QSV_CHECK_ARG(grad_blocks >= 0 && grad_blocks <= kGrad64Blocks, "grad_blocks outside [0, %d]",
                kGrad64Blocks);
  const int es = dtype_size(dtype), VN = p.vn, MB = 4, KS = 16;
  if (dtype == QSV_F32) {  // bwd64_plan admits only k_bwd64d's and k_bwd64q's geometries
    const int nbd = grad_blocks > 0 ? grad_blocks : grad64_blocks(p.n_orbits, 1);
    return p.quad ? emulate_bwd64q(p, x, ybar, nbd, dptm, ptm, xbar)
                  : emulate_grad64d(p, x, ybar, nbd, dptm, ptm, xbar);
  }
  const int nb = grad_blocks > 0 ? grad_blocks : bwd64_blocks(p, 1, 256);
  const FusedOffs fo = bwd64_offsets(p);
  // fragment table of P^T, as the kernel gathers it
  std::vector<double> As(KS * MB * 64);
  for (int i = 0; i < KS * MB * 64; ++i) {
    int o, kk;
    fused_frag_coords(64, es, 1, 1, i, o, kk);
    QSV_CHECK_ARG(o >= 0 && o < 64 && kk >= 0 && kk < 64, "bwd64: fragment map out of range");
    As[i] = ptm[int64_t(kk) * 64 + o];
  }
  const uint32_t COLS = 16 * VN, ntiles = (p.n_orbits + COLS - 1) / COLS;
  std::vector<int> written(p.n_in, 0);
  std::vector<double> bv(size_t(64) * VN * KS), acc(size_t(64) * MB * VN * 4);
  for (uint32_t t = 0; t < ntiles; ++t) {
    for (int lane = 0; lane < 64; ++lane) {
      const int n = lane & 15, j = lane >> 4;
      const uint32_t m0 = t * COLS + VN * n;
      const uint32_t m = m0 < p.n_orbits ? m0 : p.n_orbits - VN;
      int64_t bin, bout;
      orbit_base(p.g, m, bin, bout);
      for (int e = 1; e < VN; ++e) {
        int64_t bi2, bo2;
        orbit_base(p.g, m + e, bi2, bo2);
        QSV_CHECK_ARG(bi2 == bin + e && bo2 == bout + e, "bwd64: orbits of a vector not adjacent");
      }
      for (int s = 0; s < KS; ++s) {
        const int64_t a = bout + fo.in_l[j] + fo.in_u[s];
        QSV_CHECK_ARG(a % VN == 0 && a >= 0 && a + VN <= p.n_out, "bwd64: ybar load %lld",
                      (long long)a);
        for (int e = 0; e < VN; ++e) bv[(size_t(lane) * VN + e) * KS + s] = ybar[a + e];
      }
    }
    // acc[lane][mb][nb][r] = sum_s sum_kk A[row][kk] B[kk][col]
    std::fill(acc.begin(), acc.end(), 0.0);
    for (int s = 0; s < KS; ++s)
      for (int mb = 0; mb < MB; ++mb)
        for (int e = 0; e < VN; ++e)
          for (int lane = 0; lane < 64; ++lane)
            for (int r = 0; r < 4; ++r) {
              const int jj = lane >> 4, col = lane & 15;
              const int row = es == 8 ? jj + 4 * r : 4 * jj + r;
              double d = 0.0;
              for (int kk = 0; kk < 4; ++kk)
                d += As[(s * MB + mb) * 64 + row + 16 * kk] *
                     bv[(size_t(col + 16 * kk) * VN + e) * KS + s];
              acc[((size_t(lane) * MB + mb) * VN + e) * 4 + r] += d;
            }
    for (int lane = 0; lane < 64; ++lane) {
      const int n = lane & 15, j = lane >> 4;
      const uint32_t m0 = t * COLS + VN * n;
      if (m0 >= p.n_orbits) continue;
      int64_t bin, bout;
      orbit_base(p.g, m0, bin, bout);
      for (int q = 0; q < 4 * MB; ++q) {
        const int64_t a = bin + fo.out_l[j] + fo.out_u[q];
        QSV_CHECK_ARG(a % VN == 0 && a >= 0 && a + VN <= p.n_in, "bwd64: xbar store %lld",
                      (long long)a);
        for (int e = 0; e < VN; ++e) {
          xbar[a + e] = acc[((size_t(lane) * MB + (q >> 2)) * VN + e) * 4 + (q & 3)];
          ++written[a + e];
        }
      }
    }
  }
  for (int64_t a = 0; a < p.n_in; ++a)
    QSV_CHECK_ARG(written[a] == 1, "bwd64: xbar entry %lld written %d times", (long long)a,
                  written[a]);
  // part B: per tile four k-steps of 4 VN orbits, chunks strided like the tiles
  return emulate_grad64(p, es, x, ybar, nb, dptm, 4);
}
