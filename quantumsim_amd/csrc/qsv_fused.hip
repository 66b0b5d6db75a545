// Dense D x D PTM application (D = 64: three qubits, D = 256: four qubits)
// on the gfx950 matrix cores.  Index maps and the fragment table are
// described in qsv_fused.h.
//
// k_fused64: the 64 x 64 PTM (32 KB in float64) is staged in LDS once per
// block; after that every wave runs on its own, walking tiles of 16 NB
// orbits: load the tile's B fragments straight from HBM (16-byte vectors
// along orbits or along the PTM index), 16 NB x 4 MFMAs per k-step with the
// A fragments read from LDS, store the accumulators straight back.  Few
// registers (60-90 VGPRs) leave room for 3-5 waves per SIMD, which overlap
// one wave's memory wait with another's MFMAs (measured: as fast as
// explicit double-buffering in registers, which halves the occupancy).
// Memory-bound: 2 D element accesses per orbit against 2 D^2 flops.
//
// k_fused256: the 256 x 256 PTM (512 KB in float64) does not fit in LDS; a
// block of four waves takes 64 orbits (16 per wave, accumulators for all
// 256 rows in registers) and streams the PTM through LDS in chunks of 4
// (float64) or 8 (float32) k-steps, double-buffered: the next chunk and the
// next B fragments are loaded while the current chunk's MFMAs run.
// MFMA-bound.
#include "qsv_fused.h"
#include "qsv_mfma.h"

#include <stdlib.h>
#include <algorithm>

namespace qsv {

int fused_nb(int elem_bytes, int D, int vn) {
  if (D != 64) return 1;
  return vn > 1 ? vn : (elem_bytes == 4 ? 2 : 1);
}

template <typename T, int V, bool NT>
__device__ __forceinline__ void load_vec(const T* p, T* dst) {
  typedef typename VT<T, V>::t Vec;
  const Vec v = gload<NT>(reinterpret_cast<const Vec*>(p));
  if constexpr (V == 1) {
    dst[0] = v;
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) dst[e] = v[e];
  }
}

template <typename T, int V, bool NT>
__device__ __forceinline__ void store_vec(T* p, const T* vals) {
  typedef typename VT<T, V>::t Vec;
  Vec v;
  if constexpr (V == 1) {
    v = vals[0];
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) v[e] = vals[e];
  }
  gstore<NT>(reinterpret_cast<Vec*>(p), v);
}

// Memory accesses of the loops below are unconditional (the partial tail
// tile is handled separately with guards): a global access under a branch
// makes the compiler's wait-count bookkeeping conservative at the join.
//
// Column (orbit) of lane n in column block nb of a tile:
//   VN = 1: tile * 16 NB + 16 nb + n
//   VN > 1: tile * 16 VN + VN n + nb   (NB = VN: lane n owns VN consecutive
//           orbits, loaded and stored as one 16-byte vector when the
//           innermost axis is a bystander with unit stride)

// NT (nontemporal) applies to the loads; stores are nontemporal only when
// every store instruction writes whole 128-byte lines (16 consecutive
// orbits of float64, or orbit vectors): partial lines written nontemporally
// reach HBM as separate masked writes (measured: up to 2.2x the bytes).
template <typename T, int VI, int VO, int VN, bool NT>
__host__ __device__ constexpr bool fused_nt_stores() {
  return NT && (VN > 1 || (VI == 1 && VO == 1 && sizeof(T) == 8));
}

template <typename T, int VI, int VO, int VN, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_fused64(const T* __restrict__ in, T* __restrict__ out, const Geom g,
              const uint32_t n_orbits, const T* __restrict__ pf, const FusedOffs fo,
              const int64_t pstride, const int prow) {
  constexpr int MB = 4, KS = 16;
  constexpr int NB = VN > 1 ? VN : (sizeof(T) == 4 ? 2 : 1);
  constexpr int NL = VN > 1 ? 1 : NB;  // orbit bases per lane
  constexpr uint32_t COLS = 16 * NB;
  typedef typename Mfma<T>::acc Acc;
  typedef typename VT<T, VN>::t VNv;
  __shared__ T As[KS * MB * 64];
  // batch entry blockIdx.y: its state and its PTM (pstride elements apart;
  // 0 = one PTM for all).  prow: the PTM is row-major and is gathered into
  // fragment order here (batched device PTMs); otherwise pf is already in
  // fragment order.
  in += int64_t(blockIdx.y) * g.bs_in;
  out += int64_t(blockIdx.y) * g.bs_out;
  pf += int64_t(blockIdx.y) * pstride;
  if (prow) {
    for (int i = threadIdx.x; i < KS * MB * 64; i += kBlock)
      As[i] = pf[fused_frag_source(64, int(sizeof(T)), VI, VO, i)];
  } else {
    for (int i = threadIdx.x; i < KS * MB * 64; i += kBlock) As[i] = pf[i];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, n = lane & 15, j = lane >> 4;
  const uint32_t nw = gridDim.x * (kBlock / 64);
  const uint32_t wid = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const uint32_t full = n_orbits / COLS;
  const int64_t li = fo.in_l[j], lo = fo.out_l[j];

  T b[NB][KS];
  Acc acc[MB][NB];
  auto load = [&](uint32_t tile) {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      uint32_t m = fused64_orbit_of(NB, VN, tile, l, n);
      if (m >= n_orbits) m = n_orbits - VN;  // tail tile only: loaded, never stored
      int64_t bin, bout;
      orbit_base(g, m, bin, bout);
      const T* p = in + bin + li;
      if constexpr (VN > 1) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const VNv v = gload<NT>(reinterpret_cast<const VNv*>(p + fo.in_u[s]));
#pragma unroll
          for (int e = 0; e < VN; ++e) b[e][s] = v[e];
        }
      } else {
#pragma unroll
        for (int u = 0; u < KS / VI; ++u) load_vec<T, VI, NT>(p + fo.in_u[u], &b[l][u * VI]);
      }
    }
  };
  auto compute = [&]() {
    // the LDS offset is made opaque so the compiler re-reads the PTM from
    // LDS every tile instead of pinning all 64 fragments in VGPRs (which
    // would halve the waves per SIMD)
    int abase = lane;
    asm volatile("" : "+v"(abase));
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = Acc(0);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const T a = As[(s * MB + mb) * 64 + abase];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = Mfma<T>::op(a, b[nb][s], acc[mb][nb]);
      }
  };
  auto store = [&](int l, uint32_t m) {
    int64_t bin, bout;
    orbit_base(g, m, bin, bout);
    T* p = out + bout + lo;
    if constexpr (VN > 1) {
#pragma unroll
      for (int q = 0; q < 4 * MB; ++q) {
        VNv v;
#pragma unroll
        for (int e = 0; e < VN; ++e) v[e] = acc[q >> 2][e][q & 3];
        gstore<fused_nt_stores<T, VI, VO, VN, NT>()>(reinterpret_cast<VNv*>(p + fo.out_u[q]), v);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4 * MB / VO; ++v) {
        T vals[VO];
#pragma unroll
        for (int e = 0; e < VO; ++e) {
          const int q = v * VO + e;
          vals[e] = acc[q >> 2][l][q & 3];
        }
        store_vec<T, VO, fused_nt_stores<T, VI, VO, VN, NT>()>(p + fo.out_u[v], vals);
      }
    }
  };

  for (uint32_t t = wid; t < full; t += nw) {
    load(t);
    compute();
#pragma unroll
    for (int l = 0; l < NL; ++l) store(l, fused64_orbit_of(NB, VN, t, l, n));
  }
  if (full * COLS < n_orbits && wid == full % nw) {
    load(full);
    compute();
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const uint32_t m = fused64_orbit_of(NB, VN, full, l, n);
      if (m < n_orbits) store(l, m);
    }
  }
}

template <typename T, int VI, int VO, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_fused256(const T* __restrict__ in, T* __restrict__ out, const Geom g,
               const uint32_t n_orbits, const T* __restrict__ pf, const FusedOffs fo) {
  constexpr int MB = 16, KS = 64, CS = fused_chunk_steps(sizeof(T)), NCH = KS / CS;
  constexpr int CHUNK = CS * MB * 64;
  constexpr int PER = CHUNK / kBlock;
  constexpr uint32_t COLS = 16 * (kBlock / 64);
  static_assert(CS % VI == 0, "chunk must hold whole vectors");
  typedef typename Mfma<T>::acc Acc;
  __shared__ T As[2][CHUNK];

  const int tid = threadIdx.x, lane = tid & 63, n = lane & 15, j = lane >> 4, w = tid >> 6;
  const uint32_t ntiles = (n_orbits + COLS - 1) / COLS;
  const int64_t li = fo.in_l[j], lo = fo.out_l[j];

  // chunk 0 of the PTM and the first B fragments of a tile are loaded
  // during the previous tile's last chunk (NCH is even, so the prefetched
  // chunk 0 lands in As[0], where chunk 0 is read)
  static_assert(NCH % 2 == 0, "chunk 0 must come back to buffer 0");
  auto tile_base = [&](uint32_t t, uint32_t& mm, int64_t& bo) -> const T* {
    mm = t * COLS + 16 * w + n;
    int64_t bi;
    orbit_base(g, mm < n_orbits ? mm : n_orbits - 1, bi, bo);
    return in + bi + li;
  };
  T ar[PER], bc[CS], bn[CS];
  uint32_t m = 0;
  int64_t bout = 0;
  const T* pin = in;
  if (blockIdx.x < ntiles) {
    pin = tile_base(blockIdx.x, m, bout);
#pragma unroll
    for (int i = 0; i < PER; ++i) ar[i] = pf[i * kBlock + tid];
#pragma unroll
    for (int uu = 0; uu < CS / VI; ++uu) load_vec<T, VI, NT>(pin + fo.in_u[uu], &bn[uu * VI]);
#pragma unroll
    for (int i = 0; i < PER; ++i) As[0][i * kBlock + tid] = ar[i];
  }
  __syncthreads();

  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    // the next tile (this one again past the last: loaded, never used)
    const uint32_t tn = tile + gridDim.x < ntiles ? tile + gridDim.x : tile;
    uint32_t mn;
    int64_t boutn;
    const T* pinn = tile_base(tn, mn, boutn);

    Acc acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) acc[mb] = Acc(0);

#pragma unroll 1
    for (int c = 0; c < NCH; ++c) {
#pragma unroll
      for (int ss = 0; ss < CS; ++ss) bc[ss] = bn[ss];
      // next chunk of this tile, or chunk 0 and the first B fragments of
      // the next tile
      const bool last = c + 1 == NCH;
      const int cn = last ? 0 : c + 1;
      const T* src = pf + int64_t(cn) * CHUNK + tid;
      const T* pb = last ? pinn : pin;
#pragma unroll
      for (int i = 0; i < PER; ++i) ar[i] = src[i * kBlock];
#pragma unroll
      for (int uu = 0; uu < CS / VI; ++uu)
        load_vec<T, VI, NT>(pb + fo.in_u[cn * (CS / VI) + uu], &bn[uu * VI]);
      const T* A = As[c & 1];
#pragma unroll
      for (int ss = 0; ss < CS; ++ss)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[mb] = Mfma<T>::op(A[(ss * MB + mb) * 64 + lane], bc[ss], acc[mb]);
      // the other buffer: nobody reads it in this iteration
      T* dst = As[(c + 1) & 1] + tid;
#pragma unroll
      for (int i = 0; i < PER; ++i) dst[i * kBlock] = ar[i];
      __syncthreads();
    }

    const uint32_t mcur = m;
    const int64_t boutcur = bout;
    m = mn;
    bout = boutn;
    pin = pinn;
    if (mcur < n_orbits) {
      T* p = out + boutcur + lo;
#pragma unroll
      for (int v = 0; v < 4 * MB / VO; ++v) {
        T vals[VO];
#pragma unroll
        for (int e = 0; e < VO; ++e) {
          const int q = v * VO + e;
          vals[e] = acc[q >> 2][q & 3];
        }
        store_vec<T, VO, fused_nt_stores<T, VI, VO, 1, NT>()>(p + fo.out_u[v], vals);
      }
    }
  }
}

// Blocks that fit on the device at once (occupancy x compute units).
static uint32_t resident_blocks(const void* kern) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  (void)hipGetLastError();
  return uint32_t(per_cu) * uint32_t(cus);
}

// Batched launches (batch > 1) split the resident blocks over the states;
// every state gets at least one block.
template <typename T, int VI, int VO, int VN, bool NT>
static int launch64(const FusedPlan& fp, const Geom& g, const void* in, void* out, const void* pf,
                    hipStream_t s, uint32_t batch, int64_t pstride, int prow) {
  constexpr int NB = VN > 1 ? VN : (sizeof(T) == 4 ? 2 : 1);
  const auto kern = k_fused64<T, VI, VO, VN, NT>;
  static const uint32_t cap = resident_blocks(reinterpret_cast<const void*>(kern));
  const uint32_t ntiles = (fp.n_orbits + 16 * NB - 1) / (16 * NB);
  const uint32_t want = (ntiles + kBlock / 64 - 1) / (kBlock / 64);
  const uint32_t per_state = std::max<uint32_t>(1, cap / std::max<uint32_t>(batch, 1));
  hipLaunchKernelGGL(kern, dim3(want < per_state ? want : per_state, batch), dim3(kBlock), 0, s,
                     static_cast<const T*>(in), static_cast<T*>(out), g, fp.n_orbits,
                     static_cast<const T*>(pf), fp.offs, pstride, prow);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename T, int VI, int VO, bool NT>
static int launch256(const FusedPlan& fp, const Geom& g, const void* in, void* out,
                     const void* pf, hipStream_t s) {
  const auto kern = k_fused256<T, VI, VO, NT>;
  static const uint32_t cap = resident_blocks(reinterpret_cast<const void*>(kern));
  const uint32_t want = (fp.n_orbits + 63) / 64;
  hipLaunchKernelGGL(kern, dim3(want < cap ? want : cap), dim3(kBlock), 0, s,
                     static_cast<const T*>(in), static_cast<T*>(out), g, fp.n_orbits,
                     static_cast<const T*>(pf), fp.offs);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename T, int VL, bool NT>
static int dispatch(const FusedPlan& fp, const Geom& g, const void* in, void* out,
                    const void* pf, hipStream_t s, uint32_t batch = 1, int64_t pstride = 0,
                    int prow = 0) {
  const int key = (fp.vi == VL ? 2 : 0) | (fp.vo == VL ? 1 : 0);
  if (fp.D == 64) {
    if (fp.vn > 1) return launch64<T, 1, 1, VL, NT>(fp, g, in, out, pf, s, batch, pstride, prow);
    switch (key) {
      case 0: return launch64<T, 1, 1, 1, NT>(fp, g, in, out, pf, s, batch, pstride, prow);
      case 1: return launch64<T, 1, VL, 1, NT>(fp, g, in, out, pf, s, batch, pstride, prow);
      case 2: return launch64<T, VL, 1, 1, NT>(fp, g, in, out, pf, s, batch, pstride, prow);
      default: return launch64<T, VL, VL, 1, NT>(fp, g, in, out, pf, s, batch, pstride, prow);
    }
  }
  switch (key) {
    case 0: return launch256<T, 1, 1, NT>(fp, g, in, out, pf, s);
    case 1: return launch256<T, 1, VL, NT>(fp, g, in, out, pf, s);
    case 2: return launch256<T, VL, 1, NT>(fp, g, in, out, pf, s);
    default: return launch256<T, VL, VL, NT>(fp, g, in, out, pf, s);
  }
}

// Device form of fused_fragments: pf[slot] = ptm[o * D + k] for the slot's
// (k-step, row block, lane), both arrays in the state's element type.  One
// thread per slot; D * D slots.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_fused_pack(const T* __restrict__ ptm, T* __restrict__ pf, int D, int vi, int vo) {
  const int64_t slot = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (slot >= int64_t(D) * D) return;
  pf[slot] = ptm[fused_frag_source(D, int(sizeof(T)), vi, vo, slot)];
}

int launch_fused_pack(int dtype, const FusedPlan& fp, const void* ptm_dev, void* pf,
                      hipStream_t s) {
  const int64_t n = int64_t(fp.D) * fp.D;
  const dim3 grid(uint32_t((n + kBlock - 1) / kBlock));
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_fused_pack<double>, grid, dim3(kBlock), 0, s,
                       static_cast<const double*>(ptm_dev), static_cast<double*>(pf), fp.D,
                       fp.vi, fp.vo);
  else
    hipLaunchKernelGGL(k_fused_pack<float>, grid, dim3(kBlock), 0, s,
                       static_cast<const float*>(ptm_dev), static_cast<float*>(pf), fp.D, fp.vi,
                       fp.vo);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

int launch_fused64_batched(int dtype, const FusedPlan& fp, const Geom& g, bool nt,
                           const void* in, void* out, const void* ptm_rows, int64_t pstride,
                           uint32_t batch, hipStream_t s) {
  if (fp.D != 64) {
    set_error("internal: batched MFMA apply is 64 x 64 only, got D = %d", fp.D);
    return QSV_ENOTSUP;
  }
  if (fp.n_orbits == 0 || batch == 0) return QSV_OK;
  // float32 runs of four on whole tiles: the direct-to-LDS kernel (5-30 %
  // faster than k_fused64<float, 4, 4, 1> on 8-12-qubit batches, same bits)
  if (dtype == QSV_F32 && apply64q_applies(4, fp))
    return launch_apply64q(fp, g, nt, in, out, ptm_rows, pstride, batch, s);
  if (dtype == QSV_F64)
    return nt ? dispatch<double, 2, true>(fp, g, in, out, ptm_rows, s, batch, pstride, 1)
              : dispatch<double, 2, false>(fp, g, in, out, ptm_rows, s, batch, pstride, 1);
  return nt ? dispatch<float, 4, true>(fp, g, in, out, ptm_rows, s, batch, pstride, 1)
            : dispatch<float, 4, false>(fp, g, in, out, ptm_rows, s, batch, pstride, 1);
}

int launch_fused(int dtype, const FusedPlan& fp, const Geom& g, bool nt, const void* in,
                 void* out, const void* pf, hipStream_t s) {
  if (fp.D != 64 && fp.D != 256) {
    set_error("internal: no MFMA kernel for D = %d", fp.D);
    return QSV_ENOTSUP;
  }
  if (fp.n_orbits == 0) return QSV_OK;
  if (dtype == QSV_F64)
    return nt ? dispatch<double, 2, true>(fp, g, in, out, pf, s)
              : dispatch<double, 2, false>(fp, g, in, out, pf, s);
  return nt ? dispatch<float, 4, true>(fp, g, in, out, pf, s)
            : dispatch<float, 4, false>(fp, g, in, out, pf, s);
}

}  // namespace qsv
