// Dense D x D PTM application (D = 64: three qubits, D = 256: four qubits)
// on the gfx950 matrix cores.
//
//   out[o, g] = sum_k P[o, k] in[k, g],   g over the N / D orbits
//
// A block owns a tile of G = 64 orbits.  The D x G operand tile is gathered
// into LDS (lanes walk consecutive orbits, so each row of the tile is one
// coalesced run when the innermost axis is a bystander).  Each of the four
// waves computes D / 4 output rows for all G columns with
// v_mfma_{f64,f32}_16x16x4: the PTM arrives pre-arranged in A-fragment order
// (one contiguous 64-lane load per k-step, L2 resident), the tile supplies
// the B fragments, and the accumulators are stored straight to HBM.
#include "qsv_fused.h"

namespace qsv {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  __device__ static acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  __device__ static acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T, int D>
__global__ void __launch_bounds__(kBlock)
    k_apply_fused(const T* __restrict__ in, T* __restrict__ out, const Geom g,
                  const uint32_t n_orbits, const T* __restrict__ pf,
                  const int64_t* __restrict__ off_in, const int64_t* __restrict__ off_out) {
  constexpr int G = fused_cols(D);
  constexpr int LD = G + 4;      // padded LDS row
  constexpr int MB = D / 64;     // 16-row blocks per wave
  constexpr int NB = G / 16;     // 16-column blocks
  constexpr int KS = D / 4;      // k-steps
  typedef typename Mfma<T>::acc Acc;
  extern __shared__ __align__(16) unsigned char qsv_fused_smem[];
  T* Bs = reinterpret_cast<T*>(qsv_fused_smem);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t g0 = blockIdx.x * uint32_t(G);

  // gather the D x G operand tile
  {
    const int c = tid % G;
    const uint32_t gi = g0 + c;
    const bool ok = gi < n_orbits;
    int64_t bin = 0, bout = 0;
    if (ok) orbit_base(g, gi, bin, bout);
    for (int k = tid / G; k < D; k += kBlock / G) Bs[k * LD + c] = ok ? in[bin + off_in[k]] : T(0);
  }
  __syncthreads();

  Acc acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = Acc(0);

  const T* pw = pf + int64_t(w * MB) * KS * 64 + lane;
#pragma unroll 4
  for (int s = 0; s < KS; ++s) {
    T b[NB];
    const int krow = (4 * s + (lane >> 4)) * LD + (lane & 15);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) b[nb] = Bs[krow + 16 * nb];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const T a = pw[(int64_t(mb) * KS + s) * 64];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = Mfma<T>::op(a, b[nb], acc[mb][nb]);
    }
  }

  // accumulators -> HBM
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const uint32_t gi = g0 + 16 * nb + (lane & 15);
    if (gi < n_orbits) {
      int64_t bin, bout;
      orbit_base(g, gi, bin, bout);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = w * (D / 4) + 16 * mb + mfma_acc_row(sizeof(T), lane, r);
          out[bout + off_out[row]] = acc[mb][nb][r];
        }
    }
  }
}

template <typename T, int D>
static int launch_t(const Geom& g, uint32_t n_orbits, const void* in, void* out, const void* pf,
                    const int64_t* off_in, const int64_t* off_out, hipStream_t s) {
  constexpr int G = fused_cols(D);
  const size_t lds = size_t(D) * (G + 4) * sizeof(T);
  static bool attr_set = false;
  if (!attr_set) {
    QSV_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_apply_fused<T, D>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    attr_set = true;
  }
  const uint32_t tiles = (n_orbits + G - 1) / G;
  if (tiles == 0) return QSV_OK;
  hipLaunchKernelGGL((k_apply_fused<T, D>), dim3(tiles), dim3(kBlock), lds, s,
                     static_cast<const T*>(in), static_cast<T*>(out), g, n_orbits,
                     static_cast<const T*>(pf), off_in, off_out);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

int launch_fused(int dtype, int D, const Geom& g, uint32_t n_orbits, const void* in, void* out,
                 const void* pf, const int64_t* off_in, const int64_t* off_out, hipStream_t s) {
  if (dtype == QSV_F64) {
    if (D == 64) return launch_t<double, 64>(g, n_orbits, in, out, pf, off_in, off_out, s);
    if (D == 256) return launch_t<double, 256>(g, n_orbits, in, out, pf, off_in, off_out, s);
  } else {
    if (D == 64) return launch_t<float, 64>(g, n_orbits, in, out, pf, off_in, off_out, s);
    if (D == 256) return launch_t<float, 256>(g, n_orbits, in, out, pf, off_in, off_out, s);
  }
  set_error("internal: no MFMA kernel for D = %d", D);
  return QSV_ENOTSUP;
}

}  // namespace qsv
