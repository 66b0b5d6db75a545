// Gathers, reductions and scaling for Pauli vectors on gfx950, plus the
// library's error reporting.
#include "qsv_common.h"

#include <string.h>
#include <type_traits>

namespace qsv {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

constexpr int kBlock = 256;

// ------------------------------------------------------------ multitake
// out = in[np.ix_(idx_0..idx_{r-1})] (quantumsim/pauli_vectors/primitives.cu:92-130);
// an index of -1 yields 0 (levels absent from a subbasis, numpy.py:66-80).
struct TakeGeom {
  int32_t rank;
  uint32_t n_out;
  FastDiv out_dim[QSV_MAX_RANK];   // output dims
  int64_t in_stride[QSV_MAX_RANK];  // input strides
  int32_t off[QSV_MAX_RANK];        // start of axis r's list in idx
};

template <typename TI, typename TO>
__global__ void __launch_bounds__(kBlock)
    k_multitake(const TI* __restrict__ in, TO* __restrict__ out,
                const int32_t* __restrict__ idx, const TakeGeom g) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= g.n_out) return;
  uint32_t rem = j;
  int64_t addr = 0;
  bool zero = false;
  for (int r = g.rank - 1; r >= 0; --r) {
    const uint32_t q = fdiv(rem, g.out_dim[r]);
    const uint32_t digit = rem - q * g.out_dim[r].d;
    rem = q;
    const int32_t ix = idx[g.off[r] + digit];
    zero |= ix < 0;
    addr += int64_t(ix) * g.in_stride[r];
  }
  out[j] = zero ? TO(0) : TO(in[addr]);
}

// ------------------------------------------------------------- axis sum
// out[j] = sum_{o,i} in[o, j, i]; per-block float64 partials in a fixed
// order, then one block sums the partials in a fixed order.
constexpr int kMaxD = 64;
constexpr int kSumBlocks = 512;

struct SumGeom {
  uint64_t n;      // total elements
  uint64_t inner;  // n_inner
  int32_t d;       // middle dim
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_axis_sum_partial(const T* __restrict__ in, const SumGeom g, double* __restrict__ partial) {
  __shared__ double red[kBlock / 64][kMaxD];
  double acc[8];
  // d <= 8: per-thread accumulators; larger d loops over slices of 8
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = 0; j0 < g.d; j0 += 8) {
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < g.n; i += stride) {
      const uint64_t row = g.inner == 1 ? i : (i / g.inner);
      const uint32_t jj = uint32_t(row % uint32_t(g.d));
      const double x = double(in[i]);
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (jj == uint32_t(j0 + c)) acc[c] += x;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (j0 + c < g.d) {
        const double w = wave_sum(acc[c]);
        if (lane == 0) red[wave][j0 + c] = w;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < g.d) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
    partial[int64_t(blockIdx.x) * g.d + threadIdx.x] = s;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_axis_sum_final(const double* __restrict__ partial, int nblocks, int d,
                     double* __restrict__ out) {
  __shared__ double red[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = 0; j < d; ++j) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[int64_t(b) * d + j];
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) t += red[w];
      out[j] = t;
    }
    __syncthreads();
  }
}

static int sum_blocks(uint64_t n) {
  const uint64_t b = (n + kBlock * 8 - 1) / (kBlock * 8);
  return int(b < 1 ? 1 : (b > kSumBlocks ? kSumBlocks : b));
}

// ----------------------------------------------------------------- scale
template <typename T>
__global__ void __launch_bounds__(kBlock) k_scale(T* __restrict__ data, uint64_t n, double s) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
    data[i] = T(double(data[i]) * s);
}

}  // namespace qsv

using namespace qsv;

extern "C" const char* qsv_last_error(void) { return g_err; }

extern "C" int qsv_abi_version(void) { return 100; }

extern "C" int qsv_multitake(int in_dtype, int out_dtype, const void* in, void* out, int rank,
                             const int64_t* in_dims, const int64_t* out_dims,
                             const int32_t* idx_dev, const int32_t* offsets, void* stream) {
  return qsv_multitake_strided(in_dtype, out_dtype, in, out, rank, in_dims, nullptr, out_dims,
                               idx_dev, offsets, stream);
}

extern "C" int qsv_multitake_strided(int in_dtype, int out_dtype, const void* in, void* out,
                                     int rank, const int64_t* in_dims,
                                     const int64_t* in_strides, const int64_t* out_dims,
                                     const int32_t* idx_dev, const int32_t* offsets,
                                     void* stream) {
  QSV_CHECK_ARG(dtype_size(in_dtype) && dtype_size(out_dtype), "unknown dtype");
  QSV_CHECK_ARG(!(in_dtype == QSV_F64 && out_dtype == QSV_F32),
                "f64 -> f32 narrowing is not a gather");
  QSV_CHECK_ARG(rank >= 0 && rank <= QSV_MAX_RANK, "rank %d outside [0, %d]", rank,
                QSV_MAX_RANK);
  QSV_CHECK_ARG(in && out, "null data pointer");
  QSV_CHECK_ARG(rank == 0 || (in_dims && out_dims && idx_dev && offsets), "null argument");
  TakeGeom g;
  memset(&g, 0, sizeof(g));
  g.rank = rank;
  int64_t n_out = 1, stride = 1;
  for (int r = rank - 1; r >= 0; --r) {
    QSV_CHECK_ARG(in_dims[r] >= 1 && out_dims[r] >= 0, "bad dims on axis %d", r);
    if (in_strides) {
      QSV_CHECK_ARG(in_strides[r] >= stride, "strides of axis %d overlap the axes below", r);
      stride = in_strides[r];
    }
    g.in_stride[r] = stride;
    stride *= in_dims[r];
    if (out_dims[r] == 0) n_out = 0;
    n_out *= (out_dims[r] > 0 ? out_dims[r] : 1);
    g.out_dim[r] = make_fastdiv(uint32_t(out_dims[r] > 0 ? out_dims[r] : 1));
    g.off[r] = offsets[r];
    QSV_CHECK_ARG(n_out < (int64_t(1) << 32) - kBlock, "gather output too large");
  }
  if (n_out == 0) return QSV_OK;
  g.n_out = uint32_t(n_out);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(ceil_div_u32(g.n_out, kBlock));
  if (in_dtype == QSV_F64)
    hipLaunchKernelGGL((k_multitake<double, double>), grid, dim3(kBlock), 0, s,
                       static_cast<const double*>(in), static_cast<double*>(out), idx_dev, g);
  else if (out_dtype == QSV_F64)
    hipLaunchKernelGGL((k_multitake<float, double>), grid, dim3(kBlock), 0, s,
                       static_cast<const float*>(in), static_cast<double*>(out), idx_dev, g);
  else
    hipLaunchKernelGGL((k_multitake<float, float>), grid, dim3(kBlock), 0, s,
                       static_cast<const float*>(in), static_cast<float*>(out), idx_dev, g);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

extern "C" int64_t qsv_axis_sum_workspace_bytes(int64_t n_outer, int d, int64_t n_inner) {
  if (n_outer < 0 || n_inner < 1 || d < 1 || d > kMaxD) {
    set_error("qsv_axis_sum_workspace_bytes: bad shape");
    return QSV_EINVAL;
  }
  const uint64_t n = uint64_t(n_outer) * d * n_inner;
  return int64_t(sum_blocks(n)) * d * 8;
}

extern "C" int qsv_axis_sum(int dtype, const void* in, int64_t n_outer, int d, int64_t n_inner,
                            double* out_f64, void* workspace, int64_t workspace_bytes,
                            void* stream) {
  QSV_CHECK_ARG(dtype_size(dtype), "unknown dtype %d", dtype);
  QSV_CHECK_ARG(d >= 1 && d <= kMaxD, "axis size %d outside [1, %d]", d, kMaxD);
  QSV_CHECK_ARG(n_outer >= 0 && n_inner >= 1, "bad shape");
  QSV_CHECK_ARG(out_f64 && (in || n_outer == 0), "null pointer");
  const uint64_t n = uint64_t(n_outer) * d * n_inner;
  QSV_CHECK_ARG(n < (uint64_t(1) << 40), "too many elements");
  QSV_CHECK_ARG(uint64_t(n_outer) * d < (uint64_t(1) << 32), "too many rows");
  const int nb = sum_blocks(n);
  QSV_CHECK_ARG(workspace && workspace_bytes >= int64_t(nb) * d * 8,
                "axis sum needs a workspace of %lld bytes", (long long)(int64_t(nb) * d * 8));
  SumGeom g;
  memset(&g, 0, sizeof(g));
  g.n = n;
  g.d = d;
  g.inner = uint64_t(n_inner);
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* partial = static_cast<double*>(workspace);
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_axis_sum_partial<double>, dim3(nb), dim3(kBlock), 0, s,
                       static_cast<const double*>(in), g, partial);
  else
    hipLaunchKernelGGL(k_axis_sum_partial<float>, dim3(nb), dim3(kBlock), 0, s,
                       static_cast<const float*>(in), g, partial);
  QSV_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_axis_sum_final, dim3(1), dim3(kBlock), 0, s, partial, nb, d, out_f64);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

extern "C" int qsv_scale(int dtype, void* data, int64_t n, double s_, void* stream) {
  QSV_CHECK_ARG(dtype_size(dtype), "unknown dtype %d", dtype);
  QSV_CHECK_ARG(n >= 0, "negative size");
  if (n == 0) return QSV_OK;
  QSV_CHECK_ARG(data, "null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t blocks = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (blocks > 65536) blocks = 65536;
  if (dtype == QSV_F64)
    hipLaunchKernelGGL(k_scale<double>, dim3(blocks), dim3(kBlock), 0, s,
                       static_cast<double*>(data), uint64_t(n), s_);
  else
    hipLaunchKernelGGL(k_scale<float>, dim3(blocks), dim3(kBlock), 0, s,
                       static_cast<float*>(data), uint64_t(n), s_);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

extern "C" int64_t qsv_selftest_fastdiv(int32_t n_divisors, uint64_t seed) {
  uint64_t x = seed | 1;
  auto rnd = [&x]() {  // xorshift64*
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    return uint32_t((x * 0x2545F4914F6CDD1DULL) >> 32);
  };
  int64_t bad = 0;
  auto check_div = [&](uint32_t d) {
    const FastDiv f = make_fastdiv(d);
    const uint32_t fixed[] = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 0x7fffffffu, 0x80000000u,
                              0xfffffffeu, 0xffffffffu};
    for (uint32_t n : fixed)
      if (fdiv(n, f) != n / d) ++bad;
    for (int i = 0; i < 256; ++i) {
      const uint32_t n = rnd();
      if (fdiv(n, f) != n / d) ++bad;
      const uint32_t m = n % (d < 65536 ? d * 65536u : 0xffffffffu);
      if (fdiv(m, f) != m / d) ++bad;
    }
  };
  for (uint32_t d = 1; d <= uint32_t(n_divisors > 0 ? n_divisors : 0); ++d) check_div(d);
  const uint32_t bases_[] = {2, 3, 4, 9};
  for (uint32_t b : bases_)
    for (uint64_t p = b; p < (uint64_t(1) << 32); p *= b) check_div(uint32_t(p));
  for (int i = 0; i < n_divisors; ++i) {
    const uint32_t d = rnd();
    if (d) check_div(d);
  }
  check_div(0xffffffffu);
  check_div(0x80000001u);
  return bad;
}
