// Adjoint of PTM application with respect to the PTM:
//
//   dP[b, o, i] = sum over orbits g of  ybar[b, g, o] * x[b, g, i]
//
// where x is the state the PTM was applied to and ybar the gradient of the
// loss with respect to the result.  (The gradient with respect to x is the
// same contraction with the transposed PTM and reuses the apply kernels.)
//
// Blocks stage chunks of orbits in LDS and accumulate the outer products in
// float64 registers; blocks write float64 partials, a second kernel sums the
// partials in a fixed order (deterministic).
#include "qsv_geom.h"

#include <string.h>
#include <vector>

namespace qsv {

constexpr int kGradBlocks = 256;

struct GradOffs {
  int64_t off_in[16];
  int64_t off_out[16];
};

// A block walks chunks of CHUNK orbits: thread t gathers orbit t's D_in
// entries of x and D_out entries of ybar into LDS (orbit-major rows, so the
// gathers of neighbouring lanes hit neighbouring orbits), then wave w
// accumulates the outer products of its quarter of the chunk, lane l owning
// dP[o = l / 4][4 (l % 4) .. 4 (l % 4) + 3] in float64.  x and ybar are read
// once.  The four waves' sums meet in LDS at the end.
template <typename T, int CHUNK>
__global__ void __launch_bounds__(kBlock)
    k_ptm_grad(const T* __restrict__ x, const T* __restrict__ ybar, const Geom g,
               const GradOffs offs, const uint32_t n_orbits, double* __restrict__ partial) {
  constexpr int LD = 17;  // padded row of 16
  __shared__ T X[CHUNK * LD];
  __shared__ T Y[CHUNK * LD];
  __shared__ double red[kBlock / 64][64 * 4];
  const uint32_t b = blockIdx.y;
  x += int64_t(b) * g.bs_in;
  ybar += int64_t(b) * g.bs_out;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int o = lane >> 2, i0 = (lane & 3) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  constexpr int PER_WAVE = CHUNK / (kBlock / 64);

  for (uint32_t c0 = blockIdx.x * uint32_t(CHUNK); c0 < n_orbits;
       c0 += gridDim.x * uint32_t(CHUNK)) {
    for (int j = tid; j < CHUNK; j += kBlock) {
      const uint32_t m = c0 + j;
      if (m < n_orbits) {
        int64_t bin, bout;
        orbit_base(g, m, bin, bout);
#pragma unroll
        for (int i = 0; i < 16; ++i) X[j * LD + i] = i < g.din ? x[bin + offs.off_in[i]] : T(0);
#pragma unroll
        for (int r = 0; r < 16; ++r)
          Y[j * LD + r] = r < g.dout ? ybar[bout + offs.off_out[r]] : T(0);
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) X[j * LD + i] = Y[j * LD + i] = T(0);
      }
    }
    __syncthreads();
    const int j0 = w * PER_WAVE;
#pragma unroll 4
    for (int j = j0; j < j0 + PER_WAVE; ++j) {
      const double yv = double(Y[j * LD + o]);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += yv * double(X[j * LD + i0 + c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) red[w][lane * 4 + c] = acc[c];
  __syncthreads();
  // thread t < 256 sums entry t = (o, i) over the four waves
  {
    const int oo = tid >> 4, ii = tid & 15;
    const int src = (oo * 4 + (ii >> 2)) * 4 + (ii & 3);  // lane (oo, ii / 4), slot ii % 4
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < kBlock / 64; ++ww) s += red[ww][src];
    if (oo < g.dout && ii < g.din)
      partial[((int64_t(b) * gridDim.x + blockIdx.x) * g.dout + oo) * g.din + ii] = s;
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
  int64_t n_in, n_out;
};

// Geometry of the contraction without vector widths (one orbit per step).
static int make_grad_plan(int rank, const int64_t* dims, int k, const int32_t* targets,
                          const int32_t* dims_out_t, GradPlan& p) {
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
  QSV_CHECK_ARG(din <= 16 && dout <= 16, "PTM gradient supports up to 16 x 16, got %lld x %lld",
                (long long)dout, (long long)din);
  p.din = int(din);
  p.dout = int(dout);
  memset(&p.offs, 0, sizeof(p.offs));
  for (int64_t i = 0; i < din; ++i) {
    int64_t r = i, off = 0;
    for (int j = k - 1; j >= 0; --j) {
      off += (r % dims[targets[j]]) * sin[targets[j]];
      r /= dims[targets[j]];
    }
    p.offs.off_in[i] = off;
  }
  for (int64_t o = 0; o < dout; ++o) {
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
  return QSV_OK;
}

static int grad_blocks(uint32_t n_orbits) {
  const uint32_t b = (n_orbits + kBlock * 2 - 1) / (kBlock * 2);
  return int(b < 1 ? 1 : (b > kGradBlocks ? kGradBlocks : b));
}

}  // namespace qsv

using namespace qsv;

extern "C" int64_t qsv_ptm_grad_workspace_bytes(int batch, int rank, const int64_t* dims, int k,
                                                const int32_t* targets,
                                                const int32_t* dims_out_t) {
  GradPlan p;
  const int rc = make_grad_plan(rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  return int64_t(batch) * grad_blocks(p.n_orbits) * p.din * p.dout * 8;
}

extern "C" int qsv_ptm_grad(int dtype, const void* x, const void* ybar, int batch, int rank,
                            const int64_t* dims, int k, const int32_t* targets,
                            const int32_t* dims_out_t, void* dptm, int accumulate,
                            void* workspace, int64_t workspace_bytes, void* stream) {
  QSV_CHECK_ARG(dtype == QSV_F32 || dtype == QSV_F64, "unknown dtype %d", dtype);
  QSV_CHECK_ARG(batch >= 0 && batch <= 65535, "batch %d outside [0, 65535]", batch);
  GradPlan p;
  const int rc = make_grad_plan(rank, dims, k, targets, dims_out_t, p);
  if (rc) return rc;
  if (batch == 0) return QSV_OK;
  QSV_CHECK_ARG(x && ybar && dptm, "null pointer argument");
  const int nb = grad_blocks(p.n_orbits);
  const int64_t need = int64_t(batch) * nb * p.din * p.dout * 8;
  QSV_CHECK_ARG(workspace && workspace_bytes >= need, "PTM gradient needs %lld workspace bytes",
                (long long)need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* partial = static_cast<double*>(workspace);
  const dim3 grid(nb, batch);
  if (dtype == QSV_F64)
    hipLaunchKernelGGL((k_ptm_grad<double, 128>), grid, dim3(kBlock), 0, s,
                       static_cast<const double*>(x), static_cast<const double*>(ybar), p.g,
                       p.offs, p.n_orbits, partial);
  else
    hipLaunchKernelGGL((k_ptm_grad<float, 256>), grid, dim3(kBlock), 0, s,
                       static_cast<const float*>(x), static_cast<const float*>(ybar), p.g,
                       p.offs, p.n_orbits, partial);
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
