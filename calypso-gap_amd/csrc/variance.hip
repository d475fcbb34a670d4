// Predictive variance of the GAP.  The reference declares it and leaves it
// out (FGAP_CALC sets VARIANCE = 0 and keeps the formula commented,
// gappy/libgap/gap_calc.f90:205-210; FGAP_READ sets INVCMM = 0,
// gap_calc.f90:355-358):
//   sigma2_i = delta - k_i Q k_i^T,
//   k_ik = delta exp(-1/2 sum_d ((x_id - MM_kd) / theta_d)^2)   (gap_calc.f90:270-287),
// delta = 1 (gap_calc.f90:8), Q = qmm (nsparse x nsparse), and the
// structure's VARIANCE = mean_i sigma2_i.
//
// The quadratic form needs whole rows of K, so K is materialised for a block
// of atoms at a time (FP64, [rows][Mp]):
//   k_row_norms       |xt_i|^2 (and the FP64 copy of a mixed run's features),
//   dgemm             S = xt mt^T (rocBLAS),
//   k_kernel_from_dot K = exp(-(|xt_i|^2 + |mt_k|^2 - 2 S_ik) / 2) in place,
//   dgemm             U = K Q (rocBLAS),
//   k_row_quadratic   sigma2_i = delta - sum_k U_ik K_ik,
//   k_struct_mean     per-structure mean.
// Padding points (k >= nsparse) have zero rows and columns in Q and so
// contribute nothing.  Without Q (none uploaded) sigma2_i = delta, as the
// reference's INVCMM = 0 gives.
#include <rocblas/rocblas.h>

#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

namespace {

constexpr double kDelta = 1.0;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wavefront per row: xn_i = |xt_i|^2 in FP64, and an FP64 copy of FP32 rows.
template <typename T>
__global__ __launch_bounds__(256) void k_row_norms(int N, int Dp, const T *__restrict__ xt, double *__restrict__ xw,
                                                   double *__restrict__ xn) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= N) return;
  const size_t r = (size_t)i * Dp;
  double s = 0.0;
  for (int d = lane; d < Dp; d += 64) {
    const double v = (double)xt[r + d];
    if (xw) xw[r + d] = v;
    s += v * v;
  }
  s = wave_sum(s);
  if (lane == 0) xn[i] = s;
}

__global__ void k_kernel_from_dot(int rows, int Mp, const double *__restrict__ xn, const double *__restrict__ mn,
                                  double *__restrict__ S) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= (size_t)rows * Mp) return;
  const int i = (int)(o / Mp), k = (int)(o - (size_t)i * Mp);
  S[o] = kDelta * exp(-0.5 * (xn[i] + mn[k] - 2.0 * S[o]));
}

__global__ __launch_bounds__(256) void k_row_quadratic(int rows, int Mp, const double *__restrict__ U,
                                                       const double *__restrict__ K, double *__restrict__ sigma2) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= rows) return;
  const size_t r = (size_t)i * Mp;
  double s = 0.0;
  for (int k = lane; k < Mp; k += 64) s += U[r + k] * K[r + k];
  s = wave_sum(s);
  if (lane == 0) sigma2[i] = kDelta - s;
}

__global__ void k_fill_delta(int N, double *__restrict__ sigma2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) sigma2[i] = kDelta;
}

__global__ __launch_bounds__(256) void k_struct_mean(int nstruct, const StructParams *__restrict__ sp,
                                                     const double *__restrict__ sigma2, double *__restrict__ var) {
  const int s = blockIdx.x, t = threadIdx.x;
  if (s >= nstruct) return;
  const int base = sp[s].atom_base, n = sp[s].natoms;
  double v = 0.0;
  for (int a = t; a < n; a += 256) v += sigma2[base + a];
  v = wave_sum(v);
  __shared__ double part[4];
  if ((t & 63) == 0) part[t >> 6] = v;
  __syncthreads();
  if (t == 0) var[s] = n > 0 ? ((part[0] + part[1]) + (part[2] + part[3])) / n : 0.0;
}

int blas_check(rocblas_status st, const char *what) {
  if (st == rocblas_status_success) return GAPX_OK;
  return fail(GAPX_ERR_HIP, std::string("rocBLAS ") + what + " failed: " + rocblas_status_to_string(st));
}

}  // namespace

int variance_create_handle(void **handle, hipStream_t s) {
  rocblas_handle h = nullptr;
  int rc;
  if ((rc = blas_check(rocblas_create_handle(&h), "create_handle"))) return rc;
  if ((rc = blas_check(rocblas_set_stream(h, s), "set_stream"))) {
    rocblas_destroy_handle(h);
    return rc;
  }
  *handle = h;
  return GAPX_OK;
}

void variance_destroy_handle(void *handle) {
  if (handle) rocblas_destroy_handle((rocblas_handle)handle);
}

int enqueue_variance(void *handle, hipStream_t s, const VarianceArgs &v) {
  if (v.N <= 0 || v.Dp <= 0 || v.Mp <= 0 || v.rows <= 0 || (!v.xt64 && !v.xt32))
    return fail(GAPX_ERR_ARG, "variance: invalid workspace");
  if (!v.qmm) {
    k_fill_delta<<<(v.N + 255) / 256, 256, 0, s>>>(v.N, v.sigma2);
    GAPX_HIP(hipGetLastError());
  } else {
    const rocblas_handle h = (rocblas_handle)handle;
    const double one = 1.0, zero = 0.0;
    const double *xt = v.xt64;
    if (v.xt64)
      k_row_norms<double><<<(v.N + 3) / 4, 256, 0, s>>>(v.N, v.Dp, v.xt64, nullptr, v.xn);
    else
      k_row_norms<float><<<(v.N + 3) / 4, 256, 0, s>>>(v.N, v.Dp, v.xt32, v.xw, v.xn);
    GAPX_HIP(hipGetLastError());
    if (!xt) xt = v.xw;
    int rc;
    for (int r0 = 0; r0 < v.N; r0 += v.rows) {
      const int rows = std::min(v.rows, v.N - r0);
      // row-major S [rows][Mp] = xt mt^T, i.e. column-major S^T = mt xt^T
      if ((rc = blas_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, v.Mp, rows, v.Dp,
                                         &one, v.mt, v.Dp, xt + (size_t)r0 * v.Dp, v.Dp, &zero, v.K, v.Mp),
                           "dgemm (xt mt^T)")))
        return rc;
      const size_t nk = (size_t)rows * v.Mp;
      k_kernel_from_dot<<<(unsigned)((nk + 255) / 256), 256, 0, s>>>(rows, v.Mp, v.xn + r0, v.mn, v.K);
      GAPX_HIP(hipGetLastError());
      // row-major U [rows][Mp] = K Q, i.e. column-major U^T = Q^T K^T (Q row-major)
      if ((rc = blas_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, v.Mp, rows, v.Mp, &one,
                                         v.qmm, v.Mp, v.K, v.Mp, &zero, v.U, v.Mp),
                           "dgemm (K Q)")))
        return rc;
      k_row_quadratic<<<(rows + 3) / 4, 256, 0, s>>>(rows, v.Mp, v.U, v.K, v.sigma2 + r0);
      GAPX_HIP(hipGetLastError());
    }
  }
  k_struct_mean<<<v.nstruct, 256, 0, s>>>(v.nstruct, v.sp, v.sigma2, v.var);
  GAPX_HIP(hipGetLastError());
  return GAPX_OK;
}

}  // namespace gapx
