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
// of atoms at a time (FP64, [rows][Mp]), by hand-written FP64 MFMA kernels
// (v_mfma_f64_16x16x4f64):
//   k_row_norms     |xt_i|^2 (and the FP64 copy of a mixed run's features),
//   k_var_kernel    K = exp(-(|xt_i|^2 + |mt_k|^2 - 2 xt_i . mt_k) / 2) for a
//                   64-row x 64-point tile per workgroup, the points staged in
//                   LDS 32 features at a time,
//   k_var_quad      U = K Q for a 64 x 64 tile per workgroup (K and Q staged
//                   in LDS 32 points at a time), contracted with K in the
//                   epilogue: partial[column tile][i] = sum over the tile's
//                   columns c of U_ic K_ic, so U never reaches memory,
//   k_var_finish    sigma2_i = delta - sum of the partials in tile order,
//   k_struct_mean   per-structure mean.
// Padding points (k >= nsparse) have zero rows and columns in Q and so
// contribute nothing.  Without Q (none uploaded) sigma2_i = delta, as the
// reference's INVCMM = 0 gives.
//
// MFMA layout (as in gpr.hip): lane l (l16 = l % 16, lq = l / 16) holds
// A[l16][lq] and B[lq][l16] of a 16x16x4 step and the results D[lq + 4 r][l16].
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

// K[ib][k] for block rows ib = tile rows (atoms r0 + ib; rows past N read the
// last atom, their values are never used) and sparse points k of the tile.
// Wave w: 16 rows (A operand, from global, one double per lane and step),
// 64 points in four 16-point blocks (B operands from the LDS tile).
__global__ __launch_bounds__(256) void k_var_kernel(int N, int r0, int Dp, int Mp, const double *__restrict__ xt,
                                                    const double *__restrict__ xn, const double *__restrict__ mt,
                                                    const double *__restrict__ mn, double *__restrict__ K) {
  constexpr int KC = 32, SM = KC + 1;  // features per LDS chunk, odd row stride (doubles)
  __shared__ double s_m[64 * SM];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, l16 = lane & 15, lq = lane >> 4;
  const int ib0 = blockIdx.x * 64 + wv * 16;  // first block row of the wave
  const int p0 = blockIdx.y * 64;             // first point of the tile
  const int arow = min(r0 + ib0 + l16, N - 1);
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = v4d{0, 0, 0, 0};
  for (int f0 = 0; f0 < Dp; f0 += KC) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 64 * KC / 256; ++u) {  // the tile's 64 points x 32 features
      const int e = tid + 256 * u, p = e / KC, f = e - p * KC;
      s_m[p * SM + f] = mt[(size_t)(p0 + p) * Dp + f0 + f];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < KC / 4; ++t) {
      const double a = xt[(size_t)arow * Dp + f0 + 4 * t + lq];  // A[row l16][feature 4t + lq]
#pragma unroll
      for (int b = 0; b < 4; ++b)  // B[feature 4t + lq][point 16 b + l16]
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, s_m[(16 * b + l16) * SM + 4 * t + lq], acc[b], 0, 0, 0);
    }
  }
  // D[row lq + 4 r][point 16 b + l16] = S
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ib = ib0 + lq + 4 * r;
    const double xni = xn[min(r0 + ib, N - 1)];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int k = p0 + 16 * b + l16;
      K[(size_t)ib * Mp + k] = kDelta * exp(-0.5 * (xni + mn[k] - 2.0 * acc[b][r]));
    }
  }
}

// partial[ct][r0 + ib] = sum over the tile's columns c of (K Q)[ib][c] K[ib][c]
// for block rows ib and column tile ct = blockIdx.y (64 columns).  Wave w:
// 16 rows x 64 columns (four 16-column blocks), K-steps of 4 points.
__global__ __launch_bounds__(256) void k_var_quad(int N, int r0, int Mp, const double *__restrict__ K,
                                                  const double *__restrict__ Q, double *__restrict__ partial) {
  constexpr int KC = 32, SA = KC + 1, SB = 64 + 1;
  __shared__ double s_a[64 * SA];  // K tile: 64 rows x 32 points
  __shared__ double s_b[KC * SB];  // Q tile: 32 points x 64 columns
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, l16 = lane & 15, lq = lane >> 4;
  const int rb = blockIdx.x * 64;  // first block row of the tile
  const int c0 = blockIdx.y * 64;  // first column of the tile
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = v4d{0, 0, 0, 0};
  for (int k0 = 0; k0 < Mp; k0 += KC) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 64 * KC / 256; ++u) {
      const int e = tid + 256 * u;
      const int ar = e / KC, ac = e - ar * KC;  // K tile: row ar, point ac
      s_a[ar * SA + ac] = K[(size_t)(rb + ar) * Mp + k0 + ac];
      const int br = e / 64, bc = e - br * 64;  // Q tile: point br, column bc
      s_b[br * SB + bc] = Q[(size_t)(k0 + br) * Mp + c0 + bc];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < KC / 4; ++t) {
      const double a = s_a[(wv * 16 + l16) * SA + 4 * t + lq];  // A[row l16][point 4t + lq]
#pragma unroll
      for (int b = 0; b < 4; ++b)  // B[point 4t + lq][column 16 b + l16]
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, s_b[(4 * t + lq) * SB + 16 * b + l16], acc[b], 0, 0, 0);
    }
  }
  // D[row lq + 4 r][column 16 b + l16] = U; contract with K over the columns
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ib = rb + wv * 16 + lq + 4 * r;
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) s += acc[b][r] * K[(size_t)ib * Mp + c0 + 16 * b + l16];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);  // over the 16 lanes of this row
    if (l16 == 0 && r0 + ib < N) partial[(size_t)blockIdx.y * N + r0 + ib] = s;
  }
}

// sigma2_i = delta - sum over column tiles of the partials, in tile order.
__global__ __launch_bounds__(256) void k_var_finish(int N, int ntiles, const double *__restrict__ partial,
                                                    double *__restrict__ sigma2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  sigma2[i] = kDelta - split_sum(ntiles, [&](int t) { return partial[(size_t)t * N + i]; });
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

}  // namespace

int enqueue_variance(hipStream_t s, const VarianceArgs &v) {
  if (v.N <= 0 || v.Dp <= 0 || v.Mp <= 0 || v.rows <= 0 || (!v.xt64 && !v.xt32))
    return fail(GAPX_ERR_ARG, "variance: invalid workspace");
  if (v.Mp % 64 != 0 || v.Dp % 32 != 0 || v.rows % 64 != 0)
    return fail(GAPX_ERR_ARG, "variance: Mp, Dp and the row block must be multiples of 64, 32 and 64");
  if (!v.qmm) {
    k_fill_delta<<<(v.N + 255) / 256, 256, 0, s>>>(v.N, v.sigma2);
    GAPX_HIP(hipGetLastError());
  } else {
    const double *xt = v.xt64;
    if (v.xt64)
      k_row_norms<double><<<(v.N + 3) / 4, 256, 0, s>>>(v.N, v.Dp, v.xt64, nullptr, v.xn);
    else
      k_row_norms<float><<<(v.N + 3) / 4, 256, 0, s>>>(v.N, v.Dp, v.xt32, v.xw, v.xn);
    GAPX_HIP(hipGetLastError());
    if (!xt) xt = v.xw;
    const int ntiles = v.Mp / 64;
    for (int r0 = 0; r0 < v.N; r0 += v.rows) {
      const int rows = std::min(v.rows, (v.N - r0 + 63) / 64 * 64);
      const dim3 grid(rows / 64, ntiles);
      k_var_kernel<<<grid, 256, 0, s>>>(v.N, r0, v.Dp, v.Mp, xt, v.xn, v.mt, v.mn, v.K);
      GAPX_HIP(hipGetLastError());
      k_var_quad<<<grid, 256, 0, s>>>(v.N, r0, v.Mp, v.K, v.qmm, v.partial);
      GAPX_HIP(hipGetLastError());
    }
    k_var_finish<<<(v.N + 255) / 256, 256, 0, s>>>(v.N, ntiles, v.partial, v.sigma2);
    GAPX_HIP(hipGetLastError());
  }
  k_struct_mean<<<v.nstruct, 256, 0, s>>>(v.nstruct, v.sp, v.sigma2, v.var);
  GAPX_HIP(hipGetLastError());
  return GAPX_OK;
}

}  // namespace gapx
