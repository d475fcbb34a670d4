// GPR contraction of the descriptors against the sparse set on MFMA.
//
// Reference: GET_COV (gappy/libgap/gap_calc.f90:268-288), atomic energies
// (:152-154) and dE/dG (:157-166):
//   K_ik = exp(-1/2 sum_d ((x_id - m_kd) / theta_d)^2),   E_i = sum_k K_ik c_k,
//   g_id = -sum_k (x_id - m_kd) / theta_d^2 K_ik c_k.
// With xt = x / theta - mu and mt = m / theta - mu (mu = column mean of
// m / theta; distances do not change) the exponent is
//   d2_ik = |xt_i|^2 + |mt_k|^2 - 2 xt_i . mt_k,
// a dense N x D . D x M product, and with W = K o c
//   E_i = sum_k W_ik,   g_id = -(xt_id E_i - (W mt)_id) / theta_d,
// a second dense N x M . M x D product.  One kernel computes a 64-atom row
// tile against a range of sparse points: S = xt mt^T on MFMA, the exp and the
// alpha weights in registers, W staged through LDS, G += W mt on MFMA.  The
// N x M kernel matrix never reaches HBM.  The sparse points are split over
// grid.y to fill the GPU; partial E and G are summed by gpr_finish in a fixed
// order (deterministic).
#include "gapx_internal.h"

namespace gapx {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kPC = 32;  // sparse points per LDS chunk

// FP64: v_mfma_f64_16x16x4_f64.  Operand layout per lane l (l16 = l % 16,
// lq = l / 16): A[l16][lq], B[lq][l16]; result D[lq + 4 r][l16], r = 0..3
// (rows strided by 4, unlike the FP16/BF16 forms; measured on gfx950).
template <int DP>
__global__ __launch_bounds__(256) void k_gpr_f64(GprArgs a) {
  __shared__ double s_m[kPC][DP + 4];
  __shared__ double s_mn[kPC], s_c[kPC];
  __shared__ double s_w[4][16][kPC + 1];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  const int row0 = blockIdx.x * 64 + wv * 16;
  const int split = blockIdx.y;
  const int pb = split * a.pts_per_split;
  const int pe = min(a.Mp, pb + a.pts_per_split);

  double xa[DP / 4];
#pragma unroll
  for (int t = 0; t < DP / 4; ++t) xa[t] = a.xt64[(size_t)(row0 + l16) * DP + 4 * t + lq];
  double xnr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xnr[r] = a.xn[row0 + lq + 4 * r];

  v4d G[DP / 16];
#pragma unroll
  for (int b = 0; b < DP / 16; ++b) G[b] = v4d{0, 0, 0, 0};
  double er[4] = {0, 0, 0, 0};

  for (int p0 = pb; p0 < pe; p0 += kPC) {
    __syncthreads();
    for (int e = tid; e < kPC * DP; e += 256) {
      const int pc = e / DP, d = e - pc * DP;
      s_m[pc][d] = a.mt[(size_t)(p0 + pc) * DP + d];
    }
    if (tid < kPC) {
      s_mn[tid] = a.mn[p0 + tid];
      s_c[tid] = a.coeff[p0 + tid];
    }
    __syncthreads();
    v4d S0 = {0, 0, 0, 0}, S1 = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < DP / 4; ++t) {
      S0 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[t], s_m[l16][4 * t + lq], S0, 0, 0, 0);
      S1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[t], s_m[16 + l16][4 * t + lq], S1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = lq + 4 * r;
      const double w0 = exp(-0.5 * (xnr[r] + s_mn[l16] - 2.0 * S0[r])) * s_c[l16];
      const double w1 = exp(-0.5 * (xnr[r] + s_mn[16 + l16] - 2.0 * S1[r])) * s_c[16 + l16];
      er[r] += w0 + w1;
      s_w[wv][rr][l16] = w0;
      s_w[wv][rr][16 + l16] = w1;
    }
    __syncthreads();
    if (a.grad) {
#pragma unroll
      for (int k4 = 0; k4 < kPC / 4; ++k4) {
        const double aw = s_w[wv][l16][4 * k4 + lq];
#pragma unroll
        for (int b = 0; b < DP / 16; ++b)
          G[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(aw, s_m[4 * k4 + lq][16 * b + l16], G[b], 0, 0, 0);
      }
    }
  }

  const size_t base = (size_t)split * a.Np;
  if (a.grad) {
#pragma unroll
    for (int b = 0; b < DP / 16; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) a.Gpart[(base + row0 + lq + 4 * r) * DP + 16 * b + l16] = G[b][r];
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double v = er[r];
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    if (l16 == 0) a.Epart[base + row0 + lq + 4 * r] = v;
  }
}

// E_i = sum of split partials; g_id = -(xt_id E_i - G_id) theta_inv_d.
__global__ __launch_bounds__(256) void k_gpr_finish(GprArgs a, int fp64) {
  const int i = blockIdx.x;
  if (i >= a.N) return;
  double E = 0.0;
  for (int s = 0; s < a.splits; ++s) E += a.Epart[(size_t)s * a.Np + i];
  if (threadIdx.x == 0) a.eatom[i] = E;
  if (!a.grad) return;
  for (int d = threadIdx.x; d < a.Dp; d += blockDim.x) {
    double G = 0.0;
    for (int s = 0; s < a.splits; ++s) G += a.Gpart[((size_t)s * a.Np + i) * a.Dp + d];
    const size_t idx = (size_t)i * a.Dp + d;
    const double xt = fp64 ? a.xt64[idx] : (double)a.xt32[idx];
    const double g = -(xt * E - G) * a.theta_inv[d];
    if (fp64)
      a.g64[idx] = g;
    else
      a.g32[idx] = (float)g;
  }
}

}  // namespace

hipError_t launch_gpr_fp64(hipStream_t s, const GprArgs &a) {
  dim3 grid(a.Np / 64, a.splits);
  switch (a.Dp) {
    case 32: k_gpr_f64<32><<<grid, 256, 0, s>>>(a); break;
    case 64: k_gpr_f64<64><<<grid, 256, 0, s>>>(a); break;
    case 96: k_gpr_f64<96><<<grid, 256, 0, s>>>(a); break;
    case 128: k_gpr_f64<128><<<grid, 256, 0, s>>>(a); break;
    case 160: k_gpr_f64<160><<<grid, 256, 0, s>>>(a); break;
    case 192: k_gpr_f64<192><<<grid, 256, 0, s>>>(a); break;
    case 224: k_gpr_f64<224><<<grid, 256, 0, s>>>(a); break;
    case 256: k_gpr_f64<256><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gpr_finish(hipStream_t s, const GprArgs &a, bool fp64) {
  k_gpr_finish<<<a.N, 128, 0, s>>>(a, fp64 ? 1 : 0);
  return hipGetLastError();
}

}  // namespace gapx
