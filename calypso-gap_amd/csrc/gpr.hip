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
// tile against a range of sparse points: S^T = mt xt^T on MFMA (sparse points
// from LDS as the A operand, the wave's 16 descriptor rows in registers as B),
// the exp and the alpha weights in registers, then G^T += mt^T W^T on MFMA
// with W taken straight from the S^T accumulators: in the f64 16x16x4 layout
// the result rows lq + 4 r of a 16-point block are the K index lq of the
// r-th 4-point step, so no transpose through LDS is needed.  The N x M
// kernel matrix never reaches HBM.  The sparse points are split over grid.y
// to fill the GPU; partial E and G are summed by gpr_finish in a fixed order
// (deterministic).
#include "gapx_internal.h"

namespace gapx {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// sparse points per LDS chunk: 64 up to D = 128 (fewer barriers and chunk
// loads per MFMA), 32 above (LDS)
constexpr int pc_f64(int DP) { return DP <= 128 ? 64 : 32; }

// FP64: v_mfma_f64_16x16x4_f64.  Operand layout per lane l (l16 = l % 16,
// lq = l / 16): A[l16][lq], B[lq][l16]; result D[lq + 4 r][l16], r = 0..3
// (rows strided by 4, unlike the FP16/BF16 forms; measured on gfx950).
template <int DP>
__global__ __launch_bounds__(256, DP <= 160 ? 2 : 1) void k_gpr_f64(GprArgs a) {
  constexpr int kPC = pc_f64(DP);
  constexpr int SM = DP + 1;  // odd row stride (doubles): conflict-free column reads
  __shared__ double s_m[kPC * SM];
  __shared__ double s_mn[kPC], s_c[kPC];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  int tile, split;
  xcd_tile_split(tile, split);
  const int row0 = (a.tile0 + tile) * 64 + wv * 16;
  const int pb = split * a.pts_per_split;
  const int pe = min(a.Mp, pb + a.pts_per_split);

  // B operand of S^T: B[k = lq][row l16] = xt[row][4 t + lq]
  double xa[DP / 4];
#pragma unroll
  for (int t = 0; t < DP / 4; ++t) xa[t] = a.xt64[(size_t)(row0 + l16) * DP + 4 * t + lq];
  const double xnr = a.xn[row0 + l16];

  v4d G[DP / 16];  // G^T: features 16 b + lq + 4 r of row l16
#pragma unroll
  for (int b = 0; b < DP / 16; ++b) G[b] = v4d{0, 0, 0, 0};
  double er = 0.0;

  // 16-byte pieces of the chunk per thread, loaded KB at a time with the
  // loads of a batch in flight together (KB bounded by the register budget)
  constexpr int NPT = kPC * DP / 512, KB = 4;
  static_assert(kPC * DP % 512 == 0, "whole pieces per thread");
  for (int p0 = pb; p0 < pe; p0 += kPC) {
    __syncthreads();
#pragma unroll
    for (int b0 = 0; b0 < NPT; b0 += KB) {
      double2 v[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (b0 + u < NPT) {
          const int e = tid + 256 * (b0 + u), pc = e / (DP / 2), d2 = e - pc * (DP / 2);
          v[u] = *(const double2 *)(a.mt + (size_t)(p0 + pc) * DP + 2 * d2);
        }
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (b0 + u < NPT) {
          const int e = tid + 256 * (b0 + u), pc = e / (DP / 2), d2 = e - pc * (DP / 2);
          s_m[pc * SM + 2 * d2] = v[u].x;
          s_m[pc * SM + 2 * d2 + 1] = v[u].y;
        }
    }
    if (tid < kPC) {
      s_mn[tid] = a.mn[p0 + tid];
      s_c[tid] = a.coeff[p0 + tid];
    }
    __syncthreads();
    // the chunk in 32-point halves (registers as for one 32-point chunk)
#pragma unroll 1
    for (int q0 = 0; q0 < kPC; q0 += 32) {
      // S^T[p = lq + 4 r][row l16] for both 16-point blocks: four independent
      // accumulator chains (even and odd feature steps) keep the MFMA pipe full
      v4d Sa[2], Sb[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) Sa[h] = Sb[h] = v4d{0, 0, 0, 0};
#pragma unroll
      for (int t = 0; t < DP / 4; t += 2)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double *mp = s_m + (q0 + 16 * h + l16) * SM + 4 * t + lq;
          Sa[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(mp[0], xa[t], Sa[h], 0, 0, 0);
          Sb[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(mp[4], xa[t + 1], Sb[h], 0, 0, 0);
        }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double w[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int p = q0 + 16 * h + lq + 4 * r;
          w[r] = exp(-0.5 * (xnr + s_mn[p] - 2.0 * (Sa[h][r] + Sb[h][r]))) * s_c[p];
          er += w[r];
        }
        if (a.grad) {
          // K step r: points q0 + 16 h + 4 r + lq; W is lane (row, lq)'s w[r]
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int b = 0; b < DP / 16; ++b)
              G[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(s_m[(q0 + 16 * h + 4 * r + lq) * SM + 16 * b + l16], w[r],
                                                          G[b], 0, 0, 0);
        }
      }
    }
  }

  const size_t base = (size_t)split * a.Np;
  const size_t row = (size_t)(row0 + l16);
  if (a.grad) {
#pragma unroll
    for (int b = 0; b < DP / 16; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) a.Gpart[(base + row) * DP + 16 * b + lq + 4 * r] = G[b][r];
  }
  er += __shfl_xor(er, 16, 64);
  er += __shfl_xor(er, 32, 64);
  if (lq == 0) a.Epart[base + row] = er;
}

// E_i = sum of split partials; g_id = -(xt_id E_i - G_id) theta_inv_d.
__global__ __launch_bounds__(256) void k_gpr_finish(GprArgs a, int fp64) {
  const int i = blockIdx.x;
  if (i >= a.N) return;
  const double E = split_sum(a.splits, [&](int s) { return a.Epart[(size_t)s * a.Np + i]; });
  if (threadIdx.x == 0) a.eatom[i] = E;
  if (!a.grad) return;
  for (int d = threadIdx.x; d < a.Dp; d += blockDim.x) {
    const double G =
        fp64 ? split_sum(a.splits, [&](int s) { return a.Gpart[((size_t)s * a.Np + i) * a.Dp + d]; })
             : split_sum(a.splits, [&](int s) { return (double)a.Gpart32[((size_t)s * a.Np + i) * a.Dp + d]; });
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
  if (a.pts_per_split % pc_f64(a.Dp) != 0 || a.Np % 64 != 0) return hipErrorInvalidValue;
  dim3 grid(a.ntiles > 0 ? a.ntiles : a.Np / 64, a.splits);
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
