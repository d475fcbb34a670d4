// Mixed-precision GPR contraction: FP16 (or BF16) hi + lo split operands on
// v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation, FP64 norms and FP64
// energy partial sums.
//
// Same mathematics as k_gpr_f64 (gpr.hip): for a 64-atom row tile against a
// range of sparse points,
//   S = xt mt^T,  d2 = |xt|^2 + |mt|^2 - 2 S  (norms in FP64),
//   W = exp(-d2 / 2) c,  E_i += sum_k W_ik,  G += W mt.
// Every operand x is carried as hi = round(x), lo = round(x - hi) in the
// 16-bit type and each product as hi*hi + hi*lo + lo*hi (the lo*lo term is
// below FP32 rounding).  W is scaled by a power of two chosen from the
// largest |alpha| (GprArgs::wscale) before its split, so that hi cannot
// overflow and lo stays in the normal FP16 range; G is unscaled on output.
//
// Operand layout of the 16x16x32 forms per lane l (l16 = l % 16, lq = l / 16):
// A[l16][8 lq + e], B[8 lq + e][l16] (e = 0..7, one 16-byte register group),
// result D[4 lq + r][l16], r = 0..3.
#include "gapx_internal.h"

namespace gapx {

namespace {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <bool BF>
struct Half;

template <>
struct Half<false> {
  typedef v8h vec;
  typedef _Float16 scalar;
  static __device__ __forceinline__ v4f mfma(vec a, vec b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Half<true> {
  typedef v8b vec;
  typedef __bf16 scalar;
  static __device__ __forceinline__ v4f mfma(vec a, vec b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <bool BF, int DP, int PC>
__global__ __launch_bounds__(256) void k_gpr_mixed(GprArgs a) {
  typedef typename Half<BF>::vec vec;
  typedef typename Half<BF>::scalar hs;
  constexpr int SM = DP + 8;   // row stride (halves) of the point-major tile
  constexpr int ST = PC + 8;   // row stride (halves) of the feature-major tile and of W
  __shared__ __align__(16) uint16_t s_mh[PC * SM], s_ml[PC * SM];
  __shared__ __align__(16) uint16_t s_th[DP * ST], s_tl[DP * ST];
  __shared__ __align__(16) uint16_t s_wh[4][16 * ST], s_wl[4][16 * ST];
  __shared__ double s_mn[PC];
  __shared__ float s_c[PC];

  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  const int row0 = blockIdx.x * 64 + wv * 16;
  const int split = blockIdx.y;
  const int pb = split * a.pts_per_split;
  const int pe = min(a.Mp, pb + a.pts_per_split);

  // A operands (this wave's 16 rows), 8 features per lane per 32-wide step
  vec xh[DP / 32], xl[DP / 32];
#pragma unroll
  for (int t = 0; t < DP / 32; ++t) {
    const size_t off = (size_t)(row0 + l16) * DP + 32 * t + 8 * lq;
    xh[t] = *(const vec *)(a.xt_hi + off);
    xl[t] = *(const vec *)(a.xt_lo + off);
  }
  double xnr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xnr[r] = a.xn[row0 + 4 * lq + r];

  v4f G[DP / 16];
#pragma unroll
  for (int b = 0; b < DP / 16; ++b) G[b] = v4f{0.f, 0.f, 0.f, 0.f};
  double er[4] = {0, 0, 0, 0};

  for (int p0 = pb; p0 < pe; p0 += PC) {
    __syncthreads();
    // point-major tile: PC rows of DP halves (16-byte pieces)
    for (int e = tid; e < PC * DP / 8; e += 256) {
      const int pc = e / (DP / 8), q = e - pc * (DP / 8);
      const size_t g = (size_t)(p0 + pc) * DP + 8 * q;
      *(uint4 *)(s_mh + pc * SM + 8 * q) = *(const uint4 *)(a.mt_hi + g);
      *(uint4 *)(s_ml + pc * SM + 8 * q) = *(const uint4 *)(a.mt_lo + g);
    }
    // feature-major tile: DP rows of PC halves
    for (int e = tid; e < DP * PC / 8; e += 256) {
      const int d = e / (PC / 8), q = e - d * (PC / 8);
      const size_t g = (size_t)d * a.Mp + p0 + 8 * q;
      *(uint4 *)(s_th + d * ST + 8 * q) = *(const uint4 *)(a.mtT_hi + g);
      *(uint4 *)(s_tl + d * ST + 8 * q) = *(const uint4 *)(a.mtT_lo + g);
    }
    if (tid < PC) {
      s_mn[tid] = a.mn[p0 + tid];
      s_c[tid] = (float)a.coeff[p0 + tid];
    }
    __syncthreads();

    // S = xt mt^T for PC/16 column blocks, then W into LDS (hi/lo)
#pragma unroll
    for (int cb = 0; cb < PC / 16; ++cb) {
      v4f S = {0.f, 0.f, 0.f, 0.f};
      const int pt = 16 * cb + l16;
#pragma unroll
      for (int t = 0; t < DP / 32; ++t) {
        const vec mh = *(const vec *)(s_mh + pt * SM + 32 * t + 8 * lq);
        const vec ml = *(const vec *)(s_ml + pt * SM + 32 * t + 8 * lq);
        S = Half<BF>::mfma(xh[t], mh, S);
        S = Half<BF>::mfma(xh[t], ml, S);
        S = Half<BF>::mfma(xl[t], mh, S);
      }
      const double mnp = s_mn[pt];
      const float cp = s_c[pt];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float d2 = (float)(xnr[r] + mnp - 2.0 * (double)S[r]);
        const float w = __expf(-0.5f * d2) * cp;
        er[r] += (double)w;
        const float ws = w * a.wscale;
        const hs hi = (hs)ws;
        const hs lo = (hs)(ws - (float)hi);
        const int o = (4 * lq + r) * ST + pt;
        s_wh[wv][o] = __builtin_bit_cast(uint16_t, hi);
        s_wl[wv][o] = __builtin_bit_cast(uint16_t, lo);
      }
    }
    __syncthreads();
    if (a.grad) {
#pragma unroll
      for (int s = 0; s < PC / 32; ++s) {
        const vec wh = *(const vec *)(&s_wh[wv][l16 * ST + 32 * s + 8 * lq]);
        const vec wl = *(const vec *)(&s_wl[wv][l16 * ST + 32 * s + 8 * lq]);
#pragma unroll
        for (int b = 0; b < DP / 16; ++b) {
          const int d = 16 * b + l16;
          const vec th = *(const vec *)(s_th + d * ST + 32 * s + 8 * lq);
          const vec tl = *(const vec *)(s_tl + d * ST + 32 * s + 8 * lq);
          G[b] = Half<BF>::mfma(wh, th, G[b]);
          G[b] = Half<BF>::mfma(wh, tl, G[b]);
          G[b] = Half<BF>::mfma(wl, th, G[b]);
        }
      }
    }
  }

  const size_t base = (size_t)split * a.Np;
  if (a.grad) {
    const double inv = 1.0 / (double)a.wscale;
#pragma unroll
    for (int b = 0; b < DP / 16; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        a.Gpart[(base + row0 + 4 * lq + r) * DP + 16 * b + l16] = (double)G[b][r] * inv;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double v = er[r];
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    if (l16 == 0) a.Epart[base + row0 + 4 * lq + r] = v;
  }
}

template <bool BF>
hipError_t launch_typed(hipStream_t s, const GprArgs &a) {
  dim3 grid(a.Np / 64, a.splits);
  switch (a.Dp) {
    case 32: k_gpr_mixed<BF, 32, 64><<<grid, 256, 0, s>>>(a); break;
    case 64: k_gpr_mixed<BF, 64, 64><<<grid, 256, 0, s>>>(a); break;
    case 96: k_gpr_mixed<BF, 96, 64><<<grid, 256, 0, s>>>(a); break;
    case 128: k_gpr_mixed<BF, 128, 64><<<grid, 256, 0, s>>>(a); break;
    case 160: k_gpr_mixed<BF, 160, 32><<<grid, 256, 0, s>>>(a); break;
    case 192: k_gpr_mixed<BF, 192, 32><<<grid, 256, 0, s>>>(a); break;
    case 224: k_gpr_mixed<BF, 224, 32><<<grid, 256, 0, s>>>(a); break;
    case 256: k_gpr_mixed<BF, 256, 32><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

// Points per LDS chunk of the mixed kernel for a padded feature count.
int gpr_mixed_chunk(int Dp) { return Dp <= 128 ? 64 : 32; }

hipError_t launch_gpr_mixed(hipStream_t s, const GprArgs &a) {
  if (a.pts_per_split % gpr_mixed_chunk(a.Dp) != 0) return hipErrorInvalidValue;
  return a.bf16 ? launch_typed<true>(s, a) : launch_typed<false>(s, a);
}

}  // namespace gapx
