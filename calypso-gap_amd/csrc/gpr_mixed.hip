// Mixed-precision GPR contraction: FP16 (or BF16) hi + lo split operands on
// v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation, FP64 norms and FP64
// energy partial sums.
//
// Same mathematics as k_gpr_f64 (gpr.hip): for a row tile against a range of
// sparse points,
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
//
// The kernel computes S^T = mt xt^T (A = sparse points from LDS, B = the
// wave's descriptor rows, held in registers for the whole kernel).  Lane
// (l16, lq) then holds S^T for row l16 and points 4 lq + r of each 16-point
// block, which is, for a 32-point step, exactly the B operand of
// G^T = mt^T W^T when the K index runs over the points in the order
// {4 lq .. 4 lq + 3, 16 + 4 lq .. 16 + 4 lq + 3} for lane group lq.  The
// feature-major copy of mt is stored in that order (api.hip), so W goes from
// the exp straight into MFMA operands without touching LDS.  Each wave owns R
// 16-row blocks, so every LDS fragment feeds 3 R MFMAs.
#include "gapx_internal.h"

namespace gapx {

namespace {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));  // 16 bytes of halves

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

// One chunk of PC sparse points goes global -> registers -> LDS in two
// batches, each with all its loads in flight together: the point-major tile
// (PC rows of DP halves), then the feature-major tile (DP rows of PC halves,
// the permuted K order), both split halves, 16 bytes per piece and NPT pieces
// per thread (indices past the tile are clamped for the load and skipped for
// the store).  Two batches rather than one keep the staging registers within
// the budget of two waves per SIMD.
template <int DP, int PC, int SM, int ST>
__device__ __forceinline__ void load_chunk(const GprArgs &a, int p0, int tid, uint16_t *s_mh, uint16_t *s_ml,
                                           uint16_t *s_th, uint16_t *s_tl) {
  constexpr int NE = PC * DP / 8;
  constexpr int NPT = (NE + 255) / 256;
  v4u h[NPT], l[NPT];
#pragma unroll
  for (int u = 0; u < NPT; ++u) {
    const int e = min(tid + 256 * u, NE - 1);
    const int pc = e / (DP / 8), q = e - pc * (DP / 8);
    const size_t g = (size_t)(p0 + pc) * DP + 8 * q;
    h[u] = *(const v4u *)(a.mt_hi + g);
    l[u] = *(const v4u *)(a.mt_lo + g);
  }
#pragma unroll
  for (int u = 0; u < NPT; ++u) {
    const int e = tid + 256 * u;
    if (NE % 256 == 0 || e < NE) {
      const int pc = e / (DP / 8), q = e - pc * (DP / 8);
      *(v4u *)(s_mh + pc * SM + 8 * q) = h[u];
      *(v4u *)(s_ml + pc * SM + 8 * q) = l[u];
    }
  }
#pragma unroll
  for (int u = 0; u < NPT; ++u) {
    const int e = min(tid + 256 * u, NE - 1);
    const int d = e / (PC / 8), r = e - d * (PC / 8);
    const size_t g = (size_t)d * a.Mp + p0 + 8 * r;
    h[u] = *(const v4u *)(a.mtT_hi + g);
    l[u] = *(const v4u *)(a.mtT_lo + g);
  }
#pragma unroll
  for (int u = 0; u < NPT; ++u) {
    const int e = tid + 256 * u;
    if (NE % 256 == 0 || e < NE) {
      const int d = e / (PC / 8), r = e - d * (PC / 8);
      *(v4u *)(s_th + d * ST + 8 * r) = h[u];
      *(v4u *)(s_tl + d * ST + 8 * r) = l[u];
    }
  }
}

template <bool BF, int DP, int PC, int R>
__global__ __launch_bounds__(256, 2) void k_gpr_mixed(GprArgs a) {
  typedef typename Half<BF>::vec vec;
  typedef typename Half<BF>::scalar hs;
  constexpr int SM = DP + 8;  // row stride (halves) of the point-major tile
  constexpr int ST = PC + 8;  // row stride (halves) of the feature-major tile
  __shared__ __align__(16) uint16_t s_mh[PC * SM], s_ml[PC * SM];
  __shared__ __align__(16) uint16_t s_th[DP * ST], s_tl[DP * ST];
  __shared__ double s_mn[PC];
  __shared__ float s_c[PC];

  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  const int row0 = (a.tile0 + blockIdx.x) * (64 * R) + wv * (16 * R);
  const int split = blockIdx.y;
  const int pb = split * a.pts_per_split;
  const int pe = min(a.Mp, pb + a.pts_per_split);

  // B operands of S^T: row l16 of each row block, 8 features per lane per
  // 32-wide step
  vec xh[R][DP / 32], xl[R][DP / 32];
  double xnr[R];
#pragma unroll
  for (int rb = 0; rb < R; ++rb) {
    const size_t row = (size_t)(row0 + 16 * rb + l16);
#pragma unroll
    for (int t = 0; t < DP / 32; ++t) {
      const size_t off = row * DP + 32 * t + 8 * lq;
      xh[rb][t] = *(const vec *)(a.xt_hi + off);
      xl[rb][t] = *(const vec *)(a.xt_lo + off);
    }
    xnr[rb] = a.xn[row];
  }

  v4f G[R][DP / 16];  // G^T: features 16 b + 4 lq + r of row l16
#pragma unroll
  for (int rb = 0; rb < R; ++rb)
#pragma unroll
    for (int b = 0; b < DP / 16; ++b) G[rb][b] = v4f{0.f, 0.f, 0.f, 0.f};
  double er[R];
#pragma unroll
  for (int rb = 0; rb < R; ++rb) er[rb] = 0.0;

  for (int p0 = pb; p0 < pe; p0 += PC) {
    __syncthreads();
    load_chunk<DP, PC, SM, ST>(a, p0, tid, s_mh, s_ml, s_th, s_tl);
    if (tid < PC) {
      s_mn[tid] = a.mn[p0 + tid];
      s_c[tid] = (float)a.coeff[p0 + tid];
    }
    __syncthreads();

#pragma unroll
    for (int sc = 0; sc < PC / 32; ++sc) {
      // S^T for two 16-point blocks x R row blocks
      v4f S[2][R];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int rb = 0; rb < R; ++rb) S[h][rb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t < DP / 32; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int pt = 32 * sc + 16 * h + l16;
          const vec mh = *(const vec *)(s_mh + pt * SM + 32 * t + 8 * lq);
          const vec ml = *(const vec *)(s_ml + pt * SM + 32 * t + 8 * lq);
#pragma unroll
          for (int rb = 0; rb < R; ++rb) {
            S[h][rb] = Half<BF>::mfma(mh, xh[rb][t], S[h][rb]);
            S[h][rb] = Half<BF>::mfma(ml, xh[rb][t], S[h][rb]);
            S[h][rb] = Half<BF>::mfma(mh, xl[rb][t], S[h][rb]);
          }
        }
      // W in registers, already in the B layout of the G step
      vec wh[R], wl[R];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int pl = 32 * sc + 16 * h + 4 * lq + r;
          const double mnp = s_mn[pl];
          const float cp = s_c[pl];
#pragma unroll
          for (int rb = 0; rb < R; ++rb) {
            const float d2 = (float)(xnr[rb] + mnp - 2.0 * (double)S[h][rb][r]);
            const float w = __expf(-0.5f * d2) * cp;
            er[rb] += (double)w;
            const float ws = w * a.wscale;
            const hs hi = (hs)ws;
            wh[rb][4 * h + r] = hi;
            wl[rb][4 * h + r] = (hs)(ws - (float)hi);
          }
        }
      if (a.grad) {
#pragma unroll
        for (int b = 0; b < DP / 16; ++b) {
          const int d = 16 * b + l16;
          const vec th = *(const vec *)(s_th + d * ST + 32 * sc + 8 * lq);
          const vec tl = *(const vec *)(s_tl + d * ST + 32 * sc + 8 * lq);
#pragma unroll
          for (int rb = 0; rb < R; ++rb) {
            G[rb][b] = Half<BF>::mfma(th, wh[rb], G[rb][b]);
            G[rb][b] = Half<BF>::mfma(tl, wh[rb], G[rb][b]);
            G[rb][b] = Half<BF>::mfma(th, wl[rb], G[rb][b]);
          }
        }
      }
    }
  }

  const size_t base = (size_t)split * a.Np;
#pragma unroll
  for (int rb = 0; rb < R; ++rb) {
    const size_t row = (size_t)(row0 + 16 * rb + l16);
    if (a.grad) {
      const float inv = 1.0f / a.wscale;  // a power of two: exact
#pragma unroll
      for (int b = 0; b < DP / 16; ++b) {
        const float4 v = make_float4(G[rb][b][0] * inv, G[rb][b][1] * inv, G[rb][b][2] * inv, G[rb][b][3] * inv);
        *(float4 *)(a.Gpart32 + (base + row) * DP + 16 * b + 4 * lq) = v;
      }
    }
    double v = er[rb];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (lq == 0) a.Epart[base + row] = v;
  }
}

template <bool BF, int DP, int PC>
hipError_t launch_rows(hipStream_t s, const GprArgs &a) {
  if constexpr (DP <= 128) {
    if (a.rows == 128) {
      k_gpr_mixed<BF, DP, PC, 2><<<dim3(a.ntiles > 0 ? a.ntiles : a.Np / 128, a.splits), 256, 0, s>>>(a);
      return hipGetLastError();
    }
  }
  k_gpr_mixed<BF, DP, PC, 1><<<dim3(a.ntiles > 0 ? a.ntiles : a.Np / 64, a.splits), 256, 0, s>>>(a);
  return hipGetLastError();
}

template <bool BF>
hipError_t launch_typed(hipStream_t s, const GprArgs &a) {
  switch (a.Dp) {
    case 32: return launch_rows<BF, 32, 64>(s, a);
    case 64: return launch_rows<BF, 64, 64>(s, a);
    case 96: return launch_rows<BF, 96, 64>(s, a);
    case 128: return launch_rows<BF, 128, 64>(s, a);
    case 160: return launch_rows<BF, 160, 32>(s, a);
    case 192: return launch_rows<BF, 192, 32>(s, a);
    case 224: return launch_rows<BF, 224, 32>(s, a);
    case 256: return launch_rows<BF, 256, 32>(s, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int gpr_points_per_chunk(int Dp) { return Dp <= 128 ? 64 : 32; }

hipError_t launch_gpr_mixed(hipStream_t s, const GprArgs &a) {
  if (a.pts_per_split % gpr_points_per_chunk(a.Dp) != 0 || (a.rows != 64 && a.rows != 128) ||
      a.Np % a.rows != 0)
    return hipErrorInvalidValue;
  return a.bf16 ? launch_typed<true>(s, a) : launch_typed<false>(s, a);
}

}  // namespace gapx
