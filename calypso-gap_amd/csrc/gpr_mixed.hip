// Mixed-precision GPR contraction: FP16 (or BF16) hi + lo split operands on
// v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation, FP64 norms and FP64
// energy partial sums.
//
// Same mathematics as k_gpr_f64 (gpr.hip): for a row tile against a range of
// sparse points,
//   S = xt mt^T,  d2 = |xt|^2 + |mt|^2 - 2 S  (norms in FP64),
//   W = exp(-d2 / 2) c,  E_i += sum_k W_ik,  G += W mt,
// with the exponential taken as exp2 of log2(e) S - (|xt|^2 + |mt|^2) log2(e) / 2
// (the per-point term and c times the W scale tabulated at model load).
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
//
// Sparse points arrive in chunks of 32, double-buffered in LDS: while the
// waves work on one chunk, the next one travels global -> LDS by direct
// loads (global_load_lds_dwordx4: no registers, counted by vmcnt), so the
// matrix cores do not wait for the chunk loads.  A chunk buffer holds the
// point-major hi/lo tiles (32 rows of DP halves), the feature-major hi/lo
// tiles (DP rows of 32 halves) and, per wave, the chunk's norm terms and
// scaled alphas.
// A direct load writes 64 consecutive 16-byte pieces per instruction, so the
// tiles have no row padding; instead piece q of row r sits at q ^ swz(r),
// which spreads the 16 rows a 16-lane group reads over all LDS banks.
#include "gapx_internal.h"

#include <type_traits>

namespace gapx {

namespace {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));  // 16 bytes of halves
typedef __attribute__((address_space(3))) void lds_void;

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

constexpr int kPC = 32;  // sparse points per chunk

// Piece swizzle of a tile row of P 16-byte pieces: P >= 16 (a row covers
// all banks): XOR with row % 16; P = 8 or 4 (two or four rows per 256 bytes
// of banks): XOR with the row's index among the rows that share bank
// positions.  Rows of other widths are left in order.
template <int P>
__device__ __forceinline__ int swz(int row) {
  if constexpr (P >= 16 && (P & (P - 1)) == 0) return row & 15;
  else if constexpr (P == 8) return (row >> 1) & 7;
  else if constexpr (P == 4) return (row >> 2) & 3;
  else return 0;
}

// One chunk buffer and how its operand pieces are spread over the waves:
// piece s of [mh | ml | th | tl] goes with DMA instruction s / (64 NW) of
// wave (s / 64) % NW, lane s % 64; instructions past the operands fill an
// unused tail (they reload a valid piece).
template <int DP, int NW>
struct Chunk {
  static constexpr int PM = DP / 8;         // pieces per point-major row
  static constexpr int PF = kPC / 8;        // pieces per feature-major row
  static constexpr int NE = kPC * PM;       // pieces per operand tile
  static constexpr int NT = 4 * NE;
  static constexpr int PER = 64 * NW;       // pieces per instruction over the workgroup
  static constexpr int NI = (NT + PER - 1) / PER;
  static constexpr int seg(int s) { return s < NT ? s / NE : 4; }
  static constexpr bool one_seg(int u) { return seg(u * PER) == seg(u * PER + PER - 1); }
  v4u op[NI * PER];
  double ex[NW][128];  // per wave: the chunk's norm terms (kPC doubles), scaled alphas (kPC floats), unused
};

// Offset (halves) of operand piece s from its tile's chunk origin, with the
// tile index in bits 29..31 (4: the unused tail).
template <int DP, int NW>
__device__ __forceinline__ unsigned piece_src(int s, int Mp) {
  typedef Chunk<DP, NW> C;
  const int g = C::seg(s);
  if (g == 4) return 4u << 29;
  const int e = s - g * C::NE;
  unsigned off;
  if (g < 2) {
    const int row = e / C::PM, q = (e % C::PM) ^ swz<C::PM>(row);
    off = (unsigned)(row * DP + 8 * q);
  } else {
    const int d = e / C::PF, r = (e % C::PF) ^ swz<C::PF>(d);
    off = (unsigned)(d * Mp + 8 * r);
  }
  return off | ((unsigned)g << 29);
}

template <int U, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (U < N) {
    f(std::integral_constant<int, U>{});
    static_for<U + 1, N>(f);
  }
}

// vmcnt(v) with the other counters left alone (gfx9 encoding)
constexpr int vm_wait(int v) { return (v & 15) | ((v >> 4) << 14) | (7 << 4) | (15 << 8); }

// This wave's direct loads except the v most recent have landed; after the
// barrier every wave's have.
template <int V>
__device__ __forceinline__ void chunk_landed() {
  __builtin_amdgcn_s_waitcnt(vm_wait(V));
  asm volatile("s_barrier" ::: "memory");
}

// Every wave has finished reading the buffer it just worked on.
__device__ __forceinline__ void chunk_consumed() {
  __builtin_amdgcn_s_waitcnt(vm_wait(63) & ~(15 << 8));  // lgkmcnt(0)
  asm volatile("s_barrier" ::: "memory");
}

template <bool BF, int DP, int R, int NW>
__global__ __launch_bounds__(64 * NW, 8 / NW) void k_gpr_mixed(GprArgs a) {
  typedef typename Half<BF>::vec vec;
  typedef typename Half<BF>::scalar hs;
  typedef Chunk<DP, NW> C;
  __shared__ __align__(16) C t0;
  __shared__ __align__(16) C t1;

  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  int tile, split;
  xcd_tile_split(tile, split);
  const int row0 = (a.tile0 + tile) * (16 * R * NW) + wv * (16 * R);
  const int pb = split * a.pts_per_split;
  const int pe = min(a.Mp, pb + a.pts_per_split);
  const int nch = (pe - pb) / kPC;

  // B operands of S^T: row l16 of each row block, 8 features per lane per
  // 32-wide step
  vec xh[R][DP / 32], xl[R][DP / 32];
  double xq[R];  // -|xt|^2 log2(e) / 2 of the row
#pragma unroll
  for (int rb = 0; rb < R; ++rb) {
    const size_t row = (size_t)(row0 + 16 * rb + l16);
#pragma unroll
    for (int t = 0; t < DP / 32; ++t) {
      const size_t off = row * DP + 32 * t + 8 * lq;
      xh[rb][t] = *(const vec *)(a.xt_hi + off);
      xl[rb][t] = *(const vec *)(a.xt_lo + off);
    }
    xq[rb] = -0.5 * kLog2e * a.xn[row];
  }

  // this lane's operand pieces, one per DMA instruction
  unsigned src[C::NI];
#pragma unroll
  for (int u = 0; u < C::NI; ++u) src[u] = piece_src<DP, NW>((u * NW + wv) * 64 + lane, a.Mp);

  // the points' norm terms (FP64, two per lane) and scaled alphas (FP32, four
  // per lane); lanes past them reload a norm term into the unused part (the
  // loads are issued under a full lane mask)
  const char *exsrc = lane < kPC / 2       ? (const char *)(a.mnq + 2 * lane)
                      : lane < 3 * kPC / 4 ? (const char *)(a.cq + 4 * (lane - kPC / 2))
                                           : (const char *)a.mnq;
  const int exstride = (lane >= kPC / 2 && lane < 3 * kPC / 4) ? 4 : 8;  // bytes per point

  // chunk at p0 -> buffer B: NI + 1 direct loads per wave
  auto issue = [&](C &B, int p0) {
    const uint16_t *base[4] = {a.mt_hi + (size_t)p0 * DP, a.mt_lo + (size_t)p0 * DP, a.mtT_hi + p0, a.mtT_lo + p0};
    static_for<0, C::NI>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const uint16_t *p;
      if constexpr (C::one_seg(u)) {
        constexpr int g = C::seg(u * C::PER);
        p = g == 4 ? base[0] : base[g] + (src[u] & ((1u << 29) - 1));
      } else {
        const unsigned g = src[u] >> 29;
        p = (g == 4 ? base[0] : (g == 0 ? base[0] : g == 1 ? base[1] : g == 2 ? base[2] : base[3]) +
                                    (src[u] & ((1u << 29) - 1)));
      }
      __builtin_amdgcn_global_load_lds((const void *)p, (lds_void *)&B.op[(u * NW + wv) * 64], 16, 0, 0);
    });
    __builtin_amdgcn_global_load_lds((const void *)(exsrc + (size_t)p0 * exstride), (lds_void *)&B.ex[wv][0], 16, 0,
                                     0);
  };

  v4f G[R][DP / 16];  // G^T: features 16 b + 4 lq + r of row l16
#pragma unroll
  for (int rb = 0; rb < R; ++rb)
#pragma unroll
    for (int b = 0; b < DP / 16; ++b) G[rb][b] = v4f{0.f, 0.f, 0.f, 0.f};
  double er[R];
#pragma unroll
  for (int rb = 0; rb < R; ++rb) er[rb] = 0.0;

  auto compute = [&](const C &B) {
    // per 16-point block: S^T for the R row blocks, then its exp; the next
    // block's products do not depend on this block's exp, so the matrix
    // cores can run them while the vector units work on the exp
    vec wh[R], wl[R];  // W in registers, already in the B layout of the G step
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v4f S[R];
#pragma unroll
      for (int rb = 0; rb < R; ++rb) S[rb] = v4f{0.f, 0.f, 0.f, 0.f};
      const int pt = 16 * h + l16;
      // operands of step t + 1 read from LDS while step t's products run
      auto op_at = [&](int t) { return pt * C::PM + ((4 * t + lq) ^ swz<C::PM>(pt)); };
      vec mh = *(const vec *)&B.op[op_at(0)];
      vec ml = *(const vec *)&B.op[C::NE + op_at(0)];
#pragma unroll
      for (int t = 0; t < DP / 32; ++t) {
        vec mh_n = mh, ml_n = ml;
        if (t + 1 < DP / 32) {
          mh_n = *(const vec *)&B.op[op_at(t + 1)];
          ml_n = *(const vec *)&B.op[C::NE + op_at(t + 1)];
        }
#pragma unroll
        for (int rb = 0; rb < R; ++rb) {
          S[rb] = Half<BF>::mfma(mh, xh[rb][t], S[rb]);
          S[rb] = Half<BF>::mfma(ml, xh[rb][t], S[rb]);
          S[rb] = Half<BF>::mfma(mh, xl[rb][t], S[rb]);
        }
        mh = mh_n;
        ml = ml_n;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int pl = 16 * h + 4 * lq + r;
        const double mq = B.ex[wv][pl];
        const float cq = ((const float *)&B.ex[wv][kPC])[pl];
#pragma unroll
        for (int rb = 0; rb < R; ++rb) {
          // -d2 log2(e) / 2 in FP64 (norms minus twice the product), then
          // ws = exp(-d2 / 2) c wscale
          const float e2 = (float)fma((double)S[rb][r], kLog2e, xq[rb] + mq);
          const float ws = __builtin_amdgcn_exp2f(e2) * cq;
          er[rb] += (double)ws;
          const hs hi = (hs)ws;
          wh[rb][4 * h + r] = hi;
          wl[rb][4 * h + r] = (hs)(ws - (float)hi);
        }
      }
    }
    if (a.grad) {
      // operands of feature block b + 1 read while block b's products run
      auto top_at = [&](int b) {
        const int d = 16 * b + l16;
        return 2 * C::NE + d * C::PF + (lq ^ swz<C::PF>(d));
      };
      vec th = *(const vec *)&B.op[top_at(0)];
      vec tl = *(const vec *)&B.op[C::NE + top_at(0)];
#pragma unroll
      for (int b = 0; b < DP / 16; ++b) {
        vec th_n = th, tl_n = tl;
        if (b + 1 < DP / 16) {
          th_n = *(const vec *)&B.op[top_at(b + 1)];
          tl_n = *(const vec *)&B.op[C::NE + top_at(b + 1)];
        }
#pragma unroll
        for (int rb = 0; rb < R; ++rb) {
          G[rb][b] = Half<BF>::mfma(th, wh[rb], G[rb][b]);
          G[rb][b] = Half<BF>::mfma(tl, wh[rb], G[rb][b]);
          G[rb][b] = Half<BF>::mfma(th, wl[rb], G[rb][b]);
        }
        th = th_n;
        tl = tl_n;
      }
    }
  };

  // Two chunks per trip, one per buffer, so that each buffer is named
  // statically (the compiler then waits for the loads into the buffer being
  // read, not for the prefetch into the other one).  Every step prefetches,
  // so that the wait is always "all but the newest NI + 1 loads": past the
  // last chunk it loads the last chunk again into the idle buffer.
  if (nch > 0) {
    issue(t0, pb);
    for (int c = 0; c < nch; c += 2) {
      const int p0 = pb + c * kPC;
      issue(t1, p0 + kPC * min(1, nch - 1 - c));
      chunk_landed<C::NI + 1>();
      compute(t0);
      chunk_consumed();
      if (c + 1 < nch) {
        issue(t0, p0 + kPC * min(2, nch - 1 - c));
        chunk_landed<C::NI + 1>();
        compute(t1);
        chunk_consumed();
      }
    }
    __builtin_amdgcn_s_waitcnt(vm_wait(0));
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
    if (lq == 0) a.Epart[base + row] = v * (1.0 / (double)a.wscale);  // a power of two: exact
  }
}

// rows per workgroup: 64 (four waves of one 16-row block) or 128 (four
// waves of two blocks up to DP = 128; eight waves of one block above, where
// two chunk buffers fill the LDS of a CU)
template <bool BF, int DP>
hipError_t launch_rows(hipStream_t s, const GprArgs &a) {
  const int tiles = a.ntiles > 0 ? a.ntiles : a.Np / a.rows;
  if (a.rows == 128) {
    if constexpr (DP <= 128)
      k_gpr_mixed<BF, DP, 2, 4><<<dim3(tiles, a.splits), 256, 0, s>>>(a);
    else
      k_gpr_mixed<BF, DP, 1, 8><<<dim3(tiles, a.splits), 512, 0, s>>>(a);
  } else {
    k_gpr_mixed<BF, DP, 1, 4><<<dim3(tiles, a.splits), 256, 0, s>>>(a);
  }
  return hipGetLastError();
}

template <bool BF>
hipError_t launch_typed(hipStream_t s, const GprArgs &a) {
  switch (a.Dp) {
    case 32: return launch_rows<BF, 32>(s, a);
    case 64: return launch_rows<BF, 64>(s, a);
    case 96: return launch_rows<BF, 96>(s, a);
    case 128: return launch_rows<BF, 128>(s, a);
    case 160: return launch_rows<BF, 160>(s, a);
    case 192: return launch_rows<BF, 192>(s, a);
    case 224: return launch_rows<BF, 224>(s, a);
    case 256: return launch_rows<BF, 256>(s, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int gpr_points_per_chunk(int Dp) { return Dp <= 128 ? 64 : 32; }

hipError_t launch_gpr_mixed(hipStream_t s, const GprArgs &a) {
  if (a.pts_per_split % kPC != 0 || a.Mp % kPC != 0 || (a.rows != 64 && a.rows != 128) || a.Np % a.rows != 0)
    return hipErrorInvalidValue;
  return a.bf16 ? launch_typed<true>(s, a) : launch_typed<false>(s, a);
}

}  // namespace gapx
