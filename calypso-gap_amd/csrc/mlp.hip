// Per-atom MLP surrogate on the standardised wACSF features (no reference
// counterpart): e_i = w3 . s(s(xt_i W1 + b1) W2 + b2) + b3 with SiLU s, D = H =
// 128, FP32 master weights, BF16 MFMA products with FP32 accumulation (or, with
// `split`, BF16 hi + lo operands and three products per MFMA step: FP32-level
// products, used for gradient checks).
//
// Kernels:
//   k_mlp_gemm     Y[rows x 128] = X[rows x 128] . W (or W^T), W 128 x 128, with a
//                  fused epilogue (bias; SiLU storing Z and s(Z); or a product
//                  with s'(aux)).  One wave per 16 rows, W^T staged in LDS.
//   k_mlp_gemm_tn  C = A^T B over rows (weight gradients), per-block partials.
//   k_reduce_rows  deterministic sums of per-block partials.
//   k_mlp_head     e = A2 . w3 + b3 and delta2 = s'(Z2) o w3.
//   k_mlp_loss / k_mlp_q2 / k_mlp_q1 / k_adam: the elementwise parts of the
//                  energy + input-gradient loss and its (double) backward.
//
// Layout of the flat parameter vector: W1[128][128] (row = input feature),
// b1[128], W2[128][128], b2[128], w3[128], b3.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

namespace {

constexpr int kH = 128;  // input features = hidden units

typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float d2silu(float z) {
  const float s = sigm(z);
  return s * (1.0f - s) * (2.0f + z * (1.0f - 2.0f * s));
}

__device__ __forceinline__ void split_bf16(float v, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

enum Epi { kEpiBias = 0, kEpiSilu = 1, kEpiMulDsilu = 2 };

// Y = X . op(W) (+ epilogue).  op(W) = W (W[k][n]) or W^T (W[n][k]).  Rows are
// a multiple of 64 (callers pad).  MFMA 16x16x32 bf16 layout per lane (l16,
// lq): A[l16][8 lq + e], B[8 lq + e][l16], D[4 lq + r][l16].
template <bool TRANS, bool SPLIT, int EPI>
__global__ __launch_bounds__(256) void k_mlp_gemm(const float *__restrict__ X, const float *__restrict__ W,
                                                  const float *__restrict__ bias, const float *__restrict__ aux,
                                                  float *__restrict__ Y, float *__restrict__ Y2) {
  constexpr int ST = kH + 8;  // row stride (bf16) of the staged W^T
  __shared__ __align__(16) __bf16 s_wh[kH * ST];
  __shared__ __align__(16) __bf16 s_wl[SPLIT ? kH * ST : 8];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  // stage Bt[n][k] = op(W)[k][n]
  for (int e = tid; e < kH * kH; e += 256) {
    const int a = e / kH, b = e - a * kH;  // W[a][b], read coalesced
    const float v = W[e];
    const int n = TRANS ? a : b, k = TRANS ? b : a;
    __bf16 hi, lo;
    split_bf16(v, hi, lo);
    s_wh[n * ST + k] = hi;
    if (SPLIT) s_wl[n * ST + k] = lo;
  }
  __syncthreads();
  const size_t row = (size_t)blockIdx.x * 64 + wv * 16 + l16;
  v8b ah[kH / 32], al[kH / 32];
#pragma unroll
  for (int t = 0; t < kH / 32; ++t) {
    const float4 p = *(const float4 *)(X + row * kH + 32 * t + 8 * lq);
    const float4 q = *(const float4 *)(X + row * kH + 32 * t + 8 * lq + 4);
    const float v[8] = {p.x, p.y, p.z, p.w, q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      __bf16 hi, lo;
      split_bf16(v[e], hi, lo);
      ah[t][e] = hi;
      al[t][e] = lo;
    }
  }
  const size_t orow0 = (size_t)blockIdx.x * 64 + wv * 16 + 4 * lq;
#pragma unroll
  for (int nb = 0; nb < kH / 16; ++nb) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    const int n = 16 * nb + l16;
#pragma unroll
    for (int t = 0; t < kH / 32; ++t) {
      const v8b bh = *(const v8b *)(s_wh + n * ST + 32 * t + 8 * lq);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bh, acc, 0, 0, 0);
      if (SPLIT) {
        const v8b bl = *(const v8b *)(s_wl + n * ST + 32 * t + 8 * lq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[t], bh, acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const size_t o = (orow0 + r) * kH + n;
      float z = acc[r];
      if (EPI == kEpiBias) {
        if (bias) z += bias[n];
        Y[o] = z;
      } else if (EPI == kEpiSilu) {
        z += bias[n];
        Y[o] = z;
        Y2[o] = silu(z);
      } else {
        Y[o] = z * dsilu(aux[o]);
        if (Y2) Y2[o] = z;  // the product before the factor
      }
    }
  }
}

// Per-block partial of C[i][j] = sum_rows A[row][i] B[row][j]; a block takes
// rows [blockIdx.x * rpb, +rpb) in chunks of 32 staged transposed in LDS.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_mlp_gemm_tn(int rows, int rpb, const float *__restrict__ A,
                                                     const float *__restrict__ B, float *__restrict__ part) {
  constexpr int SC = 32 + 8;
  __shared__ __align__(16) __bf16 s_ah[kH * SC], s_bh[kH * SC];
  __shared__ __align__(16) __bf16 s_al[SPLIT ? kH * SC : 8], s_bl[SPLIT ? kH * SC : 8];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  const int r0 = blockIdx.x * rpb, r1 = min(rows, r0 + rpb);
  v4f C[2][kH / 16];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int jb = 0; jb < kH / 16; ++jb) C[ib][jb] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int c0 = r0; c0 < r1; c0 += 32) {
    __syncthreads();
    for (int e = tid; e < 32 * kH; e += 256) {
      const int r = e / kH, i = e - r * kH;
      const bool in = c0 + r < r1;
      const float a = in ? A[(size_t)(c0 + r) * kH + i] : 0.f;
      const float b = in ? B[(size_t)(c0 + r) * kH + i] : 0.f;
      __bf16 hi, lo;
      split_bf16(a, hi, lo);
      s_ah[i * SC + r] = hi;
      if (SPLIT) s_al[i * SC + r] = lo;
      split_bf16(b, hi, lo);
      s_bh[i * SC + r] = hi;
      if (SPLIT) s_bl[i * SC + r] = lo;
    }
    __syncthreads();
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      const int i = 32 * wv + 16 * ib + l16;
      const v8b aH = *(const v8b *)(s_ah + i * SC + 8 * lq);
      v8b aL;
      if (SPLIT) aL = *(const v8b *)(s_al + i * SC + 8 * lq);
#pragma unroll
      for (int jb = 0; jb < kH / 16; ++jb) {
        const int j = 16 * jb + l16;
        const v8b bH = *(const v8b *)(s_bh + j * SC + 8 * lq);
        C[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH, bH, C[ib][jb], 0, 0, 0);
        if (SPLIT) {
          const v8b bL = *(const v8b *)(s_bl + j * SC + 8 * lq);
          C[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH, bL, C[ib][jb], 0, 0, 0);
          C[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aL, bH, C[ib][jb], 0, 0, 0);
        }
      }
    }
  }
  float *out = part + (size_t)blockIdx.x * kH * kH;
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int jb = 0; jb < kH / 16; ++jb)
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(32 * wv + 16 * ib + 4 * lq + r) * kH + 16 * jb + l16] = C[ib][jb][r];
}

// out[k] (+)= sum_b part[b][k] for k < n, b < nb, in block order.
__global__ void k_reduce_rows(int nb, int n, const float *__restrict__ part, float *__restrict__ out, int accumulate) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * n + k];
  out[k] = accumulate ? out[k] + (float)s : (float)s;
}

// Column sums of M[rows x 128] per block of rpb rows: part[block][128].
__global__ __launch_bounds__(128) void k_colsum(int rows, int rpb, const float *__restrict__ M,
                                                float *__restrict__ part) {
  const int j = threadIdx.x;
  const int r0 = blockIdx.x * rpb, r1 = min(rows, r0 + rpb);
  float s = 0.f;
  for (int r = r0; r < r1; ++r) s += M[(size_t)r * kH + j];
  part[(size_t)blockIdx.x * kH + j] = s;
}

// e = A2 . w3 + b3 (rows < n; padding rows get 0) and delta2 = s'(Z2) o w3.
__global__ __launch_bounds__(256) void k_mlp_head(int n, int rows, const float *__restrict__ A2,
                                                  const float *__restrict__ Z2, const float *__restrict__ w3,
                                                  const float *__restrict__ b3, float *__restrict__ e,
                                                  float *__restrict__ delta2) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wv;
  if (row >= rows) return;
  const size_t o = (size_t)row * kH + 2 * lane;
  const float w0 = w3[2 * lane], w1 = w3[2 * lane + 1];
  float s = A2[o] * w0 + A2[o + 1] * w1;
  if (delta2) {
    delta2[o] = (row < n) ? dsilu(Z2[o]) * w0 : 0.f;
    delta2[o + 1] = (row < n) ? dsilu(Z2[o + 1]) * w1 : 0.f;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) e[row] = (row < n) ? s + b3[0] : 0.f;
}

// Loss residuals: de = cE (e - et) / n, R = cG (G - Gt) / n; per-block sums of
// the two squared errors (FP64).
__global__ __launch_bounds__(256) void k_mlp_loss(int n, int rows, const float *__restrict__ e,
                                                  const float *__restrict__ et, const float *__restrict__ G,
                                                  const float *__restrict__ Gt, float cE, float cG,
                                                  float *__restrict__ de, float *__restrict__ R,
                                                  double *__restrict__ lpart) {
  __shared__ double red[2][256];
  const int t = threadIdx.x;
  const int row = blockIdx.x * 2 + (t >> 7), j = t & 127;
  double le = 0.0, lg = 0.0;
  if (row < rows) {
    const size_t o = (size_t)row * kH + j;
    if (row < n) {
      const float rg = G[o] - Gt[o];
      R[o] = cG * rg / (float)n;
      lg = (double)rg * rg;
      if (j == 0) {
        const float re = e[row] - et[row];
        de[row] = cE * re / (float)n;
        le = (double)re * re;
      }
    } else {
      R[o] = 0.f;
      if (j == 0) de[row] = 0.f;
    }
  }
  red[0][t] = le;
  red[1][t] = lg;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      red[0][t] += red[0][t + off];
      red[1][t] += red[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
    lpart[2 * blockIdx.x] = red[0][0];
    lpart[2 * blockIdx.x + 1] = red[1][0];
  }
}

// q2 = U2 o w3 o s''(Z2) + de w3 o s'(Z2);  P = U2 o s'(Z2) + de A2 (dw3 terms).
__global__ void k_mlp_q2(size_t total, const float *__restrict__ U2, const float *__restrict__ Z2,
                         const float *__restrict__ A2, const float *__restrict__ w3, const float *__restrict__ de,
                         float *__restrict__ q2, float *__restrict__ P) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= total) return;
  const int j = (int)(o % kH);
  const size_t row = o / kH;
  const float z = Z2[o], u = U2[o], d = de[row];
  const float s1 = dsilu(z);
  q2[o] = u * w3[j] * d2silu(z) + d * w3[j] * s1;
  P[o] = u * s1 + d * A2[o];
}

// q1 = V o s'(Z1) + U1 o T1 o s''(Z1).
__global__ void k_mlp_q1(size_t total, const float *__restrict__ V, const float *__restrict__ Z1,
                         const float *__restrict__ U1, const float *__restrict__ T1, float *__restrict__ q1) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= total) return;
  const float z = Z1[o];
  q1[o] = V[o] * dsilu(z) + U1[o] * T1[o] * d2silu(z);
}

__global__ void k_sum_scalar(int n, const float *__restrict__ v, float *__restrict__ out, int accumulate) {
  __shared__ double red[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) s += v[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = accumulate ? out[0] + (float)red[0] : (float)red[0];
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// Gather training rows: dst[r] = src[idx[r]] for the three label arrays.
__global__ void k_gather_rows(int rows, int n_valid, uint64_t seed, uint64_t step, int n_data,
                              const float *__restrict__ X, const float *__restrict__ E,
                              const float *__restrict__ G, float *__restrict__ bx, float *__restrict__ be,
                              float *__restrict__ bg) {
  const int r = blockIdx.x;
  const int j = threadIdx.x;
  if (r >= rows) return;
  // SplitMix64 of (seed, step, r): a uniform row of the data set
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (step * 1000003ull + (uint64_t)r + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const size_t src = (size_t)(z % (uint64_t)n_data);
  const size_t o = (size_t)r * kH + j;
  if (r < n_valid) {
    bx[o] = X[src * kH + j];
    bg[o] = G[src * kH + j];
    if (j == 0) be[r] = E[src];
  } else {
    bx[o] = 0.f;
    bg[o] = 0.f;
    if (j == 0) be[r] = 0.f;
  }
}

// Labels from a GAP run: xt (FP32 or FP64), per-atom energy, G = g / theta_inv.
__global__ void k_export_labels(int n, const float *__restrict__ xt32, const double *__restrict__ xt64,
                                const float *__restrict__ g32, const double *__restrict__ g64,
                                const double *__restrict__ eatom, const double *__restrict__ theta_inv, int Dp,
                                float *__restrict__ X, float *__restrict__ E, float *__restrict__ G) {
  const int i = blockIdx.x, j = threadIdx.x;
  if (i >= n) return;
  const size_t s = (size_t)i * Dp + j, o = (size_t)i * kH + j;
  X[o] = xt32 ? xt32[s] : (float)xt64[s];
  const double g = g32 ? (double)g32[s] : g64[s];
  G[o] = (float)(g / theta_inv[j]);
  if (j == 0) E[i] = (float)eatom[i];
}

// out[0..1] = sums over blocks of the per-block pairs (block order).
__global__ void k_reduce_pair(int nb, const double *__restrict__ part, double *__restrict__ out) {
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nb; ++k) {
      a += part[2 * k];
      b += part[2 * k + 1];
    }
    out[0] = a;
    out[1] = b;
  }
}

__global__ void k_f64_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (float)a[k];
}

// Surrogate outputs in the GAP path's buffers: eatom = e (rows < n), and the
// descriptor-space gradient g = G o theta_inv (FP32 or FP64).
__global__ void k_mlp_to_gap(int n, int rows, int Dp, const float *__restrict__ e, const float *__restrict__ G,
                             const double *__restrict__ theta_inv, double *__restrict__ eatom,
                             float *__restrict__ g32, double *__restrict__ g64) {
  const int i = blockIdx.x, j = threadIdx.x;
  if (i >= rows) return;
  if (j == 0) eatom[i] = (i < n) ? (double)e[i] : 0.0;
  if (G && j < Dp) {
    const size_t o = (size_t)i * Dp + j;
    const double g = (i < n && j < kH) ? (double)G[(size_t)i * kH + j] * theta_inv[j] : 0.0;
    if (g32) g32[o] = (float)g;
    if (g64) g64[o] = g;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

constexpr int kParams = kH * kH + kH + kH * kH + kH + kH + 1;
constexpr int kOffW1 = 0, kOffB1 = kH * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH;

struct MlpBuf {
  void *p = nullptr;
  size_t cap = 0;
};

struct Mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *params = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  long long adam_t = 0;
  std::map<std::string, MlpBuf> mem;
  // device-resident training set
  float *dx = nullptr, *de = nullptr, *dg = nullptr;
  long long n_data = 0, cap_data = 0;

  template <typename T>
  int get(const char *name, size_t count, T **out) {
    MlpBuf &b = mem[name];
    const size_t need = std::max<size_t>(count, 1) * sizeof(T);
    if (b.cap < need) {
      if (b.p) (void)hipFree(b.p);
      b.p = nullptr;
      b.cap = 0;
      if (hipMalloc(&b.p, need) != hipSuccess) return fail(GAPX_ERR_NOMEM, std::string("hipMalloc failed for ") + name);
      b.cap = need;
    }
    *out = (T *)b.p;
    return GAPX_OK;
  }
};

namespace {

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double normal(uint64_t &s) {
  const double u1 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.141592653589793 * u2);
}

template <bool TRANS, int EPI>
hipError_t gemm(hipStream_t s, bool split, int rows, const float *X, const float *W, const float *bias,
                const float *aux, float *Y, float *Y2) {
  const dim3 grid(rows / 64);
  if (split)
    k_mlp_gemm<TRANS, true, EPI><<<grid, 256, 0, s>>>(X, W, bias, aux, Y, Y2);
  else
    k_mlp_gemm<TRANS, false, EPI><<<grid, 256, 0, s>>>(X, W, bias, aux, Y, Y2);
  return hipGetLastError();
}

// Work buffers for `rows` (a multiple of 64) rows.
struct Pass {
  float *Z1, *A1, *Z2, *A2, *e, *D2, *T1, *D1, *G;
};

int pass_buffers(Mlp &m, int rows, Pass &p) {
  const size_t n = (size_t)rows * kH;
  int rc;
  if ((rc = m.get("Z1", n, &p.Z1)) || (rc = m.get("A1", n, &p.A1)) || (rc = m.get("Z2", n, &p.Z2)) ||
      (rc = m.get("A2", n, &p.A2)) || (rc = m.get("e", rows, &p.e)) || (rc = m.get("D2", n, &p.D2)) ||
      (rc = m.get("T1", n, &p.T1)) || (rc = m.get("D1", n, &p.D1)) || (rc = m.get("G", n, &p.G)))
    return rc;
  return GAPX_OK;
}

// Forward and (optionally) the input gradient G = de/dxt for rows of X.
int enqueue_forward(Mlp &m, hipStream_t s, bool split, int n, int rows, const float *X, bool want_g, Pass &p) {
  const float *P = m.params;
  GAPX_HIP((gemm<false, kEpiSilu>(s, split, rows, X, P + kOffW1, P + kOffB1, nullptr, p.Z1, p.A1)));
  GAPX_HIP((gemm<false, kEpiSilu>(s, split, rows, p.A1, P + kOffW2, P + kOffB2, nullptr, p.Z2, p.A2)));
  k_mlp_head<<<(rows + 3) / 4, 256, 0, s>>>(n, rows, p.A2, p.Z2, P + kOffW3, P + kOffB3, p.e, want_g ? p.D2 : nullptr);
  GAPX_HIP(hipGetLastError());
  if (want_g) {
    // T1 = D2 W2^T, D1 = T1 o s'(Z1) (T1 kept), G = D1 W1^T
    GAPX_HIP((gemm<true, kEpiMulDsilu>(s, split, rows, p.D2, P + kOffW2, nullptr, p.Z1, p.D1, p.T1)));
    GAPX_HIP((gemm<true, kEpiBias>(s, split, rows, p.D1, P + kOffW1, nullptr, nullptr, p.G, nullptr)));
  }
  return GAPX_OK;
}

// Reduce the rows-over partial sums of a TN product into grad[off ...].
int tn_product(Mlp &m, bool split, int rows, const float *A, const float *B, float *dst, bool accumulate) {
  // rows per workgroup: enough workgroups to cover the GPU, at most 64 partials
  const int rpb = std::max(64, (rows + 63) / 64 / 32 * 32);
  const int nb = (rows + rpb - 1) / rpb;
  float *part;
  int rc;
  if ((rc = m.get("tn_part", (size_t)nb * kH * kH, &part))) return rc;
  if (split)
    k_mlp_gemm_tn<true><<<nb, 256, 0, m.stream>>>(rows, rpb, A, B, part);
  else
    k_mlp_gemm_tn<false><<<nb, 256, 0, m.stream>>>(rows, rpb, A, B, part);
  GAPX_HIP(hipGetLastError());
  k_reduce_rows<<<(kH * kH + 255) / 256, 256, 0, m.stream>>>(nb, kH * kH, part, dst, accumulate ? 1 : 0);
  GAPX_HIP(hipGetLastError());
  return GAPX_OK;
}

int colsum(Mlp &m, int rows, const float *M, float *dst) {
  const int rpb = 64;
  const int nb = (rows + rpb - 1) / rpb;
  float *part;
  int rc;
  if ((rc = m.get("cs_part", (size_t)nb * kH, &part))) return rc;
  k_colsum<<<nb, 128, 0, m.stream>>>(rows, rpb, M, part);
  GAPX_HIP(hipGetLastError());
  k_reduce_rows<<<1, 128, 0, m.stream>>>(nb, kH, part, dst, 0);
  GAPX_HIP(hipGetLastError());
  return GAPX_OK;
}

// Loss L = cE/(2n) sum (e - et)^2 + cG/(2n) sum |G - Gt|^2 and dL/dparams
// (into m.grad) for rows [0, n) of X (rows padded to a multiple of 64 with
// zero input rows that do not contribute).  lsum[0..1] receive the two sums
// of squares (device, FP64).
int enqueue_loss_grad(Mlp &m, bool split, int n, int rows, const float *X, const float *et, const float *Gt,
                      float cE, float cG, double *lsum_dev) {
  Pass p;
  int rc;
  if ((rc = pass_buffers(m, rows, p)) || (rc = enqueue_forward(m, m.stream, split, n, rows, X, true, p))) return rc;
  const size_t nn = (size_t)rows * kH;
  float *de, *R, *U1, *dT1, *U2, *q2, *P, *V, *q1;
  double *lpart;
  const int lblocks = (rows + 1) / 2;
  if ((rc = m.get("de", rows, &de)) || (rc = m.get("R", nn, &R)) || (rc = m.get("U1", nn, &U1)) ||
      (rc = m.get("dT1", nn, &dT1)) || (rc = m.get("U2", nn, &U2)) || (rc = m.get("q2", nn, &q2)) ||
      (rc = m.get("P", nn, &P)) || (rc = m.get("V", nn, &V)) || (rc = m.get("q1", nn, &q1)) ||
      (rc = m.get("lpart", 2 * (size_t)lblocks, &lpart)))
    return rc;
  hipStream_t s = m.stream;
  const float *W = m.params;
  float *g = m.grad;
  k_mlp_loss<<<lblocks, 256, 0, s>>>(n, rows, p.e, et, p.G, Gt, cE, cG, de, R, lpart);
  GAPX_HIP(hipGetLastError());
  // U1 = R W1, dT1 = U1 o s'(Z1) (U1 kept)
  GAPX_HIP((gemm<false, kEpiMulDsilu>(s, split, rows, R, W + kOffW1, nullptr, p.Z1, dT1, U1)));
  // U2 = dT1 W2
  GAPX_HIP((gemm<false, kEpiBias>(s, split, rows, dT1, W + kOffW2, nullptr, nullptr, U2, nullptr)));
  k_mlp_q2<<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(nn, U2, p.Z2, p.A2, W + kOffW3, de, q2, P);
  GAPX_HIP(hipGetLastError());
  // V = q2 W2^T, q1 = V o s'(Z1) + U1 o T1 o s''(Z1)
  GAPX_HIP((gemm<true, kEpiBias>(s, split, rows, q2, W + kOffW2, nullptr, nullptr, V, nullptr)));
  k_mlp_q1<<<(unsigned)((nn + 255) / 256), 256, 0, s>>>(nn, V, p.Z1, U1, p.T1, q1);
  GAPX_HIP(hipGetLastError());
  // weight gradients (padding rows are zero in every factor that matters)
  if ((rc = tn_product(m, split, rows, X, q1, g + kOffW1, false)) ||
      (rc = tn_product(m, split, rows, R, p.D1, g + kOffW1, true)) ||
      (rc = tn_product(m, split, rows, p.A1, q2, g + kOffW2, false)) ||
      (rc = tn_product(m, split, rows, dT1, p.D2, g + kOffW2, true)) || (rc = colsum(m, rows, q1, g + kOffB1)) ||
      (rc = colsum(m, rows, q2, g + kOffB2)) || (rc = colsum(m, rows, P, g + kOffW3)))
    return rc;
  k_sum_scalar<<<1, 256, 0, s>>>(rows, de, g + kOffB3, 0);
  GAPX_HIP(hipGetLastError());
  // loss sums
  k_reduce_pair<<<1, 64, 0, s>>>(lblocks, lpart, lsum_dev);
  GAPX_HIP(hipGetLastError());
  return GAPX_OK;
}

}  // namespace


// ---- used by the context (api.hip) -----------------------------------------
int mlp_eval_into_gap(Mlp &m, hipStream_t s, bool split, int n, int rows, const float *xt, bool grad, int Dp,
                      const double *theta_inv, double *eatom, float *g32, double *g64) {
  if (rows % 64 != 0) return fail(GAPX_ERR_ARG, "surrogate rows must be a multiple of 64");
  Pass p;
  int rc;
  if ((rc = pass_buffers(m, rows, p)) || (rc = enqueue_forward(m, s, split, n, rows, xt, grad, p))) return rc;
  k_mlp_to_gap<<<rows, std::max(Dp, 64), 0, s>>>(n, rows, Dp, p.e, grad ? p.G : nullptr, theta_inv, eatom, g32, g64);
  GAPX_HIP(hipGetLastError());
  return GAPX_OK;
}

hipError_t launch_f64_to_f32(hipStream_t s, size_t n, const double *a, float *b) {
  if (n == 0) return hipSuccess;
  k_f64_to_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, a, b);
  return hipGetLastError();
}

static int reserve_data(Mlp &m, long long need) {
  if (need <= m.cap_data) return GAPX_OK;
  long long cap = std::max<long long>(need, m.cap_data + m.cap_data / 2);
  cap = std::max<long long>(cap, 4096);
  float *x = nullptr, *e = nullptr, *g = nullptr;
  if (hipMalloc(&x, sizeof(float) * kH * (size_t)cap) != hipSuccess || hipMalloc(&e, sizeof(float) * (size_t)cap) != hipSuccess ||
      hipMalloc(&g, sizeof(float) * kH * (size_t)cap) != hipSuccess) {
    if (x) (void)hipFree(x);
    if (e) (void)hipFree(e);
    if (g) (void)hipFree(g);
    return fail(GAPX_ERR_NOMEM, "hipMalloc failed for the surrogate training set");
  }
  if (m.n_data > 0) {
    GAPX_HIP(hipMemcpyAsync(x, m.dx, sizeof(float) * kH * (size_t)m.n_data, hipMemcpyDeviceToDevice, m.stream));
    GAPX_HIP(hipMemcpyAsync(e, m.de, sizeof(float) * (size_t)m.n_data, hipMemcpyDeviceToDevice, m.stream));
    GAPX_HIP(hipMemcpyAsync(g, m.dg, sizeof(float) * kH * (size_t)m.n_data, hipMemcpyDeviceToDevice, m.stream));
    GAPX_HIP(hipStreamSynchronize(m.stream));
  }
  if (m.dx) (void)hipFree(m.dx);
  if (m.de) (void)hipFree(m.de);
  if (m.dg) (void)hipFree(m.dg);
  m.dx = x;
  m.de = e;
  m.dg = g;
  m.cap_data = cap;
  return GAPX_OK;
}

int mlp_add_labels(Mlp &m, hipStream_t s, int n, const float *xt32, const double *xt64, const float *g32,
                   const double *g64, const double *eatom, const double *theta_inv, int Dp) {
  if (Dp != kH) return fail(GAPX_ERR_MODEL, "the surrogate takes 128 descriptor columns");
  GAPX_HIP(hipStreamSynchronize(s));
  int rc;
  if ((rc = reserve_data(m, m.n_data + n))) return rc;
  k_export_labels<<<n, kH, 0, s>>>(n, xt32, xt64, g32, g64, eatom, theta_inv, Dp, m.dx + (size_t)kH * m.n_data,
                                   m.de + m.n_data, m.dg + (size_t)kH * m.n_data);
  GAPX_HIP(hipGetLastError());
  GAPX_HIP(hipStreamSynchronize(s));
  m.n_data += n;
  return GAPX_OK;
}

}  // namespace gapx

struct gapx_mlp : gapx::Mlp {};

namespace gapx {
Mlp *mlp_of(gapx_mlp *h) { return h; }
}  // namespace gapx

using namespace gapx;

namespace {

int round64(int n) { return (n + 63) / 64 * 64; }

// Host rows -> padded device rows (zero padding).
int upload_rows(Mlp &m, const char *name, int n, int rows, int width, const float *host, float **dev) {
  int rc;
  if ((rc = m.get(name, (size_t)rows * width, dev))) return rc;
  GAPX_HIP(hipMemsetAsync(*dev, 0, sizeof(float) * (size_t)rows * width, m.stream));
  if (host && n > 0) GAPX_HIP(hipMemcpyAsync(*dev, host, sizeof(float) * (size_t)n * width, hipMemcpyHostToDevice, m.stream));
  return GAPX_OK;
}

}  // namespace

extern "C" {

int gapx_mlp_param_count(int *n) {
  if (!n) return fail(GAPX_ERR_ARG, "null argument");
  *n = kParams;
  return GAPX_OK;
}

int gapx_mlp_create(unsigned long long seed, gapx_mlp **out) {
  if (!out) return fail(GAPX_ERR_ARG, "null argument");
  gapx_mlp *m = new gapx_mlp();
  if (hipGetDevice(&m->device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return fail(GAPX_ERR_HIP, "could not create a HIP stream");
  }
  if (hipMalloc(&m->params, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->grad, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->adam_m, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc(&m->adam_v, sizeof(float) * kParams) != hipSuccess) {
    gapx_mlp_free(m);
    return fail(GAPX_ERR_NOMEM, "hipMalloc failed for the surrogate parameters");
  }
  // hidden layers N(0, 1/fan_in), zero output layer and biases: the first
  // prediction is 0, so training starts at the label variance
  std::vector<float> p(kParams, 0.0f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 12345;
  const double s1 = std::sqrt(1.0 / kH);
  for (int k = 0; k < kH * kH; ++k) p[kOffW1 + k] = (float)(s1 * normal(st));
  for (int k = 0; k < kH * kH; ++k) p[kOffW2 + k] = (float)(s1 * normal(st));
  std::vector<float> z(kParams, 0.0f);
  if (hipMemcpy(m->params, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->adam_m, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->adam_v, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess) {
    gapx_mlp_free(m);
    return fail(GAPX_ERR_HIP, "surrogate initialisation copy failed");
  }
  *out = m;
  return GAPX_OK;
}

void gapx_mlp_free(gapx_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (auto &it : m->mem)
    if (it.second.p) (void)hipFree(it.second.p);
  float *bufs[] = {m->params, m->grad, m->adam_m, m->adam_v, m->dx, m->de, m->dg};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int gapx_mlp_get_params(gapx_mlp *m, float *host) {
  if (!m || !host) return fail(GAPX_ERR_ARG, "null argument");
  GAPX_HIP(hipSetDevice(m->device));
  GAPX_HIP(hipStreamSynchronize(m->stream));
  GAPX_HIP(hipMemcpy(host, m->params, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return GAPX_OK;
}

int gapx_mlp_set_params(gapx_mlp *m, const float *host) {
  if (!m || !host) return fail(GAPX_ERR_ARG, "null argument");
  GAPX_HIP(hipSetDevice(m->device));
  GAPX_HIP(hipStreamSynchronize(m->stream));
  GAPX_HIP(hipMemcpy(m->params, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
  std::vector<float> z(kParams, 0.0f);
  GAPX_HIP(hipMemcpy(m->adam_m, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  GAPX_HIP(hipMemcpy(m->adam_v, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  m->adam_t = 0;
  return GAPX_OK;
}

int gapx_mlp_forward(gapx_mlp *m, int n, const float *xt, int flags, float *e, float *G) {
  if (!m || n <= 0 || !xt || !e) return fail(GAPX_ERR_ARG, "gapx_mlp_forward: invalid argument");
  GAPX_HIP(hipSetDevice(m->device));
  const bool split = (flags & GAPX_MLP_SPLIT) != 0;
  const int rows = round64(n);
  float *X;
  int rc;
  if ((rc = upload_rows(*m, "in_x", n, rows, kH, xt, &X))) return rc;
  Pass p;
  if ((rc = pass_buffers(*m, rows, p)) || (rc = enqueue_forward(*m, m->stream, split, n, rows, X, G != nullptr, p)))
    return rc;
  GAPX_HIP(hipMemcpyAsync(e, p.e, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  if (G) GAPX_HIP(hipMemcpyAsync(G, p.G, sizeof(float) * (size_t)n * kH, hipMemcpyDeviceToHost, m->stream));
  GAPX_HIP(hipStreamSynchronize(m->stream));
  return GAPX_OK;
}

int gapx_mlp_loss_grad(gapx_mlp *m, int n, const float *xt, const float *e_t, const float *G_t, float cE, float cG,
                       int flags, double *sums, float *grad) {
  if (!m || n <= 0 || !xt || !e_t || !G_t) return fail(GAPX_ERR_ARG, "gapx_mlp_loss_grad: invalid argument");
  GAPX_HIP(hipSetDevice(m->device));
  const bool split = (flags & GAPX_MLP_SPLIT) != 0;
  const int rows = round64(n);
  float *X, *E, *Gt;
  double *ls;
  int rc;
  if ((rc = upload_rows(*m, "in_x", n, rows, kH, xt, &X)) || (rc = upload_rows(*m, "in_e", n, rows, 1, e_t, &E)) ||
      (rc = upload_rows(*m, "in_g", n, rows, kH, G_t, &Gt)) || (rc = m->get("lsum", 2, &ls)))
    return rc;
  if ((rc = enqueue_loss_grad(*m, split, n, rows, X, E, Gt, cE, cG, ls))) return rc;
  if (sums) GAPX_HIP(hipMemcpyAsync(sums, ls, 2 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (grad) GAPX_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
  GAPX_HIP(hipStreamSynchronize(m->stream));
  return GAPX_OK;
}

int gapx_mlp_data_clear(gapx_mlp *m) {
  if (!m) return fail(GAPX_ERR_ARG, "null argument");
  m->n_data = 0;
  return GAPX_OK;
}

int gapx_mlp_data_count(gapx_mlp *m, long long *n) {
  if (!m || !n) return fail(GAPX_ERR_ARG, "null argument");
  *n = m->n_data;
  return GAPX_OK;
}

int gapx_mlp_data_add(gapx_mlp *m, int n, const float *xt, const float *e_t, const float *G_t) {
  if (!m || n <= 0 || !xt || !e_t || !G_t) return fail(GAPX_ERR_ARG, "gapx_mlp_data_add: invalid argument");
  GAPX_HIP(hipSetDevice(m->device));
  int rc;
  if ((rc = reserve_data(*m, m->n_data + n))) return rc;
  GAPX_HIP(hipMemcpyAsync(m->dx + (size_t)kH * m->n_data, xt, sizeof(float) * kH * (size_t)n, hipMemcpyHostToDevice, m->stream));
  GAPX_HIP(hipMemcpyAsync(m->de + m->n_data, e_t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream));
  GAPX_HIP(hipMemcpyAsync(m->dg + (size_t)kH * m->n_data, G_t, sizeof(float) * kH * (size_t)n, hipMemcpyHostToDevice, m->stream));
  GAPX_HIP(hipStreamSynchronize(m->stream));
  m->n_data += n;
  return GAPX_OK;
}

int gapx_mlp_train(gapx_mlp *m, int steps, int batch, float lr, float cE, float cG, unsigned long long seed,
                   int flags, double *sums) {
  if (!m || steps < 0 || batch <= 0) return fail(GAPX_ERR_ARG, "gapx_mlp_train: invalid argument");
  if (m->n_data <= 0) return fail(GAPX_ERR_ARG, "gapx_mlp_train: empty training set");
  GAPX_HIP(hipSetDevice(m->device));
  const bool split = (flags & GAPX_MLP_SPLIT) != 0;
  const int rows = round64(batch);
  float *bx, *be, *bg;
  double *ls;
  int rc;
  if ((rc = m->get("b_x", (size_t)rows * kH, &bx)) || (rc = m->get("b_e", rows, &be)) ||
      (rc = m->get("b_g", (size_t)rows * kH, &bg)) || (rc = m->get("lsum", 2, &ls)))
    return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int step = 0; step < steps; ++step) {
    k_gather_rows<<<rows, kH, 0, m->stream>>>(rows, batch, seed, (uint64_t)step, (int)std::min<long long>(m->n_data, 0x7fffffff),
                                              m->dx, m->de, m->dg, bx, be, bg);
    GAPX_HIP(hipGetLastError());
    if ((rc = enqueue_loss_grad(*m, split, batch, rows, bx, be, bg, cE, cG, ls))) return rc;
    m->adam_t += 1;
    const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
    k_adam<<<(kParams + 255) / 256, 256, 0, m->stream>>>(kParams, m->params, m->grad, m->adam_m, m->adam_v, lr, b1, b2,
                                                        c1, c2, eps);
    GAPX_HIP(hipGetLastError());
  }
  if (sums) GAPX_HIP(hipMemcpyAsync(sums, ls, 2 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  GAPX_HIP(hipStreamSynchronize(m->stream));
  return GAPX_OK;
}

}  // extern "C"
