// wACSF descriptors on the GPU: one wavefront (one 64-thread workgroup) per
// centre atom, in FP64 or FP32.
//
// Forward (CAR2ACSF's xx, gappy/libgap/wacsf.f90:65-796): for every symmetry
// function line s the plain feature x[s] and the weighted one x[s + nsf]:
//   radial  (types 1, 3): sum_j exp(-a (r_ij - rs)^2) fc(r_ij)          (* w_j)
//   angular (types 2, 4): sum_{j<k} pref (1 + lam cos)^zeta
//                         exp(-eta (r_ij^2 + r_ik^2 + r_jk^2)) fc fc fc  (* w_j w_k)
// with fc(r) = (cos(pi r / rc) + 1) / 2, pi = 3.141592654, a term counted
// only when every distance is <= the line's cutoff, (a, rs) = (alpha, 0) for
// type 1 and (4, Rs) for type 3, lam = +1 (type 2) / -1 (type 4), and
// pref = 2^(1 - zeta) (= 1 at zeta = 1, the reference's only case).
//
// Backward (replaces the dense dxdy/strs of wacsf.f90 and the force and
// stress loops of gap_calc.f90:172-203): with g = dE/dx of the centre atom,
// every term depends on the positions only through the distances r_ij, r_ik,
// r_jk, so the kernel accumulates T = sum_s g_eff,s d(term_s)/d(distance)
// (g_eff = g_s + w g_{s+nsf}) and turns each T into force vectors along the
// distance and the virial T d d^T / r.  Nothing of size D x neighbours is
// stored.
//
// Work layout: the atom's neighbours live in LDS (positions relative to the
// centre, distance, weight, cutoff function per cutoff group).  Radial sums
// put one neighbour on each lane.  Angular sums read the atom's triplet list,
// built once per evaluation by k_build<true> (neighbour pairs j < k with all
// three distances within the largest angular cutoff, compacted with a
// ballot), 64 triplets per step, so every lane holds a live triplet and
// forward and backward share one enumeration.  The forward keeps NCH lines'
// partial sums in registers per lane and reduces them across the wavefront
// after each chunk of lines.  Lines are sorted by (cutoff group, eta, lambda,
// zeta), so exponentials, cutoff products and the powers (1 + lambda cos)^zeta
// are updated incrementally along a chunk (zeta = 2, 4, 16 by squaring).
#include <atomic>
#include <mutex>
#include <set>
#include <utility>

#include "gapx_internal.h"

namespace gapx {

namespace {

template <typename T>
struct Math;

template <>
struct Math<double> {
  static __device__ __forceinline__ double ex(double x) { return exp(x); }
  static __device__ __forceinline__ double cs(double x) { return cos(x); }
  static __device__ __forceinline__ double sn(double x) { return sin(x); }
  // sine and cosine of one argument with one argument reduction
  static __device__ __forceinline__ void sc(double x, double &s, double &c) { sincos(x, &s, &c); }
  static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
  static __device__ __forceinline__ double pw(double b, double e) { return pow(b, e); }
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
};

template <>
struct Math<float> {
  static __device__ __forceinline__ float ex(float x) { return __expf(x); }
  static __device__ __forceinline__ float cs(float x) { return __cosf(x); }
  static __device__ __forceinline__ float sn(float x) { return __sinf(x); }
  static __device__ __forceinline__ void sc(float x, float &s, float &c) {
    s = __sinf(x);
    c = __cosf(x);
  }
  static __device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float pw(float b, float e) { return powf(b, e); }
  // hardware approximations (about 1 ulp) for the mixed path's inner loops
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

// Two values of T read as one LDS access (the backward's per-neighbour pairs).
template <typename T>
struct Pair;
template <>
struct Pair<float> {
  typedef float2 type;
};
template <>
struct Pair<double> {
  typedef double2 type;
};

// A radial-table entry in the kernel's precision: the FP32 copy for FP32
// kernels (the same rounding as converting the FP64 entry, without a
// conversion per line and lane).
template <typename T>
__device__ __forceinline__ T rad_table(const double *t64, const float *t32, int s) {
  if constexpr (sizeof(T) == 4)
    return t32[s];
  else
    return t64[s];
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Powers (1 + lam cos)^zeta along lines sorted by (lam, zeta) within an eta
// group.  pw = base^z and pm = base^(z-1) for the current z; a new lambda or
// a smaller zeta restarts at z = 1; larger integer zetas are reached by
// squaring (z -> 2z) or one more factor (z -> z+1), decisions that are the
// same on every lane.
template <typename T>
struct PowState {
  T lam = T(0), base = T(0), pw = T(0), pm = T(1);
  int z = 0;
  __device__ __forceinline__ void at(T new_lam, int zi, T cth) {
    if (new_lam != lam || zi < z || z == 0) {
      lam = new_lam;
      base = T(1) + lam * cth;
      pw = base;
      pm = T(1);
      z = 1;
    }
    while (z < zi) {
      if (2 * z <= zi) {
        pm = pw * pm;
        pw = pw * pw;
        z *= 2;
      } else {
        pm = pw;
        pw = pw * base;
        z += 1;
      }
    }
  }
};

// P = pref base^zeta and dP/dcos = pref zeta lam base^(zeta-1) of line s.
template <typename T>
__device__ __forceinline__ void line_power(const SFDev &sf, int s, T cth, PowState<T> &ps, T &P, T &dP) {
  const int zi = sf.ang_zeta[s];
  const T lam = (T)sf.ang_lam[s];
  if (zi == 1) {
    ps.at(lam, 1, cth);
    P = ps.pw;
    dP = lam;
  } else if (zi > 1) {
    ps.at(lam, zi, cth);
    const T pref = (T)sf.ang_pref[s];
    P = pref * ps.pw;
    dP = pref * (T)zi * lam * ps.pm;
  } else {  // non-integer zeta
    const T base = T(1) + lam * cth, zr = (T)sf.ang_zeta_real[s], pref = (T)sf.ang_pref[s];
    const T b1 = (base > T(0)) ? Math<T>::pw(base, zr - T(1)) : T(0);
    P = pref * b1 * base;
    dP = pref * zr * lam * b1;
  }
}

// A feature's FP16 contraction operand: xt clamped to [-2^15, 2^15], so the
// hi part is finite and the lo part exact.  The host keeps every scaled
// sparse-point component within 2^14 on this path, so a clamped component
// still differs from each of them by more than 2^14: the kernel value is
// exp(-d2 / 2) with d2 > 2^28 whether the norm of the clamped or of the true
// row enters d2, i.e. 0 in FP32 and in FP64 alike.
__device__ __forceinline__ double fp16_operand(double xt) { return fmin(fmax(xt, -32768.0), 32768.0); }

// Weight of a neighbour slot: the caller's (CAR2ACSF's neighbour table), or
// that of the neighbour's species.
__device__ __forceinline__ double slot_weight(const NeighborBufs &nb, const double *wz, const int *species,
                                              size_t slot) {
  return nb.nb_w ? nb.nb_w[slot] : wz[species[nb.nb_j[slot]]];
}

// Launch geometry of the descriptor kernels: LDS layout (neighbour slots,
// cutoff groups, padded features) and the first atom of the launch (one
// workgroup per atom from atom0 on).
struct Lds {
  int KP, G, Dp;
  int atom0;
};

// Triplet indices of wave steps t, t + stride, .. (4 steps) for one lane; the
// loads are issued together, one iteration ahead of their use, so a wave
// waits for global memory once per four steps at most.  Inactive entries
// point at neighbours 0 and 1 (valid whenever the list is not empty).
__device__ __forceinline__ void load_q4(const uint32_t *tl, int nt, int t, int stride, uint32_t (&q)[4]) {
#pragma unroll
  for (int u = 0; u < 4; ++u) q[u] = (t + u * stride < nt) ? tl[t + u * stride] : (1u << 16);
}

// Front of the queue, the rest moved up (static register indices: the step's
// body stays a single copy, so the kernel keeps its register budget).
__device__ __forceinline__ uint32_t pop_q4(uint32_t (&q)[4]) {
  const uint32_t v = q[0];
  q[0] = q[1];
  q[1] = q[2];
  q[2] = q[3];
  return v;
}

// ---------------------------------------------------------------------------
// triplet lists
// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <typename T, int NCH>
__device__ __forceinline__ void ang_fwd_term(uint32_t q, bool active, int c0, const SFDev &sf, const T *s_d,
                                             const T *s_w, const T *s_fc, int KP, T (&acc)[NCH],
                                             T (&accw)[NCH]) {
  const int a = q & 0xffff, b = q >> 16;
  const T ax = s_d[4 * a], ay = s_d[4 * a + 1], az = s_d[4 * a + 2], ra = s_d[4 * a + 3];
  const T bx = s_d[4 * b], by = s_d[4 * b + 1], bz = s_d[4 * b + 2], rb = s_d[4 * b + 3];
  const T ex = bx - ax, ey = by - ay, ez = bz - az;
  const T e2 = ex * ex + ey * ey + ez * ez;
  const T e = Math<T>::sq(e2);
  const T cth = (ra * ra + rb * rb - e2) / (T(2) * ra * rb);
  const T S2 = ra * ra + rb * rb + e2;
  const T wab = s_w[a] * s_w[b];
  const T on = active ? T(1) : T(0);
  int cur_g = -1;
  T F3 = 0, E = 0, cur_eta = T(-1);
  PowState<T> ps;
#pragma unroll
  for (int l = 0; l < NCH; ++l) {
    const int s = c0 + l;
    if (s < sf.n_ang) {
      const int g = sf.ang_group[s];
      if (g != cur_g) {
        cur_g = g;
        cur_eta = T(-1);  // E below carries this group's cutoff product
        const T rc = (T)sf.group_rc[g];
        const T fe = (e <= rc) ? T(0.5) * (Math<T>::cs((T)kPi * e / rc) + T(1)) : T(0);
        F3 = s_fc[g * KP + a] * s_fc[g * KP + b] * fe * on;
      }
      const T eta = (T)sf.ang_eta[s];
      if (eta != cur_eta) {
        cur_eta = eta;
        E = Math<T>::ex(-eta * S2) * F3;
      }
      T P, dP;
      line_power<T>(sf, s, cth, ps, P, dP);
      const T t = P * E;
      acc[l] += t;
      accw[l] += t * wab;
    }
  }
}

// W wavefronts per atom (workgroup of 64 W threads): neighbours, triplets
// and radial terms are split over the waves; each chunk's per-lane sums are
// reduced within a wave and then across the waves through LDS.
template <typename T, int NCH, int W>
__device__ __forceinline__ void chunk_reduce(T (&acc)[NCH], T (&accw)[NCH], int wv, int lane, int tid, int c0,
                                             int n, const int *__restrict__ feat, int nsf, T *s_part, T *s_feat) {
#pragma unroll
  for (int l = 0; l < NCH; ++l) {
    const T v = wave_sum(acc[l]);
    const T vw = wave_sum(accw[l]);
    if (lane == 0) {
      s_part[wv * 2 * NCH + l] = v;
      s_part[wv * 2 * NCH + NCH + l] = vw;
    }
  }
  __syncthreads();
  if (tid < 2 * NCH) {
    const int l = tid % NCH;
    if (c0 + l < n) {
      T sum = T(0);
#pragma unroll
      for (int w = 0; w < W; ++w) sum += s_part[w * 2 * NCH + tid];
      s_feat[feat[c0 + l] + (tid >= NCH ? nsf : 0)] = sum;
    }
  }
  __syncthreads();
}

// FG: the per-group cutoff tables live in the global buffer nb.fc_scratch
// (row of atom i) instead of LDS, for neighbour lists too long for the LDS
// layout (launched with one wavefront per atom).
template <typename T, int NCH, int W, bool FG>
__global__ __launch_bounds__(256) void k_desc_fwd(int N, const int *__restrict__ species,
                                                  const double *__restrict__ wz, NeighborBufs nb, TripletBufs tb,
                                                  SFDev sf, Lds L, const double *__restrict__ theta_inv,
                                                  const double *__restrict__ mu, int split_mode, DescOut out) {
  extern __shared__ __align__(16) unsigned char smem[];
  if ((int)blockIdx.x >= N) return;  // one workgroup per atom of [atom0, atom0 + N)
  const int i = L.atom0 + blockIdx.x;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int KP = L.KP, Dp = L.Dp;
  double *s_xn = (double *)smem;  // [W]
  T *s_d = (T *)(s_xn + W);
  T *s_w = s_d + 4 * KP;
  T *s_fc = FG ? (T *)nb.fc_scratch + (size_t)i * L.G * KP : s_w + KP;
  T *s_feat = FG ? s_w + KP : s_fc + L.G * KP;
  T *s_part = s_feat + Dp;  // [W * 2 * NCH]

  const int start = i * nb.kcap;
  const int K = nb.nb_count[i];
  if (K > KP) {  // the host sizes KP from the capacity; never index LDS past it
    if (tid == 0) atomicAdd(&nb.total[3], 1);
    return;
  }
  for (int a = tid; a < K; a += 64 * W) {
    const double *dd = nb.nb_d + 4 * (size_t)(start + a);
    const double r = dd[3];
    s_d[4 * a] = (T)dd[0];
    s_d[4 * a + 1] = (T)dd[1];
    s_d[4 * a + 2] = (T)dd[2];
    s_d[4 * a + 3] = (T)r;
    s_w[a] = (T)slot_weight(nb, wz, species, start + a);
    for (int g = 0; g < sf.n_groups; ++g) {
      const double rc = sf.group_rc[g];
      s_fc[g * KP + a] = (r <= rc) ? (T)(0.5 * (cos(kPi * r / rc) + 1.0)) : T(0);
    }
  }
  for (int d = tid; d < Dp; d += 64 * W) s_feat[d] = T(0);
  __syncthreads();

  // radial lines, NCH at a time
  for (int c0 = 0; c0 < sf.n_rad; c0 += NCH) {
    T acc[NCH], accw[NCH];
#pragma unroll
    for (int l = 0; l < NCH; ++l) acc[l] = accw[l] = T(0);
    for (int a = tid; a < K; a += 64 * W) {
      const T r = s_d[4 * a + 3], w = s_w[a];
#pragma unroll
      for (int l = 0; l < NCH; ++l) {
        const int s = c0 + l;
        if (s < sf.n_rad) {
          const T dr = r - rad_table<T>(sf.rad_rs, sf.rad_rs32, s);
          const T t = Math<T>::ex(-rad_table<T>(sf.rad_a, sf.rad_a32, s) * dr * dr) * s_fc[sf.rad_group[s] * KP + a];
          acc[l] += t;
          accw[l] += t * w;
        }
      }
    }
    chunk_reduce<T, NCH, W>(acc, accw, wv, lane, tid, c0, sf.n_rad, sf.rad_feat, sf.nsf, s_part, s_feat);
  }

  // angular lines, NCH at a time, over the atom's triplet list
  const uint32_t *tl = tb.list + (size_t)i * tb.tcap;
  const int nt = tb.count[i];
  for (int c0 = 0; c0 < sf.n_ang; c0 += NCH) {
    T acc[NCH], accw[NCH];
#pragma unroll
    for (int l = 0; l < NCH; ++l) acc[l] = accw[l] = T(0);
    constexpr int ST = 64 * W;
    uint32_t qc[4], qn[4];
    load_q4(tl, nt, 64 * wv + lane, ST, qc);
    for (int t0 = 64 * wv; t0 < nt; t0 += 4 * ST) {
      load_q4(tl, nt, t0 + 4 * ST + lane, ST, qn);
#pragma unroll 1
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = pop_q4(qc);
        ang_fwd_term<T, NCH>(q, t0 + u * ST + lane < nt, c0, sf, s_d, s_w, s_fc, KP, acc, accw);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) qc[u] = qn[u];
    }
    chunk_reduce<T, NCH, W>(acc, accw, wv, lane, tid, c0, sf.n_ang, sf.ang_feat, sf.nsf, s_part, s_feat);
  }

  // epilogue: raw features, centred/scaled features, FP64 norm, split operands
  double xn = 0.0;
  const size_t row = (size_t)i * Dp;
  for (int d = tid; d < Dp; d += 64 * W) {
    const T x = s_feat[d];
    if (out.x64) out.x64[row + d] = (double)x;
    if (out.x32) out.x32[row + d] = (float)x;
    const double xt = (double)x * theta_inv[d] - mu[d];
    xn += xt * xt;
    if (out.xt64) out.xt64[row + d] = xt;
    if (out.xt32) out.xt32[row + d] = (float)xt;
    if (split_mode == 1) {  // FP16 hi + lo of xt clamped to the FP16 range (see fp16_operand)
      const double xc = fp16_operand(xt);
      const _Float16 hi = (_Float16)xc;
      const _Float16 lo = (_Float16)(xc - (double)hi);
      out.xt_hi[row + d] = __builtin_bit_cast(uint16_t, hi);
      out.xt_lo[row + d] = __builtin_bit_cast(uint16_t, lo);
    } else if (split_mode == 2) {  // BF16 hi + lo
      const __bf16 hi = (__bf16)(float)xt;
      const __bf16 lo = (__bf16)(float)(xt - (double)(float)hi);
      out.xt_hi[row + d] = __builtin_bit_cast(uint16_t, hi);
      out.xt_lo[row + d] = __builtin_bit_cast(uint16_t, lo);
    }
  }
  xn = wave_sum(xn);
  if (lane == 0) s_xn[wv] = xn;
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) v += s_xn[w];
    out.xn[i] = v;
  }
}

// ---------------------------------------------------------------------------
// grid form of the angular table: eta x (lambda = +1, -1) x zeta, zeta in
// {1, 2, 4, 8, 16}, one cutoff group (SFDev::grid).  Powers come from one
// squaring chain per lambda, (lambda, zeta) products are unrolled, and only
// the eta loop is a (uniform) runtime loop.
// ---------------------------------------------------------------------------
// With pref_k = 2^(1 - zeta_k) (the only prefactor the tables hold), the
// angular factor is pref_k b^zeta_k = 2 u^zeta_k for u = b / 2 = (1 + lam cos)
// / 2, lam = +1 (p) and -1 (m).  This returns the halves P_k = u^zeta_k and
// dP_k/dcos = lam zeta_k u^(zeta_k - 1) / 2; the callers put the factor 2 on
// their per-triplet products once.  zeta_k = 2^zlog[k] with zlog ascending:
// one squaring chain, advanced by uniform trip counts (registers only; an
// indexed table of powers would go to scratch memory).
// Z16: the zeta list is {1, 2, 4, 16} (zlog {0, 1, 2, 4}, the usual grid),
// known at compile time, so the chain is straight-line code with the same
// multiplications in the same order.
template <typename T, int NZ, bool DERIV, bool Z16>
__device__ __forceinline__ void grid_powers(T cth, const SFDev &sf, T (&Pp)[NZ], T (&Pm)[NZ], T (&dPp)[NZ],
                                            T (&dPm)[NZ]) {
  static_assert(!Z16 || NZ == 4, "the {1, 2, 4, 16} list has four zetas");
  const T hc = T(0.5) * cth;
  T p = T(0.5) + hc, m = T(0.5) - hc;  // u^(2^t)
  T qp = T(1), qm = T(1);              // u^(2^t - 1)
  int t = 0;
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    const int j = Z16 ? (k < 3 ? k : 4) : sf.zlog[k];
    if constexpr (Z16) {
#pragma unroll
      for (int u = 0; u < 4; ++u)  // at most 4 squarings per zeta; t and j are constants here
        if (t < j) {
          if (DERIV) {
            qp *= p;
            qm *= m;
          }
          p *= p;
          m *= m;
          ++t;
        }
    } else {
      for (; t < j; ++t) {
        if (DERIV) {
          qp *= p;
          qm *= m;
        }
        p *= p;
        m *= m;
      }
    }
    Pp[k] = p;
    Pm[k] = m;
    if (DERIV) {
      const T hz = T(0.5) * (T)sf.zeta_k[k];  // zeta_k / 2, exact
      dPp[k] = hz * qp;
      dPm[k] = -hz * qm;
    }
  }
}

// exp(-eta x) as the grid kernels evaluate it: in FP32 the hardware 2^x with
// the coefficient -eta log2(e) formed once per eta (one multiplication per
// term instead of two); in FP64 exp itself.
template <typename T>
__device__ __forceinline__ T eta_coef(T eta);
template <>
__device__ __forceinline__ float eta_coef<float>(float eta) { return -eta * 1.44269504088896341f; }
template <>
__device__ __forceinline__ double eta_coef<double>(double eta) { return -eta; }
__device__ __forceinline__ float exp_coef(float c, float x) { return __builtin_amdgcn_exp2f(c * x); }
__device__ __forceinline__ double exp_coef(double c, double x) { return exp(c * x); }

// Sum NV per-lane values over the workgroup: s_out[j] = total of v[j].  Within
// a wavefront a butterfly reduce-scatter: at offset m (32, 16, .., 1) every
// lane keeps the half of its live values selected by its own bit m and adds
// the partner's copy of that half, so lane l ends with the total of value l
// (NV padded to 32 or 64; for 32 the butterfly runs within each half-wave and
// the halves' totals are added last).  NV - 1 exchanges instead of NV full
// wave reductions; the offsets 32 and 16 use gfx950's lane swaps (below).
// The waves' totals are then added in wave order (deterministic).
// One butterfly level per instantiation, so every register index is a
// compile-time constant (a loop over M left rolled by the compiler would
// index the array at run time).
// gfx950's lane swaps between the halves of a wave (permlane32) or between
// the 16-lane rows of each half (permlane16), applied to a pair (a, b): a
// becomes {a, b} of the lower halves (rows 0 and 2) and b becomes {a, b} of
// the upper ones (rows 1 and 3), as the lanes see them.  Then a + b gives every
// lower lane its own a plus its partner's a, and every upper lane its own b
// plus its partner's b: one butterfly step without a shuffle and two selects.
template <int M>
__device__ __forceinline__ void swap_halves(unsigned &a, unsigned &b) {
  static_assert(M == 32 || M == 16, "lane swaps exist for offsets 32 and 16");
  if constexpr (M == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = r[0];
    b = r[1];
  }
}

template <int M>
__device__ __forceinline__ float swap_sum(float a, float b) {
  unsigned u = __float_as_uint(a), v = __float_as_uint(b);
  swap_halves<M>(u, v);
  return __uint_as_float(u) + __uint_as_float(v);
}

template <int M>
__device__ __forceinline__ double swap_sum(double a, double b) {
  unsigned long long u = __double_as_longlong(a), v = __double_as_longlong(b);
  unsigned ul = (unsigned)u, uh = (unsigned)(u >> 32), vl = (unsigned)v, vh = (unsigned)(v >> 32);
  swap_halves<M>(ul, vl);
  swap_halves<M>(uh, vh);
  return __longlong_as_double((long long)(((unsigned long long)uh << 32) | ul)) +
         __longlong_as_double((long long)(((unsigned long long)vh << 32) | vl));
}

template <typename T, int M, int NVP>
__device__ __forceinline__ void butterfly(T (&x)[NVP], int lane) {
  if constexpr (M == 32 || M == 16) {
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = swap_sum<M>(x[j], x[j + M]);
    butterfly<T, M / 2, NVP>(x, lane);
  } else if constexpr (M >= 1) {
    const bool upper = (lane & M) != 0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const T send = upper ? x[j] : x[j + M];
      const T keep = upper ? x[j + M] : x[j];
      x[j] = keep + __shfl_xor(send, M, 64);
    }
    butterfly<T, M / 2, NVP>(x, lane);
  }
}

template <typename T, int NV, int W>
__device__ __forceinline__ void group_sum(T (&v)[NV], int wv, int lane, int tid, T *s_part, T *s_out) {
  constexpr int NVP = NV <= 32 ? 32 : 64;
  static_assert(NV <= 64, "at most 64 values");
  T x[NVP];
#pragma unroll
  for (int j = 0; j < NVP; ++j) x[j] = (j < NV) ? v[j] : T(0);
  butterfly<T, NVP / 2, NVP>(x, lane);
  // 32 values: the butterfly (offsets 16 .. 1) stays within each half of
  // the wave; lanes l and l + 32 then add their halves' totals of value l
  if (NVP == 32) x[0] += __shfl_xor(x[0], 32, 64);
  if (W == 1) {
    if (lane < NV) s_out[lane] = x[0];
    __syncthreads();
    return;
  }
  if (lane < NV) s_part[wv * NV + lane] = x[0];
  __syncthreads();
  for (int j = tid; j < NV; j += 64 * W) {
    T sum = T(0);
#pragma unroll
    for (int w = 0; w < W; ++w) sum += s_part[w * NV + j];
    s_out[j] = sum;
  }
  __syncthreads();
}

template <typename T, int NZ>
struct GridShape {
  static constexpr int CAP = 32;  // (plain, weighted) pairs per pass
  static constexpr int NE = (CAP / (2 * NZ)) > 0 ? CAP / (2 * NZ) : 1;  // etas per pass
  static constexpr int NV = NE * 2 * NZ * 2;
};

template <typename T, int NZ, int W, bool Z16, bool FG>
__global__ __launch_bounds__(256) void k_desc_fwd_grid(int N, const int *__restrict__ species,
                                                       const double *__restrict__ wz, NeighborBufs nb,
                                                       TripletBufs tb, SFDev sf, Lds L,
                                                       const double *__restrict__ theta_inv,
                                                       const double *__restrict__ mu, int split_mode, DescOut out) {
  constexpr int NE = GridShape<T, NZ>::NE, NV = GridShape<T, NZ>::NV;
  constexpr int RCH = 16;
  extern __shared__ __align__(16) unsigned char smem[];
  if ((int)blockIdx.x >= N) return;  // one workgroup per atom of [atom0, atom0 + N)
  const int i = L.atom0 + blockIdx.x;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int KP = L.KP, Dp = L.Dp;
  double *s_xn = (double *)smem;  // [W], padded to 16 bytes
  T *s_d = (T *)(s_xn + ((W + 1) & ~1));
  T *s_rw = s_d + 4 * KP;       // [KP] pairs (1 / r, species weight)
  T *s_fc = FG ? (T *)nb.fc_scratch + (size_t)i * L.G * KP : s_rw + 2 * KP;
  T *s_feat = FG ? s_rw + 2 * KP : s_fc + L.G * KP;
  T *s_part = s_feat + Dp;      // [W * NV]
  T *s_sum = s_part + W * NV;   // [NV]

  const int start = i * nb.kcap;
  const int K = nb.nb_count[i];
  if (K > KP) {
    if (tid == 0) atomicAdd(&nb.total[3], 1);
    return;
  }
  for (int a = tid; a < K; a += 64 * W) {
    const double *dd = nb.nb_d + 4 * (size_t)(start + a);
    const double r = dd[3];
    s_d[4 * a] = (T)dd[0];
    s_d[4 * a + 1] = (T)dd[1];
    s_d[4 * a + 2] = (T)dd[2];
    s_d[4 * a + 3] = (T)r;
    s_rw[2 * a] = (T)(1.0 / r);
    s_rw[2 * a + 1] = (T)slot_weight(nb, wz, species, start + a);
    for (int g = 0; g < sf.n_groups; ++g) {
      const double rc = sf.group_rc[g];
      s_fc[g * KP + a] = (r <= rc) ? (T)(0.5 * (cos(kPi * r / rc) + 1.0)) : T(0);
    }
  }
  for (int d = tid; d < Dp; d += 64 * W) s_feat[d] = T(0);
  __syncthreads();

  // radial lines, RCH at a time: plain sums in acc[0, RCH), weighted ones in
  // acc[RCH, 2 RCH), reduced over the group with the angular pass's exchange
  static_assert(2 * RCH <= NV, "radial sums reuse the angular reduction buffers");
  for (int c0 = 0; c0 < sf.n_rad; c0 += RCH) {
    T acc[2 * RCH];
#pragma unroll
    for (int l = 0; l < 2 * RCH; ++l) acc[l] = T(0);
    for (int a = tid; a < K; a += 64 * W) {
      const T r = s_d[4 * a + 3], w = s_rw[2 * a + 1];
#pragma unroll
      for (int l = 0; l < RCH; ++l) {
        const int s = c0 + l;
        if (s < sf.n_rad) {
          const T dr = r - rad_table<T>(sf.rad_rs, sf.rad_rs32, s);
          const T t = Math<T>::ex(-rad_table<T>(sf.rad_a, sf.rad_a32, s) * dr * dr) * s_fc[sf.rad_group[s] * KP + a];
          acc[l] += t;
          acc[RCH + l] += t * w;
        }
      }
    }
    group_sum<T, 2 * RCH, W>(acc, wv, lane, tid, s_part, s_sum);
    for (int j = tid; j < 2 * RCH; j += 64 * W) {
      const int l = j % RCH;
      if (c0 + l < sf.n_rad) s_feat[sf.rad_feat[c0 + l] + (j >= RCH ? sf.nsf : 0)] = s_sum[j];
    }
    __syncthreads();
  }

  // angular lines in grid form, NE etas per pass
  const uint32_t *tl = tb.list + (size_t)i * tb.tcap;
  const int nt = tb.count[i];
  const int g = sf.g_ang;
  const T rc = (T)sf.group_rc[g];
  const T pirc = (T)(kPi / sf.group_rc[g]);
  for (int h0 = 0; h0 < sf.n_eta; h0 += NE) {
    T acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = T(0);
    T ceta_h[NE];  // this pass's exp coefficients (uniform), formed once
#pragma unroll
    for (int hh = 0; hh < NE; ++hh)
      ceta_h[hh] = eta_coef<T>((h0 + hh < sf.n_eta) ? (T)sf.ang_eta[(h0 + hh) * 2 * NZ] : T(0));
    constexpr int ST = 64 * W;
    uint32_t qc[4], qn[4];
    load_q4(tl, nt, 64 * wv + lane, ST, qc);
    for (int t00 = 64 * wv; t00 < nt; t00 += 4 * ST) {
      load_q4(tl, nt, t00 + 4 * ST + lane, ST, qn);
#pragma unroll 1
      for (int u = 0; u < 4; ++u) {
        const bool act = t00 + u * ST + lane < nt;
        const uint32_t q = pop_q4(qc);
        const int a = q & 0xffff, b = q >> 16;
        const T ax = s_d[4 * a], ay = s_d[4 * a + 1], az = s_d[4 * a + 2], ra = s_d[4 * a + 3];
        const T bx = s_d[4 * b], by = s_d[4 * b + 1], bz = s_d[4 * b + 2], rb = s_d[4 * b + 3];
        const T ex = bx - ax, ey = by - ay, ez = bz - az;
        const T e2 = ex * ex + ey * ey + ez * ez;
        const T e = Math<T>::sq(e2);
        typedef typename Pair<T>::type P2;
        const P2 rwa = ((const P2 *)s_rw)[a], rwb = ((const P2 *)s_rw)[b];  // (1/r, w)
        const T cth = (ra * ra + rb * rb - e2) * (T(0.5) * rwa.x * rwb.x);
        const T S2 = ra * ra + rb * rb + e2;
        const T wab = rwa.y * rwb.y;
        // 2 fc(r_jk): the factor 2 of the halved powers (grid_powers) rides here
        const T fe2 = (e <= rc) ? Math<T>::cs(pirc * e) + T(1) : T(0);
        const T F3 = act ? s_fc[g * KP + a] * s_fc[g * KP + b] * fe2 : T(0);
        T Pp[NZ], Pm[NZ], dPp[NZ], dPm[NZ];
        grid_powers<T, NZ, false, Z16>(cth, sf, Pp, Pm, dPp, dPm);
#pragma unroll
        for (int hh = 0; hh < NE; ++hh) {
          const int h = h0 + hh;
          if (h < sf.n_eta) {
            // plain and weighted sums as four FMAs per zeta (E w_ab formed once)
            const T E = exp_coef(ceta_h[hh], S2) * F3, Ew = E * wab;
#pragma unroll
            for (int k = 0; k < NZ; ++k) {
              acc[(hh * 2 + 0) * NZ + k] += E * Pp[k];
              acc[(hh * 2 + 1) * NZ + k] += E * Pm[k];
              acc[NV / 2 + (hh * 2 + 0) * NZ + k] += Ew * Pp[k];
              acc[NV / 2 + (hh * 2 + 1) * NZ + k] += Ew * Pm[k];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) qc[u] = qn[u];
    }
    group_sum<T, NV, W>(acc, wv, lane, tid, s_part, s_sum);
    for (int j = tid; j < NV; j += 64 * W) {
      const int jl = j % (NV / 2);            // (hh, li, k)
      const int h = h0 + jl / (2 * NZ);
      if (h < sf.n_eta) s_feat[sf.ang_feat[h0 * 2 * NZ + jl] + (j >= NV / 2 ? sf.nsf : 0)] = s_sum[j];
    }
    __syncthreads();
  }

  // epilogue
  double xn = 0.0;
  const size_t row = (size_t)i * Dp;
  for (int d = tid; d < Dp; d += 64 * W) {
    const T x = s_feat[d];
    if (out.x64) out.x64[row + d] = (double)x;
    if (out.x32) out.x32[row + d] = (float)x;
    const double xt = (double)x * theta_inv[d] - mu[d];
    xn += xt * xt;
    if (out.xt64) out.xt64[row + d] = xt;
    if (out.xt32) out.xt32[row + d] = (float)xt;
    if (split_mode == 1) {
      const double xc = fp16_operand(xt);
      const _Float16 hi = (_Float16)xc;
      const _Float16 lo = (_Float16)(xc - (double)hi);
      out.xt_hi[row + d] = __builtin_bit_cast(uint16_t, hi);
      out.xt_lo[row + d] = __builtin_bit_cast(uint16_t, lo);
    } else if (split_mode == 2) {
      const __bf16 hi = (__bf16)(float)xt;
      const __bf16 lo = (__bf16)(float)(xt - (double)(float)hi);
      out.xt_hi[row + d] = __builtin_bit_cast(uint16_t, hi);
      out.xt_lo[row + d] = __builtin_bit_cast(uint16_t, lo);
    }
  }
  xn = wave_sum(xn);
  if (lane == 0) s_xn[wv] = xn;
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) v += s_xn[w];
    out.xn[i] = v;
  }
}

// Values per eta record of the grid-form backward (see k_desc_bwd): 4 NZ g
// values, eta and its exp coefficient, rounded up to whole 16-byte groups.
template <int NZ>
struct EtaRec {
  static constexpr int RS = NZ > 0 ? (4 * NZ + 2 + 3) & ~3 : 0;
};

template <typename T, int NZ, bool Z16>
__device__ __forceinline__ void ang_bwd_grid(uint32_t q, bool active, const SFDev &sf, const T *s_d, const T *s_rw,
                                             const T *s_fd, const T *s_er, int KP, T rc, T pirc, double *s_T,
                                             double *s_F) {
  if (!active) return;
  const int a = q & 0xffff, b = q >> 16;
  const T ax = s_d[4 * a], ay = s_d[4 * a + 1], az = s_d[4 * a + 2], ra = s_d[4 * a + 3];
  const T bx = s_d[4 * b], by = s_d[4 * b + 1], bz = s_d[4 * b + 2], rb = s_d[4 * b + 3];
  typedef typename Pair<T>::type P2;
  const P2 rwa = ((const P2 *)s_rw)[a], rwb = ((const P2 *)s_rw)[b];  // (1/r, w)
  const T ria = rwa.x, rib = rwb.x;
  const T ex = bx - ax, ey = by - ay, ez = bz - az;
  const T e2 = ex * ex + ey * ey + ez * ez;
  const T e = Math<T>::sq(e2);
  const T inv2ab = T(0.5) * ria * rib;
  const T cth = (ra * ra + rb * rb - e2) * inv2ab;
  const T S2 = ra * ra + rb * rb + e2;
  const T dca = rib - cth * ria, dcb = ria - cth * rib, dce = -e * (T(2) * inv2ab);
  const T wab = rwa.y * rwb.y;
  const int g = sf.g_ang;
  const P2 fda = ((const P2 *)s_fd)[g * KP + a], fdb = ((const P2 *)s_fd)[g * KP + b];  // (fc, dfc)
  const T fa = fda.x, dfa = fda.y, fb = fdb.x, dfb = fdb.y;
  const bool in = e <= rc;
  T sine, cosine;
  Math<T>::sc(pirc * e, sine, cosine);
  // 2 fc(r_jk) and its derivative: every term below carries exactly one of
  // them, which puts the factor 2 of the halved powers (grid_powers) on all
  const T fe = in ? cosine + T(1) : T(0);
  const T dfe = in ? -sine * pirc : T(0);
  T Pp[NZ], Pm[NZ], dPp[NZ], dPm[NZ];
  grid_powers<T, NZ, true, Z16>(cth, sf, Pp, Pm, dPp, dPm);
  T SA = 0, SB = 0, SC = 0;
  // one eta: the lines' g contracted with the powers (sb) and with their
  // cos-derivatives (sa), each sum started from its first term
  auto eta_terms = [&](int h) {
    const T *rec = s_er + h * EtaRec<NZ>::RS;
    const T eta = rec[4 * NZ];
    const T E = exp_coef(rec[4 * NZ + 1], S2);
    T gA[2 * NZ], gW[2 * NZ];
    if constexpr (sizeof(T) == 4 && NZ % 2 == 0) {
      // 16-byte reads: records and both blocks are aligned
      const float4 *a4 = (const float4 *)rec, *w4 = (const float4 *)(rec + 2 * NZ);
#pragma unroll
      for (int v = 0; v < NZ / 2; ++v) {
        const float4 x = a4[v], y = w4[v];
        gA[4 * v] = x.x, gA[4 * v + 1] = x.y, gA[4 * v + 2] = x.z, gA[4 * v + 3] = x.w;
        gW[4 * v] = y.x, gW[4 * v + 1] = y.y, gW[4 * v + 2] = y.z, gW[4 * v + 3] = y.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 2 * NZ; ++j) {
        gA[j] = rec[j];
        gW[j] = rec[2 * NZ + j];
      }
    }
    // one fused multiply-add per term (a sum of term pairs would cost a
    // multiply, an fma and an add per pair)
    const T gp0 = gA[0] + wab * gW[0], gm0 = gA[NZ] + wab * gW[NZ];
    T sa = gp0 * dPp[0];
    T sb = gp0 * Pp[0];
    sa = fma(gm0, dPm[0], sa);
    sb = fma(gm0, Pm[0], sb);
#pragma unroll
    for (int k = 1; k < NZ; ++k) {
      const T gp = gA[k] + wab * gW[k];
      const T gm = gA[NZ + k] + wab * gW[NZ + k];
      sa = fma(gp, dPp[k], sa);
      sb = fma(gp, Pp[k], sb);
      sa = fma(gm, dPm[k], sa);
      sb = fma(gm, Pm[k], sb);
    }
    SA += sa * E;
    SB += sb * E;
    SC += eta * sb * E;
  };
  for (int h = 0; h < sf.n_eta; ++h) eta_terms(h);
  const T F = fa * fb * fe;
  const T Ta = F * SA * dca - T(2) * ra * F * SC + SB * dfa * fb * fe;
  const T Tb = F * SA * dcb - T(2) * rb * F * SC + SB * fa * dfb * fe;
  const T Te = F * SA * dce - T(2) * e * F * SC + SB * fa * fb * dfe;
  atomicAdd(&s_T[a], (double)Ta);
  atomicAdd(&s_T[b], (double)Tb);
  if (e > T(0)) {
    const T sc = Te * Math<T>::rcp(e);
    const T ux = sc * ex, uy = sc * ey, uz = sc * ez;
    atomicAdd(&s_F[3 * b], (double)ux);
    atomicAdd(&s_F[3 * b + 1], (double)uy);
    atomicAdd(&s_F[3 * b + 2], (double)uz);
    atomicAdd(&s_F[3 * a], (double)-ux);
    atomicAdd(&s_F[3 * a + 1], (double)-uy);
    atomicAdd(&s_F[3 * a + 2], (double)-uz);
  }
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void ang_bwd_term(uint32_t q, bool active, const SFDev &sf, const T *s_d,
                                             const T *s_rw, const T *s_fd, const T *s_gA, const T *s_gW, int KP,
                                             double *s_T, double *s_F) {
  if (!active) return;
  const int a = q & 0xffff, b = q >> 16;
  const T ax = s_d[4 * a], ay = s_d[4 * a + 1], az = s_d[4 * a + 2], ra = s_d[4 * a + 3];
  const T bx = s_d[4 * b], by = s_d[4 * b + 1], bz = s_d[4 * b + 2], rb = s_d[4 * b + 3];
  const T ex = bx - ax, ey = by - ay, ez = bz - az;
  const T e2 = ex * ex + ey * ey + ez * ez;
  const T e = Math<T>::sq(e2);
  const T inv2ab = T(1) / (T(2) * ra * rb);
  const T cth = (ra * ra + rb * rb - e2) * inv2ab;
  const T S2 = ra * ra + rb * rb + e2;
  const T dca = T(1) / rb - cth / ra, dcb = T(1) / ra - cth / rb, dce = -e * (T(2) * inv2ab);
  const T wab = s_rw[2 * a + 1] * s_rw[2 * b + 1];
  T Ta = 0, Tb = 0, Te = 0;
  T SA = 0, SB = 0, SC = 0;
  T fa = 0, dfa = 0, fb = 0, dfb = 0, fe = 0, dfe = 0;
  int cur_g = -1;
  T E = 0, cur_eta = T(-1);
  PowState<T> ps;
  auto flush = [&]() {
    const T F = fa * fb * fe;
    Ta += F * SA * dca - T(2) * ra * F * SC + SB * dfa * fb * fe;
    Tb += F * SA * dcb - T(2) * rb * F * SC + SB * fa * dfb * fe;
    Te += F * SA * dce - T(2) * e * F * SC + SB * fa * fb * dfe;
  };
  for (int s = 0; s < sf.n_ang; ++s) {
    const int g = sf.ang_group[s];
    if (g != cur_g) {
      if (cur_g >= 0) flush();
      cur_g = g;
      SA = SB = SC = T(0);
      const T rc = (T)sf.group_rc[g];
      fa = s_fd[2 * (g * KP + a)];
      dfa = s_fd[2 * (g * KP + a) + 1];
      fb = s_fd[2 * (g * KP + b)];
      dfb = s_fd[2 * (g * KP + b) + 1];
      const bool in = e <= rc;
      T sine, cosine;
      Math<T>::sc((T)kPi * e / rc, sine, cosine);
      fe = in ? T(0.5) * (cosine + T(1)) : T(0);
      dfe = in ? T(0.5) * (-sine) * ((T)kPi / rc) : T(0);
    }
    const T eta = (T)sf.ang_eta[s];
    if (eta != cur_eta) {
      cur_eta = eta;
      E = Math<T>::ex(-eta * S2);
    }
    const T gE = (s_gA[s] + wab * s_gW[s]) * E;
    T P, dP;
    line_power<T>(sf, s, cth, ps, P, dP);
    const T B = gE * P;
    SA += gE * dP;
    SB += B;
    SC += eta * B;
  }
  if (cur_g >= 0) flush();
  atomicAdd(&s_T[a], (double)Ta);
  atomicAdd(&s_T[b], (double)Tb);
  if (e > T(0)) {
    const T s = Te / e;
    const T ux = s * ex, uy = s * ey, uz = s * ez;
    atomicAdd(&s_F[3 * b], (double)ux);
    atomicAdd(&s_F[3 * b + 1], (double)uy);
    atomicAdd(&s_F[3 * b + 2], (double)uz);
    atomicAdd(&s_F[3 * a], (double)-ux);
    atomicAdd(&s_F[3 * a + 1], (double)-uy);
    atomicAdd(&s_F[3 * a + 2], (double)-uz);
  }
}

template <typename T, int W, int GNZ, bool Z16, bool FG>
__global__ __launch_bounds__(256) void k_desc_bwd(int N, const int *__restrict__ species,
                                                  const double *__restrict__ wz, NeighborBufs nb, TripletBufs tb,
                                                  SFDev sf, Lds L, GradSrc gs, double *__restrict__ slot_f,
                                                  double *__restrict__ atom_vir) {
  extern __shared__ __align__(16) unsigned char smem[];
  if ((int)blockIdx.x >= N) return;  // one workgroup per atom of [atom0, atom0 + N)
  const int i = L.atom0 + blockIdx.x;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int KP = L.KP, Dp = L.Dp, G = L.G;
  // per-wave accumulators in FP64 for both precisions (LDS FP64 atomics are
  // also the faster ones here), summed in wave order at the end: deterministic
  double *s_vir = (double *)smem;  // [W * 6]
  double *s_T = s_vir + 6 * W;     // [W][KP]
  double *s_F = s_T + W * KP;      // [W][3 KP]
  T *s_d = (T *)(s_F + 3 * W * KP);
  T *s_rw = s_d + 4 * KP;    // [KP] pairs (1 / r, species weight)
  // [G][KP] pairs (fc, dfc/dr) per cutoff group, in LDS or (FG) in the
  // atom's row of the global buffer
  T *s_fd = FG ? (T *)nb.fc_scratch + (size_t)i * 2 * G * KP : s_rw + 2 * KP;
  T *s_g = FG ? s_rw + 2 * KP : s_fd + 2 * G * KP;
  // generic form: per angular line, g of the plain and of the weighted
  // feature; s_gW starts on a 4-value boundary (n_ang <= Dp / 2)
  T *s_gA = s_g + Dp;
  T *s_gW = s_gA + ((sf.n_ang + 3) & ~3);
  // grid form: one record per eta, [g plain (2 NZ) | g weighted (2 NZ) | eta |
  // exp coefficient | pad], so that a pass reads its eta's values from one
  // base address; 16-byte aligned records (every region above has a multiple
  // of 4 values), at most 2 Dp values (n_eta <= Dp / (4 NZ))
  constexpr int RS = EtaRec<GNZ>::RS;
  T *s_er = s_g + Dp;

  const int start = i * nb.kcap;
  const int K = nb.nb_count[i];
  if (K > KP) {
    if (tid == 0) atomicAdd(&nb.total[3], 1);
    return;
  }
  for (int a = tid; a < K; a += 64 * W) {
    const double *dd = nb.nb_d + 4 * (size_t)(start + a);
    const double r = dd[3];
    s_d[4 * a] = (T)dd[0];
    s_d[4 * a + 1] = (T)dd[1];
    s_d[4 * a + 2] = (T)dd[2];
    s_d[4 * a + 3] = (T)r;
    s_rw[2 * a] = (T)(1.0 / r);
    s_rw[2 * a + 1] = (T)slot_weight(nb, wz, species, start + a);
    for (int g = 0; g < sf.n_groups; ++g) {
      const double rc = sf.group_rc[g];
      const bool in = r <= rc;
      s_fd[2 * (g * KP + a)] = in ? (T)(0.5 * (cos(kPi * r / rc) + 1.0)) : T(0);
      s_fd[2 * (g * KP + a) + 1] = in ? (T)(0.5 * (-sin(kPi * r / rc)) * (kPi / rc)) : T(0);
    }
  }
  for (int k = tid; k < 4 * W * KP; k += 64 * W) s_T[k] = 0.0;  // s_T and s_F
  if (GNZ > 0)
    for (int h = tid; h < sf.n_eta; h += 64 * W) {
      s_er[h * RS + 4 * GNZ] = (T)sf.ang_eta[h * 2 * GNZ];
      s_er[h * RS + 4 * GNZ + 1] = eta_coef<T>((T)sf.ang_eta[h * 2 * GNZ]);
    }
  const size_t row = (size_t)i * Dp;
  if (gs.splits > 0) {
    // the contraction's finish for this atom (same order as k_gpr_finish)
    const double E = split_sum(gs.splits, [&](int sp) { return gs.Epart[(size_t)sp * gs.Np + i]; });
    if (tid == 0) gs.eatom[i] = E;
    for (int d = tid; d < Dp; d += 64 * W) {
      const double G =
          gs.Gpart ? split_sum(gs.splits, [&](int sp) { return gs.Gpart[((size_t)sp * gs.Np + i) * Dp + d]; })
                   : split_sum(gs.splits,
                               [&](int sp) { return (double)gs.Gpart32[((size_t)sp * gs.Np + i) * Dp + d]; });
      const double xt = gs.xt64 ? gs.xt64[row + d] : (double)gs.xt32[row + d];
      const double g = -(xt * E - G) * gs.theta_inv[d];
      if (gs.g64_out) {
        gs.g64_out[row + d] = g;
        s_g[d] = (T)g;
      } else {
        const float gf = (float)g;
        gs.g32_out[row + d] = gf;
        s_g[d] = (T)gf;
      }
    }
  } else {
    for (int d = tid; d < Dp; d += 64 * W) s_g[d] = gs.g64 ? (T)gs.g64[row + d] : (T)gs.g32[row + d];
  }
  __syncthreads();
  for (int s = tid; s < sf.n_ang; s += 64 * W) {
    const int f = sf.ang_feat[s];
    if constexpr (GNZ > 0) {
      const int h = s / (2 * GNZ), k = s - h * 2 * GNZ;
      s_er[h * RS + k] = s_g[f];
      s_er[h * RS + 2 * GNZ + k] = s_g[f + sf.nsf];
    } else {
      s_gA[s] = s_g[f];
      s_gW[s] = s_g[f + sf.nsf];
    }
  }
  __syncthreads();

  // radial: one neighbour per thread
  for (int a = tid; a < K; a += 64 * W) {
    const T r = s_d[4 * a + 3], w = s_rw[2 * a + 1];
    T acc = 0;
    for (int s = 0; s < sf.n_rad; ++s) {
      const int f = sf.rad_feat[s];
      const int g = sf.rad_group[s];
      const T A = rad_table<T>(sf.rad_a, sf.rad_a32, s);
      const T dr = r - rad_table<T>(sf.rad_rs, sf.rad_rs32, s);
      const T geff = s_g[f] + w * s_g[f + sf.nsf];
      acc += geff * Math<T>::ex(-A * dr * dr) * (T(-2) * A * dr * s_fd[2 * (g * KP + a)] + s_fd[2 * (g * KP + a) + 1]);
    }
    s_T[a] = (double)acc;  // wave 0's row
  }
  __syncthreads();


  if (sf.n_ang > 0) {
    const uint32_t *tl = tb.list + (size_t)i * tb.tcap;
    const int nt = tb.count[i];
    const T rcA = (T)sf.group_rc[sf.g_ang], pirc = (T)(kPi / sf.group_rc[sf.g_ang]);
    constexpr int ST = 64 * W;
    uint32_t qc[4], qn[4];
    load_q4(tl, nt, 64 * wv + lane, ST, qc);
    for (int t0 = 64 * wv; t0 < nt; t0 += 4 * ST) {
      load_q4(tl, nt, t0 + 4 * ST + lane, ST, qn);
#pragma unroll 1
      for (int u = 0; u < 4; ++u) {
        const bool act = t0 + u * ST + lane < nt;
        const uint32_t q = pop_q4(qc);
        if constexpr (GNZ > 0)
          ang_bwd_grid<T, GNZ, Z16>(q, act, sf, s_d, s_rw, s_fd, s_er, KP, rcA, pirc, s_T + wv * KP,
                                    s_F + 3 * wv * KP);
        else
          ang_bwd_term<T>(q, act, sf, s_d, s_rw, s_fd, s_gA, s_gW, KP, s_T + wv * KP, s_F + 3 * wv * KP);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) qc[u] = qn[u];
    }
  }
  __syncthreads();

  // per-slot force contributions dE_i/dR(slot atom) and the virial: the
  // radial part T d d^T / r plus F_slot d^T for the r_jk forces (each triplet's
  // u (r_k - r_j)^T is u r_k^T - u r_j^T, so summing F_slot d^T over the slots
  // gives the triplet sum with one product per slot instead of per triplet)
  double vir[6] = {0, 0, 0, 0, 0, 0};
  for (int a = tid; a < K; a += 64 * W) {
    const T r = s_d[4 * a + 3];
    double Ts = 0.0, Fx = 0.0, Fy = 0.0, Fz = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      Ts += s_T[w * KP + a];
      Fx += s_F[3 * (w * KP + a)];
      Fy += s_F[3 * (w * KP + a) + 1];
      Fz += s_F[3 * (w * KP + a) + 2];
    }
    const double sc = Ts / (double)r;
    const double dx = s_d[4 * a], dy = s_d[4 * a + 1], dz = s_d[4 * a + 2];
    if constexpr (sizeof(T) == 4) {
      // the FP32 kernels (mixed runs) store the slot forces in FP32: half the
      // bytes written here and read by the force gather
      float *o = (float *)slot_f + 3 * (size_t)(start + a);
      o[0] = (float)(sc * dx + Fx);
      o[1] = (float)(sc * dy + Fy);
      o[2] = (float)(sc * dz + Fz);
    } else {
      double *o = slot_f + 3 * (size_t)(start + a);
      o[0] = sc * dx + Fx;
      o[1] = sc * dy + Fy;
      o[2] = sc * dz + Fz;
    }
    vir[0] += sc * dx * dx + Fx * dx;
    vir[1] += sc * dy * dy + Fy * dy;
    vir[2] += sc * dz * dz + Fz * dz;
    vir[3] += sc * dx * dy + Fx * dy;
    vir[4] += sc * dy * dz + Fy * dz;
    vir[5] += sc * dx * dz + Fx * dz;
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double v = wave_sum(vir[c]);
    if (lane == 0) s_vir[6 * wv + c] = v;
  }
  __syncthreads();
  if (tid < 6) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) v += s_vir[6 * w + tid];
    atom_vir[6 * (size_t)i + tid] = v;
  }
}

// LDS bytes of each kernel's layout; FG: the cutoff tables in global memory.
template <typename T, int NCH, int W, bool FG>
size_t fwd_lds(const Lds &L) {
  return W * sizeof(double) + (size_t)(4 * L.KP + L.KP + (FG ? 0 : L.G * L.KP) + L.Dp + W * 2 * NCH) * sizeof(T);
}

template <typename T, int W, bool FG>
size_t bwd_lds(const Lds &L) {
  return (6 * W + 4 * W * (size_t)L.KP) * sizeof(double) +
         (size_t)(6 * L.KP + (FG ? 0 : 2 * L.G * L.KP) + 3 * L.Dp) * sizeof(T);
}

template <typename T, int NZ, int W, bool FG>
size_t fwd_grid_lds(const Lds &L) {
  constexpr int NV = GridShape<T, NZ>::NV;
  return ((W + 1) & ~1) * sizeof(double) +
         (size_t)(6 * L.KP + (FG ? 0 : L.G * L.KP) + L.Dp + W * NV + NV) * sizeof(T);
}

// Layouts with several wavefronts per atom stay within 64 KB, so that several
// workgroups share a CU.  A one-wave layout may take the device's whole
// per-workgroup LDS (160 KB on gfx950) for long neighbour lists; past that the
// cutoff tables move to global memory (FG).
constexpr size_t kLdsBudget = 64 * 1024;

// Per-workgroup LDS limit of the current device, queried once per device.
size_t lds_limit() {
  static std::atomic<long long> cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return kLdsBudget;
  long long v = cache[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    int b = 0;
    if (hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        b < (int)kLdsBudget)
      b = (int)kLdsBudget;
    v = b;
    cache[dev].store(v, std::memory_order_relaxed);
  }
  return (size_t)v;
}

// A kernel launched with more than 64 KB of dynamic LDS is first allowed the
// device's limit (once per kernel and device, outside any stream capture:
// the eager run that precedes a capture makes the same launches).
hipError_t allow_lds(const void *kernel, size_t bytes) {
  if (bytes <= kLdsBudget) return hipSuccess;
  static std::mutex mu;
  static std::set<std::pair<int, const void *>> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(mu);
  if (done.count({dev, kernel})) return hipSuccess;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit());
  if (e == hipSuccess) done.insert({dev, kernel});
  return e;
}

// Launch shape of a descriptor kernel: the most wavefronts per atom (up to
// `waves`, at most `wmax`) whose layout fits 64 KB, else one wavefront within
// the device's limit, else one wavefront with the cutoff tables in global
// memory.  bytes(W, fg) is the layout's LDS size.  False if nothing fits
// (the host caps the list length so that the last form always does).
template <typename F>
bool pick_shape(int waves, int wmax, F bytes, int &W, bool &fg) {
  fg = false;
  for (int w = 4; w >= 2; w /= 2)
    if (w <= wmax && waves >= w && bytes(w, false) <= kLdsBudget) {
      W = w;
      return true;
    }
  W = 1;
  if (bytes(1, false) <= lds_limit()) return true;
  fg = true;
  return bytes(1, true) <= lds_limit();
}

template <typename T, int NCH, int W, bool FG>
hipError_t fwd_launch(hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                      const TripletBufs &tb, const SFDev &sf, Lds L, const double *theta_inv, const double *mu,
                      int split_mode, DescOut out) {
  const size_t lds = fwd_lds<T, NCH, W, FG>(L);
  const hipError_t e = allow_lds((const void *)k_desc_fwd<T, NCH, W, FG>, lds);
  if (e != hipSuccess) return e;
  k_desc_fwd<T, NCH, W, FG><<<N, 64 * W, lds, s>>>(N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  return hipGetLastError();
}

template <typename T, int NCH>
hipError_t fwd_waves(int waves, hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                     const TripletBufs &tb, const SFDev &sf, Lds L, const double *theta_inv, const double *mu,
                     int split_mode, DescOut out) {
  int W;
  bool fg;
  auto bytes = [&](int w, bool g) {
    if (w == 4) return fwd_lds<T, NCH, 4, false>(L);
    if (w == 2) return fwd_lds<T, NCH, 2, false>(L);
    return g ? fwd_lds<T, NCH, 1, true>(L) : fwd_lds<T, NCH, 1, false>(L);
  };
  if (!pick_shape(waves, 4, bytes, W, fg) || (fg && !nb.fc_scratch)) return hipErrorInvalidValue;
  if (fg) return fwd_launch<T, NCH, 1, true>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  if (W == 1) return fwd_launch<T, NCH, 1, false>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  if (W == 2) return fwd_launch<T, NCH, 2, false>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  return fwd_launch<T, NCH, 4, false>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
}

// zeta list {1, 2, 4, 16}: the grid kernels' specialisation with a compile-time chain
bool zeta_1_2_4_16(const SFDev &sf) {
  return sf.nz == 4 && sf.zlog[0] == 0 && sf.zlog[1] == 1 && sf.zlog[2] == 2 && sf.zlog[3] == 4;
}

template <typename T, int W, int GNZ, bool Z16, bool FG>
hipError_t bwd_launch(hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                      const TripletBufs &tb, const SFDev &sf, Lds L, const GradSrc &gs, double *slot_f,
                      double *atom_vir) {
  const size_t lds = bwd_lds<T, W, FG>(L);
  const hipError_t e = allow_lds((const void *)k_desc_bwd<T, W, GNZ, Z16, FG>, lds);
  if (e != hipSuccess) return e;
  k_desc_bwd<T, W, GNZ, Z16, FG><<<N, 64 * W, lds, s>>>(N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  return hipGetLastError();
}

template <typename T, int W, bool FG>
hipError_t bwd_grid(hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                    const TripletBufs &tb, const SFDev &sf, Lds L, const GradSrc &gs, double *slot_f,
                    double *atom_vir) {
  if (!sf.grid) return bwd_launch<T, W, 0, false, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  switch (sf.nz) {
    case 1: return bwd_launch<T, W, 1, false, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
    case 2: return bwd_launch<T, W, 2, false, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
    case 3: return bwd_launch<T, W, 3, false, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
    case 4:
      if (zeta_1_2_4_16(sf))
        return bwd_launch<T, W, 4, true, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
      return bwd_launch<T, W, 4, false, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
    default: return bwd_launch<T, W, 5, false, FG>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  }
}

template <typename T, int NZ, int W, bool Z16, bool FG>
hipError_t fwd_grid_launch(hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                           const TripletBufs &tb, const SFDev &sf, Lds L, const double *theta_inv, const double *mu,
                           int split_mode, DescOut out) {
  const size_t lds = fwd_grid_lds<T, NZ, W, FG>(L);
  const hipError_t e = allow_lds((const void *)k_desc_fwd_grid<T, NZ, W, Z16, FG>, lds);
  if (e != hipSuccess) return e;
  k_desc_fwd_grid<T, NZ, W, Z16, FG><<<N, 64 * W, lds, s>>>(N, species, wz, nb, tb, sf, L, theta_inv, mu,
                                                            split_mode, out);
  return hipGetLastError();
}

template <typename T, int NZ, bool Z>
hipError_t fwd_grid_nz(int waves, hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                       const TripletBufs &tb, const SFDev &sf, Lds L, const double *theta_inv, const double *mu,
                       int split_mode, DescOut out) {
  int W;
  bool fg;
  auto bytes = [&](int w, bool g) {
    if (w == 2) return fwd_grid_lds<T, NZ, 2, false>(L);
    return g ? fwd_grid_lds<T, NZ, 1, true>(L) : fwd_grid_lds<T, NZ, 1, false>(L);
  };
  if (!pick_shape(waves, 2, bytes, W, fg) || (fg && !nb.fc_scratch)) return hipErrorInvalidValue;
  if (fg) return fwd_grid_launch<T, NZ, 1, Z, true>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  if (W == 2)
    return fwd_grid_launch<T, NZ, 2, Z, false>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  return fwd_grid_launch<T, NZ, 1, Z, false>(s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
}

template <typename T>
hipError_t fwd_grid(int waves, hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                    const TripletBufs &tb, const SFDev &sf, Lds L, const double *theta_inv, const double *mu,
                    int split_mode, DescOut out) {
  switch (sf.nz) {
    case 1: return fwd_grid_nz<T, 1, false>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
    case 2: return fwd_grid_nz<T, 2, false>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
    case 3: return fwd_grid_nz<T, 3, false>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
    case 4:
      if (zeta_1_2_4_16(sf))
        return fwd_grid_nz<T, 4, true>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
      return fwd_grid_nz<T, 4, false>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
    default: return fwd_grid_nz<T, 5, false>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  }
}

// Whether some descriptor kernel of this table needs its cutoff tables in
// global memory at KP slots (its one-wave LDS layout exceeds the device's limit).
template <typename T>
bool needs_global_tables(const SFDev &sf, const Lds &L) {
  const size_t lim = lds_limit();
  if (bwd_lds<T, 1, false>(L) > lim) return true;
  if (!sf.grid) return fwd_lds<T, 16, 1, false>(L) > lim;
  return fwd_grid_lds<T, 5, 1, false>(L) > lim;  // the widest grid layout
}

// Whether every descriptor kernel of this table has a layout for KP slots.
template <typename T>
bool fits(const SFDev &sf, const Lds &L) {
  const size_t lim = lds_limit();
  if (bwd_lds<T, 1, true>(L) > lim) return false;
  if (!sf.grid) return fwd_lds<T, 16, 1, true>(L) <= lim;
  return fwd_grid_lds<T, 5, 1, true>(L) <= lim;
}

template <typename T>
hipError_t bwd_fit(int waves, hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                   const TripletBufs &tb, const SFDev &sf, Lds L, const GradSrc &gs, double *slot_f,
                   double *atom_vir) {
  int W;
  bool fg;
  auto bytes = [&](int w, bool g) {
    if (w == 4) return bwd_lds<T, 4, false>(L);
    if (w == 2) return bwd_lds<T, 2, false>(L);
    return g ? bwd_lds<T, 1, true>(L) : bwd_lds<T, 1, false>(L);
  };
  if (!pick_shape(waves, 4, bytes, W, fg) || (fg && !nb.fc_scratch)) return hipErrorInvalidValue;
  if (fg) return bwd_grid<T, 1, true>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  if (W == 4) return bwd_grid<T, 4, false>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  if (W == 2) return bwd_grid<T, 2, false>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  return bwd_grid<T, 1, false>(s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
}

}  // namespace

// Largest per-atom neighbour count the descriptor kernels take (every
// kernel of the table then has a layout within the device's LDS limit, with
// the cutoff tables in global memory where needed), capped at kMaxNeighbors.
int max_neighbors_for_lds(bool fp64, const SFDev &sf, int Dp) {
  const int G = sf.n_groups > 0 ? sf.n_groups : 1;
  int kp = kMaxNeighbors;
  for (; kp > 8; kp -= 8) {
    const Lds L{kp, G, Dp, 0};
    if (fp64 ? fits<double>(sf, L) : fits<float>(sf, L)) break;
  }
  return kp;
}

size_t desc_table_scratch_bytes(bool fp64, const SFDev &sf, int Dp, int KP, int N) {
  const int G = sf.n_groups > 0 ? sf.n_groups : 1;
  const Lds L{KP, G, Dp, 0};
  if (!(fp64 ? needs_global_tables<double>(sf, L) : needs_global_tables<float>(sf, L))) return 0;
  return (size_t)N * 2 * G * KP * (fp64 ? sizeof(double) : sizeof(float));
}

hipError_t launch_descriptors_fwd(hipStream_t s, bool fp64, int N, const int *species, const double *wz,
                                  const NeighborBufs &nb, const TripletBufs &tb, const SFDev &sf,
                                  const double *theta_inv, const double *mu, int split_mode, DescOut out, int chunk,
                                  int waves, int KP, int atom0) {
  Lds L{KP, sf.n_groups > 0 ? sf.n_groups : 1, out.Dp, atom0};
  if (N <= 0) return hipSuccess;
  if (sf.grid) {
    if (fp64) return fwd_grid<double>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
    return fwd_grid<float>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  }
  if (fp64) {
    if (chunk <= 8) return fwd_waves<double, 8>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
    return fwd_waves<double, 16>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  }
  if (chunk <= 8) return fwd_waves<float, 8>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
  return fwd_waves<float, 16>(waves, s, N, species, wz, nb, tb, sf, L, theta_inv, mu, split_mode, out);
}

hipError_t launch_descriptors_bwd(hipStream_t s, bool fp64, int N, const int *species, const double *wz,
                                  const NeighborBufs &nb, const TripletBufs &tb, const SFDev &sf, int Dp,
                                  const GradSrc &gs, double *slot_f, double *atom_vir, int waves, int KP,
                                  int atom0) {
  Lds L{KP, sf.n_groups > 0 ? sf.n_groups : 1, Dp, atom0};
  if (N <= 0) return hipSuccess;
  // the LDS layout's 16-byte reads of g need KP (slots per atom) a multiple of 4
  if (KP % 4 != 0) return hipErrorInvalidValue;
  if (fp64) return bwd_fit<double>(waves, s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
  return bwd_fit<float>(waves, s, N, species, wz, nb, tb, sf, L, gs, slot_f, atom_vir);
}

}  // namespace gapx
