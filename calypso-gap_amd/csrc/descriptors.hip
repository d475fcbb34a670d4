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
// put one neighbour on each lane.  Angular sums enumerate neighbour pairs
// j < k, compact the pairs that pass the cutoff test into an LDS queue with a
// ballot, and process them 64 at a time, so every lane holds a live triplet.
// The forward keeps NCH lines' partial sums in registers per lane and reduces
// them across the wavefront after each chunk of lines.
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
  static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
  static __device__ __forceinline__ double pw(double b, double e) { return pow(b, e); }
};

template <>
struct Math<float> {
  static __device__ __forceinline__ float ex(float x) { return __expf(x); }
  static __device__ __forceinline__ float cs(float x) { return __cosf(x); }
  static __device__ __forceinline__ float sn(float x) { return __sinf(x); }
  static __device__ __forceinline__ float sq(float x) { return sqrtf(x); }
  static __device__ __forceinline__ float pw(float b, float e) { return powf(b, e); }
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// base^n for an integer n >= 0 that is the same on every lane.
template <typename T>
__device__ __forceinline__ T ipow(T base, int n) {
  T r = T(1);
  while (n) {
    if (n & 1) r *= base;
    base *= base;
    n >>= 1;
  }
  return r;
}

// (a, b), a < b, of the p-th pair in row-major order of the upper triangle.
__device__ __forceinline__ void pair_index(int p, int K, int &a, int &b) {
  const int n2 = 2 * K - 1;
  double disc = (double)n2 * n2 - 8.0 * p;
  int r = (int)((n2 - sqrt(disc)) * 0.5);
  if (r < 0) r = 0;
  while (r > 0 && r * (2 * K - r - 1) / 2 > p) --r;
  while ((r + 1) * (2 * K - r - 2) / 2 <= p) ++r;
  a = r;
  b = p - r * (2 * K - r - 1) / 2 + r + 1;
}

// Angular power P = pref base^zeta and dP/dcos = pref zeta lam base^(zeta-1).
template <typename T>
__device__ __forceinline__ void ang_pow(T base, T lam, int zi, T zr, T pref, T &P, T &dP) {
  if (zi == 1) {
    P = base;
    dP = lam;
  } else if (zi > 1) {
    T b1 = ipow(base, zi - 1);
    P = pref * b1 * base;
    dP = pref * (T)zi * lam * b1;
  } else {
    T b1 = (base > T(0)) ? Math<T>::pw(base, zr - T(1)) : T(0);
    P = pref * b1 * base;
    dP = pref * zr * lam * b1;
  }
}

struct Lds {
  int KP, G, Dp;
};

// ---------------------------------------------------------------------------
// shared neighbour staging
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void stage_neighbors(int start, int K, int lane, const NeighborBufs &nb,
                                                const int *__restrict__ species, const double *__restrict__ wz,
                                                const SFDev &sf, int KP, T *s_d, T *s_w, T *s_fc, T *s_dfc) {
  for (int a = lane; a < K; a += 64) {
    const double *dd = nb.nb_d + 4 * (size_t)(start + a);
    const double r = dd[3];
    s_d[4 * a] = (T)dd[0];
    s_d[4 * a + 1] = (T)dd[1];
    s_d[4 * a + 2] = (T)dd[2];
    s_d[4 * a + 3] = (T)r;
    s_w[a] = (T)wz[species[nb.nb_j[start + a]]];
    for (int g = 0; g < sf.n_groups; ++g) {
      const double rc = sf.group_rc[g];
      const bool in = r <= rc;
      s_fc[g * KP + a] = in ? (T)(0.5 * (cos(kPi * r / rc) + 1.0)) : T(0);
      if (s_dfc) s_dfc[g * KP + a] = in ? (T)(0.5 * (-sin(kPi * r / rc)) * (kPi / rc)) : T(0);
    }
  }
}

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
#pragma unroll
  for (int l = 0; l < NCH; ++l) {
    const int s = c0 + l;
    if (s < sf.n_ang) {
      const int g = sf.ang_group[s];
      if (g != cur_g) {
        cur_g = g;
        const T rc = (T)sf.group_rc[g];
        const T fe = (e <= rc) ? T(0.5) * (Math<T>::cs((T)kPi * e / rc) + T(1)) : T(0);
        F3 = s_fc[g * KP + a] * s_fc[g * KP + b] * fe * on;
      }
      const T eta = (T)sf.ang_eta[s];
      if (eta != cur_eta) {
        cur_eta = eta;
        E = Math<T>::ex(-eta * S2);
      }
      T P, dP;
      ang_pow<T>(T(1) + (T)sf.ang_lam[s] * cth, (T)sf.ang_lam[s], sf.ang_zeta[s], (T)sf.ang_zeta_real[s],
                 (T)sf.ang_pref[s], P, dP);
      const T t = P * E * F3;
      acc[l] += t;
      accw[l] += t * wab;
    }
  }
}

template <typename T, int NCH>
__global__ __launch_bounds__(64) void k_desc_fwd(int N, const int *__restrict__ species,
                                                 const double *__restrict__ wz, NeighborBufs nb, SFDev sf,
                                                 Lds L, const double *__restrict__ theta_inv,
                                                 const double *__restrict__ mu, int split_mode, DescOut out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int KP = L.KP, Dp = L.Dp;
  T *s_d = (T *)smem;
  T *s_w = s_d + 4 * KP;
  T *s_fc = s_w + KP;
  T *s_feat = s_fc + L.G * KP;
  uint32_t *s_q = (uint32_t *)(s_feat + Dp);

  const int start = i * nb.kcap;
  const int K = nb.nb_count[i];
  if (K > KP) {  // the host sizes KP from the largest count; never index LDS past it
    if (lane == 0) atomicAdd(&nb.total[3], 1);
    return;
  }
  stage_neighbors<T>(start, K, lane, nb, species, wz, sf, KP, s_d, s_w, s_fc, (T *)nullptr);
  for (int d = lane; d < Dp; d += 64) s_feat[d] = T(0);
  __syncthreads();

  // radial lines, NCH at a time
  for (int c0 = 0; c0 < sf.n_rad; c0 += NCH) {
    T acc[NCH], accw[NCH];
#pragma unroll
    for (int l = 0; l < NCH; ++l) acc[l] = accw[l] = T(0);
    for (int a = lane; a < K; a += 64) {
      const T r = s_d[4 * a + 3], w = s_w[a];
#pragma unroll
      for (int l = 0; l < NCH; ++l) {
        const int s = c0 + l;
        if (s < sf.n_rad) {
          const T dr = r - (T)sf.rad_rs[s];
          const T t = Math<T>::ex(-(T)sf.rad_a[s] * dr * dr) * s_fc[sf.rad_group[s] * KP + a];
          acc[l] += t;
          accw[l] += t * w;
        }
      }
    }
#pragma unroll
    for (int l = 0; l < NCH; ++l) {
      const T v = wave_sum(acc[l]);
      const T vw = wave_sum(accw[l]);
      if (lane == 0 && c0 + l < sf.n_rad) {
        const int f = sf.rad_feat[c0 + l];
        s_feat[f] = v;
        s_feat[f + sf.nsf] = vw;
      }
    }
  }

  // angular lines, NCH at a time, over compacted triplets
  const int npairs = K * (K - 1) / 2;
  const T rcA = (T)sf.rc_max_ang;
  unsigned long long ntrip = 0;
  for (int c0 = 0; c0 < sf.n_ang; c0 += NCH) {
    T acc[NCH], accw[NCH];
#pragma unroll
    for (int l = 0; l < NCH; ++l) acc[l] = accw[l] = T(0);
    int qn = 0;
    for (int p0 = 0; p0 < npairs; p0 += 64) {
      const int p = p0 + lane;
      bool ok = false;
      int a = 0, b = 0;
      if (p < npairs) {
        pair_index(p, K, a, b);
        if (s_d[4 * a + 3] <= rcA && s_d[4 * b + 3] <= rcA) {
          const T ex = s_d[4 * b] - s_d[4 * a], ey = s_d[4 * b + 1] - s_d[4 * a + 1],
                  ez = s_d[4 * b + 2] - s_d[4 * a + 2];
          ok = Math<T>::sq(ex * ex + ey * ey + ez * ez) <= rcA;
        }
      }
      const unsigned long long m = __ballot(ok);
      if (ok) s_q[qn + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)a | ((uint32_t)b << 16);
      qn += __popcll(m);
      ntrip += __popcll(m);
      __syncthreads();
      if (qn >= 64) {
        ang_fwd_term<T, NCH>(s_q[lane], true, c0, sf, s_d, s_w, s_fc, KP, acc, accw);
        const int rest = qn - 64;
        const uint32_t mv = (lane < rest) ? s_q[64 + lane] : 0u;
        __syncthreads();
        if (lane < rest) s_q[lane] = mv;
        qn = rest;
        __syncthreads();
      }
    }
    if (qn > 0) {
      const bool act = lane < qn;
      ang_fwd_term<T, NCH>(act ? s_q[lane] : (1u << 16), act, c0, sf, s_d, s_w, s_fc, KP, acc, accw);
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < NCH; ++l) {
      const T v = wave_sum(acc[l]);
      const T vw = wave_sum(accw[l]);
      if (lane == 0 && c0 + l < sf.n_ang) {
        const int f = sf.ang_feat[c0 + l];
        s_feat[f] = v;
        s_feat[f + sf.nsf] = vw;
      }
    }
  }
  __syncthreads();
  if (lane == 0 && out.triplets && ntrip) {
    const int chunks = (sf.n_ang + NCH - 1) / NCH;
    atomicAdd(out.triplets, ntrip / (unsigned long long)chunks);
  }

  // epilogue: raw features, centred/scaled features, FP64 norm, split operands
  double xn = 0.0;
  const size_t row = (size_t)i * Dp;
  for (int d = lane; d < Dp; d += 64) {
    const T x = s_feat[d];
    if (out.x64) out.x64[row + d] = (double)x;
    if (out.x32) out.x32[row + d] = (float)x;
    const double xt = (double)x * theta_inv[d] - mu[d];
    xn += xt * xt;
    if (out.xt64) out.xt64[row + d] = xt;
    if (out.xt32) out.xt32[row + d] = (float)xt;
    if (split_mode == 1) {  // FP16 hi + lo
      const _Float16 hi = (_Float16)xt;
      const _Float16 lo = (_Float16)(xt - (double)hi);
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
  if (lane == 0) out.xn[i] = xn;
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void ang_bwd_term(uint32_t q, bool active, const SFDev &sf, const T *s_d,
                                             const T *s_w, const T *s_fc, const T *s_dfc, const T *s_g,
                                             int KP, T *s_T, T *s_F, T (&vir)[6]) {
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
  const T wab = s_w[a] * s_w[b];
  T Ta = 0, Tb = 0, Te = 0;
  T SA = 0, SB = 0, SC = 0;
  T fa = 0, dfa = 0, fb = 0, dfb = 0, fe = 0, dfe = 0;
  int cur_g = -1;
  T E = 0, cur_eta = T(-1);
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
      fa = s_fc[g * KP + a];
      dfa = s_dfc[g * KP + a];
      fb = s_fc[g * KP + b];
      dfb = s_dfc[g * KP + b];
      const bool in = e <= rc;
      fe = in ? T(0.5) * (Math<T>::cs((T)kPi * e / rc) + T(1)) : T(0);
      dfe = in ? T(0.5) * (-Math<T>::sn((T)kPi * e / rc)) * ((T)kPi / rc) : T(0);
    }
    const T eta = (T)sf.ang_eta[s];
    if (eta != cur_eta) {
      cur_eta = eta;
      E = Math<T>::ex(-eta * S2);
    }
    const int f = sf.ang_feat[s];
    const T geff = s_g[f] + wab * s_g[f + sf.nsf];
    T P, dP;
    ang_pow<T>(T(1) + (T)sf.ang_lam[s] * cth, (T)sf.ang_lam[s], sf.ang_zeta[s], (T)sf.ang_zeta_real[s],
               (T)sf.ang_pref[s], P, dP);
    const T B = geff * P * E;
    SA += geff * dP * E;
    SB += B;
    SC += eta * B;
  }
  if (cur_g >= 0) flush();
  atomicAdd(&s_T[a], Ta);
  atomicAdd(&s_T[b], Tb);
  if (e > T(0)) {
    const T s = Te / e;
    const T ux = s * ex, uy = s * ey, uz = s * ez;
    atomicAdd(&s_F[3 * b], ux);
    atomicAdd(&s_F[3 * b + 1], uy);
    atomicAdd(&s_F[3 * b + 2], uz);
    atomicAdd(&s_F[3 * a], -ux);
    atomicAdd(&s_F[3 * a + 1], -uy);
    atomicAdd(&s_F[3 * a + 2], -uz);
    vir[0] += ux * ex;
    vir[1] += uy * ey;
    vir[2] += uz * ez;
    vir[3] += ux * ey;
    vir[4] += uy * ez;
    vir[5] += ux * ez;
  }
}

template <typename T>
__global__ __launch_bounds__(64) void k_desc_bwd(int N, const int *__restrict__ species,
                                                 const double *__restrict__ wz, NeighborBufs nb, SFDev sf,
                                                 Lds L, const double *__restrict__ g64,
                                                 const float *__restrict__ g32, double *__restrict__ slot_f,
                                                 double *__restrict__ atom_vir) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int KP = L.KP, Dp = L.Dp, G = L.G;
  T *s_d = (T *)smem;
  T *s_w = s_d + 4 * KP;
  T *s_fc = s_w + KP;
  T *s_dfc = s_fc + G * KP;
  T *s_g = s_dfc + G * KP;
  T *s_T = s_g + Dp;
  T *s_F = s_T + KP;
  uint32_t *s_q = (uint32_t *)(s_F + 3 * KP);

  const int start = i * nb.kcap;
  const int K = nb.nb_count[i];
  if (K > KP) {
    if (lane == 0) atomicAdd(&nb.total[3], 1);
    return;
  }
  stage_neighbors<T>(start, K, lane, nb, species, wz, sf, KP, s_d, s_w, s_fc, s_dfc);
  const size_t row = (size_t)i * Dp;
  for (int d = lane; d < Dp; d += 64) s_g[d] = g64 ? (T)g64[row + d] : (T)g32[row + d];
  for (int a = lane; a < K; a += 64) {
    s_F[3 * a] = s_F[3 * a + 1] = s_F[3 * a + 2] = T(0);
  }
  __syncthreads();

  // radial: one neighbour per lane
  for (int a = lane; a < K; a += 64) {
    const T r = s_d[4 * a + 3], w = s_w[a];
    T acc = 0;
    for (int s = 0; s < sf.n_rad; ++s) {
      const int f = sf.rad_feat[s];
      const int g = sf.rad_group[s];
      const T A = (T)sf.rad_a[s];
      const T dr = r - (T)sf.rad_rs[s];
      const T geff = s_g[f] + w * s_g[f + sf.nsf];
      acc += geff * Math<T>::ex(-A * dr * dr) * (T(-2) * A * dr * s_fc[g * KP + a] + s_dfc[g * KP + a]);
    }
    s_T[a] = acc;
  }
  __syncthreads();

  T vir[6] = {0, 0, 0, 0, 0, 0};
  const int npairs = K * (K - 1) / 2;
  const T rcA = (T)sf.rc_max_ang;
  if (sf.n_ang > 0) {
    int qn = 0;
    for (int p0 = 0; p0 < npairs; p0 += 64) {
      const int p = p0 + lane;
      bool ok = false;
      int a = 0, b = 0;
      if (p < npairs) {
        pair_index(p, K, a, b);
        if (s_d[4 * a + 3] <= rcA && s_d[4 * b + 3] <= rcA) {
          const T ex = s_d[4 * b] - s_d[4 * a], ey = s_d[4 * b + 1] - s_d[4 * a + 1],
                  ez = s_d[4 * b + 2] - s_d[4 * a + 2];
          ok = Math<T>::sq(ex * ex + ey * ey + ez * ez) <= rcA;
        }
      }
      const unsigned long long m = __ballot(ok);
      if (ok) s_q[qn + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)a | ((uint32_t)b << 16);
      qn += __popcll(m);
      __syncthreads();
      if (qn >= 64) {
        ang_bwd_term<T>(s_q[lane], true, sf, s_d, s_w, s_fc, s_dfc, s_g, KP, s_T, s_F, vir);
        const int rest = qn - 64;
        const uint32_t mv = (lane < rest) ? s_q[64 + lane] : 0u;
        __syncthreads();
        if (lane < rest) s_q[lane] = mv;
        qn = rest;
        __syncthreads();
      }
    }
    if (qn > 0) {
      const bool act = lane < qn;
      ang_bwd_term<T>(act ? s_q[lane] : (1u << 16), act, sf, s_d, s_w, s_fc, s_dfc, s_g, KP, s_T, s_F, vir);
    }
  }
  __syncthreads();

  // per-slot force contributions dE_i/dR(slot atom) and the radial virial
  for (int a = lane; a < K; a += 64) {
    const T r = s_d[4 * a + 3];
    const T s = s_T[a] / r;
    const T dx = s_d[4 * a], dy = s_d[4 * a + 1], dz = s_d[4 * a + 2];
    const T fx = s * dx + s_F[3 * a], fy = s * dy + s_F[3 * a + 1], fz = s * dz + s_F[3 * a + 2];
    double *o = slot_f + 3 * (size_t)(start + a);
    o[0] = (double)fx;
    o[1] = (double)fy;
    o[2] = (double)fz;
    vir[0] += s * dx * dx;
    vir[1] += s * dy * dy;
    vir[2] += s * dz * dz;
    vir[3] += s * dx * dy;
    vir[4] += s * dy * dz;
    vir[5] += s * dx * dz;
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double v = wave_sum((double)vir[c]);
    if (lane == 0) atom_vir[6 * (size_t)i + c] = v;
  }
}

template <typename T>
size_t fwd_lds(const Lds &L) {
  return (size_t)(4 * L.KP + L.KP + L.G * L.KP + L.Dp) * sizeof(T) + 128 * sizeof(uint32_t);
}

template <typename T>
size_t bwd_lds(const Lds &L) {
  return (size_t)(4 * L.KP + L.KP + 2 * L.G * L.KP + L.Dp + L.KP + 3 * L.KP) * sizeof(T) +
         128 * sizeof(uint32_t);
}

template <typename T, int NCH>
hipError_t fwd_launch(hipStream_t s, int N, const int *species, const double *wz, const NeighborBufs &nb,
                      const SFDev &sf, Lds L, const double *theta_inv, const double *mu, int split_mode,
                      DescOut out) {
  const size_t lds = fwd_lds<T>(L);
  k_desc_fwd<T, NCH><<<N, 64, lds, s>>>(N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
  return hipGetLastError();
}

}  // namespace

// Largest neighbour count the LDS layout allows (checked by the host).
int max_neighbors_for_lds(bool fp64, int G, int Dp) {
  const size_t budget = 64 * 1024;
  const size_t bytes = fp64 ? 8 : 4;
  // backward is the larger layout: (9 + 2G) KP + Dp values + queue
  long kp = (long)((budget - 512) / bytes - Dp) / (9 + 2 * G);
  return (int)(kp / 8 * 8);
}

hipError_t launch_descriptors_fwd(hipStream_t s, bool fp64, int N, const int *species, const double *wz,
                                  const NeighborBufs &nb, const SFDev &sf, const double *theta_inv,
                                  const double *mu, int split_mode, DescOut out, int chunk, int KP) {
  Lds L{KP, sf.n_groups > 0 ? sf.n_groups : 1, out.Dp};
  if (N <= 0) return hipSuccess;
  if (fp64) {
    if (chunk <= 8) return fwd_launch<double, 8>(s, N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
    if (chunk <= 16) return fwd_launch<double, 16>(s, N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
    return fwd_launch<double, 32>(s, N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
  }
  if (chunk <= 8) return fwd_launch<float, 8>(s, N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
  if (chunk <= 16) return fwd_launch<float, 16>(s, N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
  return fwd_launch<float, 32>(s, N, species, wz, nb, sf, L, theta_inv, mu, split_mode, out);
}

hipError_t launch_descriptors_bwd(hipStream_t s, bool fp64, int N, const int *species, const double *wz,
                                  const NeighborBufs &nb, const SFDev &sf, int Dp, const double *g64,
                                  const float *g32, double *slot_f, double *atom_vir, int KP) {
  Lds L{KP, sf.n_groups > 0 ? sf.n_groups : 1, Dp};
  if (N <= 0) return hipSuccess;
  if (fp64) {
    k_desc_bwd<double><<<N, 64, bwd_lds<double>(L), s>>>(N, species, wz, nb, sf, L, g64, g32, slot_f, atom_vir);
  } else {
    k_desc_bwd<float><<<N, 64, bwd_lds<float>(L), s>>>(N, species, wz, nb, sf, L, g64, g32, slot_f, atom_vir);
  }
  return hipGetLastError();
}

}  // namespace gapx
