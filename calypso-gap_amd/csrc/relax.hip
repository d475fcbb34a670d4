// Batched structure relaxation on the device: LBFGS as ASE runs it (below),
// the reference's optimiser, and FIRE (Bitzek et al., PRL 97,
// 170201 (2006)) with the parameters of ASE's FIRE optimiser (dt 0.1,
// dtmax 1.0, Nmin 5, finc 1.1, fdec 0.5, astart 0.1, fa 0.99, maxstep 0.2 A),
// positions only at fixed cell, as the reference's relaxation driver
// (gappy/example/ASE-GAPPY/run_opt.py:108-109, LBFGS with fmax = -EDIFFG)
// relaxes one structure per process (tools/cgg2.py:100-137).
//
// Per step, after an E+F evaluation of the whole batch:
//   k_fire_reduce  one workgroup per structure: F.v, |F|^2, |v|^2, max |F_i|,
//                  then the FIRE state update (dt, a, steps since a downhill
//                  restart) and the velocity mixing coefficients;
//   k_fire_update  one thread per atom: v = cv v + cf F, x += s v.
// Converged structures (max |F_i| <= fmax) stop moving.
#include <cmath>

#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

namespace {

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_max_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ __launch_bounds__(256) void k_fire_reduce(int nstruct, const StructParams *__restrict__ sp,
                                                     const double *__restrict__ force, const double *__restrict__ vel,
                                                     FireState *__restrict__ st, FireParams fp, StepCtl ctl) {
  const int s = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (s >= nstruct) return;
  if (ctl.skip && *ctl.skip) {
    if (s == 0 && t == 0 && ctl.flag) atomicOr(ctl.flag, 2);
    return;
  }
  const StructParams &P = sp[s];
  __shared__ double red[4][4];
  double vf = 0, ff = 0, vv = 0, fm = 0;
  for (int a = t; a < P.natoms; a += 256) {
    const size_t i = (size_t)(P.atom_base + a);
    double f2 = 0;
    for (int c = 0; c < 3; ++c) {
      const double f = force[3 * i + c], v = vel[3 * i + c];
      vf += f * v;
      f2 += f * f;
      vv += v * v;
    }
    ff += f2;
    fm = fmax(fm, f2);
  }
  vf = wave_sum_d(vf);
  ff = wave_sum_d(ff);
  vv = wave_sum_d(vv);
  fm = wave_max_d(fm);
  if (lane == 0) {
    red[wv][0] = vf;
    red[wv][1] = ff;
    red[wv][2] = vv;
    red[wv][3] = fm;
  }
  __syncthreads();
  if (t != 0) return;
  fm = fmax(fmax(red[0][3], red[1][3]), fmax(red[2][3], red[3][3]));
  vf = red[0][0] + red[1][0] + red[2][0] + red[3][0];
  ff = red[0][1] + red[1][1] + red[2][1] + red[3][1];
  vv = red[0][2] + red[1][2] + red[2][2] + red[3][2];
  FireState S = st[s];
  S.fmax = sqrt(fm);
  if (S.converged || S.fmax <= fp.fmax) {  // no move from here on
    S.converged = 1;
    S.cv = 0;
    S.cf = 0;
    S.step = 0;
    st[s] = S;
    return;
  }
  double cv = 1.0, cf = 0.0;
  if (S.started) {
    if (vf > 0.0) {
      cv = 1.0 - S.a;
      cf = (ff > 0.0) ? S.a * sqrt(vv / ff) : 0.0;
      if (S.nsteps > fp.nmin) {
        S.dt = fmin(S.dt * fp.finc, fp.dtmax);
        S.a *= fp.fa;
      }
      S.nsteps += 1;
    } else {
      cv = 0.0;  // v = 0
      cf = 0.0;
      S.a = fp.astart;
      S.dt *= fp.fdec;
      S.nsteps = 0;
    }
  }
  S.started = 1;
  if (ctl.flag) atomicOr(ctl.flag, 1);
  // v' = cv v + (cf + dt) F;  dr = dt v', capped at maxstep in norm
  cf += S.dt;
  const double vn2 = cv * cv * vv + 2.0 * cv * cf * vf + cf * cf * ff;
  const double dr = S.dt * sqrt(fmax(vn2, 0.0));
  S.cv = cv;
  S.cf = cf;
  S.step = (dr > fp.maxstep) ? S.dt * fp.maxstep / dr : S.dt;
  st[s] = S;
}

__global__ void k_fire_update(int N, const int *__restrict__ atom_struct, const double *__restrict__ force,
                              double *__restrict__ vel, double *__restrict__ pos, const FireState *__restrict__ st,
                              const int *__restrict__ skip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || (skip && *skip)) return;
  const FireState &S = st[atom_struct[i]];
  if (S.converged) return;
  for (int c = 0; c < 3; ++c) {
    const double v = S.cv * vel[3 * i + c] + S.cf * force[3 * i + c];
    vel[3 * i + c] = v;
    pos[3 * i + c] += S.step * v;
  }
}

// ---------------------------------------------------------------------------
// LBFGS as ASE's optimiser runs it (ase/optimize/lbfgs.py of ASE 3.18.1, the
// version the reference's example data were made with; the reference relaxes
// with LBFGS(atoms) at its defaults, gappy/example/ASE-GAPPY/run_opt.py:108):
// memory 100, H0 = 1 / alpha with alpha 70, damping 1, no line search,
// maxstep 0.2 A.  Per step, with forces f at positions r:
//   converged when max_i |f_i|^2 < fmax^2 (no move from then on);
//   after the first step, append s = r - r0, y = f0 - f, rho = 1 / (y . s)
//   to the history (the oldest dropped beyond `memory` entries);
//   two-loop recursion: q = -f; for i = L-1 .. 0: a_i = rho_i s_i . q,
//   q -= a_i y_i; z = H0 q; for i = 0 .. L-1: b = rho_i y_i . z,
//   z += (a_i - b) s_i; the step p = -z, scaled by maxstep / (longest
//   per-atom step) when that is at least maxstep; r0 = r, f0 = f, r += p.
// One workgroup per structure; every dot product is a workgroup reduction
// in a fixed order (lanes, then waves), so repeated runs agree bit for bit.
// The history is a ring of `memory` slots per structure: S[slot][3 N] and
// Y[slot][3 N] over the whole batch, rho[structure][slot].
// ---------------------------------------------------------------------------

// Sum (or maximum) of one value per thread over the 256-thread workgroup,
// returned to every thread.
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v, double *s_red) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  v = MAX ? wave_max_d(v) : wave_sum_d(v);
  __syncthreads();  // s_red free (the previous reduction's readers are done)
  if (lane == 0) s_red[wv] = v;
  __syncthreads();
  return MAX ? fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3]))
             : (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
}

__global__ __launch_bounds__(256) void k_lbfgs_step(int nstruct, const StructParams *__restrict__ sp, int n3tot,
                                                    const double *__restrict__ force, double *__restrict__ pos,
                                                    double *__restrict__ r0, double *__restrict__ f0,
                                                    double *__restrict__ S, double *__restrict__ Y,
                                                    double *__restrict__ rho, double *__restrict__ q,
                                                    LbfgsState *__restrict__ st, LbfgsParams lp, StepCtl ctl) {
  __shared__ double s_red[4];
  __shared__ double s_a[kLbfgsMaxMemory];
  const int s = blockIdx.x, t = threadIdx.x;
  if (s >= nstruct) return;
  if (ctl.skip && *ctl.skip) {
    if (s == 0 && t == 0 && ctl.flag) atomicOr(ctl.flag, 2);
    return;
  }
  const StructParams &P = sp[s];
  const int n3 = 3 * P.natoms;
  const size_t b0 = 3 * (size_t)P.atom_base;
  LbfgsState L = st[s];
  // largest |f_i|^2
  double fm = 0.0;
  for (int a = t; a < P.natoms; a += 256) {
    const size_t i = (size_t)(P.atom_base + a);
    const double fx = force[3 * i], fy = force[3 * i + 1], fz = force[3 * i + 2];
    fm = fmax(fm, fx * fx + fy * fy + fz * fz);
  }
  fm = block_reduce<true>(fm, s_red);
  if (L.converged || fm < lp.fmax * lp.fmax || lp.check_only) {
    if (t == 0) {
      L.converged = L.converged || fm < lp.fmax * lp.fmax;
      L.fmax = sqrt(fm);
      st[s] = L;
    }
    return;
  }
  const int mem = lp.memory;
  // history update from the previous step
  if (L.iter > 0) {
    const size_t slot = (size_t)L.head * n3tot + b0;
    double ys = 0.0;
    for (int e = t; e < n3; e += 256) {
      const double sv = pos[b0 + e] - r0[b0 + e], yv = f0[b0 + e] - force[b0 + e];
      S[slot + e] = sv;
      Y[slot + e] = yv;
      ys += yv * sv;
    }
    ys = block_reduce<false>(ys, s_red);
    if (t == 0) rho[(size_t)s * mem + L.head] = 1.0 / ys;
    L.head = (L.head + 1) % mem;
    L.count = min(L.count + 1, mem);
  }
  __syncthreads();  // rho of the new entry visible to every thread
  const int nh = L.count, oldest = (L.head - nh + mem) % mem;
  for (int e = t; e < n3; e += 256) q[b0 + e] = -force[b0 + e];
  for (int k = nh - 1; k >= 0; --k) {
    const int sl = (oldest + k) % mem;
    const double *Sk = S + (size_t)sl * n3tot + b0, *Yk = Y + (size_t)sl * n3tot + b0;
    double d = 0.0;
    for (int e = t; e < n3; e += 256) d += Sk[e] * q[b0 + e];
    const double ak = rho[(size_t)s * mem + sl] * block_reduce<false>(d, s_red);
    if (t == 0) s_a[k] = ak;
    for (int e = t; e < n3; e += 256) q[b0 + e] -= ak * Yk[e];
  }
  for (int e = t; e < n3; e += 256) q[b0 + e] *= lp.h0;  // z = H0 q
  __syncthreads();
  for (int k = 0; k < nh; ++k) {
    const int sl = (oldest + k) % mem;
    const double *Sk = S + (size_t)sl * n3tot + b0, *Yk = Y + (size_t)sl * n3tot + b0;
    double d = 0.0;
    for (int e = t; e < n3; e += 256) d += Yk[e] * q[b0 + e];
    const double bk = rho[(size_t)s * mem + sl] * block_reduce<false>(d, s_red);
    const double c = s_a[k] - bk;
    for (int e = t; e < n3; e += 256) q[b0 + e] += c * Sk[e];
  }
  // Until here every thread read and wrote only its own elements of q; the
  // per-atom lengths below read the three components of an atom, which
  // other threads wrote.
  __syncthreads();
  // step p = -z, scaled to maxstep by its longest per-atom displacement
  double lm = 0.0;
  for (int a = t; a < P.natoms; a += 256) {
    const size_t i = b0 + 3 * (size_t)a;
    lm = fmax(lm, q[i] * q[i] + q[i + 1] * q[i + 1] + q[i + 2] * q[i + 2]);
  }
  const double longest = sqrt(block_reduce<true>(lm, s_red));
  const double scale = (longest >= lp.maxstep) ? lp.maxstep / longest : 1.0;
  for (int e = t; e < n3; e += 256) {
    const double r = pos[b0 + e];
    r0[b0 + e] = r;
    f0[b0 + e] = force[b0 + e];
    pos[b0 + e] = r - q[b0 + e] * scale;
  }
  if (t == 0) {
    L.iter += 1;
    L.fmax = sqrt(fm);
    st[s] = L;
    if (ctl.flag) atomicOr(ctl.flag, 1);
  }
}

}  // namespace

hipError_t launch_lbfgs_step(hipStream_t s, int nstruct, int N, const StructParams *sp, const double *force,
                             double *pos, double *r0, double *f0, double *S, double *Y, double *rho, double *q,
                             LbfgsState *st, const LbfgsParams &lp, StepCtl ctl) {
  if (lp.memory < 1 || lp.memory > kLbfgsMaxMemory) return hipErrorInvalidValue;
  k_lbfgs_step<<<nstruct, 256, 0, s>>>(nstruct, sp, 3 * N, force, pos, r0, f0, S, Y, rho, q, st, lp, ctl);
  return hipGetLastError();
}

hipError_t launch_fire_step(hipStream_t s, int nstruct, int N, const StructParams *sp, const int *atom_struct,
                            const double *force, double *vel, double *pos, FireState *st, const FireParams &fp,
                            StepCtl ctl) {
  k_fire_reduce<<<nstruct, 256, 0, s>>>(nstruct, sp, force, vel, st, fp, ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_fire_update<<<(N + 255) / 256, 256, 0, s>>>(N, atom_struct, force, vel, pos, st, ctl.skip);
  return hipGetLastError();
}

}  // namespace gapx
