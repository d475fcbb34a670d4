// Batched structure relaxation on the device: FIRE (Bitzek et al., PRL 97,
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
                                                     FireState *__restrict__ st, FireParams fp) {
  const int s = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (s >= nstruct) return;
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
                              double *__restrict__ vel, double *__restrict__ pos, const FireState *__restrict__ st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const FireState &S = st[atom_struct[i]];
  if (S.converged) return;
  for (int c = 0; c < 3; ++c) {
    const double v = S.cv * vel[3 * i + c] + S.cf * force[3 * i + c];
    vel[3 * i + c] = v;
    pos[3 * i + c] += S.step * v;
  }
}

}  // namespace

hipError_t launch_fire_step(hipStream_t s, int nstruct, int N, const StructParams *sp, const int *atom_struct,
                            const double *force, double *vel, double *pos, FireState *st, const FireParams &fp) {
  k_fire_reduce<<<nstruct, 256, 0, s>>>(nstruct, sp, force, vel, st, fp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_fire_update<<<(N + 255) / 256, 256, 0, s>>>(N, atom_struct, force, vel, pos, st);
  return hipGetLastError();
}

}  // namespace gapx
