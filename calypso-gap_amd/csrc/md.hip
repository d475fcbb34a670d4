// Batched NVT molecular dynamics on the device: the step of ASE's
// NVTBerendsen (ASE 3.18.1, the version the reference pins), which the
// reference's MD driver runs one structure per process
// (gappy/example/ASE-GAPPY/run_md.py:10-21: NVTBerendsen with
// timestep = POTIM fs, temperature = TEBEG K, taut = 500 fs):
//
//   T      = 2 Ekin / (3 N kB),  Ekin = sum p^2 / 2m
//   s      = sqrt(1 + (T0 / T - 1) dt / taut), clipped to [0.9, 1.1]
//   p     <- s p + dt/2 F
//   p     <- p - sum(p) / N            (fixcm: total momentum removed)
//   x     <- x + dt p / m
//   F     <- forces at x               (a GAP evaluation of the batch)
//   p     <- p + dt/2 F
//
// with every structure of the batch advanced together.  taut <= 0 turns the
// scaling off (plain velocity Verlet).  Units are ASE's: positions A,
// momenta amu A / t, masses amu, forces eV/A, dt in ASE time units
// (1 fs = 0.09822694788464063 t).
#include <cmath>

#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

namespace {

__device__ __forceinline__ double wave_sum_md(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One workgroup per structure: 2 Ekin, sum p, sum F over its atoms, summed in
// a fixed order (lanes, then the four waves).
__device__ void struct_sums(const StructParams &P, const double *__restrict__ mass, const double *__restrict__ p,
                            const double *__restrict__ f, double out[7]) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int a = t; a < P.natoms; a += 256) {
    const size_t i = (size_t)(P.atom_base + a);
    const double m = mass[i];
    for (int c = 0; c < 3; ++c) {
      const double pc = p[3 * i + c];
      v[0] += pc * pc / m;
      v[1 + c] += pc;
      v[4 + c] += f[3 * i + c];
    }
  }
  __shared__ double red[4][7];
  for (int k = 0; k < 7; ++k) {
    const double s = wave_sum_md(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  for (int k = 0; k < 7; ++k) out[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_md_prepare(int nstruct, const StructParams *__restrict__ sp,
                                                    const double *__restrict__ mass, const double *__restrict__ p,
                                                    const double *__restrict__ f, MdParams mp,
                                                    double *__restrict__ coef) {
  const int s = blockIdx.x;
  if (s >= nstruct) return;
  const StructParams &P = sp[s];
  double sum[7];
  struct_sums(P, mass, p, f, sum);
  if (threadIdx.x != 0) return;
  const int n = P.natoms;
  double scale = 1.0;
  if (mp.taut > 0.0) {
    const double T = sum[0] / (3.0 * n * mp.kB);  // 2 Ekin / (3 N kB)
    scale = sqrt(1.0 + (mp.temperature / T - 1.0) * (mp.dt / mp.taut));
    // T = 0 gives +inf here and the clip to 1.1, as the division does in ASE
    if (!(scale <= 1.1)) scale = 1.1;
    if (scale < 0.9) scale = 0.9;
  }
  double *cf = coef + 4 * (size_t)s;
  cf[0] = scale;
  for (int c = 0; c < 3; ++c)
    cf[1 + c] = mp.fixcm ? (scale * sum[1 + c] + 0.5 * mp.dt * sum[4 + c]) / n : 0.0;
}

__global__ void k_md_drift(int N, const int *__restrict__ atom_struct, const double *__restrict__ mass,
                           const double *__restrict__ f, const double *__restrict__ coef, double dt,
                           double *__restrict__ p, double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double *cf = coef + 4 * (size_t)atom_struct[i];
  const double m = mass[i];
  for (int c = 0; c < 3; ++c) {
    const double pc = cf[0] * p[3 * (size_t)i + c] + 0.5 * dt * f[3 * (size_t)i + c] - cf[1 + c];
    p[3 * (size_t)i + c] = pc;
    x[3 * (size_t)i + c] += dt * pc / m;
  }
}

__global__ void k_md_kick(int N, const double *__restrict__ f, double half_dt, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * N) p[i] += half_dt * f[i];
}

// log[s] = (E_pot, E_kin, T) of structure s
__global__ __launch_bounds__(256) void k_md_log(int nstruct, const StructParams *__restrict__ sp,
                                                const double *__restrict__ mass, const double *__restrict__ p,
                                                const double *__restrict__ f, const double *__restrict__ ene,
                                                double kB, double *__restrict__ log) {
  const int s = blockIdx.x;
  if (s >= nstruct) return;
  double sum[7];
  struct_sums(sp[s], mass, p, f, sum);
  if (threadIdx.x != 0) return;
  log[3 * (size_t)s + 0] = ene[s];
  log[3 * (size_t)s + 1] = 0.5 * sum[0];
  log[3 * (size_t)s + 2] = sum[0] / (3.0 * sp[s].natoms * kB);
}

}  // namespace

hipError_t launch_md_prepare(hipStream_t s, int nstruct, const StructParams *sp, const double *mass, const double *p,
                             const double *f, const MdParams &mp, double *coef) {
  k_md_prepare<<<nstruct, 256, 0, s>>>(nstruct, sp, mass, p, f, mp, coef);
  return hipGetLastError();
}

hipError_t launch_md_drift(hipStream_t s, int N, const int *atom_struct, const double *mass, const double *f,
                           const double *coef, double dt, double *p, double *x) {
  k_md_drift<<<(N + 255) / 256, 256, 0, s>>>(N, atom_struct, mass, f, coef, dt, p, x);
  return hipGetLastError();
}

hipError_t launch_md_kick(hipStream_t s, int N, const double *f, double half_dt, double *p) {
  k_md_kick<<<(3 * N + 255) / 256, 256, 0, s>>>(N, f, half_dt, p);
  return hipGetLastError();
}

hipError_t launch_md_log(hipStream_t s, int nstruct, const StructParams *sp, const double *mass, const double *p,
                         const double *f, const double *ene, double kB, double *log) {
  k_md_log<<<nstruct, 256, 0, s>>>(nstruct, sp, mass, p, f, ene, kB, log);
  return hipGetLastError();
}

}  // namespace gapx
