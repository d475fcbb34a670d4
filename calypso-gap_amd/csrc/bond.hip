// Minimum interatomic distance over periodic images (FGET_BOND,
// gappy/libgap/get_bond.f90:4-117), used to screen unphysical candidates.
//
// Same image set as the reference: for every atom pair and every image
// n in [-n_k, n_k] with n_k = ceil(rcut |b_k| / 2 pi) (reciprocal vectors
// from the columns of LAT, get_bond.f90:40-43 and :89-98), skipping (i, i, 0)
// and distances above rcut; the result starts at 10 (get_bond.f90:32), so it
// is min(10, shortest distance within rcut).  One thread per (i, j) pair
// loops over the images; positive doubles order like their bit patterns, so
// the minimum is an integer atomicMin and independent of thread order.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

namespace {

__global__ void k_min_dist(int na, const double *__restrict__ lat, const double *__restrict__ pos, double rcut,
                           int n0, int n1, int n2, unsigned long long *__restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)na * na) return;
  const int i = (int)(t / na), j = (int)(t % na);
  double best = 10.0;
  for (int a = -n0; a <= n0; ++a)
    for (int b = -n1; b <= n1; ++b)
      for (int c = -n2; c <= n2; ++c) {
        if (a == 0 && b == 0 && c == 0 && i == j) continue;
        double d2 = 0.0;
        for (int x = 0; x < 3; ++x) {
          const double xyz = pos[3 * j + x] + (double)a * lat[x] + (double)b * lat[3 + x] + (double)c * lat[6 + x];
          const double dr = pos[3 * i + x] - xyz;
          d2 += dr * dr;
        }
        const double dis = sqrt(d2);
        if (dis > rcut) continue;
        if (dis < best) best = dis;
      }
  atomicMin(out, (unsigned long long)__double_as_longlong(best));
}

void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

// Image counts of gap_calc.f90:85-88 / get_bond.f90:40-43 for row-major lat.
void reference_image_counts(const double *lat, double rcut, int nabc[3]) {
  double col[3][3], rc[3][3];
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 3; ++k) col[c][k] = lat[3 * k + c];
  cross3(col[1], col[2], rc[0]);
  cross3(col[2], col[0], rc[1]);
  cross3(col[0], col[1], rc[2]);
  const double vol = std::fabs(col[0][0] * rc[0][0] + col[0][1] * rc[0][1] + col[0][2] * rc[0][2]);
  for (int r = 0; r < 3; ++r) {
    const double v0 = rc[0][r] / vol * kPi * 2.0, v1 = rc[1][r] / vol * kPi * 2.0, v2 = rc[2][r] / vol * kPi * 2.0;
    nabc[r] = (int)std::ceil(rcut * std::sqrt(v0 * v0 + v1 * v1 + v2 * v2) / kPi / 2.0);
  }
}

}  // namespace gapx

extern "C" int gapx_min_distance(int na, const double *lat, const double *pos, double rcut, int flags, double *dmin) {
  using namespace gapx;
  if (na <= 0 || !lat || !pos || !dmin || !(rcut > 0)) return fail(GAPX_ERR_ARG, "gapx_min_distance: invalid argument");
  double L[9];
  std::vector<double> p(3 * (size_t)na);
  if (flags & GAPX_LAYOUT_F) {
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) L[3 * k + c] = lat[k + 3 * c];
    for (int i = 0; i < na; ++i)
      for (int c = 0; c < 3; ++c) p[3 * (size_t)i + c] = pos[i + (size_t)na * c];
  } else {
    std::memcpy(L, lat, sizeof L);
    std::memcpy(p.data(), pos, sizeof(double) * 3 * na);
  }
  int n[3];
  reference_image_counts(L, rcut, n);
  if ((double)(2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1) > 1e6)
    return fail(GAPX_ERR_ARG, "gapx_min_distance: cell far smaller than rcut");
  double *d_lat = nullptr, *d_pos = nullptr;
  unsigned long long *d_out = nullptr;
  const double init = 10.0;
  unsigned long long h;
  std::memcpy(&h, &init, sizeof h);
  hipStream_t s = nullptr;
  int rc = GAPX_OK;
  if (hipMalloc(&d_lat, sizeof L) != hipSuccess || hipMalloc(&d_pos, sizeof(double) * 3 * na) != hipSuccess ||
      hipMalloc(&d_out, sizeof h) != hipSuccess) {
    rc = fail(GAPX_ERR_NOMEM, "gapx_min_distance: hipMalloc failed");
  } else if (hipMemcpy(d_lat, L, sizeof L, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_pos, p.data(), sizeof(double) * 3 * na, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_out, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(GAPX_ERR_HIP, "gapx_min_distance: copy failed");
  } else {
    const long long pairs = (long long)na * na;
    k_min_dist<<<(unsigned)((pairs + 255) / 256), 256, 0, s>>>(na, d_lat, d_pos, rcut, n[0], n[1], n[2], d_out);
    if (hipGetLastError() != hipSuccess || hipMemcpy(&h, d_out, sizeof h, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(GAPX_ERR_HIP, "gapx_min_distance: kernel failed");
    else
      std::memcpy(dmin, &h, sizeof h);
  }
  if (d_lat) (void)hipFree(d_lat);
  if (d_pos) (void)hipFree(d_pos);
  if (d_out) (void)hipFree(d_out);
  return rc;
}
