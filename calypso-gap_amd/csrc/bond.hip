// Minimum interatomic distance over periodic images (FGET_BOND,
// gappy/libgap/get_bond.f90:4-117), used to screen unphysical candidates.
//
// The reference tests every atom pair against every image n in [-n_k, n_k],
// n_k = ceil(rcut |b_k| / 2 pi) (reciprocal vectors from the columns of LAT,
// get_bond.f90:40-43 and :89-98), skipping (i, i, 0) and distances above
// rcut, from a starting value of 10 (get_bond.f90:32): min(10, shortest
// distance within rcut).  Here the structure is binned (the evaluator's
// cell-list kernels, bins at least rcut / 2 wide) and one wavefront per atom
// visits only the candidates of the neighbouring bins; each candidate within
// rcut gets its image index relative to the positions as given, is kept only
// inside the reference's image range, and its distance is recomputed in the
// reference's order of operations, so the minimum is the reference's bit for
// bit (k_min_bond in neighbors.hip).
//
// Two entry points share the pipeline: gapx_min_distance allocates its
// workspace per call; gapx_min_distance_noalloc works in a fixed file-scope
// device array (`scratch`) and allocates nothing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "../../include/gapx.h"
#include "gapx_internal.h"

// The allocation-free entry point's workspace: one 16 MiB array per GPU, kept
// until the process ends.  Host code never passes `scratch` itself (in host
// code the name is a stand-in, not a device address) but the address
// hipGetSymbolAddress gives for the current device, kept per device.
static __device__ __align__(256) unsigned char scratch[16u << 20];

namespace gapx {

void reference_image_counts(const double *lat, double rcut, int nabc[3]);

namespace {

// Device workspace of one FGET_BOND call: the staged inputs and the binning
// arrays of the cell-list kernels.
struct BondWork {
  double *pos = nullptr;       // [3 N] positions as given, C layout
  double *pos_f = nullptr;     // [3 N] the caller's Fortran-layout positions (staging)
  StructParams *sp = nullptr;  // [1]
  int *astruct = nullptr;      // [N] structure of every atom (all 0)
  int *ticket = nullptr;       // [1] the binning's last-block ticket
  int *stats = nullptr;        // [8] the binning's statistics (unused here)
  unsigned long long *out = nullptr;  // the minimum's bit pattern
  NeighborBufs nb{};
};

// Pieces of the workspace for N atoms and nbins bins, 256-byte aligned, from
// base (null: sizes only); returns the bytes needed.
size_t bond_layout(char *base, int N, int nbins, BondWork &w) {
  size_t off = 0;
  auto take = [&](auto **p, size_t count) {
    using T = typename std::remove_reference<decltype(**p)>::type;
    if (base) *p = (T *)(base + off);
    off += ((std::max<size_t>(count, 1) * sizeof(T) + 255) / 256) * 256;
  };
  NeighborBufs &nb = w.nb;
  take(&w.pos, 3 * (size_t)N);
  take(&w.pos_f, 3 * (size_t)N);
  take(&w.sp, 1);
  take(&w.astruct, N);
  take(&w.ticket, 1);
  take(&w.stats, 8);
  take(&w.out, 1);
  take(&nb.bin_count, nbins);
  take(&nb.bin_start, (size_t)nbins + 1);
  take(&nb.bin_cursor, std::max(nbins, N));
  take(&nb.bin_atoms, N);
  take(&nb.atom_bin, N);
  take(&nb.wpos, 3 * (size_t)N);
  take(&nb.shift, 3 * (size_t)N);
  take(&nb.spos, 4 * (size_t)N);
  take(&nb.sinfo, 4 * (size_t)N);
  return off;
}

void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Fortran POS(N, 3) (column-major) to C layout pos[3 i + c].
__global__ void k_positions_from_fortran(int N, const double *__restrict__ pf, double *__restrict__ pc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < 3; ++c) pc[3 * i + c] = pf[i + (size_t)N * c];
}

// The sizes of a call: lattice rows, image counts, the cell list's bins.
struct BondShape {
  double L[9];
  int nabc[3];
  double rsel;
  StructParams P;
  int nbins;
};

int bond_shape(int na, const double *lat, double rcut, int flags, BondShape &b) {
  if (flags & GAPX_LAYOUT_F) {
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) b.L[3 * k + c] = lat[k + 3 * c];
  } else {
    std::memcpy(b.L, lat, sizeof b.L);
  }
  reference_image_counts(b.L, rcut, b.nabc);
  // candidates are taken a hair beyond rcut on the wrapped positions, so a
  // distance the reference's arithmetic puts at rcut is not lost to rounding
  b.rsel = rcut * (1.0 + 1e-9) + 1e-12;
  int rc = cell_params(b.L, b.rsel, 64, b.P);
  if (rc) return rc;
  b.P.natoms = na;
  b.nbins = b.P.nb[0] * b.P.nb[1] * b.P.nb[2];
  return GAPX_OK;
}

// Stage the inputs into the workspace (base, laid out by bond_layout), bin,
// search, and read the minimum back, all on the null stream.  Everything the
// kernels need zeroed is cleared here: the workspace holds whatever the last
// call left.
int bond_run(char *base, int na, const double *pos, double rcut, int flags, const BondShape &b, double *dmin) {
  BondWork w;
  bond_layout(base, na, b.nbins, w);
  hipStream_t s = nullptr;
  GAPX_HIP(hipMemcpy(w.sp, &b.P, sizeof(StructParams), hipMemcpyHostToDevice));
  if (flags & GAPX_LAYOUT_F) {
    GAPX_HIP(hipMemcpy(w.pos_f, pos, sizeof(double) * 3 * na, hipMemcpyHostToDevice));
    k_positions_from_fortran<<<(na + 255) / 256, 256, 0, s>>>(na, w.pos_f, w.pos);
    GAPX_HIP(hipGetLastError());
  } else {
    GAPX_HIP(hipMemcpy(w.pos, pos, sizeof(double) * 3 * na, hipMemcpyHostToDevice));
  }
  NeighborBufs nb = w.nb;
  nb.total = w.stats;
  nb.overflow = w.stats + 2;
  if (na <= kFusedAtoms && fused_bin_lds_bytes(na, b.nbins) <= kFusedLdsBytes) {
    GAPX_HIP(launch_bin_structures(s, 1, w.pos, w.sp, nb, b.nbins, na, b.nbins, w.stats));
  } else {
    GAPX_HIP(launch_fill_int(s, w.astruct, na, 0));
    GAPX_HIP(launch_zero_ints(s, nb.bin_count, b.nbins, w.ticket, 1, nullptr, 0));
    GAPX_HIP(launch_wrap_scan(s, na, b.nbins, w.pos, w.astruct, w.sp, nb, w.ticket, w.stats));
    GAPX_HIP(launch_scatter_bins(s, na, b.nbins, nb));
  }
  const double init = 10.0;  // get_bond.f90:32
  unsigned long long bits;
  std::memcpy(&bits, &init, sizeof bits);
  GAPX_HIP(launch_min_bond(s, na, w.sp, w.pos, nb, b.rsel, rcut, b.nabc, w.out, bits));
  GAPX_HIP(hipMemcpy(&bits, w.out, sizeof bits, hipMemcpyDeviceToHost));
  std::memcpy(dmin, &bits, sizeof bits);
  return GAPX_OK;
}

constexpr size_t kBondScratchBytes = 16u << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(char **ptr) {
  int dev = -1;
  GAPX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GAPX_ERR_ARG, "gapx_min_distance_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !p)
      return fail(GAPX_ERR_HIP, std::string("gapx_min_distance_noalloc: hipGetSymbolAddress of the scratch array: ") +
                                    hipGetErrorString(e));
    g_scratch_ptr[dev] = (char *)p;
  }
  *ptr = g_scratch_ptr[dev];
  return GAPX_OK;
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
  BondShape b;
  int rc = bond_shape(na, lat, rcut, flags, b);
  if (rc) return rc;
  BondWork w;
  const size_t bytes = bond_layout(nullptr, na, b.nbins, w);
  char *base = nullptr;
  if (hipMalloc((void **)&base, bytes) != hipSuccess) return fail(GAPX_ERR_NOMEM, "gapx_min_distance: hipMalloc failed");
  rc = bond_run(base, na, pos, rcut, flags, b, dmin);
  (void)hipFree(base);
  return rc;
}

extern "C" int gapx_min_distance_noalloc(int na, const double *lat, const double *pos, double rcut, int flags,
                                         double *dmin) {
  using namespace gapx;
  if (na <= 0 || !lat || !pos || !dmin || !(rcut > 0))
    return fail(GAPX_ERR_ARG, "gapx_min_distance_noalloc: invalid argument");
  BondShape b;
  int rc = bond_shape(na, lat, rcut, flags, b);
  if (rc) return rc;
  BondWork w;
  const size_t bytes = bond_layout(nullptr, na, b.nbins, w);
  static_assert(sizeof(::scratch) == kBondScratchBytes, "scratch size");
  if (bytes > kBondScratchBytes) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "gapx_min_distance_noalloc: %d atoms in %d bins need %zu bytes of scratch, above its %zu "
                  "(use gapx_min_distance)",
                  na, b.nbins, bytes, kBondScratchBytes);
    return fail(GAPX_ERR_ARG, buf);
  }
  char *base = nullptr;
  if ((rc = scratch_of_current_device(&base))) return rc;
  return bond_run(base, na, pos, rcut, flags, b, dmin);
}
