// Internal definitions shared by the gapx host code and its HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

struct gapx_mlp;

namespace gapx {

constexpr double kPi = 3.141592654;       // truncated pi of wacsf.f90:31
constexpr double kLog2e = 1.4426950408889634;
constexpr double kGPa = 6.24219e-3;       // eV/A^3 per GPa, gap_calc.f90:9
constexpr int kRefNeighborCap = 1000;     // gap_calc.f90:68
constexpr int kMaxSpeciesZ = 128;

// Symmetry-function tables as the kernels read them.  Radial lines (types 1
// and 3) are exp(-a (r - rs)^2) fc_g(r); angular lines (types 2 and 4) are
// pref (1 + lam cos)^zeta exp(-eta (rij^2 + rik^2 + rjk^2)) fc_g fc_g fc_g.
// Angular lines are sorted by (group, eta, lam, zeta) so that consecutive
// lines share exponentials and cutoff products.
struct SFTables {
  int n_rad = 0, n_ang = 0, n_groups = 0;
  double rc_max_rad = 0, rc_max_ang = 0;
  // radial
  std::vector<double> rad_a, rad_rs;
  std::vector<int> rad_group, rad_feat;
  // angular
  std::vector<double> ang_eta, ang_lam, ang_pref;
  std::vector<int> ang_zeta, ang_group, ang_feat;
  std::vector<double> ang_zeta_real;   // non-integer zeta support (pow path)
  std::vector<double> group_rc;        // cutoff per group
};

// Device-side pointers of the SF tables (uniform loads inside kernels).
struct SFDev {
  int n_rad, n_ang, n_groups, nsf;
  double rc_max_rad, rc_max_ang;
  // Grid form of the angular table (set when every angular line shares one
  // cutoff group and the sorted lines are eta x (lambda = +1, -1) x zeta with
  // the same zeta list, zeta in {1, 2, 4, 8, 16}): line s = h * 2 nz + li * nz + k.
  int grid, n_eta, nz, g_ang;
  int zlog[5];             // log2 zeta_k
  double zeta_k[5], pref_k[5];
  const double *rad_a, *rad_rs;
  const float *rad_a32, *rad_rs32;  // the same, rounded to FP32
  const int *rad_group, *rad_feat;
  const double *ang_eta, *ang_lam, *ang_pref, *ang_zeta_real;
  const int *ang_zeta, *ang_group, *ang_feat;
  const double *group_rc;
};

struct Model {
  int device = 0;
  int nspecies = 0, nsf = 0, D = 0, M = 0;
  int Dp = 0, Mp = 0;                    // padded (multiples of 32 / 64)
  std::vector<int> z;
  std::vector<double> w;
  std::vector<int> sf_type;
  std::vector<double> sf_alpha, sf_cut, sf_zeta;
  std::vector<double> theta, mm, coeff;  // host copies, mm row-major [M][D]
  SFTables sf;
  // device
  void *dev_block = nullptr;             // one allocation for everything below
  SFDev sfd{};
  double *d_wz = nullptr;                // weight by atomic number [kMaxSpeciesZ]
  int *d_hasz = nullptr;                 // 1 if listed
  double *d_theta_inv = nullptr;         // [Dp] (0 in padding)
  double *d_mu = nullptr;                // [Dp] column mean of mm/theta
  double *d_mt = nullptr;                // [Mp][Dp] centred scaled sparse set
  double *d_mn = nullptr;                // [Mp] |mt|^2
  double *d_coeff = nullptr;             // [Mp] (0 in padding)
  double *d_mnq = nullptr;               // [Mp] -|mt|^2 log2(e) / 2 (mixed contraction)
  float *d_cq = nullptr;                 // [Mp] coeff * w_scale in FP32 (mixed contraction)
  uint16_t *d_mt_hi = nullptr, *d_mt_lo = nullptr;     // FP16 split [Mp][Dp]
  uint16_t *d_mt_bhi = nullptr, *d_mt_blo = nullptr;   // BF16 split [Mp][Dp]
  uint16_t *d_mtT_hi = nullptr, *d_mtT_lo = nullptr;   // FP16 split, transposed [Dp][Mp]
  uint16_t *d_mtT_bhi = nullptr, *d_mtT_blo = nullptr; // BF16 split, transposed
  double *d_qmm = nullptr;               // [Mp][Mp] variance matrix Q, zero padded (own allocation)
  float w_scale = 1024.f;                // power of 2 applied to W before its FP16 split
  double cancel_probe = 0.0;             // largest sum_k |W_jk| over the probed sparse points j
  bool mixed_f64 = false;                // mixed runs use the FP64 contraction (see upload_model)
  double desc_probe = 0.0;               // largest first-order dE_j per unit relative feature error
  bool mixed_desc_f64 = false;           // mixed runs are FP64 runs (FP32 features too coarse)
  size_t dev_bytes = 0;
  struct Ctx *cached_ctx = nullptr;      // workspace for gapx_calc
};

// Per-structure cell-list parameters (host computes, device reads).
struct StructParams {
  double L[9];       // rows = lattice vectors
  double Linv[9];    // inverse
  double vol;
  int nb[3];         // bins per axis
  int S[3];          // search half-width in bins
  int bin_base;      // offset into the global bin array
  int atom_base;     // first atom
  int natoms;
  int pad;
  // R00 R01 R02 R11 R12 R22 of the upper triangular R with R^T R = E E^T,
  // E the bin edge vectors (rows L_k / nb_k): |sum_k f_k E_k| = |R f| for a
  // displacement of f_k bins along axis k (zeros where the factor does not
  // exist numerically: no bound)
  float R[6];
};

// Cell-list parameters of one structure from its lattice rows L9: volume,
// inverse, bins per axis (at least rcut / 2 wide, at most nb_max), search
// half-widths and the bin metric factor R; GAPX_ERR_ARG for a singular
// lattice or a plane spacing below rcut / 48.  bin_base, atom_base and natoms
// are left 0.
int cell_params(const double *L9, double rcut, int nb_max, StructParams &P);

// Thread-local error message.
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define GAPX_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::gapx::fail(GAPX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- kernel launchers (kernels live in the .hip files) ----------------------
struct NeighborBufs {
  int *bin_count, *bin_start, *bin_cursor, *bin_atoms;  // [nbins_total (+1)], bin_atoms [N]
  int *atom_bin;        // [N]
  double *wpos;         // [N*3] wrapped positions
  int *shift;           // [N*3] floor(frac)
  int kcap;             // slots per atom: atom i's neighbours are slots i*kcap + [0, nb_count[i])
  int *nb_count;        // [N]
  int *nb_j;            // [N*kcap]
  unsigned long long *nb_key;  // [N*kcap] packed (j, image) for the reverse-slot search
  // nonzero: nb_key holds 32-bit keys instead (j < 2^17 and image fields of
  // 5 bits, valid while every search half-width S <= 15)
  int key32;
  double *nb_d;         // [N*kcap*4] dx dy dz r (relative to the centre)
  int *nb_rev;          // [N*kcap] slot (j -> i, -n) of slot (i -> j, n), or -1
  int *nb_raw;          // [N] neighbour count before the kcap clamp
  double *spos;         // [N*4] wrapped positions in bin order (x y z, pad)
  int *sinfo;           // [N*4] in bin order: atom index, its three cell shifts
  int *overflow;        // = total + 2
  int *total;           // [8] see launch_forces_reduce; [3] is the descriptor kernels' guard
  // per-group cutoff tables of the descriptor kernels, [N][2 G kcap] values of
  // the run's precision, for lists too long for the kernels' LDS layouts
  // (desc_table_scratch_bytes); null otherwise
  void *fc_scratch;
  // per-slot neighbour weights given by the caller (CAR2ACSF's neighbour
  // table); null: the weight of the neighbour's species from the model
  const double *nb_w;
};

// Keys the reverse-slot search reads per batch; nb_key carries that many
// keys of padding past the last slot.
constexpr int kRevBatch = 16;

// Wrap positions into their cells, bin and count the atoms, and scan the
// counts into bin_start, in one launch; needs bin_count and *ticket zero on
// entry and leaves them zero (with bin_cursor[0, nbins) and stats[0..7]).
hipError_t launch_wrap_scan(hipStream_t s, int N, int nbins, const double *pos, const int *atom_struct,
                            const StructParams *sp, NeighborBufs nb, int *ticket, int *stats);
hipError_t launch_scatter_bins(hipStream_t s, int N, int nbins, NeighborBufs nb);
// The whole binning of structures with at most kFusedAtoms atoms each, one
// workgroup per structure (also clears stats[0..7]); its LDS holds the
// largest structure's bins and atoms (fused_bin_lds_bytes, at most
// kFusedLdsBytes).  Measured on the 4096-atom workload, one workgroup takes
// longer than the three launches (wrap + count + scan, scatter, rank), so
// the one-launch path is for batches of structures up to kFusedAtoms atoms.
constexpr int kFusedAtoms = 1024;
constexpr size_t kFusedLdsBytes = 64 * 1024;
size_t fused_bin_lds_bytes(int atoms_max, int bins_max);
hipError_t launch_bin_structures(hipStream_t s, int nstruct, const double *pos, const StructParams *sp,
                                 NeighborBufs nb, int nbins_total, int atoms_max, int bins_max, int *stats);
hipError_t launch_reverse_slots(hipStream_t s, int N, NeighborBufs nb);
hipError_t launch_fill_int(hipStream_t s, int *p, int n, int v);
hipError_t launch_zero_ints(hipStream_t s, int *a, int na, int *b, int nb, int *c, int nc);
hipError_t launch_fill_double(hipStream_t s, double *p, size_t n, double v);

struct DescOut {
  int N, Dp;
  double *x64;          // [N][Dp] raw features (FP64 run) or null
  float *x32;           // [N][Dp] raw features (FP32 run) or null
  double *xt64;         // [N][Dp] centred scaled (FP64)
  float *xt32;          // [N][Dp] centred scaled (FP32, mixed run)
  uint16_t *xt_hi, *xt_lo;   // [N][Dp] split operands (mixed)
  double *xn;           // [N] |xt|^2 (FP64)
};

// Per-atom triplet lists: neighbour pairs (a, b), a < b, with r_a, r_b and
// r_ab within the largest angular cutoff, packed a | b << 16 at i * tcap.
struct TripletBufs {
  int tcap;
  uint32_t *list;       // [N*tcap]
  int *count;           // [N] kept triplets (<= tcap)
  int *raw;             // [N] triplet count before the tcap clamp
};

// Neighbour lists (one wavefront per atom); with tb also the triplet lists
// for angular cutoff rcA, from the same wave.
hipError_t launch_build_neighbors(hipStream_t s, int N, const int *atom_struct, const StructParams *sp,
                                  double rcut, NeighborBufs nb, double rcA, const TripletBufs *tb);
// FGET_BOND over the binned structure 0 of sp: *out = the bit pattern of
// min(init, every |pos_i - (pos_j + n.L)| <= rcut with |n_k| <= nabc_k and
// (j, n) != (i, 0)), distances in the reference's order of operations;
// candidates are taken within rsel (>= rcut) of the wrapped positions.
hipError_t launch_min_bond(hipStream_t s, int N, const StructParams *sp, const double *pos, NeighborBufs nb,
                           double rsel, double rcut, const int nabc[3], unsigned long long *out,
                           unsigned long long init);
// Triplet lists (as launch_build_neighbors makes them) of neighbour slots
// that are already filled: nb_count, nb_d (and kcap) are read.
hipError_t launch_triplets_from_slots(hipStream_t s, int N, NeighborBufs nb, double rcA, TripletBufs tb);

// Descriptor forward for atoms [atom0, atom0 + N).
hipError_t launch_descriptors_fwd(hipStream_t s, bool fp64, int N, const int *species,
                                  const double *wz, const NeighborBufs &nb, const TripletBufs &tb,
                                  const SFDev &sf, const double *theta_inv, const double *mu,
                                  int split_mode, DescOut out, int chunk, int waves, int KP, int atom0 = 0);
// dE/dx for the backward pass: either given (g64 or g32, splits = 0), or
// finished from the contraction's split partials inside the backward kernel
// (splits > 0): E_i = sum_s Epart[s][i] and g = -(xt E - sum_s G[s][i]) theta_inv,
// summed in split order; E_i goes to eatom and g to g64/g32 as well.
// sum_{sp < n} load(sp) in ascending sp, with the loads issued eight at a
// time (a plain loop's accumulation makes every load wait for the one
// before it).  n >= 0; load is called with indices in [0, n) only.
template <typename F>
__device__ __forceinline__ double split_sum(int n, F load) {
  double s = 0.0;
  for (int s0 = 0; s0 < n; s0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = load(min(s0 + u, n - 1));
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (s0 + u < n) s += v[u];
  }
  return s;
}

// (row tile, split) of this workgroup of a contraction grid (tiles on x,
// splits on y).  Workgroups go to the eight XCDs of the chip in turn by
// linear id (x fastest), each XCD with its own L2; left as they are, every
// XCD runs every split and reads the whole sparse set.  When the grid allows
// an exact regrouping, XCD x takes only the splits with split % 8 == x (or,
// with fewer than eight splits, a share of the tiles of split x % splits),
// so each L2 holds an eighth (or 1/splits) of the sparse points.
__device__ __forceinline__ void xcd_tile_split(int &tile, int &split) {
  const int nx = gridDim.x, ny = gridDim.y;
  const int id = blockIdx.x + nx * blockIdx.y, x = id & 7, loc = id >> 3;
  if (ny % 8 == 0) {
    split = x + 8 * (loc / nx);
    tile = loc % nx;
  } else if (8 % ny == 0 && (nx * ny) % 8 == 0) {
    split = x % ny;
    tile = loc * (8 / ny) + x / ny;
  } else {
    tile = blockIdx.x;
    split = blockIdx.y;
  }
}

// Work unit of block b of a 1-D range of nb blocks (nb a multiple of 8):
// workgroups go to the eight XCDs in turn, so block b runs on XCD b % 8;
// giving XCD x the consecutive units [x nb/8, (x + 1) nb/8) keeps the atoms
// of one structure, whose lists point at each other, on one XCD and its L2.
// Units at or past the real count are padding (the caller returns).
__device__ __forceinline__ int xcd_unit(int b, int nb) { return (b & 7) * (nb >> 3) + (b >> 3); }
constexpr int xcd_blocks(int n) { return (n + 7) & ~7; }

struct GradSrc {
  const double *g64 = nullptr;
  const float *g32 = nullptr;
  int splits = 0, Np = 0;
  const double *Epart = nullptr, *Gpart = nullptr;
  const float *Gpart32 = nullptr;
  const double *xt64 = nullptr;
  const float *xt32 = nullptr;
  const double *theta_inv = nullptr;
  double *eatom = nullptr;
  double *g64_out = nullptr;
  float *g32_out = nullptr;
};
// Backward for atoms [atom0, atom0 + N).  Per-slot forces go to slot_f as
// three values per slot: FP64 (fp64), or FP32 from the FP32 kernels.
hipError_t launch_descriptors_bwd(hipStream_t s, bool fp64, int N, const int *species,
                                  const double *wz, const NeighborBufs &nb, const TripletBufs &tb,
                                  const SFDev &sf, int Dp, const GradSrc &gs,
                                  double *slot_f, double *atom_vir, int waves, int KP, int atom0 = 0);
// Largest per-atom neighbour count the descriptor kernels take: within the
// device's per-workgroup LDS (160 KB on gfx950) with the cutoff tables moved
// to global memory where needed, and at most kMaxNeighbors.
constexpr int kMaxNeighbors = 4096;
int max_neighbors_for_lds(bool fp64, const SFDev &sf, int Dp);
// Bytes of NeighborBufs::fc_scratch the descriptor kernels need for N atoms
// with KP slots each (0: their LDS layouts hold the cutoff tables).
size_t desc_table_scratch_bytes(bool fp64, const SFDev &sf, int Dp, int KP, int N);

// Work buffers of the reduction after the backward pass.  Per-structure sums
// run in parts of at most kReduceAtoms atoms, one workgroup each; the last
// part of a structure to finish (a per-structure ticket) adds the parts'
// partial sums in part order, so the result does not depend on timing.  The
// list statistics are folded by ceil(N / kReduceAtoms) workgroups with integer
// atomics (exact, so also independent of order); the last of them writes the
// output and clears the accumulators.  ticket and stats_acc are zero between
// runs (cleared once when allocated, then by the kernel itself).
constexpr int kReduceAtoms = 1024;
struct ReduceBufs {
  const int2 *part_map;  // [nparts] (structure, part index); a structure's parts are consecutive
  int nparts;
  double *part_sum;      // [nparts][8]
  int *ticket;           // [nstruct]
  int *stats_acc;        // [8], [7] the statistics ticket
  // [8] or null: the elementwise maximum of every run's statistics since the
  // caller last cleared it (a loop of runs checks its capacities once)
  int *sticky;
  // [1] or null: 1 when this run's lists overflowed (kcap, tcap or the
  // descriptor kernels' layout), else 0; written by every run
  int *overflow_now;
};

// After the backward pass, one launch: per-atom forces (zeros without grad),
// per-structure energy and stress, and the list statistics: out[0..7] = kept
// pairs, largest raw pair count, atoms above kcap, descriptor guard
// (nb.total[3]), kept triplets, largest raw triplet count, atoms above tcap, 0.
// stats_out may be pinned host memory.  slot_f holds FP32 slot forces when
// slot32 (the FP32 backward writes them so), else FP64; sums are FP64.
hipError_t launch_forces_reduce(hipStream_t s, int N, int grad, const NeighborBufs &nb, const void *slot_f,
                                bool slot32, double *force, int nstruct, const StructParams *sp,
                                const double *eatom, const double *atom_vir, double *ene, double *stress6,
                                const TripletBufs &tb, const ReduceBufs &rb, int *stats_out);

// Contraction: E_i = sum_k exp(-|xt_i - mt_k|^2 / 2) c_k and
// g_i = -(xt_i E_i - sum_k W_ik mt_k) * theta_inv.
struct GprArgs {
  int N, Np, Dp, Mp, splits, pts_per_split;
  int rows;                       // rows per workgroup (64; the mixed kernel also 128)
  int tile0 = 0, ntiles = 0;      // row tiles [tile0, tile0 + ntiles) of this launch (0: all)
  const double *xt64; const double *xn;
  const uint16_t *xt_hi, *xt_lo;
  const float *xt32;
  const double *mt, *mn, *coeff;
  const double *mnq;              // mixed: -|mt|^2 log2(e) / 2
  const float *cq;                // mixed: coeff * wscale
  const uint16_t *mt_hi, *mt_lo, *mtT_hi, *mtT_lo;
  const double *theta_inv;
  double *Gpart; double *Epart;   // [splits][Np][Dp], [splits][Np]
  float *Gpart32;                 // [splits][Np][Dp]: the mixed path's partials
  double *eatom;                  // [N]
  double *g64; float *g32;        // [N][Dp]
  int grad;
  int bf16;
  float wscale;                   // W multiplier before the 16-bit split (power of 2)
};
hipError_t launch_gpr_fp64(hipStream_t s, const GprArgs &a);
hipError_t launch_gpr_mixed(hipStream_t s, const GprArgs &a);
int gpr_points_per_chunk(int Dp);  // sparse points per LDS chunk (both precisions)
hipError_t launch_gpr_finish(hipStream_t s, const GprArgs &a, bool fp64);

// MLP surrogate (mlp.hip), as used by the context.
struct Mlp;
Mlp *mlp_of(gapx_mlp *h);
// Surrogate energies (eatom, FP64) and, with grad, the descriptor-space
// gradient g = dE/dx (g32 or g64) for rows of standardised features xt
// (rows a multiple of 64, rows >= n zero).
int mlp_eval_into_gap(Mlp &m, hipStream_t s, bool split, int n, int rows, const float *xt, bool grad, int Dp,
                      const double *theta_inv, double *eatom, float *g32, double *g64);
// Append (xt, eatom, g / theta_inv) of n atoms to the surrogate's training set.
int mlp_add_labels(Mlp &m, hipStream_t s, int n, const float *xt32, const double *xt64, const float *g32,
                   const double *g64, const double *eatom, const double *theta_inv, int Dp);
hipError_t launch_f64_to_f32(hipStream_t s, size_t n, const double *a, float *b);

// Batched FIRE relaxation (relax.hip): per-structure state and parameters.
struct FireState {
  double dt, a;        // time step, mixing parameter
  double cv, cf, step; // this step: v' = cv v + cf F, x += step v'
  double fmax;         // largest |F_i| at the last evaluation
  int nsteps;          // downhill steps since the last restart
  int started, converged, pad;
};
struct FireParams {
  double dt0, dtmax, finc, fdec, astart, fa, maxstep, fmax;
  int nmin, pad;
};
// Device-side control of a relaxation step (both optimisers): when *skip is
// set (the evaluation the step would use overflowed its lists) the step
// changes nothing and sets bit 2 of *flag; a step that moves any structure
// sets bit 1.  Either pointer may be null.
struct StepCtl {
  const int *skip;
  int *flag;
};
hipError_t launch_fire_step(hipStream_t s, int nstruct, int N, const StructParams *sp, const int *atom_struct,
                            const double *force, double *vel, double *pos, FireState *st, const FireParams &fp,
                            StepCtl ctl);

// Batched LBFGS (relax.hip): per-structure state and parameters.
constexpr int kLbfgsMaxMemory = 1024;
struct LbfgsState {
  int iter;       // steps taken
  int count;      // history entries held, min(iter - 1, memory) once started
  int head;       // ring slot of the next entry
  int converged;
  double fmax;    // largest |F_i| at the last evaluation
};
struct LbfgsParams {
  double maxstep, h0, fmax;
  int memory;
  int check_only;  // only the convergence test (after the last step)
};
// S, Y: [memory][3 N]; rho: [nstruct][memory]; r0, f0, q: [3 N].
hipError_t launch_lbfgs_step(hipStream_t s, int nstruct, int N, const StructParams *sp, const double *force,
                             double *pos, double *r0, double *f0, double *S, double *Y, double *rho, double *q,
                             LbfgsState *st, const LbfgsParams &lp, StepCtl ctl);

// Batched NVT MD (md.hip): ASE NVTBerendsen steps in ASE units.
struct MdParams {
  double dt, temperature, taut, kB;
  int fixcm, pad;
};
// coef[4 s + 0] = velocity scale of structure s, [1..3] = momentum shift per atom
hipError_t launch_md_prepare(hipStream_t s, int nstruct, const StructParams *sp, const double *mass, const double *p,
                             const double *f, const MdParams &mp, double *coef);
hipError_t launch_md_drift(hipStream_t s, int N, const int *atom_struct, const double *mass, const double *f,
                           const double *coef, double dt, double *p, double *x);
hipError_t launch_md_kick(hipStream_t s, int N, const double *f, double half_dt, double *p);
hipError_t launch_md_log(hipStream_t s, int nstruct, const StructParams *sp, const double *mass, const double *p,
                         const double *f, const double *ene, double kB, double *log);

// Predictive variance (variance.hip): sigma2_i = 1 - k_i Q k_i^T over the
// last run's standardised features, in blocks of `rows` atoms.
struct VarianceArgs {
  int N = 0, Dp = 0, Mp = 0, nstruct = 0;
  int rows = 0;                   // atoms per block of K (a multiple of 64; K holds rows x Mp)
  const double *xt64 = nullptr;   // [N][Dp] the run's features: FP64 run ...
  const float *xt32 = nullptr;    // ... or mixed run
  const double *mt = nullptr, *mn = nullptr;
  const double *qmm = nullptr;    // [Mp][Mp] row-major, zero padded; null: Q = 0
  const StructParams *sp = nullptr;
  double *xw = nullptr;           // [N][Dp] FP64 copy of xt32
  double *xn = nullptr;           // [N]
  double *K = nullptr;            // [rows][Mp]
  double *partial = nullptr;      // [Mp / 64][N] per-column-tile sums of the quadratic form
  double *sigma2 = nullptr;       // [N]
  double *var = nullptr;          // [nstruct]
};
int enqueue_variance(hipStream_t s, const VarianceArgs &v);

}  // namespace gapx
