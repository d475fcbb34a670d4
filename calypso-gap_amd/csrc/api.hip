// gapx host code: model upload, device workspaces, the evaluation pipeline
// and the C ABI declared in include/gapx.h.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------

static int build_tables(Model &m) {
  SFTables &t = m.sf;
  t = SFTables{};
  std::vector<double> cuts;
  for (int s = 0; s < m.nsf; ++s) {
    int ty = m.sf_type[s];
    if (ty < 1 || ty > 4) continue;
    bool seen = false;
    for (double c : cuts) seen = seen || (c == m.sf_cut[s]);
    if (!seen) cuts.push_back(m.sf_cut[s]);
  }
  std::sort(cuts.begin(), cuts.end(), [](double a, double b) { return a > b; });
  if ((int)cuts.size() > 16) return fail(GAPX_ERR_MODEL, "more than 16 distinct symmetry-function cutoffs");
  t.group_rc = cuts;
  t.n_groups = (int)cuts.size();
  auto group_of = [&](double c) {
    for (int g = 0; g < (int)cuts.size(); ++g)
      if (cuts[g] == c) return g;
    return 0;
  };
  struct Ang {
    int g, feat, zi;
    double eta, lam, zr, pref;
  };
  std::vector<Ang> ang;
  for (int s = 0; s < m.nsf; ++s) {
    const int ty = m.sf_type[s];
    const double al = m.sf_alpha[s], rc = m.sf_cut[s];
    if (rc <= 0) return fail(GAPX_ERR_MODEL, "symmetry-function cutoff must be positive");
    if (ty == 1 || ty == 3) {
      t.rad_a.push_back(ty == 1 ? al : 4.0);   // type 3 has alpha fixed at 4 (wacsf.f90:439)
      t.rad_rs.push_back(ty == 1 ? 0.0 : al);  // and the file's alpha column is Rs (:438)
      t.rad_group.push_back(group_of(rc));
      t.rad_feat.push_back(s);
      t.rc_max_rad = std::max(t.rc_max_rad, rc);
    } else if (ty == 2 || ty == 4) {
      const double z = m.sf_zeta.empty() ? 1.0 : m.sf_zeta[s];
      Ang a;
      a.g = group_of(rc);
      a.feat = s;
      a.eta = al;
      a.lam = (ty == 2) ? 1.0 : -1.0;
      a.zr = z;
      a.zi = (z == std::floor(z) && z >= 1.0 && z <= 64.0) ? (int)z : -1;
      a.pref = std::pow(2.0, 1.0 - z);
      ang.push_back(a);
      t.rc_max_ang = std::max(t.rc_max_ang, rc);
    } else {
      std::fprintf(stderr, "gapx: unknown symmetry-function type %d on line %d; its features are 0\n", ty, s + 1);
    }
  }
  std::stable_sort(ang.begin(), ang.end(), [](const Ang &x, const Ang &y) {
    if (x.g != y.g) return x.g < y.g;
    if (x.eta != y.eta) return x.eta < y.eta;
    if (x.lam != y.lam) return x.lam > y.lam;
    return x.zr < y.zr;
  });
  for (const Ang &a : ang) {
    t.ang_eta.push_back(a.eta);
    t.ang_lam.push_back(a.lam);
    t.ang_pref.push_back(a.pref);
    t.ang_zeta.push_back(a.zi);
    t.ang_zeta_real.push_back(a.zr);
    t.ang_group.push_back(a.g);
    t.ang_feat.push_back(a.feat);
  }
  t.n_rad = (int)t.rad_a.size();
  t.n_ang = (int)t.ang_eta.size();
  return GAPX_OK;
}

// Recognise an angular table of the form eta x (lambda = +1, -1) x zeta
// (one cutoff group, the same zeta list for every eta, zeta a power of two
// up to 16); the descriptor kernels then use their unrolled grid form.
static void detect_grid(const SFTables &t, SFDev &sd) {
  sd.grid = 0;
  if (t.n_ang == 0 || std::getenv("GAPX_NO_GRID")) return;
  const int g = t.ang_group[0];
  for (int s = 0; s < t.n_ang; ++s)
    if (t.ang_group[s] != g) return;
  int nz = 0;
  while (nz < t.n_ang && t.ang_eta[nz] == t.ang_eta[0] && t.ang_lam[nz] == 1.0) ++nz;
  if (nz == 0 || nz > 5 || t.n_ang % (2 * nz) != 0) return;
  const int ne = t.n_ang / (2 * nz);
  for (int h = 0; h < ne; ++h)
    for (int li = 0; li < 2; ++li)
      for (int k = 0; k < nz; ++k) {
        const int s = h * 2 * nz + li * nz + k;
        if (t.ang_eta[s] != t.ang_eta[h * 2 * nz]) return;
        if (t.ang_lam[s] != (li == 0 ? 1.0 : -1.0)) return;
        if (t.ang_zeta[s] != t.ang_zeta[k]) return;
      }
  for (int k = 0; k < nz; ++k) {
    const int z = t.ang_zeta[k];
    if (k > 0 && z < t.ang_zeta[k - 1]) return;  // the kernels advance one power chain
    if (z != 1 && z != 2 && z != 4 && z != 8 && z != 16) return;
    sd.zlog[k] = (z == 1) ? 0 : (z == 2) ? 1 : (z == 4) ? 2 : (z == 8) ? 3 : 4;
    sd.zeta_k[k] = z;
    sd.pref_k[k] = t.ang_pref[k];
  }
  sd.nz = nz;
  sd.n_eta = ne;
  sd.g_ang = g;
  sd.grid = 1;
}

static uint16_t f16_bits(double v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
static double f16_val(uint16_t b) { return (double)__builtin_bit_cast(_Float16, b); }
static uint16_t bf16_bits(double v) {
  float f = (float)v;
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  uint32_t r = 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)((u + r) >> 16);
}
static double bf16_val(uint16_t b) { return (double)__builtin_bit_cast(float, (uint32_t)b << 16); }

struct Arena {
  std::vector<std::pair<void **, size_t>> items;
  template <typename T>
  void add(T **p, size_t count) { items.push_back({(void **)p, ((count * sizeof(T) + 255) / 256) * 256}); }
  size_t total() const {
    size_t s = 0;
    for (auto &it : items) s += it.second;
    return s;
  }
};

static int upload_model(Model &m) {
  const int D = m.D, M = m.M, Dp = m.Dp, Mp = m.Mp;
  const SFTables &t = m.sf;
  // host-side derived arrays
  std::vector<double> wz(kMaxSpeciesZ, 0.0);
  std::vector<int> hasz(kMaxSpeciesZ, 0);
  for (int k = 0; k < m.nspecies; ++k) {
    if (m.z[k] < 0 || m.z[k] >= kMaxSpeciesZ) return fail(GAPX_ERR_MODEL, "atomic number out of range in model header");
    wz[m.z[k]] = m.w[k];  // later lines win, as in gap_calc.f90:77-82
    hasz[m.z[k]] = 1;
  }
  std::vector<double> tinv(Dp, 0.0), mu(Dp, 0.0), mt((size_t)Mp * Dp, 0.0), mn(Mp, 0.0), cf(Mp, 0.0);
  for (int d = 0; d < D; ++d) {
    if (!(m.theta[d] != 0.0)) return fail(GAPX_ERR_MODEL, "theta must be non-zero");
    tinv[d] = 1.0 / m.theta[d];
  }
  for (int d = 0; d < D; ++d) {
    double s = 0.0;
    for (int k = 0; k < M; ++k) s += m.mm[(size_t)k * D + d] * tinv[d];
    mu[d] = s / M;
  }
  for (int k = 0; k < M; ++k) {
    double n = 0.0;
    for (int d = 0; d < D; ++d) {
      double v = m.mm[(size_t)k * D + d] * tinv[d] - mu[d];
      mt[(size_t)k * Dp + d] = v;
      n += v * v;
    }
    mn[k] = n;
    cf[k] = m.coeff[k];
  }
  // W = K c with 0 < K <= 1: scale it by a power of two that keeps |W| below
  // 2^14 so its FP16 hi part cannot overflow and its lo part stays normal
  double cmax = 0.0;
  for (int k = 0; k < M; ++k) cmax = std::max(cmax, std::fabs(m.coeff[k]));
  int e2 = (cmax > 0) ? (int)std::floor(std::log2(16384.0 / cmax)) : 10;
  e2 = std::max(-24, std::min(e2, 10));
  m.w_scale = std::ldexp(1.0f, e2);
  // Whether the mixed contraction can hold the energy: E_i = sum_k W_ik is
  // computed from FP32 kernel values (relative error about 2^-21 each), so
  // its error grows with A_i = sum_k |W_ik|, not with |E_i|.  A_i is probed
  // on the model's own sparse points (the descriptor distribution the model
  // was fitted on), up to 64 of them, in FP64 with direct differences.  A
  // model whose terms cancel (A * 2^-21 above 1e-4 eV, a tenth of the
  // 1 meV/atom bar) or whose scaled points leave the FP16 range runs its
  // contraction in FP64 on the mixed path.
  // The FP32 descriptors themselves: a relative error u in the raw feature
  // x_d moves xt_d by u |x_d| / theta_d, and E_j to first order by
  // B_j u with B_j = sum_d |x_jd / theta_d| |dE_j / dxt_d| and
  // dE_j / dxt_d = sum_k c_k K_jk (mt_kd - mt_jd).  Features that are large
  // against their length scale theta (sums over about 10^3 neighbours in a
  // dense cell, with theta fitted on a narrow pool) make B large; at
  // B * 2^-24 above 1e-4 eV (measured: the FP32 sums err by up to 4 times
  // that figure, 1.35 meV/atom on a model with B * 2^-24 = 3.4e-4) mixed runs
  // of the model are FP64 runs.
  double mt_abs = 0.0;
  for (size_t q = 0; q < mt.size(); ++q) mt_abs = std::max(mt_abs, std::fabs(mt[q]));
  double amax = 0.0, bmax = 0.0;
  const int nprobe = std::min(M, 64);
  std::vector<double> gj(D);
  for (int p = 0; p < nprobe; ++p) {
    const int j = (nprobe > 1) ? (int)((long long)p * (M - 1) / (nprobe - 1)) : 0;
    double a = 0.0;
    std::fill(gj.begin(), gj.end(), 0.0);
    for (int k = 0; k < M; ++k) {
      double d2 = 0.0;
      for (int d = 0; d < D; ++d) {
        const double v = mt[(size_t)j * Dp + d] - mt[(size_t)k * Dp + d];
        d2 += v * v;
      }
      const double kv = std::exp(-0.5 * d2);
      a += std::fabs(m.coeff[k]) * kv;
      const double w = m.coeff[k] * kv;
      if (w != 0.0)
        for (int d = 0; d < D; ++d) gj[d] += w * (mt[(size_t)k * Dp + d] - mt[(size_t)j * Dp + d]);
    }
    double b = 0.0;
    for (int d = 0; d < D; ++d) b += std::fabs(mt[(size_t)j * Dp + d] + mu[d]) * std::fabs(gj[d]);
    amax = std::max(amax, a);
    bmax = std::max(bmax, b);
  }
  m.cancel_probe = amax;
  m.mixed_f64 = (amax * std::ldexp(1.0, -21) > 1e-4) || !(mt_abs <= 16384.0);
  m.desc_probe = bmax;
  m.mixed_desc_f64 = !(bmax * std::ldexp(1.0, -24) <= 1e-4);
  std::vector<uint16_t> hi((size_t)Mp * Dp), lo((size_t)Mp * Dp), bhi((size_t)Mp * Dp), blo((size_t)Mp * Dp);
  std::vector<uint16_t> thi((size_t)Mp * Dp), tlo((size_t)Mp * Dp), tbhi((size_t)Mp * Dp), tblo((size_t)Mp * Dp);
  // Feature-major copies: within every group of 32 points, point p sits at
  // position 8 ((p % 16) / 4) + p % 4 + 4 (p / 16), the K order in which the
  // mixed contraction holds W (see gpr_mixed.hip).
  for (int k = 0; k < Mp; ++k)
    for (int d = 0; d < Dp; ++d) {
      const int p = k & 31;
      const int kq = (k & ~31) + 8 * ((p & 15) >> 2) + (p & 3) + 4 * (p >> 4);
      const size_t a = (size_t)k * Dp + d, b = (size_t)d * Mp + kq;
      const double v = mt[a];
      hi[a] = thi[b] = f16_bits(v);
      lo[a] = tlo[b] = f16_bits(v - f16_val(hi[a]));
      bhi[a] = tbhi[b] = bf16_bits(v);
      blo[a] = tblo[b] = bf16_bits(v - bf16_val(bhi[a]));
    }
  // device layout
  Arena ar;
  double *rad_a, *rad_rs, *ang_eta, *ang_lam, *ang_pref, *ang_zr, *group_rc;
  float *rad_a32, *rad_rs32;
  int *rad_group, *rad_feat, *ang_zeta, *ang_group, *ang_feat;
  const int nr = std::max(1, t.n_rad), na = std::max(1, t.n_ang), ng = std::max(1, t.n_groups);
  ar.add(&rad_a, nr); ar.add(&rad_rs, nr); ar.add(&rad_group, nr); ar.add(&rad_feat, nr);
  ar.add(&rad_a32, nr); ar.add(&rad_rs32, nr);
  ar.add(&ang_eta, na); ar.add(&ang_lam, na); ar.add(&ang_pref, na); ar.add(&ang_zr, na);
  ar.add(&ang_zeta, na); ar.add(&ang_group, na); ar.add(&ang_feat, na); ar.add(&group_rc, ng);
  ar.add(&m.d_wz, kMaxSpeciesZ); ar.add(&m.d_hasz, kMaxSpeciesZ);
  ar.add(&m.d_theta_inv, Dp); ar.add(&m.d_mu, Dp); ar.add(&m.d_mt, (size_t)Mp * Dp); ar.add(&m.d_mn, Mp);
  ar.add(&m.d_coeff, Mp);
  ar.add(&m.d_mnq, Mp); ar.add(&m.d_cq, Mp);
  ar.add(&m.d_mt_hi, (size_t)Mp * Dp); ar.add(&m.d_mt_lo, (size_t)Mp * Dp);
  ar.add(&m.d_mt_bhi, (size_t)Mp * Dp); ar.add(&m.d_mt_blo, (size_t)Mp * Dp);
  ar.add(&m.d_mtT_hi, (size_t)Mp * Dp); ar.add(&m.d_mtT_lo, (size_t)Mp * Dp);
  ar.add(&m.d_mtT_bhi, (size_t)Mp * Dp); ar.add(&m.d_mtT_blo, (size_t)Mp * Dp);
  const size_t bytes = ar.total();
  char *blk = nullptr;
  GAPX_HIP(hipGetDevice(&m.device));
  if (hipMalloc((void **)&blk, bytes) != hipSuccess) return fail(GAPX_ERR_NOMEM, "hipMalloc of the model failed");
  m.dev_block = blk;
  m.dev_bytes = bytes;
  size_t off = 0;
  for (auto &it : ar.items) {
    *it.first = blk + off;
    off += it.second;
  }
  auto put = [&](void *dst, const void *src, size_t n) { return hipMemcpy(dst, src, n, hipMemcpyHostToDevice); };
  std::vector<double> zero(1, 0.0);
  std::vector<int> izero(1, 0);
#define PUT(dst, vec, T)                                                                     \
  GAPX_HIP(put(dst, (vec).empty() ? (const void *)zero.data() : (const void *)(vec).data(), \
               std::max<size_t>(1, (vec).size()) * sizeof(T)))
  PUT(rad_a, t.rad_a, double); PUT(rad_rs, t.rad_rs, double);
  {
    // the same tables rounded to FP32 for the FP32 kernels (no conversion per line)
    std::vector<float> a32(t.rad_a.begin(), t.rad_a.end()), rs32(t.rad_rs.begin(), t.rad_rs.end());
    std::vector<float> fzero(1, 0.f);
    GAPX_HIP(put(rad_a32, a32.empty() ? fzero.data() : a32.data(), std::max<size_t>(1, a32.size()) * 4));
    GAPX_HIP(put(rad_rs32, rs32.empty() ? fzero.data() : rs32.data(), std::max<size_t>(1, rs32.size()) * 4));
  }
  if (!t.rad_group.empty()) { GAPX_HIP(put(rad_group, t.rad_group.data(), t.rad_group.size() * 4)); GAPX_HIP(put(rad_feat, t.rad_feat.data(), t.rad_feat.size() * 4)); }
  PUT(ang_eta, t.ang_eta, double); PUT(ang_lam, t.ang_lam, double); PUT(ang_pref, t.ang_pref, double);
  PUT(ang_zr, t.ang_zeta_real, double);
  if (!t.ang_group.empty()) {
    GAPX_HIP(put(ang_zeta, t.ang_zeta.data(), t.ang_zeta.size() * 4));
    GAPX_HIP(put(ang_group, t.ang_group.data(), t.ang_group.size() * 4));
    GAPX_HIP(put(ang_feat, t.ang_feat.data(), t.ang_feat.size() * 4));
  }
  PUT(group_rc, t.group_rc, double);
#undef PUT
  GAPX_HIP(put(m.d_wz, wz.data(), wz.size() * 8));
  GAPX_HIP(put(m.d_hasz, hasz.data(), hasz.size() * 4));
  GAPX_HIP(put(m.d_theta_inv, tinv.data(), Dp * 8));
  GAPX_HIP(put(m.d_mu, mu.data(), Dp * 8));
  GAPX_HIP(put(m.d_mt, mt.data(), mt.size() * 8));
  {
    // the mixed contraction's per-point terms: -|mt|^2 log2(e) / 2, so that
    // exp(-d2 / 2) = exp2(log2(e) S + xq + mnq) with the row's xq formed the
    // same way, and the alpha times the W scale (a power of two) in FP32
    std::vector<double> mnq(Mp);
    std::vector<float> cq(Mp);
    for (int k = 0; k < Mp; ++k) {
      mnq[k] = -0.5 * kLog2e * mn[k];
      cq[k] = (float)cf[k] * m.w_scale;
    }
    GAPX_HIP(put(m.d_mnq, mnq.data(), Mp * 8));
    GAPX_HIP(put(m.d_cq, cq.data(), Mp * 4));
  }
  GAPX_HIP(put(m.d_mn, mn.data(), Mp * 8));
  GAPX_HIP(put(m.d_coeff, cf.data(), Mp * 8));
  GAPX_HIP(put(m.d_mt_hi, hi.data(), hi.size() * 2));
  GAPX_HIP(put(m.d_mt_lo, lo.data(), lo.size() * 2));
  GAPX_HIP(put(m.d_mt_bhi, bhi.data(), bhi.size() * 2));
  GAPX_HIP(put(m.d_mt_blo, blo.data(), blo.size() * 2));
  GAPX_HIP(put(m.d_mtT_hi, thi.data(), thi.size() * 2));
  GAPX_HIP(put(m.d_mtT_lo, tlo.data(), tlo.size() * 2));
  GAPX_HIP(put(m.d_mtT_bhi, tbhi.data(), tbhi.size() * 2));
  GAPX_HIP(put(m.d_mtT_blo, tblo.data(), tblo.size() * 2));
  SFDev &sd = m.sfd;
  sd.n_rad = t.n_rad;
  sd.n_ang = t.n_ang;
  sd.n_groups = t.n_groups;
  sd.nsf = m.nsf;
  sd.rc_max_rad = t.rc_max_rad;
  sd.rc_max_ang = t.rc_max_ang;
  sd.rad_a = rad_a; sd.rad_rs = rad_rs; sd.rad_group = rad_group; sd.rad_feat = rad_feat;
  sd.rad_a32 = rad_a32; sd.rad_rs32 = rad_rs32;
  sd.ang_eta = ang_eta; sd.ang_lam = ang_lam; sd.ang_pref = ang_pref; sd.ang_zeta_real = ang_zr;
  sd.ang_zeta = ang_zeta; sd.ang_group = ang_group; sd.ang_feat = ang_feat; sd.group_rc = group_rc;
  detect_grid(t, sd);
  return GAPX_OK;
}

// ---------------------------------------------------------------------------
// context (device workspace for a batch of structures)
// ---------------------------------------------------------------------------

struct DevMem {
  void *p = nullptr;
  size_t cap = 0;
  unsigned gen = 0;  // bumped on every (re)allocation
};

enum Stage { kNeighbors, kDescFwd, kContraction, kFinish, kDescBwd, kForces, kNumStages };
static const char *kStageNames[kNumStages] = {"neighbor_list", "descriptors_fwd", "gpr_contraction",
                                              "gpr_finish", "descriptors_bwd", "forces_stress"};

// Launch shapes chosen per context and precision (defaults, environment
// overrides, or gapx_ctx_autotune).
struct Tuning {
  int waves_fwd = 4, waves_bwd = 4;  // wavefronts per atom in the descriptor kernels
  int gpr_rows = 64;                 // rows per workgroup of the contraction (64 or 128)
  int gpr_blocks = 512;              // target workgroups of the contraction (sets the splits)
  int halves = 1;                    // the two-half schedule where the batch allows it (1) or one piece (0)
};

struct Ctx {
  Model *m = nullptr;
  hipStream_t stream = nullptr;
  // second stream for the reverse-slot map, which only the final force gather
  // reads: it runs beside the contraction (fork/join events, captured as
  // parallel graph branches)
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // third stream for the second half of a two-half run (see enqueue_run)
  hipStream_t aux2 = nullptr;
  hipEvent_t half_ev = nullptr, half_done_ev = nullptr;
  int device = 0;
  int nstruct = 0, N = 0, Np = 0, nbins = 0;
  bool fused_bins = false;  // every structure small enough for k_bin_struct
  int fused_atoms = 0, fused_bins_max = 0;  // its LDS sizes: largest structure and bin count
  double rcut = 6.0;
  std::vector<int> natoms, offsets;
  std::vector<double> lat;
  std::vector<StructParams> sp;
  std::map<std::string, DevMem> mem;
  size_t bytes = 0;
  int kcap = 64, KP = 64, tcap = 1024;
  int *host_stats = nullptr;      // pinned: the last run's list statistics
  int *host_stats_dev = nullptr;  // its device-side address
  int last_flags = -1;
  bool ran = false;
  int chunk = 0;  // lines per forward pass of the generic kernels (0: 16)
  Tuning tune[2];  // launch shapes for FP64 [0] and mixed [1] runs
  Mlp *mlp = nullptr;  // set while the last run used the surrogate
  bool md_ready = false;  // md_p holds the momenta of the last MD run
  unsigned long long bins_clean = ~0ULL;  // allocation generations of bin_count/bin_ticket last cleared
  int red_parts = 0;                      // reduction parts over all structures (see ReduceBufs)
  unsigned long long red_clean = ~0ULL;   // allocation generations of red_ticket/stats_acc last cleared
  unsigned long long pad_clean = ~0ULL;   // key of the contraction operands whose padding rows were cleared
  // timing
  bool timing = false;
  hipEvent_t ev[2 * kNumStages] = {};
  double stage_ms[kNumStages] = {};
  int stage_launches[kNumStages] = {};
  bool stage_pending[kNumStages] = {};
  // graph replay
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  int graph_flags = -1;
  int graph_kcap = -1, graph_tcap = -1, graph_N = -1;
  // launch parameters of the binning that the graph fixes at capture: the
  // bin count (grid of the rank kernel), which binning kernel runs and the
  // one-launch kernel's LDS sizes (largest structure and bin count)
  int graph_nbins = -1, graph_fused = -1, graph_fused_atoms = -1, graph_fused_bins_max = -1;
  // the reverse-slot search's key width (32-bit keys: N <= 2^17 and every
  // search half-width <= 15), set with the cell parameters
  int key32 = 0, graph_key32 = -1;
  bool graph_matches(int flags) const {
    return gexec && graph_flags == flags && graph_kcap == kcap && graph_tcap == tcap && graph_N == N &&
           graph_nbins == nbins && graph_fused == (fused_bins ? 1 : 0) && graph_fused_atoms == fused_atoms &&
           graph_fused_bins_max == fused_bins_max && graph_key32 == key32;
  }

  template <typename T>
  int get(const char *name, size_t count, T **out) {
    DevMem &d = mem[name];
    size_t need = std::max<size_t>(count, 1) * sizeof(T);
    if (d.cap < need) {
      if (d.p) {
        hipFree(d.p);
        bytes -= d.cap;
      }
      size_t alloc = need + need / 8;
      if (hipMalloc(&d.p, alloc) != hipSuccess) {
        d.p = nullptr;
        d.cap = 0;
        return fail(GAPX_ERR_NOMEM, std::string("hipMalloc failed for ") + name);
      }
      d.cap = alloc;
      ++d.gen;
      bytes += alloc;
      invalidate_graph();
    }
    *out = (T *)d.p;
    return GAPX_OK;
  }
  template <typename T>
  T *ptr(const char *name) {
    auto it = mem.find(name);
    return it == mem.end() ? nullptr : (T *)it->second.p;
  }
  void invalidate_graph() {
    if (gexec) {
      hipGraphExecDestroy(gexec);
      gexec = nullptr;
    }
    graph_flags = -1;
  }
};

static void inv3(const double *A, double *B) {
  double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
  B[0] = (A[4] * A[8] - A[5] * A[7]) / det;
  B[1] = (A[2] * A[7] - A[1] * A[8]) / det;
  B[2] = (A[1] * A[5] - A[2] * A[4]) / det;
  B[3] = (A[5] * A[6] - A[3] * A[8]) / det;
  B[4] = (A[0] * A[8] - A[2] * A[6]) / det;
  B[5] = (A[2] * A[3] - A[0] * A[5]) / det;
  B[6] = (A[3] * A[7] - A[4] * A[6]) / det;
  B[7] = (A[1] * A[6] - A[0] * A[7]) / det;
  B[8] = (A[0] * A[4] - A[1] * A[3]) / det;
}

}  // namespace gapx

// Cell-list parameters of one structure (declared in gapx_internal.h).
int gapx::cell_params(const double *L9, double rcut, int nb_max, StructParams &P) {
  std::memcpy(P.L, L9, 9 * sizeof(double));
  const double *L = P.L;
  // Det() of gap_calc.f90:290-299
  const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                     L[2] * (L[3] * L[7] - L[6] * L[4]);
  if (!(std::fabs(det) > 1e-12)) return fail(GAPX_ERR_ARG, "singular lattice");
  P.vol = std::fabs(det);
  inv3(P.L, P.Linv);
  for (int k = 0; k < 3; ++k) {
    double g = std::sqrt(P.Linv[k] * P.Linv[k] + P.Linv[3 + k] * P.Linv[3 + k] + P.Linv[6 + k] * P.Linv[6 + k]);
    double dk = 1.0 / g;
    // bins at least rcut/2 wide (search half-width S = 2 where the cell
    // allows): on the 4096-atom cell a centre tests ~300 candidates instead
    // of ~510 with rcut-wide bins, and the neighbour stage takes 6 us less
    int nbk = (int)std::floor(2.0 * dk / rcut);
    nbk = std::max(1, std::min(nbk, nb_max));
    P.nb[k] = nbk;
    P.S[k] = (int)std::ceil(rcut * nbk / dk);
    // the list kernel's FP32 index arithmetic wants fewer than 2^20 images
    if (P.S[k] > 48) return fail(GAPX_ERR_ARG, "lattice plane spacing below rcut / 48: too many periodic images");
  }
  // Cholesky factor of the Gram matrix of the bin edges E_k = L_k / nb_k
  double G[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      G[a][b] = (L[3 * a] * L[3 * b] + L[3 * a + 1] * L[3 * b + 1] + L[3 * a + 2] * L[3 * b + 2]) / (P.nb[a] * P.nb[b]);
  const double c00 = std::sqrt(G[0][0]), c10 = G[1][0] / c00, c20 = G[2][0] / c00;
  const double q11 = G[1][1] - c10 * c10;
  const double c11 = q11 > 0 ? std::sqrt(q11) : 0.0, c21 = c11 > 0 ? (G[2][1] - c20 * c10) / c11 : 0.0;
  const double q22 = G[2][2] - c20 * c20 - c21 * c21;
  const double c22 = q22 > 0 ? std::sqrt(q22) : 0.0;
  const double Rv[6] = {c00, c10, c20, c11, c21, c22};
  const bool ok = std::isfinite(c00) && c11 > 0 && c22 > 0;
  for (int k = 0; k < 6; ++k) P.R[k] = ok ? (float)Rv[k] : 0.0f;
  P.bin_base = 0;
  P.atom_base = 0;
  P.natoms = 0;
  P.pad = 0;
  return GAPX_OK;
}

namespace gapx {

static int struct_params(Ctx &c) {
  c.sp.resize(c.nstruct);
  int bin_base = 0;
  for (int s = 0; s < c.nstruct; ++s) {
    StructParams &P = c.sp[s];
    const int rc = cell_params(&c.lat[9 * s], c.rcut, 128, P);
    if (rc) return rc;
    P.bin_base = bin_base;
    bin_base += P.nb[0] * P.nb[1] * P.nb[2];
    P.atom_base = c.offsets[s];
    P.natoms = c.natoms[s];
    P.pad = 0;
  }
  c.nbins = bin_base;
  // one-launch binning when every structure fits its LDS layout
  c.fused_atoms = c.fused_bins_max = 0;
  for (int s = 0; s < c.nstruct; ++s) {
    const StructParams &P = c.sp[s];
    c.fused_atoms = std::max(c.fused_atoms, P.natoms);
    c.fused_bins_max = std::max(c.fused_bins_max, P.nb[0] * P.nb[1] * P.nb[2]);
  }
  c.fused_bins = std::getenv("GAPX_NO_FUSED_BINS") == nullptr && c.fused_atoms <= kFusedAtoms &&
                 fused_bin_lds_bytes(c.fused_atoms, c.fused_bins_max) <= kFusedLdsBytes;
  int s_max = 0;
  for (int s = 0; s < c.nstruct; ++s)
    s_max = std::max({s_max, c.sp[s].S[0], c.sp[s].S[1], c.sp[s].S[2]});
  c.key32 = (c.N <= (1 << 17) && s_max <= 15) ? 1 : 0;
  return GAPX_OK;
}

// Initial slots per atom from the densest structure of the batch.
static int initial_tcap(const Ctx &c) {
  // pairs within the angular cutoff: about (15/32) K (K - 1) / 2 at uniform density
  const double ra = c.m->sf.rc_max_ang;
  double kest = 0;
  for (int s = 0; s < c.nstruct; ++s)
    kest = std::max(kest, 4.0 / 3.0 * 3.141592653589793 * ra * ra * ra * c.natoms[s] / c.sp[s].vol);
  const double t = 15.0 / 32.0 * kest * kest / 2.0;
  return round_up(std::max(256, (int)(1.6 * t) + 128), 64);
}

static int initial_kcap(const Ctx &c) {
  double kest = 0;
  for (int s = 0; s < c.nstruct; ++s)
    kest = std::max(kest, 4.0 / 3.0 * 3.141592653589793 * c.rcut * c.rcut * c.rcut * c.natoms[s] / c.sp[s].vol);
  return round_up(std::max(32, (int)(1.25 * kest) + 24), 8);
}

static int ctx_load_batch(Ctx &c, int nstruct, const int *natoms, const int *species, const double *lat,
                          const double *pos, double rcut, bool fortran) {
  if (nstruct <= 0 || !natoms || !species || !lat || !pos) return fail(GAPX_ERR_ARG, "empty batch");
  if (!(rcut > 0)) return fail(GAPX_ERR_ARG, "rcut must be positive");
  GAPX_HIP(hipSetDevice(c.device));
  c.nstruct = nstruct;
  c.rcut = rcut;
  c.natoms.assign(natoms, natoms + nstruct);
  c.offsets.assign(nstruct + 1, 0);
  for (int s = 0; s < nstruct; ++s) {
    if (natoms[s] <= 0) return fail(GAPX_ERR_ARG, "structure with no atoms");
    c.offsets[s + 1] = c.offsets[s] + natoms[s];
  }
  c.N = c.offsets[nstruct];
  c.Np = round_up(c.N, 128);  // row tiles of the contraction kernels
  // species must be listed in the model header (the reference leaves their
  // weight undefined, gap_calc.f90:73-83)
  std::vector<int> listed(kMaxSpeciesZ, 0);
  for (int k = 0; k < c.m->nspecies; ++k) listed[c.m->z[k]] = 1;
  for (int i = 0; i < c.N; ++i)
    if (species[i] < 0 || species[i] >= kMaxSpeciesZ || !listed[species[i]]) {
      char buf[128];
      std::snprintf(buf, sizeof buf, "atom %d has atomic number %d, which the model does not list", i + 1, species[i]);
      return fail(GAPX_ERR_SPECIES, buf);
    }
  c.lat.resize(9 * nstruct);
  std::vector<double> hpos(3 * (size_t)c.N);
  if (fortran) {  // single structure: LAT(3,3) and POS(NA,3) column-major
    for (int k = 0; k < 3; ++k)
      for (int x = 0; x < 3; ++x) c.lat[3 * k + x] = lat[k + 3 * x];
    for (int i = 0; i < c.N; ++i)
      for (int x = 0; x < 3; ++x) hpos[3 * (size_t)i + x] = pos[i + (size_t)c.N * x];
  } else {
    std::memcpy(c.lat.data(), lat, 9 * sizeof(double) * nstruct);
    std::memcpy(hpos.data(), pos, 3 * sizeof(double) * c.N);
  }
  int rc = struct_params(c);
  if (rc) return rc;
  std::vector<int> astruct(c.N);
  for (int s = 0; s < nstruct; ++s)
    for (int i = c.offsets[s]; i < c.offsets[s + 1]; ++i) astruct[i] = s;
  int *d_species, *d_astruct;
  double *d_pos;
  StructParams *d_sp;
  if ((rc = c.get("species", c.N, &d_species)) || (rc = c.get("astruct", c.N, &d_astruct)) ||
      (rc = c.get("pos", 3 * (size_t)c.N, &d_pos)) || (rc = c.get("sp", nstruct, &d_sp)))
    return rc;
  GAPX_HIP(hipMemcpyAsync(d_species, species, sizeof(int) * c.N, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(hipMemcpyAsync(d_astruct, astruct.data(), sizeof(int) * c.N, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(hipMemcpyAsync(d_pos, hpos.data(), sizeof(double) * 3 * c.N, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(hipMemcpyAsync(d_sp, c.sp.data(), sizeof(StructParams) * nstruct, hipMemcpyHostToDevice, c.stream));
  // the reduction's structure parts (kReduceAtoms atoms each, consecutive per structure)
  std::vector<int2> part_map;
  for (int s = 0; s < nstruct; ++s)
    for (int p = 0; p * kReduceAtoms < std::max(1, c.natoms[s]); ++p) part_map.push_back(make_int2(s, p));
  int2 *d_map;
  if ((rc = c.get("red_map", part_map.size(), &d_map))) return rc;
  GAPX_HIP(hipMemcpyAsync(d_map, part_map.data(), sizeof(int2) * part_map.size(), hipMemcpyHostToDevice, c.stream));
  c.red_parts = (int)part_map.size();
  GAPX_HIP(hipStreamSynchronize(c.stream));
  c.kcap = initial_kcap(c);
  c.tcap = initial_tcap(c);
  c.ran = false;
  c.md_ready = false;
  c.invalidate_graph();
  return GAPX_OK;
}

// A load that fails part-way leaves the context empty (no structures), so no
// later call can run kernels sized for the new batch on the old buffers.
static int ctx_load(Ctx &c, int nstruct, const int *natoms, const int *species, const double *lat,
                    const double *pos, double rcut, bool fortran) {
  const int rc = ctx_load_batch(c, nstruct, natoms, species, lat, pos, rcut, fortran);
  if (rc) {
    c.N = 0;
    c.nstruct = 0;
    c.ran = false;
    c.md_ready = false;
    c.invalidate_graph();
  }
  return rc;
}

struct Bufs {
  NeighborBufs nb{};
  TripletBufs tb{};
  DescOut out{};
  GprArgs g{};
  double *slot_f = nullptr, *atom_vir = nullptr, *force = nullptr, *ene = nullptr, *stress = nullptr;
  int *bin_cursor = nullptr;
};

static void stage_begin(Ctx &c, int st) {
  if (c.timing) hipEventRecord(c.ev[2 * st], c.stream);
}
static void stage_end(Ctx &c, int st) {
  if (c.timing) {
    hipEventRecord(c.ev[2 * st + 1], c.stream);
    c.stage_pending[st] = true;
  }
}

static void collect_timing(Ctx &c) {
  if (!c.timing) return;
  for (int st = 0; st < kNumStages; ++st) {
    if (!c.stage_pending[st]) continue;
    float ms = 0;
    if (hipEventElapsedTime(&ms, c.ev[2 * st], c.ev[2 * st + 1]) == hipSuccess) {
      c.stage_ms[st] += ms;
      c.stage_launches[st] += 1;
    }
    c.stage_pending[st] = false;
  }
}

// Largest neighbour count the run accepts (LDS layout, or the reference's cap).
static int neighbor_cap(Ctx &c, int flags) {
  const int lds = max_neighbors_for_lds(!(flags & GAPX_MIXED), c.m->sfd, c.m->Dp);
  return (flags & GAPX_REF_NEIGHBOR_CAP) ? std::min(lds, kRefNeighborCap) : lds;
}

// Neighbour list: bins, then one wavefront per atom writing up to kcap slots.
// Nothing here waits for the device; counts above kcap are detected when the
// results are fetched (the run is then repeated with a larger kcap).
static int run_neighbors(Ctx &c, int flags, Bufs &b) {
  NeighborBufs &nb = b.nb;
  int rc;
  int *stats;
  // never size the descriptor kernels' LDS past the cap: a count above it is
  // reported by ctx_settle as GAPX_ERR_NEIGHBORS
  c.kcap = std::min(c.kcap, round_up(neighbor_cap(c, flags), 8));
  const size_t slots = (size_t)c.N * c.kcap;
  if ((rc = c.get("bin_count", c.nbins, &nb.bin_count)) || (rc = c.get("bin_start", c.nbins + 1, &nb.bin_start)) ||
      (rc = c.get("bin_cursor", std::max(c.nbins, c.N), &nb.bin_cursor)) ||
      (rc = c.get("bin_atoms", c.N, &nb.bin_atoms)) || (rc = c.get("atom_bin", c.N, &nb.atom_bin)) ||
      (rc = c.get("wpos", 3 * (size_t)c.N, &nb.wpos)) || (rc = c.get("shift", 3 * (size_t)c.N, &nb.shift)) ||
      (rc = c.get("nb_count", c.N, &nb.nb_count)) || (rc = c.get("stats", 8, &stats)) ||
      (rc = c.get("nb_raw", c.N, &nb.nb_raw)) || (rc = c.get("spos", 4 * (size_t)c.N, &nb.spos)) ||
      (rc = c.get("sinfo", 4 * (size_t)c.N, &nb.sinfo)) || (rc = c.get("nb_j", slots, &nb.nb_j)) ||
      (rc = c.get("nb_key", slots + kRevBatch, &nb.nb_key)) ||  // padding for the batched reads
      (rc = c.get("nb_d", 4 * slots, &nb.nb_d)) || (rc = c.get("nb_rev", slots, &nb.nb_rev)))
    return rc;
  nb.total = stats;
  nb.overflow = stats + 2;
  nb.kcap = c.kcap;
  nb.key32 = c.key32;
  // cutoff tables in global memory for lists too long for the descriptor
  // kernels' LDS layouts
  nb.fc_scratch = nullptr;
  const size_t tbytes = desc_table_scratch_bytes(!(flags & GAPX_MIXED), c.m->sfd, c.m->Dp, c.kcap, c.N);
  if (tbytes) {
    unsigned char *t;
    if ((rc = c.get("fc_scratch", tbytes, &t))) return rc;
    nb.fc_scratch = t;
  }
  const int *d_astruct = c.ptr<int>("astruct");
  const double *d_pos = c.ptr<double>("pos");
  const StructParams *d_sp = c.ptr<StructParams>("sp");
  if (c.fused_bins) {
    GAPX_HIP(launch_bin_structures(c.stream, c.nstruct, d_pos, d_sp, nb, c.nbins, c.fused_atoms, c.fused_bins_max,
                                   stats));
  } else {
    // k_wrap_scan keeps bin_count and its ticket at zero between runs; clear
    // them (whole capacity) once after they are (re)allocated
    int *ticket;
    if ((rc = c.get("bin_ticket", 1, &ticket))) return rc;
    const DevMem &dc = c.mem["bin_count"], &dt = c.mem["bin_ticket"];
    const unsigned long long g = ((unsigned long long)dc.gen << 32) | dt.gen;
    if (g != c.bins_clean) {
      GAPX_HIP(launch_zero_ints(c.stream, nb.bin_count, (int)(dc.cap / sizeof(int)), ticket, 1, nullptr, 0));
      c.bins_clean = g;
    }
    GAPX_HIP(launch_wrap_scan(c.stream, c.N, c.nbins, d_pos, d_astruct, d_sp, nb, ticket, stats));
    GAPX_HIP(launch_scatter_bins(c.stream, c.N, c.nbins, nb));
  }
  // triplet lists for the angular lines (forward and backward share them),
  // built by the same wavefronts as the neighbour lists
  TripletBufs &tb = b.tb;
  tb.tcap = c.tcap;
  if ((rc = c.get("trip_list", (size_t)c.N * c.tcap, &tb.list)) || (rc = c.get("trip_count", c.N, &tb.count)) ||
      (rc = c.get("trip_raw", c.N, &tb.raw)))
    return rc;
  const bool ang = c.m->sf.n_ang > 0;
  GAPX_HIP(launch_build_neighbors(c.stream, c.N, d_astruct, d_sp, c.rcut, nb, c.m->sf.rc_max_ang, ang ? &tb : nullptr));
  if (!ang) {
    GAPX_HIP(launch_fill_int(c.stream, tb.count, c.N, 0));
    tb.raw = nullptr;  // no triplet statistics
  }
  c.KP = c.kcap;
  return GAPX_OK;
}

// The reverse-slot map on the second stream, forked after the descriptor
// forward was enqueued: it then runs beside the contraction, which leaves the
// vector units mostly idle (running it beside the descriptor kernels slowed
// them by as much as it saved).  enqueue_forces joins it before the gather.
// With per-stage timing on, every stage runs alone on the work stream and
// the map is built inside the neighbour stage instead (reverse_in_timing), so
// that each stage's time is its own kernels' time.
static int fork_reverse(Ctx &c, Bufs &b) {
  if (c.timing) return GAPX_OK;
  GAPX_HIP(hipEventRecord(c.fork_ev, c.stream));
  GAPX_HIP(hipStreamWaitEvent(c.aux, c.fork_ev, 0));
  GAPX_HIP(launch_reverse_slots(c.aux, c.N, b.nb));
  GAPX_HIP(hipEventRecord(c.join_ev, c.aux));
  return GAPX_OK;
}
static int reverse_in_timing(Ctx &c, Bufs &b, bool grad) {
  if (c.timing && grad) GAPX_HIP(launch_reverse_slots(c.stream, c.N, b.nb));
  return GAPX_OK;
}

// f64_kernel: the contraction runs k_gpr_f64 (64-row tiles); tune: the launch
// shapes of the run's precision (0 FP64, 1 mixed).
static int gpr_layout(Ctx &c, GprArgs &g, bool f64_kernel, int tune) {
  const Model &m = *c.m;
  const Tuning &t = c.tune[tune];
  g.N = c.N;
  g.Np = c.Np;
  g.Dp = m.Dp;
  g.Mp = m.Mp;
  // the mixed kernel above D = 128 runs 128-row tiles only (eight waves
  // sharing one workgroup's two chunk buffers)
  g.rows = f64_kernel ? 64 : m.Dp > 128 ? 128 : t.gpr_rows == 128 ? 128 : 64;
  // Split the sparse points over grid.y until there are about gpr_blocks
  // workgroups (resident waves; more splits only add partial-sum traffic).
  const int tiles = c.Np / g.rows;
  const int chunk = gpr_points_per_chunk(m.Dp);
  int splits = std::max(1, (t.gpr_blocks + tiles - 1) / tiles);
  splits = std::min(splits, m.Mp / chunk);
  const int per = round_up((m.Mp + splits - 1) / splits, chunk);
  g.splits = (m.Mp + per - 1) / per;
  g.pts_per_split = per;
  return GAPX_OK;
}

// From dE/dx to forces and stress: descriptor backward pass, force gather,
// per-structure energy and virial, list statistics.
static int forces_buffers(Ctx &c, Bufs &b, bool grad) {
  int rc;
  if ((rc = c.get("force", 3 * (size_t)c.N, &b.force)) || (rc = c.get("ene", c.nstruct, &b.ene)) ||
      (rc = c.get("stress", 6 * (size_t)c.nstruct, &b.stress)) || (rc = c.get("atom_vir", 6 * (size_t)c.N, &b.atom_vir)))
    return rc;
  if (grad && (rc = c.get("slot_f", 3 * (size_t)c.N * c.kcap, &b.slot_f))) return rc;
  return GAPX_OK;
}

// Descriptor backward for atoms [atom0, atom0 + count) on stream s.
static int enqueue_bwd(Ctx &c, Bufs &b, bool fp64, const GradSrc &gs, int waves_bwd, hipStream_t s, int atom0,
                       int count) {
  Model &m = *c.m;
  GAPX_HIP(launch_descriptors_bwd(s, fp64, count, c.ptr<int>("species"), m.d_wz, b.nb, b.tb, m.sfd, m.Dp, gs,
                                  b.slot_f, b.atom_vir, waves_bwd, c.KP, atom0));
  return GAPX_OK;
}

static int enqueue_forces(Ctx &c, Bufs &b, bool fp64, bool grad, const double *eatom, const GradSrc &gs,
                          int waves_bwd, bool bwd_done = false) {
  NeighborBufs &nb = b.nb;
  const StructParams *d_sp = c.ptr<StructParams>("sp");
  int rc;
  if ((rc = forces_buffers(c, b, grad))) return rc;
  if (grad && !bwd_done) {
    stage_begin(c, kDescBwd);
    if ((rc = enqueue_bwd(c, b, fp64, gs, waves_bwd, c.stream, 0, c.N))) return rc;
    stage_end(c, kDescBwd);
  }
  // forces, energies, stress, and the list statistics straight into pinned
  // host memory (read by ctx_settle), in one launch; the force gather needs
  // the reverse-slot map from the second stream
  ReduceBufs rb;
  double *part_sum;
  int *ticket, *stats_acc;
  if ((rc = c.get("red_sum", 8 * (size_t)c.red_parts, &part_sum)) || (rc = c.get("red_ticket", c.nstruct, &ticket)) ||
      (rc = c.get("stats_acc", 17, &stats_acc)))
    return rc;
  {
    // the tickets and accumulators stay zero between runs; clear them once
    // after they are (re)allocated
    const unsigned long long g = ((unsigned long long)c.mem["red_ticket"].gen << 32) | c.mem["stats_acc"].gen;
    if (g != c.red_clean) {
      GAPX_HIP(launch_zero_ints(c.stream, ticket, (int)(c.mem["red_ticket"].cap / sizeof(int)), stats_acc, 17,
                                nullptr, 0));
      c.red_clean = g;
    }
  }
  rb.part_map = c.ptr<int2>("red_map");
  rb.nparts = c.red_parts;
  rb.part_sum = part_sum;
  rb.ticket = ticket;
  rb.stats_acc = stats_acc;
  rb.sticky = stats_acc + 8;
  rb.overflow_now = stats_acc + 16;
  if (grad && !c.timing) GAPX_HIP(hipStreamWaitEvent(c.stream, c.join_ev, 0));
  stage_begin(c, kForces);
  GAPX_HIP(launch_forces_reduce(c.stream, c.N, grad ? 1 : 0, nb, b.slot_f, !fp64, b.force, c.nstruct, d_sp, eatom,
                                b.atom_vir, b.ene, b.stress, b.tb, rb, c.host_stats_dev));
  stage_end(c, kForces);
  return GAPX_OK;
}

// Enqueue one evaluation on the context's stream; nothing here waits for the
// device, so the sequence can be captured into a graph.
static int enqueue_run(Ctx &c, int flags) {
  Model &m = *c.m;
  const bool fp64 = !(flags & GAPX_MIXED);  // descriptor precision
  const bool c64 = fp64 || m.mixed_f64;      // contraction precision
  const bool grad = (flags & GAPX_GRAD) != 0;
  const int Dp = m.Dp;
  Bufs b;
  int rc;
  stage_begin(c, kNeighbors);
  if ((rc = run_neighbors(c, flags, b)) || (rc = reverse_in_timing(c, b, grad))) return rc;
  NeighborBufs &nb = b.nb;
  const int *d_species = c.ptr<int>("species");
  stage_end(c, kNeighbors);

  // descriptors
  DescOut &o = b.out;
  o.N = c.N;
  o.Dp = Dp;
  const size_t ND = (size_t)c.Np * Dp;
  if ((rc = c.get("xn", c.Np, &o.xn))) return rc;
  if (fp64) {
    if ((rc = c.get("x64", ND, &o.x64)) || (rc = c.get("xt64", ND, &o.xt64))) return rc;
  } else {
    if ((rc = c.get("x32", ND, &o.x32)) || (rc = c.get("xt32", ND, &o.xt32))) return rc;
    // the contraction's operands: FP64 xt (from the FP32 features) or the 16-bit split
    if (c64) {
      if ((rc = c.get("xt64", ND, &o.xt64))) return rc;
    } else if ((rc = c.get("xt_hi", ND, &o.xt_hi)) || (rc = c.get("xt_lo", ND, &o.xt_lo))) {
      return rc;
    }
  }
  // padding rows [N, Np) of the contraction operands are zero: cleared once
  // per allocation and size (the descriptor kernels write rows below N only,
  // and the other writers of these buffers reset the key)
  const unsigned long long pad_key =
      ((unsigned long long)c.mem["xn"].gen << 40) ^
      ((unsigned long long)(c64 ? c.mem["xt64"].gen : c.mem["xt_hi"].gen + 7 * c.mem["xt_lo"].gen) << 20) ^
      ((unsigned long long)c.N << 1) ^ (c64 ? 1ULL : 0ULL);
  if (c.Np > c.N && pad_key != c.pad_clean) {
    if (c64) GAPX_HIP(launch_fill_double(c.stream, o.xt64 + (size_t)c.N * Dp, (size_t)(c.Np - c.N) * Dp, 0.0));
    GAPX_HIP(launch_fill_double(c.stream, o.xn + c.N, c.Np - c.N, 0.0));
    if (!c64) {
      GAPX_HIP(launch_fill_int(c.stream, (int *)(o.xt_hi + (size_t)c.N * Dp), (c.Np - c.N) * Dp / 2, 0));
      GAPX_HIP(launch_fill_int(c.stream, (int *)(o.xt_lo + (size_t)c.N * Dp), (c.Np - c.N) * Dp / 2, 0));
    }
  }
  c.pad_clean = pad_key;
  const int split_mode = c64 ? 0 : ((flags & GAPX_BF16) ? 2 : 1);
  const int chunk = c.chunk > 0 ? c.chunk : 16;
  const Tuning &tu = c.tune[fp64 ? 0 : 1];

  // contraction arguments
  GprArgs &g = b.g;
  gpr_layout(c, g, c64, fp64 ? 0 : 1);
  g.xt64 = o.xt64;
  g.xn = o.xn;
  g.xt_hi = o.xt_hi;
  g.xt_lo = o.xt_lo;
  g.xt32 = o.xt32;
  g.mt = m.d_mt;
  g.mn = m.d_mn;
  g.coeff = m.d_coeff;
  g.mnq = m.d_mnq;
  g.cq = m.d_cq;
  const bool bf = (flags & GAPX_BF16) != 0;
  g.mt_hi = bf ? m.d_mt_bhi : m.d_mt_hi;
  g.mt_lo = bf ? m.d_mt_blo : m.d_mt_lo;
  g.mtT_hi = bf ? m.d_mtT_bhi : m.d_mtT_hi;
  g.mtT_lo = bf ? m.d_mtT_blo : m.d_mtT_lo;
  g.theta_inv = m.d_theta_inv;
  g.grad = grad ? 1 : 0;
  g.bf16 = bf ? 1 : 0;
  g.wscale = m.w_scale;
  if ((rc = c.get("Epart", (size_t)g.splits * c.Np, &g.Epart)) || (rc = c.get("eatom", c.Np, &g.eatom))) return rc;
  if (grad) {
    if (c64) {
      if ((rc = c.get("Gpart", (size_t)g.splits * ND, &g.Gpart))) return rc;
    } else {
      if ((rc = c.get("Gpart32", (size_t)g.splits * ND, &g.Gpart32))) return rc;
    }
    if (fp64) {
      if ((rc = c.get("g64", ND, &g.g64))) return rc;
    } else {
      if ((rc = c.get("g32", ND, &g.g32))) return rc;
    }
  }
  GradSrc gs;
  if (grad) {
    // the finish (sum of split partials, dE/dx) runs inside the backward
    // kernel, with the xt the contraction used (FP64 whenever it ran in FP64)
    gs.splits = g.splits;
    gs.Np = g.Np;
    gs.Epart = g.Epart;
    gs.Gpart = c64 ? g.Gpart : nullptr;
    gs.Gpart32 = c64 ? nullptr : g.Gpart32;
    gs.xt64 = o.xt64;
    gs.xt32 = o.xt32;
    gs.theta_inv = m.d_theta_inv;
    gs.eatom = g.eatom;
    gs.g64_out = g.g64;
    gs.g32_out = g.g32;
  }
  auto contraction = [&](hipStream_t st, const GprArgs &ga) -> int {
    if (c64)
      GAPX_HIP(launch_gpr_fp64(st, ga));
    else
      GAPX_HIP(launch_gpr_mixed(st, ga));
    return GAPX_OK;
  };

  // Two-half schedule (forces runs, per-stage timing off, both halves at
  // least kHalfMinAtoms): the atoms split at a row-tile boundary S, half A on
  // the work stream, half B on the third stream once A's forward is done, so
  // that B's forward runs beside A's contraction and A's backward beside B's
  // contraction (vector units and matrix cores busy together); B joins before
  // the force gather.  Smaller batches run in one piece: their descriptor
  // kernels are latency-bound at low occupancy, and a half-size launch takes
  // about as long as a whole one (4096 atoms: 12 % slower split).  Whether it
  // pays on a given batch is the autotuner's call (Tuning::halves): the
  // contraction's workgroups fill their CUs' registers, so the overlap can
  // also just share the CUs out.
  constexpr int kHalfMinAtoms = 8192;
  const int S = (c.N / 2) / g.rows * g.rows;
  if (grad && !c.timing && tu.halves && S >= kHalfMinAtoms && c.N - S >= kHalfMinAtoms) {
    GAPX_HIP(launch_descriptors_fwd(c.stream, fp64, S, d_species, m.d_wz, nb, b.tb, m.sfd, m.d_theta_inv, m.d_mu,
                                    split_mode, o, chunk, tu.waves_fwd, c.KP, 0));
    if ((rc = fork_reverse(c, b))) return rc;
    GAPX_HIP(hipEventRecord(c.half_ev, c.stream));
    GAPX_HIP(hipStreamWaitEvent(c.aux2, c.half_ev, 0));
    GprArgs ga = g, gb = g;
    ga.tile0 = 0;
    ga.ntiles = S / g.rows;
    gb.tile0 = S / g.rows;
    gb.ntiles = c.Np / g.rows - S / g.rows;
    if ((rc = contraction(c.stream, ga))) return rc;
    GAPX_HIP(launch_descriptors_fwd(c.aux2, fp64, c.N - S, d_species, m.d_wz, nb, b.tb, m.sfd, m.d_theta_inv,
                                    m.d_mu, split_mode, o, chunk, tu.waves_fwd, c.KP, S));
    if ((rc = contraction(c.aux2, gb))) return rc;
    if ((rc = forces_buffers(c, b, grad)) || (rc = enqueue_bwd(c, b, fp64, gs, tu.waves_bwd, c.stream, 0, S)) ||
        (rc = enqueue_bwd(c, b, fp64, gs, tu.waves_bwd, c.aux2, S, c.N - S)))
      return rc;
    GAPX_HIP(hipEventRecord(c.half_done_ev, c.aux2));
    GAPX_HIP(hipStreamWaitEvent(c.stream, c.half_done_ev, 0));
    return enqueue_forces(c, b, fp64, grad, g.eatom, gs, tu.waves_bwd, true);
  }

  stage_begin(c, kDescFwd);
  GAPX_HIP(launch_descriptors_fwd(c.stream, fp64, c.N, d_species, m.d_wz, nb, b.tb, m.sfd, m.d_theta_inv, m.d_mu,
                                  split_mode, o, chunk, tu.waves_fwd, c.KP));
  stage_end(c, kDescFwd);
  if (grad && (rc = fork_reverse(c, b))) return rc;
  stage_begin(c, kContraction);
  if ((rc = contraction(c.stream, g))) return rc;
  stage_end(c, kContraction);
  if (!grad) {
    stage_begin(c, kFinish);
    GAPX_HIP(launch_gpr_finish(c.stream, g, c64));
    stage_end(c, kFinish);
  }
  return enqueue_forces(c, b, fp64, grad, g.eatom, gs, tu.waves_bwd);
}

// The surrogate in place of the contraction: same neighbour lists and
// descriptors, MLP energies and dE/dx, same backward pass.  The MLP time is
// reported as the contraction stage.
static int enqueue_surrogate(Ctx &c, Mlp &mlp, int flags) {
  Model &m = *c.m;
  if (m.D != 128 || m.Dp != 128) return fail(GAPX_ERR_MODEL, "the surrogate takes models with 128 descriptor columns");
  const bool fp64 = !(flags & GAPX_MIXED);
  const bool grad = (flags & GAPX_GRAD) != 0;
  const int Dp = m.Dp;
  Bufs b;
  int rc;
  stage_begin(c, kNeighbors);
  if ((rc = run_neighbors(c, flags, b)) || (rc = reverse_in_timing(c, b, grad))) return rc;
  stage_end(c, kNeighbors);
  DescOut &o = b.out;
  o.N = c.N;
  o.Dp = Dp;
  const size_t ND = (size_t)c.Np * Dp;
  if ((rc = c.get("xn", c.Np, &o.xn))) return rc;
  c.pad_clean = ~0ULL;  // this path writes the operand buffers its own way
  if (fp64) {
    if ((rc = c.get("x64", ND, &o.x64)) || (rc = c.get("xt64", ND, &o.xt64))) return rc;
  } else {
    if ((rc = c.get("x32", ND, &o.x32)) || (rc = c.get("xt32", ND, &o.xt32))) return rc;
  }
  const Tuning &tu = c.tune[fp64 ? 0 : 1];
  stage_begin(c, kDescFwd);
  GAPX_HIP(launch_descriptors_fwd(c.stream, fp64, c.N, c.ptr<int>("species"), m.d_wz, b.nb, b.tb, m.sfd,
                                  m.d_theta_inv, m.d_mu, 0, o, c.chunk > 0 ? c.chunk : 16, tu.waves_fwd, c.KP));
  stage_end(c, kDescFwd);
  if (grad && (rc = fork_reverse(c, b))) return rc;
  float *xin = o.xt32;
  if (fp64) {
    if ((rc = c.get("xtf", ND, &xin))) return rc;
    GAPX_HIP(launch_f64_to_f32(c.stream, (size_t)c.N * Dp, o.xt64, xin));
  }
  if (c.Np > c.N) GAPX_HIP(launch_fill_int(c.stream, (int *)(xin + (size_t)c.N * Dp), (c.Np - c.N) * Dp, 0));
  double *eatom = nullptr, *g64 = nullptr;
  float *g32 = nullptr;
  if ((rc = c.get("eatom", c.Np, &eatom))) return rc;
  if (grad) {
    if (fp64) {
      if ((rc = c.get("g64", ND, &g64))) return rc;
    } else {
      if ((rc = c.get("g32", ND, &g32))) return rc;
    }
  }
  stage_begin(c, kContraction);
  const bool exact = (flags & GAPX_MLP_EXACT) != 0;
  if ((rc = mlp_eval_into_gap(mlp, c.stream, exact, c.N, c.Np, xin, grad, Dp, m.d_theta_inv, eatom, g32, g64)))
    return rc;
  stage_end(c, kContraction);
  GradSrc gs;
  gs.g64 = g64;
  gs.g32 = g32;
  return enqueue_forces(c, b, fp64, grad, eatom, gs, tu.waves_bwd);
}

// One evaluation.  With graphs enabled (and timing off) the enqueued sequence
// is captured once per (flags, atom count, list capacities) and replayed: an
// eager run first brings every workspace buffer to its final size, so the
// capture itself allocates nothing.
// The flags a run of this model executes with: a model whose FP32 features
// cannot hold the mixed path's bar (upload_model) runs GAPX_MIXED as FP64.
static int run_flags(const Model &m, int flags) {
  return m.mixed_desc_f64 ? (flags & ~(GAPX_MIXED | GAPX_BF16)) : flags;
}

static int ctx_run(Ctx &c, int flags, Mlp *mlp = nullptr) {
  if (c.N <= 0) return fail(GAPX_ERR_ARG, "no structures loaded");
  GAPX_HIP(hipSetDevice(c.device));
  int rc;
  c.mlp = mlp;
  if (!mlp) flags = run_flags(*c.m, flags);
  if (c.mlp) {  // surrogate runs are eager
    if ((rc = enqueue_surrogate(c, *c.mlp, flags))) return rc;
  } else if (c.use_graph && !c.timing) {
    if (!c.graph_matches(flags)) {
      c.invalidate_graph();
      if ((rc = enqueue_run(c, flags))) return rc;
      hipGraph_t graph = nullptr;
      GAPX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_run(c, flags);
      const hipError_t e = hipStreamEndCapture(c.stream, &graph);
      if (rc || e != hipSuccess || !graph) {
        if (graph) hipGraphDestroy(graph);
        c.invalidate_graph();
        return rc ? rc : fail(GAPX_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
      }
      const hipError_t ei = hipGraphInstantiate(&c.gexec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        c.gexec = nullptr;
        return fail(GAPX_ERR_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
      }
      c.graph_flags = flags;
      c.graph_kcap = c.kcap;
      c.graph_tcap = c.tcap;
      c.graph_N = c.N;
      c.graph_nbins = c.nbins;
      c.graph_fused = c.fused_bins ? 1 : 0;
      c.graph_fused_atoms = c.fused_atoms;
      c.graph_fused_bins_max = c.fused_bins_max;
      c.graph_key32 = c.key32;
    }
    GAPX_HIP(hipGraphLaunch(c.gexec, c.stream));
  } else {
    if ((rc = enqueue_run(c, flags))) return rc;
  }
  c.last_flags = flags;
  c.ran = true;
  if (c.timing) {  // one event pair per stage: read them before the next run
    GAPX_HIP(hipStreamSynchronize(c.stream));
    collect_timing(c);
  }
  return GAPX_OK;
}

// Wait for the last run and check its neighbour counts: above kcap the run
// is repeated with room for the largest count; above the cap it is an error.
// At most two reruns are needed: the raw neighbour counts do not depend on
// kcap, but the triplet counts of a run whose lists were cut at kcap do, so
// the rerun with the full lists can still overflow tcap once.
// Verdict on list statistics h (the layout of launch_forces_reduce): GAPX_OK
// with *grow false when every list fitted; GAPX_OK with *grow true after
// raising kcap and/or tcap to the largest counts plus the given headroom
// (the graph is invalidated); an error when a count is above the cap.
static int capacity_verdict(Ctx &c, const int *h, int kroom, int troom, bool *grow) {
  *grow = false;
  const int cap = neighbor_cap(c, c.last_flags);
  if (h[1] > cap) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "an atom has %d neighbours within rcut, above the limit of %d", h[1], cap);
    return fail(GAPX_ERR_NEIGHBORS, buf);
  }
  if (h[3]) return fail(GAPX_ERR_NEIGHBORS, "descriptor kernels saw more neighbours than their LDS layout holds");
  if (h[2] == 0 && h[6] == 0) return GAPX_OK;
  if (h[2]) c.kcap = round_up(std::min(h[1] + kroom, std::max(cap, h[1] + 8)), 8);
  if (h[6]) c.tcap = round_up(h[5] + troom, 64);
  c.invalidate_graph();
  *grow = true;
  return GAPX_OK;
}

static int ctx_settle(Ctx &c) {
  GAPX_HIP(hipStreamSynchronize(c.stream));
  collect_timing(c);
  for (int attempt = 0; attempt < 4; ++attempt) {
    bool grow;
    int rc = capacity_verdict(c, c.host_stats, 8, 64, &grow);
    if (rc || !grow) return rc;
    if ((rc = ctx_run(c, c.last_flags, c.mlp))) return rc;
    GAPX_HIP(hipStreamSynchronize(c.stream));
    collect_timing(c);
  }
  return fail(GAPX_ERR_NEIGHBORS, "neighbour capacity did not converge");
}

// ---------------------------------------------------------------------------
// autotuner: coordinate descent over the launch shapes, each candidate timed
// as whole evaluations of the loaded batch, back to back (HIP events).  Results are
// cached per process for the shape class (padded D, M, table form, precision,
// atom and structure counts rounded up to powers of two).
// ---------------------------------------------------------------------------
static std::mutex g_tune_mu;
static std::map<std::string, Tuning> g_tune_cache;

static int pow2_bucket(int n) {
  int b = 1;
  while (b < n && b < (1 << 30)) b <<= 1;
  return b;
}

static std::string tune_key(const Ctx &c, int flags) {
  char buf[200];
  std::snprintf(buf, sizeof buf, "D%d M%d G%d grid%d mixed%d bf%d N%d S%d", c.m->Dp, c.m->Mp, c.m->sf.n_groups,
                c.m->sfd.grid, (flags & GAPX_MIXED) ? 1 : 0, (flags & GAPX_BF16) ? 1 : 0, pow2_bucket(c.N),
                pow2_bucket(c.nstruct));
  return buf;
}

static void reset_timing(Ctx &c) {
  for (int s = 0; s < kNumStages; ++s) {
    c.stage_ms[s] = 0;
    c.stage_launches[s] = 0;
    c.stage_pending[s] = false;
  }
}

static int ctx_autotune(Ctx &c, int flags, bool force) {
  if (c.N <= 0) return fail(GAPX_ERR_ARG, "no structures loaded");
  flags = run_flags(*c.m, flags);
  const int p = (flags & GAPX_MIXED) ? 1 : 0;
  const std::string key = tune_key(c, flags);
  if (!force) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      c.tune[p] = it->second;
      c.invalidate_graph();
      return GAPX_OK;
    }
  }
  const bool was_timing = c.timing, was_graph = c.use_graph;
  c.use_graph = false;
  c.timing = false;
  c.invalidate_graph();
  int rc = ctx_run(c, flags);
  if (!rc) rc = ctx_settle(c);  // capacities first
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(GAPX_ERR_HIP, "autotune: could not create events");
  // whole evaluations, back to back on the stream, replayed from a graph
  // captured for each candidate (so host launch costs, which do not depend
  // on the launch shapes, stay out of the comparison): the mean of `reps`
  // runs between two events, best of `trials` trials
  c.use_graph = true;
  auto measure = [&](double &out, int reps, int trials) -> int {
    c.invalidate_graph();
    int r = ctx_run(c, flags);  // capture, warm-up (buffers, caches)
    if (r) return r;
    out = 1e300;
    for (int trial = 0; trial < trials; ++trial) {
      GAPX_HIP(hipEventRecord(e0, c.stream));
      for (int k = 0; k < reps; ++k)
        if ((r = ctx_run(c, flags))) return r;
      GAPX_HIP(hipEventRecord(e1, c.stream));
      GAPX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      GAPX_HIP(hipEventElapsedTime(&ms, e0, e1));
      out = std::min(out, (double)ms / reps);
    }
    return GAPX_OK;
  };
  // Run-to-run noise is about 1 %: a candidate replaces the current choice
  // only when it is faster by more than kMargin, and each search starts
  // from the current choice, measured first.
  constexpr double kMargin = 0.01;
  Tuning &t = c.tune[p];
  const Tuning start = t;
  struct Knob {
    int *field;
    std::vector<int> values;
  };
  std::vector<Knob> knobs;
  knobs.push_back({&t.waves_fwd, {1, 2, 4}});
  if (flags & GAPX_GRAD) knobs.push_back({&t.waves_bwd, {1, 2, 4}});
  // the descriptor kernels' waves one knob at a time
  auto descend = [&]() {
    for (Knob &k : knobs) {
      if (rc) break;
      const int cur = *k.field;
      double best_ms = 0;
      if ((rc = measure(best_ms, 10, 3))) break;
      int best = cur;
      for (int v : k.values) {
        if (v == cur) continue;
        *k.field = v;
        double ms = 0;
        if ((rc = measure(ms, 10, 3))) break;
        if (ms < best_ms * (1.0 - kMargin)) {
          best_ms = ms;
          best = v;
        }
      }
      *k.field = best;
    }
  };
  if (!rc) descend();
  // then the contraction's row tile and workgroup target together (they
  // interact: the target sets the split of the sparse points, the tile the
  // number of row tiles)
  if (!rc) {
    const std::vector<int> rows = (p == 1 && c.m->Dp <= 128) ? std::vector<int>{64, 128} : std::vector<int>{t.gpr_rows};
    const std::vector<int> blocks = {256, 512, 768, 1024, 1536, 2048, 3072};
    const int cur_r = t.gpr_rows, cur_b = t.gpr_blocks;
    double best_ms = 0;
    rc = measure(best_ms, 10, 3);
    int best_r = cur_r, best_b = cur_b;
    for (int r : rows) {
      for (int b : blocks) {
        if (rc || (r == cur_r && b == cur_b)) continue;
        t.gpr_rows = r;
        t.gpr_blocks = b;
        double ms = 0;
        if ((rc = measure(ms, 10, 3))) break;
        if (ms < best_ms * (1.0 - kMargin)) {
          best_ms = ms;
          best_r = r;
          best_b = b;
        }
      }
    }
    t.gpr_rows = best_r;
    t.gpr_blocks = best_b;
  }
  if (!rc) descend();  // the waves again, now beside the chosen contraction shape
  if (!rc && (flags & GAPX_GRAD)) {  // the two-half schedule, where the batch is large enough for it
    const int cur = t.halves;
    double cur_ms = 0, other_ms = 0;
    if (!(rc = measure(cur_ms, 10, 3))) {
      t.halves = 1 - cur;
      if (!(rc = measure(other_ms, 10, 3))) t.halves = (other_ms < cur_ms * (1.0 - kMargin)) ? 1 - cur : cur;
    }
  }
  // last, the result against the shapes the search started from, measured
  // longer: the search keeps its result only if it is faster
  if (!rc) {
    const Tuning found = t;
    double found_ms = 0, start_ms = 0;
    if (!(rc = measure(found_ms, 20, 5))) {
      t = start;
      if (!(rc = measure(start_ms, 20, 5))) t = (found_ms < start_ms) ? found : start;
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  c.timing = was_timing;
  reset_timing(c);
  c.use_graph = was_graph;
  c.invalidate_graph();
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = t;
  return GAPX_OK;
}

// Relax every structure of the batch at fixed cell (FIRE, relax.hip): each
// step is one E+F evaluation (graph replay when enabled) and two small
// kernels that update velocities and positions in place; the step waits for
// the device once to check list capacities and convergence.
// Relaxation loops (FIRE and LBFGS) without a host wait per step.  Each
// iteration is the optimiser's step on the forces at hand (device kernels),
// then an E+F evaluation of the new positions; the host reads the step flags,
// the running maximum of the list statistics and the convergence states once
// per block of kRelaxBlock iterations.  A step whose evaluation overflowed
// its lists changes nothing on the device (StepCtl::skip) and is repeated
// after the host has grown the capacities, so the steps are those a check
// after every evaluation would give.  The capacities get the same headroom as
// in MD.  After the last block the final evaluation is settled as any run.
static constexpr int kRelaxBlock = 8;

// Headroom for a trajectory: an eighth of the largest counts of the last
// settled run (no rerun: that run fitted).
static void trajectory_headroom(Ctx &c) {
  const int *h = c.host_stats;
  const int cap = neighbor_cap(c, c.last_flags);
  const int k = round_up(std::min(h[1] + std::max(8, h[1] / 8), std::max(cap, h[1])), 8);
  const int t = round_up(h[5] + std::max(64, h[5] / 8), 64);
  if (k > c.kcap || t > c.tcap) {
    c.kcap = std::max(c.kcap, k);
    c.tcap = std::max(c.tcap, t);
    c.invalidate_graph();
  }
}

// The blocked loop shared by both optimisers.  step(ctl) enqueues one
// optimiser step; conv() reads the number of converged structures from the
// states the caller copied back.  Returns the steps that moved a structure.
template <typename Step, typename Fetch>
static int relax_blocks(Ctx &c, int flags, int max_steps, Step step, Fetch fetch_converged, int *steps_done,
                        int *n_converged) {
  int rc, done = 0, conv = 0;
  int *flagbuf;
  if ((rc = c.get("relax_flags", kRelaxBlock, &flagbuf))) return rc;
  bool all_converged = false;
  int hf[kRelaxBlock], hs[8];
  while (done < max_steps && !all_converged) {
    const int nb = std::min(kRelaxBlock, max_steps - done);
    GAPX_HIP(launch_fill_int(c.stream, flagbuf, kRelaxBlock, 0));
    GAPX_HIP(launch_fill_int(c.stream, c.ptr<int>("stats_acc") + 8, 8, 0));
    for (int k = 0; k < nb; ++k) {
      GAPX_HIP(step(StepCtl{c.ptr<int>("stats_acc") + 16, flagbuf + k}));
      if ((rc = ctx_run(c, flags))) return rc;
    }
    GAPX_HIP(hipMemcpyAsync(hf, flagbuf, sizeof hf, hipMemcpyDeviceToHost, c.stream));
    GAPX_HIP(hipMemcpyAsync(hs, c.ptr<int>("stats_acc") + 8, sizeof hs, hipMemcpyDeviceToHost, c.stream));
    if ((rc = fetch_converged(&conv))) return rc;  // enqueues its copy and synchronises
    collect_timing(c);
    for (int k = 0; k < nb; ++k) {
      if (hf[k] & 2) continue;  // skipped: its evaluation overflowed
      if (hf[k] & 1) {
        ++done;
      } else {  // nothing moved: every structure had converged
        all_converged = true;
        break;
      }
    }
    bool grow;
    if ((rc = capacity_verdict(c, hs, std::max(8, hs[1] / 8), std::max(64, hs[5] / 8), &grow))) return rc;
  }
  if ((rc = ctx_settle(c))) return rc;  // the final evaluation
  if (steps_done) *steps_done = done;
  if (n_converged) *n_converged = conv;
  return GAPX_OK;
}

static int ctx_relax(Ctx &c, int flags, int max_steps, double fmax, double maxstep, int *steps_done,
                     int *n_converged) {
  if (c.N <= 0) return fail(GAPX_ERR_ARG, "no structures loaded");
  if (!(fmax > 0) || !(maxstep > 0) || max_steps < 0) return fail(GAPX_ERR_ARG, "gapx_ctx_relax: invalid argument");
  GAPX_HIP(hipSetDevice(c.device));
  flags = (flags | GAPX_GRAD) & ~GAPX_LAYOUT_F;
  double *vel;
  FireState *st;
  int rc;
  if ((rc = c.get("fire_vel", 3 * (size_t)c.N, &vel)) || (rc = c.get("fire_state", c.nstruct, &st))) return rc;
  GAPX_HIP(launch_fill_double(c.stream, vel, 3 * (size_t)c.N, 0.0));
  FireParams fp{0.1, 1.0, 1.1, 0.5, 0.1, 0.99, maxstep, fmax, 5, 0};
  std::vector<FireState> hs(c.nstruct);
  for (FireState &S : hs) {
    S = FireState{};
    S.dt = fp.dt0;
    S.a = fp.astart;
  }
  GAPX_HIP(hipMemcpyAsync(st, hs.data(), sizeof(FireState) * c.nstruct, hipMemcpyHostToDevice, c.stream));
  if ((rc = ctx_run(c, flags)) || (rc = ctx_settle(c))) return rc;  // forces at the start
  trajectory_headroom(c);
  auto step = [&](StepCtl ctl) {
    return launch_fire_step(c.stream, c.nstruct, c.N, c.ptr<StructParams>("sp"), c.ptr<int>("astruct"),
                            c.ptr<double>("force"), vel, c.ptr<double>("pos"), st, fp, ctl);
  };
  auto fetch = [&](int *conv) {
    GAPX_HIP(hipMemcpyAsync(hs.data(), st, sizeof(FireState) * c.nstruct, hipMemcpyDeviceToHost, c.stream));
    GAPX_HIP(hipStreamSynchronize(c.stream));
    *conv = 0;
    for (const FireState &S : hs) *conv += S.converged;
    return GAPX_OK;
  };
  return relax_blocks(c, flags, max_steps, step, fetch, steps_done, n_converged);
}

// LBFGS relaxation of the loaded batch (relax.hip has the step).
static int ctx_relax_lbfgs(Ctx &c, int flags, int max_steps, double fmax, double maxstep, int memory, double alpha,
                           int *steps_done, int *n_converged) {
  if (c.N <= 0) return fail(GAPX_ERR_ARG, "no structures loaded");
  if (!(fmax > 0) || !(maxstep > 0) || max_steps < 0 || memory < 1 || memory > kLbfgsMaxMemory || !(alpha > 0))
    return fail(GAPX_ERR_ARG, "gapx_ctx_relax_lbfgs: invalid argument");
  GAPX_HIP(hipSetDevice(c.device));
  flags = (flags | GAPX_GRAD) & ~GAPX_LAYOUT_F;
  const int mem = std::max(1, std::min(memory, max_steps));
  const size_t n3 = 3 * (size_t)c.N;
  double *r0, *f0, *S, *Y, *rho, *q;
  LbfgsState *st;
  int rc;
  if ((rc = c.get("lbfgs_r0", n3, &r0)) || (rc = c.get("lbfgs_f0", n3, &f0)) ||
      (rc = c.get("lbfgs_S", (size_t)mem * n3, &S)) || (rc = c.get("lbfgs_Y", (size_t)mem * n3, &Y)) ||
      (rc = c.get("lbfgs_rho", (size_t)mem * c.nstruct, &rho)) || (rc = c.get("lbfgs_q", n3, &q)) ||
      (rc = c.get("lbfgs_state", c.nstruct, &st)))
    return rc;
  std::vector<LbfgsState> hs(c.nstruct, LbfgsState{0, 0, 0, 0, 0.0});
  GAPX_HIP(hipMemcpyAsync(st, hs.data(), sizeof(LbfgsState) * c.nstruct, hipMemcpyHostToDevice, c.stream));
  const LbfgsParams lp{maxstep, 1.0 / alpha, fmax, mem, 0};
  if ((rc = ctx_run(c, flags)) || (rc = ctx_settle(c))) return rc;  // forces at the start
  trajectory_headroom(c);
  auto launch = [&](StepCtl ctl, const LbfgsParams &p) {
    return launch_lbfgs_step(c.stream, c.nstruct, c.N, c.ptr<StructParams>("sp"), c.ptr<double>("force"),
                             c.ptr<double>("pos"), r0, f0, S, Y, rho, q, st, p, ctl);
  };
  auto fetch = [&](int *conv) {
    GAPX_HIP(hipMemcpyAsync(hs.data(), st, sizeof(LbfgsState) * c.nstruct, hipMemcpyDeviceToHost, c.stream));
    GAPX_HIP(hipStreamSynchronize(c.stream));
    *conv = 0;
    for (const LbfgsState &L : hs) *conv += L.converged;
    return GAPX_OK;
  };
  int done = 0, conv = 0;
  if ((rc = relax_blocks(c, flags, max_steps, [&](StepCtl ctl) { return launch(ctl, lp); }, fetch, &done, &conv)))
    return rc;
  // after max_steps steps, ASE's last convergence test on the final positions
  // (the kernel moves nothing in check_only mode)
  if (done == max_steps) {
    GAPX_HIP(launch(StepCtl{nullptr, nullptr}, LbfgsParams{maxstep, 0.0, fmax, mem, 1}));
    if ((rc = fetch(&conv))) return rc;
  }
  if (steps_done) *steps_done = done;
  if (n_converged) *n_converged = conv;
  return GAPX_OK;
}

// Batched NVT MD: nsteps of ASE's NVTBerendsen over the loaded batch (md.hip
// has the step).  The forces of the starting positions come from one
// evaluation; each step is the velocity update and drift (two kernels), an
// E+F evaluation (graph replay when enabled), and the second half-kick.
// Optional log rows: [nsteps + 1][nstruct][E_pot, E_kin, T].
//
// The host does not wait for every step's list statistics.  Steps run in
// blocks of kMdCheckEvery; every evaluation folds its statistics into a
// running maximum on the device (ReduceBufs::sticky), which the host reads
// once per block.  Capacities carry headroom (an eighth of the largest
// count) so that moving atoms rarely outgrow them; when a block's lists did
// overflow, the positions, momenta, forces and energies saved at the block's
// start are restored and the block is run again with the larger capacities,
// so the trajectory is the one every-step checking would give.
static constexpr double kBoltzmannEv = 8.6173303e-5;  // eV/K, CODATA 2014 (ASE 3.18.1 units)
static constexpr int kMdCheckEvery = 16;

static int ctx_md(Ctx &c, int flags, int nsteps, double dt, double temperature, double taut, int fixcm,
                  const double *masses, const double *momenta, double *log) {
  if (c.N <= 0) return fail(GAPX_ERR_ARG, "no structures loaded");
  if (!masses || nsteps < 0 || !(dt > 0) || !(temperature >= 0) || !std::isfinite(taut))
    return fail(GAPX_ERR_ARG, "gapx_ctx_md: invalid argument");
  for (int i = 0; i < c.N; ++i)
    if (!(masses[i] > 0)) return fail(GAPX_ERR_ARG, "gapx_ctx_md: masses must be positive");
  GAPX_HIP(hipSetDevice(c.device));
  flags = (flags | GAPX_GRAD) & ~GAPX_LAYOUT_F;
  const size_t n3 = 3 * (size_t)c.N;
  double *mass, *p, *coef, *dlog = nullptr, *ck_x, *ck_p, *ck_f, *ck_e;
  int rc;
  if ((rc = c.get("md_mass", c.N, &mass)) || (rc = c.get("md_p", n3, &p)) ||
      (rc = c.get("md_coef", 4 * (size_t)c.nstruct, &coef)) || (rc = c.get("md_ck_x", n3, &ck_x)) ||
      (rc = c.get("md_ck_p", n3, &ck_p)) || (rc = c.get("md_ck_f", n3, &ck_f)) ||
      (rc = c.get("md_ck_e", c.nstruct, &ck_e)))
    return rc;
  if (log && (rc = c.get("md_log", 3 * (size_t)c.nstruct * (nsteps + 1), &dlog))) return rc;
  GAPX_HIP(hipMemcpyAsync(mass, masses, sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
  if (momenta)
    GAPX_HIP(hipMemcpyAsync(p, momenta, sizeof(double) * n3, hipMemcpyHostToDevice, c.stream));
  else
    GAPX_HIP(launch_fill_double(c.stream, p, n3, 0.0));
  const MdParams mp{dt, temperature, taut, kBoltzmannEv, fixcm ? 1 : 0, 0};
  const StructParams *d_sp = c.ptr<StructParams>("sp");
  const int *d_as = c.ptr<int>("astruct");
  if ((rc = ctx_run(c, flags)) || (rc = ctx_settle(c))) return rc;
  trajectory_headroom(c);
  if (dlog) GAPX_HIP(launch_md_log(c.stream, c.nstruct, d_sp, mass, p, c.ptr<double>("force"), c.ptr<double>("ene"),
                                   mp.kB, dlog));
  int hs[8];
  for (int b0 = 0, redo = 0; b0 < nsteps;) {
    const int nb = std::min(kMdCheckEvery, nsteps - b0);
    // the block's starting state, and a cleared running maximum
    GAPX_HIP(hipMemcpyAsync(ck_x, c.ptr<double>("pos"), sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    GAPX_HIP(hipMemcpyAsync(ck_p, p, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    GAPX_HIP(hipMemcpyAsync(ck_f, c.ptr<double>("force"), sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    GAPX_HIP(hipMemcpyAsync(ck_e, c.ptr<double>("ene"), sizeof(double) * c.nstruct, hipMemcpyDeviceToDevice,
                            c.stream));
    int *sticky = c.ptr<int>("stats_acc") + 8;
    GAPX_HIP(launch_fill_int(c.stream, sticky, 8, 0));
    for (int step = b0; step < b0 + nb; ++step) {
      double *x = c.ptr<double>("pos"), *f = c.ptr<double>("force");
      GAPX_HIP(launch_md_prepare(c.stream, c.nstruct, d_sp, mass, p, f, mp, coef));
      GAPX_HIP(launch_md_drift(c.stream, c.N, d_as, mass, f, coef, dt, p, x));
      if ((rc = ctx_run(c, flags))) return rc;
      f = c.ptr<double>("force");
      GAPX_HIP(launch_md_kick(c.stream, c.N, f, 0.5 * dt, p));
      if (dlog)
        GAPX_HIP(launch_md_log(c.stream, c.nstruct, d_sp, mass, p, f, c.ptr<double>("ene"), mp.kB,
                               dlog + 3 * (size_t)c.nstruct * (step + 1)));
    }
    sticky = c.ptr<int>("stats_acc") + 8;
    GAPX_HIP(hipMemcpyAsync(hs, sticky, sizeof hs, hipMemcpyDeviceToHost, c.stream));
    GAPX_HIP(hipStreamSynchronize(c.stream));
    collect_timing(c);
    bool grow;
    if ((rc = capacity_verdict(c, hs, std::max(8, hs[1] / 8), std::max(64, hs[5] / 8), &grow))) return rc;
    if (!grow) {
      b0 += nb;
      redo = 0;
      continue;
    }
    if (++redo > 3) return fail(GAPX_ERR_NEIGHBORS, "neighbour capacity did not converge during MD");
    // back to the block's start (buffers of the run's size class are not
    // reallocated by a capacity change, so the saved state still applies)
    GAPX_HIP(hipMemcpyAsync(c.ptr<double>("pos"), ck_x, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    GAPX_HIP(hipMemcpyAsync(p, ck_p, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    GAPX_HIP(hipMemcpyAsync(c.ptr<double>("force"), ck_f, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    GAPX_HIP(hipMemcpyAsync(c.ptr<double>("ene"), ck_e, sizeof(double) * c.nstruct, hipMemcpyDeviceToDevice,
                            c.stream));
  }
  if (log)
    GAPX_HIP(hipMemcpyAsync(log, dlog, sizeof(double) * 3 * c.nstruct * (nsteps + 1), hipMemcpyDeviceToHost,
                            c.stream));
  GAPX_HIP(hipStreamSynchronize(c.stream));
  c.md_ready = true;
  return GAPX_OK;
}

static int ctx_results(Ctx &c, double *ene, double *force, double *stress6, double *eatom, bool fortran) {
  if (!c.ran) return fail(GAPX_ERR_ARG, "no evaluation has run on this context");
  int rc = ctx_settle(c);
  if (rc) return rc;
  if (ene) GAPX_HIP(hipMemcpy(ene, c.ptr<double>("ene"), sizeof(double) * c.nstruct, hipMemcpyDeviceToHost));
  if (stress6) GAPX_HIP(hipMemcpy(stress6, c.ptr<double>("stress"), sizeof(double) * 6 * c.nstruct, hipMemcpyDeviceToHost));
  if (eatom) GAPX_HIP(hipMemcpy(eatom, c.ptr<double>("eatom"), sizeof(double) * c.N, hipMemcpyDeviceToHost));
  if (force) {
    if (!fortran) {
      GAPX_HIP(hipMemcpy(force, c.ptr<double>("force"), sizeof(double) * 3 * c.N, hipMemcpyDeviceToHost));
    } else {
      std::vector<double> f(3 * (size_t)c.N);
      GAPX_HIP(hipMemcpy(f.data(), c.ptr<double>("force"), sizeof(double) * 3 * c.N, hipMemcpyDeviceToHost));
      for (int i = 0; i < c.N; ++i)
        for (int x = 0; x < 3; ++x) force[i + (size_t)c.N * x] = f[3 * (size_t)i + x];
    }
  }
  return GAPX_OK;
}

// Predictive variance over the last run's features (after it settled).
static int ctx_variance(Ctx &c, double *var, double *sigma2) {
  if (!c.ran) return fail(GAPX_ERR_ARG, "no evaluation has run on this context");
  int rc = ctx_settle(c);
  if (rc) return rc;
  const Model &m = *c.m;
  const bool fp64 = !(c.last_flags & GAPX_MIXED);
  VarianceArgs v;
  v.N = c.N;
  v.Dp = m.Dp;
  v.Mp = m.Mp;
  v.nstruct = c.nstruct;
  // K in blocks of at most 512 MiB, a multiple of 64 rows
  v.rows = (int)std::max<size_t>(64, std::min<size_t>(round_up(c.N, 64), ((size_t)1 << 26) / m.Mp) / 64 * 64);
  v.xt64 = fp64 ? c.ptr<double>("xt64") : nullptr;
  v.xt32 = fp64 ? nullptr : c.ptr<float>("xt32");
  if (!v.xt64 && !v.xt32) return fail(GAPX_ERR_ARG, "the last run kept no features");
  v.mt = m.d_mt;
  v.mn = m.d_mn;
  v.qmm = m.d_qmm;
  v.sp = c.ptr<StructParams>("sp");
  if ((rc = c.get("var_sigma2", c.N, &v.sigma2)) || (rc = c.get("var_struct", c.nstruct, &v.var))) return rc;
  if (v.qmm) {
    if ((rc = c.get("var_xn", c.N, &v.xn)) || (rc = c.get("var_K", (size_t)v.rows * m.Mp, &v.K)) ||
        (rc = c.get("var_partial", (size_t)(m.Mp / 64) * c.N, &v.partial)))
      return rc;
    if (!fp64 && (rc = c.get("var_xw", (size_t)c.N * m.Dp, &v.xw))) return rc;
  }
  if ((rc = enqueue_variance(c.stream, v))) return rc;
  if (var) GAPX_HIP(hipMemcpyAsync(var, v.var, sizeof(double) * c.nstruct, hipMemcpyDeviceToHost, c.stream));
  if (sigma2) GAPX_HIP(hipMemcpyAsync(sigma2, v.sigma2, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.stream));
  GAPX_HIP(hipStreamSynchronize(c.stream));
  return GAPX_OK;
}

// CAR2ACSF from a caller's neighbour table (wacsf.f90:2-796): the table goes
// into the neighbour slots (image offset from the centre, the table's
// distance, weight and index), the triplet lists are built from the slots,
// the FP64 forward gives xx and, with lgrad, one backward pass per feature
// column f with dE/dx = the unit vector of column f gives the per-slot
// gradients d xx(f, i) / d pos(j) and the per-atom virials, from which dxdy
// and strs are assembled on the host in slot order.  The context holds the
// table afterwards, not a loaded batch.
static int ctx_car2acsf(Ctx &c, int na, int max_nb, int nf, const double *pos, const double *neighbor,
                        const int *count, double *xx, double *dxdy, double *strs, bool lgrad) {
  Model &m = *c.m;
  if (na <= 0 || max_nb < 0 || !pos || (!neighbor && max_nb > 0) || !count || !xx)
    return fail(GAPX_ERR_ARG, "gapx_car2acsf: invalid argument");
  if (nf != m.D) return fail(GAPX_ERR_ARG, "gapx_car2acsf: nf must be des_len = 2 nsf of the model header");
  int kmax = 0;
  for (int i = 0; i < na; ++i) {
    if (count[i] < 0 || count[i] > max_nb)
      return fail(GAPX_ERR_ARG, "gapx_car2acsf: neighbour count outside [0, max_neighbor]");
    kmax = std::max(kmax, count[i]);
  }
  const int cap = max_neighbors_for_lds(true, m.sfd, m.Dp);
  if (kmax > cap) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "gapx_car2acsf: an atom has %d neighbours, above the limit of %d", kmax, cap);
    return fail(GAPX_ERR_NEIGHBORS, buf);
  }
  GAPX_HIP(hipSetDevice(c.device));
  c.N = 0;
  c.nstruct = 0;
  c.ran = false;
  c.md_ready = false;
  c.invalidate_graph();
  const int N = na, Dp = m.Dp, kcap = round_up(std::max(8, kmax), 8);
  const size_t slots = (size_t)N * kcap;
  std::vector<int> hj(slots, 0);
  std::vector<double> hd(4 * slots, 0.0), hw(slots, 0.0);
  // NEIGHBOR(i, a, col), Fortran order; columns: image x y z, distance,
  // weight, neighbour index as a real (gap_calc.f90:113-116)
  auto tab = [&](int i, int a, int col) { return neighbor[i + (size_t)na * (a + (size_t)max_nb * col)]; };
  for (int i = 0; i < N; ++i)
    for (int a = 0; a < count[i]; ++a) {
      const size_t sl = (size_t)i * kcap + a;
      for (int x = 0; x < 3; ++x) hd[4 * sl + x] = tab(i, a, x) - pos[i + (size_t)na * x];
      hd[4 * sl + 3] = tab(i, a, 3);
      hw[sl] = tab(i, a, 4);
      const double jd = tab(i, a, 5);
      if (!(jd >= 1.0 && jd < na + 1.0)) return fail(GAPX_ERR_ARG, "gapx_car2acsf: neighbour index outside [1, NA]");
      if (!(hd[4 * sl + 3] > 0.0)) return fail(GAPX_ERR_ARG, "gapx_car2acsf: neighbour distance must be positive");
      hj[sl] = (int)jd - 1;
    }
  NeighborBufs nb{};
  TripletBufs tb{};
  DescOut o{};
  int *stats;
  double *nbw;
  int rc;
  if ((rc = c.get("nb_count", N, &nb.nb_count)) || (rc = c.get("nb_j", slots, &nb.nb_j)) ||
      (rc = c.get("nb_d", 4 * slots, &nb.nb_d)) || (rc = c.get("nb_w", slots, &nbw)) ||
      (rc = c.get("stats", 8, &stats)) || (rc = c.get("x64", (size_t)N * Dp, &o.x64)) ||
      (rc = c.get("xt64", (size_t)N * Dp, &o.xt64)) || (rc = c.get("xn", N, &o.xn)))
    return rc;
  c.pad_clean = ~0ULL;  // rows of xt64 / xn are written here for this call's atoms
  nb.nb_w = nbw;
  nb.kcap = kcap;
  nb.total = stats;
  nb.overflow = stats + 2;
  const size_t tbytes = desc_table_scratch_bytes(true, m.sfd, Dp, kcap, N);
  if (tbytes) {
    unsigned char *t;
    if ((rc = c.get("fc_scratch", tbytes, &t))) return rc;
    nb.fc_scratch = t;
  }
  GAPX_HIP(hipMemcpyAsync(nb.nb_count, count, sizeof(int) * N, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(hipMemcpyAsync(nb.nb_j, hj.data(), sizeof(int) * slots, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(hipMemcpyAsync(nb.nb_d, hd.data(), sizeof(double) * 4 * slots, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(hipMemcpyAsync(nbw, hw.data(), sizeof(double) * slots, hipMemcpyHostToDevice, c.stream));
  GAPX_HIP(launch_fill_int(c.stream, stats, 8, 0));
  // triplet lists: at most K (K - 1) / 2 pairs per atom, so no capacity retry
  tb.tcap = round_up(std::max(64, kmax * (kmax - 1) / 2), 64);
  if ((rc = c.get("trip_list", (size_t)N * tb.tcap, &tb.list)) || (rc = c.get("trip_count", N, &tb.count)) ||
      (rc = c.get("trip_raw", N, &tb.raw)))
    return rc;
  if (m.sf.n_ang > 0)
    GAPX_HIP(launch_triplets_from_slots(c.stream, N, nb, m.sf.rc_max_ang, tb));
  else
    GAPX_HIP(launch_fill_int(c.stream, tb.count, N, 0));
  o.N = N;
  o.Dp = Dp;
  GAPX_HIP(launch_descriptors_fwd(c.stream, true, N, nullptr, m.d_wz, nb, tb, m.sfd, m.d_theta_inv, m.d_mu, 0, o,
                                  c.chunk > 0 ? c.chunk : 16, c.tune[0].waves_fwd, kcap));
  std::vector<double> hx((size_t)N * Dp);
  GAPX_HIP(hipMemcpyAsync(hx.data(), o.x64, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, c.stream));
  GAPX_HIP(hipStreamSynchronize(c.stream));
  for (int i = 0; i < N; ++i)
    for (int f = 0; f < nf; ++f) xx[f + (size_t)nf * i] = hx[(size_t)i * Dp + f];
  if (dxdy) std::fill(dxdy, dxdy + (size_t)nf * na * na * 3, 0.0);
  if (strs) std::fill(strs, strs + (size_t)9 * nf * na, 0.0);
  if (lgrad && (dxdy || strs)) {
    double *g64, *slot_f, *vir;
    if ((rc = c.get("g64", (size_t)N * Dp, &g64)) || (rc = c.get("slot_f", 3 * slots, &slot_f)) ||
        (rc = c.get("atom_vir", 6 * (size_t)N, &vir)))
      return rc;
    GradSrc gs;
    gs.g64 = g64;
    std::vector<double> hg((size_t)N * Dp, 0.0), hs(3 * slots), hv(6 * (size_t)N);
    for (int f = 0; f < nf; ++f) {
      for (int i = 0; i < N; ++i) {
        if (f > 0) hg[(size_t)i * Dp + f - 1] = 0.0;
        hg[(size_t)i * Dp + f] = 1.0;
      }
      GAPX_HIP(hipMemcpyAsync(g64, hg.data(), sizeof(double) * hg.size(), hipMemcpyHostToDevice, c.stream));
      GAPX_HIP(launch_descriptors_bwd(c.stream, true, N, nullptr, m.d_wz, nb, tb, m.sfd, Dp, gs, slot_f, vir,
                                      c.tune[0].waves_bwd, kcap));
      GAPX_HIP(hipMemcpyAsync(hs.data(), slot_f, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, c.stream));
      GAPX_HIP(hipMemcpyAsync(hv.data(), vir, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, c.stream));
      GAPX_HIP(hipStreamSynchronize(c.stream));
      if (dxdy)  // d xx(f, i) / d pos(n, x): + the slot's gradient for its atom n, - it for the centre i
        for (int i = 0; i < N; ++i)
          for (int a = 0; a < count[i]; ++a) {
            const size_t sl = (size_t)i * kcap + a;
            const size_t n = (size_t)hj[sl];
            for (int x = 0; x < 3; ++x) {
              const double v = hs[3 * sl + x];
              dxdy[f + (size_t)nf * (i + (size_t)na * (n + (size_t)na * x))] += v;
              dxdy[f + (size_t)nf * (i + (size_t)na * (i + (size_t)na * x))] -= v;
            }
          }
      if (strs)  // strs(a, c, f, i) = sum over slots of delta_a d xx / d pos(j, c); symmetric
        for (int i = 0; i < N; ++i) {
          const double *v = &hv[6 * (size_t)i];
          const double S[3][3] = {{v[0], v[3], v[5]}, {v[3], v[1], v[4]}, {v[5], v[4], v[2]}};
          for (int a = 0; a < 3; ++a)
            for (int cc = 0; cc < 3; ++cc) strs[a + 3 * (cc + 3 * (f + (size_t)nf * i))] = S[a][cc];
        }
    }
  }
  int guard = 0;
  GAPX_HIP(hipMemcpy(&guard, stats + 3, sizeof(int), hipMemcpyDeviceToHost));
  if (guard) return fail(GAPX_ERR_NEIGHBORS, "descriptor kernels saw more neighbours than their LDS layout holds");
  return GAPX_OK;
}

static void ctx_free(Ctx *c);

static int ctx_create(Model *m, Ctx **out) {
  Ctx *c = new Ctx();
  c->m = m;
  c->device = m->device;
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(GAPX_ERR_HIP, "could not create a HIP stream");
  }
  for (auto &e : c->ev) hipEventCreate(&e);
  if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->aux2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->half_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->half_done_ev, hipEventDisableTiming) != hipSuccess) {
    ctx_free(c);
    return fail(GAPX_ERR_HIP, "could not create the second stream or its events");
  }
  if (hipHostMalloc((void **)&c->host_stats, 8 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->host_stats_dev, c->host_stats, 0) != hipSuccess) {
    ctx_free(c);
    return fail(GAPX_ERR_NOMEM, "pinned host allocation failed");
  }
  std::memset(c->host_stats, 0, 8 * sizeof(int));
  // Grid-form tables run fastest with one wavefront per atom (measured on the
  // 4096-atom ternary cell, D = 128); the generic kernels with four.  The
  // mixed contraction takes 128-row tiles (above D = 128 it has no other).
  for (int p = 0; p < 2; ++p) {
    Tuning &t = c->tune[p];
    if (c->m->sfd.grid) t.waves_fwd = t.waves_bwd = 1;
    t.gpr_rows = p == 1 ? 128 : 64;
    if (const char *w = std::getenv("GAPX_DESC_WAVES_FWD")) t.waves_fwd = std::atoi(w);
    if (const char *w = std::getenv("GAPX_DESC_WAVES_BWD")) t.waves_bwd = std::atoi(w);
  }
  if (const char *ch = std::getenv("GAPX_DESC_CHUNK")) c->chunk = std::atoi(ch);
  *out = c;
  return GAPX_OK;
}

static void ctx_free(Ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->aux) hipStreamSynchronize(c->aux);
  if (c->aux2) hipStreamSynchronize(c->aux2);
  c->invalidate_graph();
  for (auto &kv : c->mem)
    if (kv.second.p) hipFree(kv.second.p);
  for (auto &e : c->ev)
    if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  if (c->aux) hipStreamDestroy(c->aux);
  if (c->aux2) hipStreamDestroy(c->aux2);
  if (c->fork_ev) hipEventDestroy(c->fork_ev);
  if (c->join_ev) hipEventDestroy(c->join_ev);
  if (c->half_ev) hipEventDestroy(c->half_ev);
  if (c->half_done_ev) hipEventDestroy(c->half_done_ev);
  if (c->host_stats) hipHostFree(c->host_stats);
  delete c;
}

// ---------------------------------------------------------------------------
// model file (gap_parameters) reader, Fortran list-directed rules
// ---------------------------------------------------------------------------

struct ListReader {
  std::vector<std::string> lines;
  size_t pos = 0;
  bool read(int n, std::vector<double> &out) {
    out.clear();
    while ((int)out.size() < n) {
      if (pos >= lines.size()) return false;
      std::string s = lines[pos++];
      for (char &ch : s)
        if (ch == ',' || ch == '\t' || ch == '\r') ch = ' ';
      std::istringstream is(s);
      std::string tok;
      while (is >> tok) {
        for (char &ch : tok)
          if (ch == 'd' || ch == 'D') ch = 'e';
        char *end = nullptr;
        double v = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str()) return false;
        out.push_back(v);
      }
    }
    return true;
  }
  void skip() { ++pos; }
};

static int read_model_file(const char *path, Model &m) {
  std::ifstream f(path);
  if (!f) return fail(GAPX_ERR_FILE, std::string(path) + " does not exist");
  ListReader r;
  std::string line;
  while (std::getline(f, line)) r.lines.push_back(line);
  std::vector<double> v;
  auto need = [&](int n, const char *what) {
    if (!r.read(n, v)) {
      set_error(std::string("model file ") + path + ": cannot read " + what);
      return false;
    }
    return true;
  };
  if (!need(1, "nspecies")) return GAPX_ERR_FILE;
  const int ns = (int)v[0];
  if (ns <= 0 || ns > kMaxSpeciesZ) return fail(GAPX_ERR_FILE, "bad species count");
  for (int k = 0; k < ns; ++k) {
    if (!need(2, "species weight")) return GAPX_ERR_FILE;
    m.z.push_back((int)v[0]);
    m.w.push_back(v[1]);
  }
  m.nspecies = ns;
  if (!need(1, "nsf")) return GAPX_ERR_FILE;
  const int nsf = (int)v[0];
  if (nsf <= 0) return fail(GAPX_ERR_FILE, "bad symmetry-function count");
  for (int s = 0; s < nsf; ++s) {
    const size_t first = r.pos;
    if (!need(3, "symmetry function line")) return GAPX_ERR_FILE;
    m.sf_type.push_back((int)v[0]);
    m.sf_alpha.push_back(v[1]);
    m.sf_cut.push_back(v[2]);
    double z = 1.0;
    if (r.pos == first + 1) {  // optional 4th column on the same record: zeta
      std::istringstream is(r.lines[first]);
      std::string t[4];
      if (is >> t[0] >> t[1] >> t[2] >> t[3]) {
        for (char &ch : t[3])
          if (ch == 'd' || ch == 'D') ch = 'e';
        char *end = nullptr;
        double zz = std::strtod(t[3].c_str(), &end);
        if (end != t[3].c_str()) z = zz;
      }
    }
    m.sf_zeta.push_back(z);
  }
  m.nsf = nsf;
  if (!need(2, "nsparseX des_len")) return GAPX_ERR_FILE;
  m.M = (int)v[0];
  m.D = (int)v[1];
  if (m.M <= 0 || m.D <= 0) return fail(GAPX_ERR_FILE, "bad nsparseX / des_len");
  r.skip();
  r.skip();
  r.skip();
  if (!need(m.D, "theta")) return GAPX_ERR_FILE;
  m.theta = v;
  m.mm.resize((size_t)m.M * m.D);
  for (int k = 0; k < m.M; ++k) {
    if (!need(m.D, "sparse point")) return GAPX_ERR_FILE;
    std::copy(v.begin(), v.end(), m.mm.begin() + (size_t)k * m.D);
  }
  if (!need(m.M, "coefficients")) return GAPX_ERR_FILE;
  m.coeff = v;
  return GAPX_OK;
}

static int finish_model(Model *m) {
  if (m->D != 2 * m->nsf)
    return fail(GAPX_ERR_MODEL, "des_len must be twice the number of symmetry-function lines");
  if (m->D > 1024) return fail(GAPX_ERR_MODEL, "des_len above 1024 is not supported");
  m->Dp = round_up(m->D, 32);
  if (m->Dp > 256) return fail(GAPX_ERR_MODEL, "des_len above 256 is not supported by the contraction kernels");
  m->Mp = round_up(m->M, 64);
  int rc = build_tables(*m);
  if (rc) return rc;
  return upload_model(*m);
}

// ---------------------------------------------------------------------------
// model writer (gap_parameters text) and binary model files
// ---------------------------------------------------------------------------

// Text in the layout the reference reads (gap_calc.f90:328-362,
// wacsf.f90:46-55): species table, SF lines (a 4th column zeta only when some
// zeta differs from 1; the reference's list-directed read ignores it), sizes,
// three blank records, theta, one record per sparse point, coefficients.
// Numbers carry 17 significant digits, so a read gives the same doubles back.
static int write_model_text(const Model &m, const char *path) {
  std::FILE *f = std::fopen(path, "w");
  if (!f) return fail(GAPX_ERR_FILE, std::string("cannot write ") + path);
  bool any_zeta = false;
  for (double z : m.sf_zeta) any_zeta |= (z != 1.0);
  std::fprintf(f, "%d\n", m.nspecies);
  for (int k = 0; k < m.nspecies; ++k) std::fprintf(f, "%d %.17g\n", m.z[k], m.w[k]);
  std::fprintf(f, "%d\n", m.nsf);
  for (int s = 0; s < m.nsf; ++s) {
    std::fprintf(f, "%d %.17g %.17g", m.sf_type[s], m.sf_alpha[s], m.sf_cut[s]);
    if (any_zeta) std::fprintf(f, " %.17g", m.sf_zeta[s]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "%d %d\n\n\n\n", m.M, m.D);
  auto row = [&](const double *a, int n) {
    for (int d = 0; d < n; ++d) std::fprintf(f, d ? " %.17g" : "%.17g", a[d]);
    std::fprintf(f, "\n");
  };
  row(m.theta.data(), m.D);
  for (int k = 0; k < m.M; ++k) row(m.mm.data() + (size_t)k * m.D, m.D);
  row(m.coeff.data(), m.M);
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return fail(GAPX_ERR_FILE, std::string("write error on ") + path);
  return GAPX_OK;
}

// Binary model file: "GAPXMDL1", u32 version 1, u32 0x01020304 (byte order),
// i64 size and mtime (ns) of the text file it was made from (0 if none),
// i32 nspecies nsf D M, then z[nspecies] (i32), w, sf_type (i32), sf_alpha,
// sf_cut, sf_zeta, theta[D], mm[M][D] row-major, coeff[M] (f64), and a 64-bit
// FNV-1a of every byte after the magic.
static const char kModelMagic[8] = {'G', 'A', 'P', 'X', 'M', 'D', 'L', '1'};

struct ByteSink {
  std::vector<unsigned char> b;
  template <typename T>
  void put(const T *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    b.insert(b.end(), c, c + n * sizeof(T));
  }
  template <typename T>
  void put1(T v) { put(&v, 1); }
};

struct ByteSource {
  const unsigned char *p, *end;
  template <typename T>
  bool get(T *out, size_t n) {
    const size_t bytes = n * sizeof(T);
    if ((size_t)(end - p) < bytes) return false;
    std::memcpy(out, p, bytes);
    p += bytes;
    return true;
  }
};

static uint64_t fnv1a(const unsigned char *p, size_t n) {
  uint64_t h = 14695981039346656037ULL;
  for (size_t k = 0; k < n; ++k) {
    h ^= p[k];
    h *= 1099511628211ULL;
  }
  return h;
}

static int save_model_binary(const Model &m, const char *path, long long src_size, long long src_mtime) {
  ByteSink s;
  s.put1<uint32_t>(1);
  s.put1<uint32_t>(0x01020304u);
  s.put1<int64_t>(src_size);
  s.put1<int64_t>(src_mtime);
  s.put1<int32_t>(m.nspecies);
  s.put1<int32_t>(m.nsf);
  s.put1<int32_t>(m.D);
  s.put1<int32_t>(m.M);
  s.put(m.z.data(), m.z.size());
  s.put(m.w.data(), m.w.size());
  s.put(m.sf_type.data(), m.sf_type.size());
  s.put(m.sf_alpha.data(), m.sf_alpha.size());
  s.put(m.sf_cut.data(), m.sf_cut.size());
  s.put(m.sf_zeta.data(), m.sf_zeta.size());
  s.put(m.theta.data(), m.theta.size());
  s.put(m.mm.data(), m.mm.size());
  s.put(m.coeff.data(), m.coeff.size());
  const uint64_t h = fnv1a(s.b.data(), s.b.size());
  // written to a temporary name first, then renamed over the target, so a
  // reader never sees a partial file
  const std::string tmp = std::string(path) + ".tmp" + std::to_string((long long)getpid());
  std::FILE *f = std::fopen(tmp.c_str(), "wb");
  if (!f) return fail(GAPX_ERR_FILE, std::string("cannot write ") + tmp);
  bool ok = std::fwrite(kModelMagic, 1, 8, f) == 8 && std::fwrite(s.b.data(), 1, s.b.size(), f) == s.b.size() &&
            std::fwrite(&h, sizeof h, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(GAPX_ERR_FILE, std::string("write error on ") + path);
  }
  return GAPX_OK;
}

// Reads a binary model into m (host arrays only).  src_size/src_mtime get the
// stamp of the text file it records.
static int load_model_binary(const char *path, Model &m, long long *src_size, long long *src_mtime) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(GAPX_ERR_FILE, std::string(path) + " does not exist");
  std::vector<unsigned char> all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (all.size() < 8 + 8 || std::memcmp(all.data(), kModelMagic, 8) != 0)
    return fail(GAPX_ERR_FILE, std::string(path) + ": not a gapx model file");
  const size_t body = all.size() - 8 - 8;
  uint64_t h;
  std::memcpy(&h, all.data() + 8 + body, 8);
  if (fnv1a(all.data() + 8, body) != h) return fail(GAPX_ERR_FILE, std::string(path) + ": checksum mismatch");
  ByteSource r{all.data() + 8, all.data() + 8 + body};
  uint32_t ver = 0, order = 0;
  int64_t sz = 0, mt = 0;
  int32_t ns = 0, nsf = 0, D = 0, M = 0;
  if (!r.get(&ver, 1) || !r.get(&order, 1) || ver != 1 || order != 0x01020304u)
    return fail(GAPX_ERR_FILE, std::string(path) + ": unsupported version or byte order");
  if (!r.get(&sz, 1) || !r.get(&mt, 1) || !r.get(&ns, 1) || !r.get(&nsf, 1) || !r.get(&D, 1) || !r.get(&M, 1) ||
      ns <= 0 || ns > kMaxSpeciesZ || nsf <= 0 || D <= 0 || M <= 0 || D > 4096 || M > (1 << 24))
    return fail(GAPX_ERR_FILE, std::string(path) + ": bad header");
  m.nspecies = ns;
  m.nsf = nsf;
  m.D = D;
  m.M = M;
  m.z.resize(ns);
  m.w.resize(ns);
  m.sf_type.resize(nsf);
  m.sf_alpha.resize(nsf);
  m.sf_cut.resize(nsf);
  m.sf_zeta.resize(nsf);
  m.theta.resize(D);
  m.mm.resize((size_t)M * D);
  m.coeff.resize(M);
  if (!r.get(m.z.data(), ns) || !r.get(m.w.data(), ns) || !r.get(m.sf_type.data(), nsf) ||
      !r.get(m.sf_alpha.data(), nsf) || !r.get(m.sf_cut.data(), nsf) || !r.get(m.sf_zeta.data(), nsf) ||
      !r.get(m.theta.data(), D) || !r.get(m.mm.data(), m.mm.size()) || !r.get(m.coeff.data(), M) || r.p != r.end)
    return fail(GAPX_ERR_FILE, std::string(path) + ": truncated or oversized");
  if (src_size) *src_size = sz;
  if (src_mtime) *src_mtime = mt;
  return GAPX_OK;
}

}  // namespace gapx

// ===========================================================================
// C ABI
// ===========================================================================
using namespace gapx;

struct gapx_model : Model {};
struct gapx_ctx : Ctx {};

extern "C" {

int gapx_version(void) { return 100; }

const char *gapx_last_error(void) { return g_err.c_str(); }

int gapx_device_count(int *n) {
  if (!n) return fail(GAPX_ERR_ARG, "null pointer");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return GAPX_OK;
}

int gapx_model_create(int nspecies, const int *z, const double *w, int nsf, const int *sf_type,
                      const double *sf_alpha, const double *sf_cut, const double *sf_zeta, int des_len,
                      int nsparse, const double *theta, const double *mm, int mm_layout, int ld_mm,
                      const double *coeff, gapx_model **out) {
  if (!out || nspecies <= 0 || nsf <= 0 || des_len <= 0 || nsparse <= 0 || !z || !w || !sf_type || !sf_alpha ||
      !sf_cut || !theta || !mm || !coeff)
    return fail(GAPX_ERR_ARG, "gapx_model_create: invalid argument");
  gapx_model *m = new gapx_model();
  m->nspecies = nspecies;
  m->z.assign(z, z + nspecies);
  m->w.assign(w, w + nspecies);
  m->nsf = nsf;
  m->sf_type.assign(sf_type, sf_type + nsf);
  m->sf_alpha.assign(sf_alpha, sf_alpha + nsf);
  m->sf_cut.assign(sf_cut, sf_cut + nsf);
  if (sf_zeta)
    m->sf_zeta.assign(sf_zeta, sf_zeta + nsf);
  else
    m->sf_zeta.assign(nsf, 1.0);
  m->D = des_len;
  m->M = nsparse;
  m->theta.assign(theta, theta + des_len);
  m->mm.resize((size_t)nsparse * des_len);
  const int ld = (mm_layout == 1) ? std::max(ld_mm, nsparse) : des_len;
  for (int k = 0; k < nsparse; ++k)
    for (int d = 0; d < des_len; ++d)
      m->mm[(size_t)k * des_len + d] = (mm_layout == 1) ? mm[k + (size_t)ld * d] : mm[(size_t)k * des_len + d];
  m->coeff.assign(coeff, coeff + nsparse);
  int rc = finish_model(m);
  if (rc) {
    if (m->dev_block) hipFree(m->dev_block);
    delete m;
    return rc;
  }
  *out = m;
  return GAPX_OK;
}

int gapx_model_read(const char *path, gapx_model **out) {
  if (!path || !out) return fail(GAPX_ERR_ARG, "gapx_model_read: null argument");
  gapx_model *m = new gapx_model();
  int rc = read_model_file(path, *m);
  if (!rc) rc = finish_model(m);
  if (rc) {
    if (m->dev_block) hipFree(m->dev_block);
    delete m;
    return rc;
  }
  *out = m;
  return GAPX_OK;
}

int gapx_model_dims(const gapx_model *m, int *nspecies, int *nsf, int *des_len, int *nsparse) {
  if (!m) return fail(GAPX_ERR_ARG, "null model");
  if (nspecies) *nspecies = m->nspecies;
  if (nsf) *nsf = m->nsf;
  if (des_len) *des_len = m->D;
  if (nsparse) *nsparse = m->M;
  return GAPX_OK;
}

int gapx_model_arrays(const gapx_model *m, double *theta, double *mm, int mm_layout, int ld_mm, double *coeff) {
  if (!m) return fail(GAPX_ERR_ARG, "null model");
  if (theta) std::copy(m->theta.begin(), m->theta.end(), theta);
  if (coeff) std::copy(m->coeff.begin(), m->coeff.end(), coeff);
  if (mm) {
    const int ld = (mm_layout == 1) ? std::max(ld_mm, m->M) : m->D;
    for (int k = 0; k < m->M; ++k)
      for (int d = 0; d < m->D; ++d) {
        double v = m->mm[(size_t)k * m->D + d];
        if (mm_layout == 1)
          mm[k + (size_t)ld * d] = v;
        else
          mm[(size_t)k * m->D + d] = v;
      }
  }
  return GAPX_OK;
}

int gapx_model_header(const gapx_model *m, int *z, double *w, int *sf_type, double *sf_alpha, double *sf_cut,
                      double *sf_zeta) {
  if (!m) return fail(GAPX_ERR_ARG, "null model");
  if (z) std::copy(m->z.begin(), m->z.end(), z);
  if (w) std::copy(m->w.begin(), m->w.end(), w);
  if (sf_type) std::copy(m->sf_type.begin(), m->sf_type.end(), sf_type);
  if (sf_alpha) std::copy(m->sf_alpha.begin(), m->sf_alpha.end(), sf_alpha);
  if (sf_cut) std::copy(m->sf_cut.begin(), m->sf_cut.end(), sf_cut);
  if (sf_zeta) std::copy(m->sf_zeta.begin(), m->sf_zeta.end(), sf_zeta);
  return GAPX_OK;
}

int gapx_model_precision(const gapx_model *m, double *cancellation, int *mixed_contraction_fp64,
                         double *feature_sensitivity, int *mixed_runs_fp64) {
  if (!m) return fail(GAPX_ERR_ARG, "null model");
  if (cancellation) *cancellation = m->cancel_probe;
  if (mixed_contraction_fp64) *mixed_contraction_fp64 = (m->mixed_f64 || m->mixed_desc_f64) ? 1 : 0;
  if (feature_sensitivity) *feature_sensitivity = m->desc_probe;
  if (mixed_runs_fp64) *mixed_runs_fp64 = m->mixed_desc_f64 ? 1 : 0;
  return GAPX_OK;
}

int gapx_model_write(const gapx_model *m, const char *path) {
  if (!m || !path) return fail(GAPX_ERR_ARG, "gapx_model_write: null argument");
  return write_model_text(*m, path);
}

int gapx_model_save(const gapx_model *m, const char *path) {
  if (!m || !path) return fail(GAPX_ERR_ARG, "gapx_model_save: null argument");
  return save_model_binary(*m, path, 0, 0);
}

static int model_from_host(gapx_model *m, gapx_model **out) {
  int rc = finish_model(m);
  if (rc) {
    if (m->dev_block) hipFree(m->dev_block);
    delete m;
    return rc;
  }
  *out = m;
  return GAPX_OK;
}

int gapx_model_load(const char *path, gapx_model **out) {
  if (!path || !out) return fail(GAPX_ERR_ARG, "gapx_model_load: null argument");
  gapx_model *m = new gapx_model();
  int rc = load_model_binary(path, *m, nullptr, nullptr);
  if (rc) {
    delete m;
    return rc;
  }
  return model_from_host(m, out);
}

int gapx_model_read_cached(const char *path, const char *cache_path, gapx_model **out) {
  if (!path || !out) return fail(GAPX_ERR_ARG, "gapx_model_read_cached: null argument");
  long long size = 0, mtime = 0;
  int rc = gapx_file_stamp(path, &size, &mtime);
  if (rc) return rc;
  const std::string cache = cache_path ? std::string(cache_path) : std::string(path) + ".gapxbin";
  gapx_model *m = new gapx_model();
  long long cs = -1, cm = -1;
  if (load_model_binary(cache.c_str(), *m, &cs, &cm) == GAPX_OK && cs == size && cm == mtime)
    return model_from_host(m, out);
  // stale, missing or damaged cache: parse the text and refresh the cache
  // (a cache that cannot be written is not an error)
  delete m;
  m = new gapx_model();
  if ((rc = read_model_file(path, *m))) {
    delete m;
    return rc;
  }
  const std::string keep = g_err;
  if (save_model_binary(*m, cache.c_str(), size, mtime) != GAPX_OK) g_err = keep;
  return model_from_host(m, out);
}

void gapx_model_free(gapx_model *m) {
  if (!m) return;
  if (m->cached_ctx) ctx_free(m->cached_ctx);
  if (m->dev_block || m->d_qmm) {
    hipSetDevice(m->device);
    if (m->dev_block) hipFree(m->dev_block);
    if (m->d_qmm) hipFree(m->d_qmm);
  }
  delete m;
}

int gapx_model_set_qmm(gapx_model *m, const double *qmm, int layout, int ld) {
  if (!m) return fail(GAPX_ERR_ARG, "null model");
  GAPX_HIP(hipSetDevice(m->device));
  if (!qmm) {
    if (m->d_qmm) GAPX_HIP(hipFree(m->d_qmm));
    m->d_qmm = nullptr;
    return GAPX_OK;
  }
  const int M = m->M, Mp = m->Mp;
  if (layout != 0 && layout != 1) return fail(GAPX_ERR_ARG, "gapx_model_set_qmm: layout must be 0 or 1");
  const size_t lda = (layout == 1) ? (size_t)std::max(ld, M) : (size_t)M;
  std::vector<double> q((size_t)Mp * Mp, 0.0);
  for (int k = 0; k < M; ++k)
    for (int l = 0; l < M; ++l) {
      const double v = (layout == 1) ? qmm[k + lda * l] : qmm[(size_t)k * M + l];
      if (!std::isfinite(v)) return fail(GAPX_ERR_ARG, "gapx_model_set_qmm: non-finite entry");
      q[(size_t)k * Mp + l] = v;
    }
  if (!m->d_qmm && hipMalloc((void **)&m->d_qmm, q.size() * sizeof(double)) != hipSuccess) {
    m->d_qmm = nullptr;
    return fail(GAPX_ERR_NOMEM, "hipMalloc of Q failed");
  }
  GAPX_HIP(hipMemcpy(m->d_qmm, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  return GAPX_OK;
}

int gapx_file_stamp(const char *path, long long *size, long long *mtime_ns) {
  struct stat st;
  if (!path || stat(path, &st) != 0) return fail(GAPX_ERR_FILE, std::string(path ? path : "(null)") + " does not exist");
  if (size) *size = (long long)st.st_size;
  if (mtime_ns) *mtime_ns = (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
  return GAPX_OK;
}

static int cached(gapx_model *m, Ctx **c) {
  if (!m) return fail(GAPX_ERR_ARG, "null model");
  if (!m->cached_ctx) {
    int rc = ctx_create(m, &m->cached_ctx);
    if (rc) return rc;
  }
  *c = m->cached_ctx;
  return GAPX_OK;
}

int gapx_calc(gapx_model *m, int na, const int *species, const double *lat, const double *pos, double rcut, int flags,
              double *ene, double *force, double *stress6, double *variance) {
  Ctx *c;
  int rc = cached(m, &c);
  if (rc) return rc;
  if (na <= 0) return fail(GAPX_ERR_ARG, "na must be positive");
  const bool fortran = (flags & GAPX_LAYOUT_F) != 0;
  if ((rc = ctx_load(*c, 1, &na, species, lat, pos, rcut, fortran))) return rc;
  if ((rc = ctx_run(*c, flags & ~GAPX_VARIANCE))) return rc;
  if ((rc = ctx_results(*c, ene, force, stress6, nullptr, fortran))) return rc;
  if (variance) {
    if (flags & GAPX_VARIANCE) return ctx_variance(*c, variance, nullptr);
    *variance = 0.0;  // gap_calc.f90:206
  }
  return GAPX_OK;
}

int gapx_calc_batch(gapx_model *m, int nstruct, const int *natoms, const int *species, const double *lat,
                    const double *pos, double rcut, int flags, double *ene, double *force, double *stress6) {
  Ctx *c;
  int rc = cached(m, &c);
  if (rc) return rc;
  if ((rc = ctx_load(*c, nstruct, natoms, species, lat, pos, rcut, false))) return rc;
  if ((rc = ctx_run(*c, flags & ~GAPX_LAYOUT_F))) return rc;
  return ctx_results(*c, ene, force, stress6, nullptr, false);
}

int gapx_descriptors(gapx_model *m, int na, const int *species, const double *lat, const double *pos, double rcut,
                     int flags, double *x) {
  Ctx *c;
  int rc = cached(m, &c);
  if (rc) return rc;
  const bool fortran = (flags & GAPX_LAYOUT_F) != 0;
  if ((rc = ctx_load(*c, 1, &na, species, lat, pos, rcut, fortran))) return rc;
  if ((rc = ctx_run(*c, 0))) return rc;
  if ((rc = ctx_settle(*c))) return rc;
  std::vector<double> tmp((size_t)c->N * m->Dp);
  GAPX_HIP(hipMemcpy(tmp.data(), c->ptr<double>("x64"), tmp.size() * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < na; ++i)
    for (int d = 0; d < m->D; ++d) x[(size_t)i * m->D + d] = tmp[(size_t)i * m->Dp + d];
  return GAPX_OK;
}

int gapx_car2acsf(gapx_model *m, int na, int max_neighbor, int nf, const double *pos, const double *neighbor,
                  const int *neighbor_count, double *xx, double *dxdy, double *strs, int lgrad) {
  Ctx *c;
  int rc = cached(m, &c);
  if (rc) return rc;
  return ctx_car2acsf(*c, na, max_neighbor, nf, pos, neighbor, neighbor_count, xx, dxdy, strs, lgrad != 0);
}

int gapx_ctx_create(gapx_model *m, gapx_ctx **out) {
  if (!m || !out) return fail(GAPX_ERR_ARG, "null argument");
  Ctx *c = nullptr;
  int rc = ctx_create(m, &c);
  if (rc) return rc;
  *out = static_cast<gapx_ctx *>(c);
  return GAPX_OK;
}

int gapx_ctx_load(gapx_ctx *c, int nstruct, const int *natoms, const int *species, const double *lat,
                  const double *pos, double rcut) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  return ctx_load(*c, nstruct, natoms, species, lat, pos, rcut, false);
}

int gapx_ctx_update(gapx_ctx *c, const double *lat, const double *pos) {
  if (!c || !pos || c->N <= 0) return fail(GAPX_ERR_ARG, "gapx_ctx_update: nothing loaded");
  GAPX_HIP(hipSetDevice(c->device));
  if (lat) {
    // a rejected lattice (singular) leaves the loaded cells as they were
    const std::vector<double> lat_before = c->lat;
    const std::vector<StructParams> sp_before = c->sp;
    std::memcpy(c->lat.data(), lat, 9 * sizeof(double) * c->nstruct);
    int rc = struct_params(*c);
    if (rc) {
      c->lat = lat_before;
      c->sp = sp_before;
      struct_params(*c);  // restores nbins and the binning choice of the loaded cells
      return rc;
    }
    // a captured graph whose binning launch parameters no longer match is
    // replaced at the next run (Ctx::graph_matches)
    GAPX_HIP(hipMemcpyAsync(c->ptr<StructParams>("sp"), c->sp.data(), sizeof(StructParams) * c->nstruct,
                            hipMemcpyHostToDevice, c->stream));
  }
  GAPX_HIP(hipMemcpyAsync(c->ptr<double>("pos"), pos, sizeof(double) * 3 * c->N, hipMemcpyHostToDevice, c->stream));
  GAPX_HIP(hipStreamSynchronize(c->stream));
  return GAPX_OK;
}

int gapx_ctx_run(gapx_ctx *c, int flags) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  collect_timing(*c);
  return ctx_run(*c, flags & ~GAPX_LAYOUT_F);
}

int gapx_ctx_sync(gapx_ctx *c) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  if (c->ran) return ctx_settle(*c);
  GAPX_HIP(hipStreamSynchronize(c->stream));
  collect_timing(*c);
  return GAPX_OK;
}

int gapx_ctx_results(gapx_ctx *c, double *ene, double *force, double *stress6, double *eatom) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  return ctx_results(*c, ene, force, stress6, eatom, false);
}

int gapx_ctx_features(gapx_ctx *c, double *x, double *dedx) {
  if (!c || !c->ran) return fail(GAPX_ERR_ARG, "no evaluation has run on this context");
  int rc = ctx_settle(*c);
  if (rc) return rc;
  const Model &m = *c->m;
  const bool fp64 = !(c->last_flags & GAPX_MIXED);
  const size_t n = (size_t)c->N * m.Dp;
  std::vector<double> t(n);
  auto out = [&](double *dst) {
    for (int i = 0; i < c->N; ++i)
      for (int d = 0; d < m.D; ++d) dst[(size_t)i * m.D + d] = t[(size_t)i * m.Dp + d];
  };
  if (x) {
    if (fp64) {
      GAPX_HIP(hipMemcpy(t.data(), c->ptr<double>("x64"), n * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> f(n);
      GAPX_HIP(hipMemcpy(f.data(), c->ptr<float>("x32"), n * 4, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < n; ++k) t[k] = f[k];
    }
    out(x);
  }
  if (dedx) {
    if (!(c->last_flags & GAPX_GRAD)) return fail(GAPX_ERR_ARG, "dE/dx needs a run with GAPX_GRAD");
    if (fp64) {
      GAPX_HIP(hipMemcpy(t.data(), c->ptr<double>("g64"), n * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> f(n);
      GAPX_HIP(hipMemcpy(f.data(), c->ptr<float>("g32"), n * 4, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < n; ++k) t[k] = f[k];
    }
    out(dedx);
  }
  return GAPX_OK;
}

int gapx_ctx_md(gapx_ctx *c, int flags, int nsteps, double dt, double temperature, double taut, int fixcm,
                const double *masses, const double *momenta, double *log) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  return ctx_md(*c, flags, nsteps, dt, temperature, taut, fixcm, masses, momenta, log);
}

int gapx_ctx_momenta(gapx_ctx *c, double *p) {
  if (!c || !p || !c->md_ready) return fail(GAPX_ERR_ARG, "gapx_ctx_momenta: no MD run on this context");
  GAPX_HIP(hipSetDevice(c->device));
  GAPX_HIP(hipMemcpyAsync(p, c->ptr<double>("md_p"), sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost, c->stream));
  GAPX_HIP(hipStreamSynchronize(c->stream));
  return GAPX_OK;
}

int gapx_ctx_variance(gapx_ctx *c, double *var, double *sigma2) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  return ctx_variance(*c, var, sigma2);
}

int gapx_ctx_timing(gapx_ctx *c, int enable) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  GAPX_HIP(hipStreamSynchronize(c->stream));
  collect_timing(*c);
  c->timing = enable != 0;
  for (int s = 0; s < kNumStages; ++s) {
    c->stage_ms[s] = 0;
    c->stage_launches[s] = 0;
    c->stage_pending[s] = false;
  }
  return GAPX_OK;
}

int gapx_ctx_stage_count(gapx_ctx *c, int *n) {
  if (!c || !n) return fail(GAPX_ERR_ARG, "null argument");
  *n = kNumStages;
  return GAPX_OK;
}

int gapx_ctx_stage_info(gapx_ctx *c, int k, const char **name, double *ms, int *launches) {
  if (!c || k < 0 || k >= kNumStages) return fail(GAPX_ERR_ARG, "bad stage index");
  GAPX_HIP(hipStreamSynchronize(c->stream));
  collect_timing(*c);
  if (name) *name = kStageNames[k];
  if (ms) *ms = c->stage_ms[k];
  if (launches) *launches = c->stage_launches[k];
  return GAPX_OK;
}

int gapx_ctx_run_surrogate(gapx_ctx *c, gapx_mlp *m, int flags) {
  if (!c || !m) return fail(GAPX_ERR_ARG, "null argument");
  collect_timing(*c);
  return ctx_run(*c, flags & ~GAPX_LAYOUT_F, mlp_of(m));
}

int gapx_ctx_export_labels(gapx_ctx *c, gapx_mlp *m) {
  if (!c || !m) return fail(GAPX_ERR_ARG, "null argument");
  if (!c->ran || c->mlp || !(c->last_flags & GAPX_GRAD))
    return fail(GAPX_ERR_ARG, "gapx_ctx_export_labels: needs a GAP run with GAPX_GRAD first");
  GAPX_HIP(hipSetDevice(c->device));
  int rc;
  if ((rc = ctx_settle(*c))) return rc;
  const bool fp64 = !(c->last_flags & GAPX_MIXED);
  return mlp_add_labels(*mlp_of(m), c->stream, c->N, fp64 ? nullptr : c->ptr<float>("xt32"),
                        fp64 ? c->ptr<double>("xt64") : nullptr, fp64 ? nullptr : c->ptr<float>("g32"),
                        fp64 ? c->ptr<double>("g64") : nullptr, c->ptr<double>("eatom"), c->m->d_theta_inv, c->m->Dp);
}

int gapx_ctx_relax(gapx_ctx *c, int flags, int max_steps, double fmax, double maxstep, int *steps_done,
                   int *n_converged) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  collect_timing(*c);
  return ctx_relax(*c, flags, max_steps, fmax, maxstep, steps_done, n_converged);
}

int gapx_ctx_relax_lbfgs(gapx_ctx *c, int flags, int max_steps, double fmax, double maxstep, int memory,
                         double alpha, int *steps_done, int *n_converged) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  collect_timing(*c);
  return ctx_relax_lbfgs(*c, flags, max_steps, fmax, maxstep, memory, alpha, steps_done, n_converged);
}

int gapx_ctx_positions(gapx_ctx *c, double *pos) {
  if (!c || !pos || c->N <= 0) return fail(GAPX_ERR_ARG, "gapx_ctx_positions: nothing loaded");
  GAPX_HIP(hipSetDevice(c->device));
  GAPX_HIP(hipStreamSynchronize(c->stream));
  GAPX_HIP(hipMemcpy(pos, c->ptr<double>("pos"), sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost));
  return GAPX_OK;
}

int gapx_ctx_autotune(gapx_ctx *c, int flags, int force) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  GAPX_HIP(hipSetDevice(c->device));
  return ctx_autotune(*c, flags & ~GAPX_LAYOUT_F, force != 0);
}

int gapx_ctx_get_tuning(gapx_ctx *c, int flags, int *waves_fwd, int *waves_bwd, int *gpr_rows, int *gpr_blocks,
                        int *halves) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  const Tuning &t = c->tune[(run_flags(*c->m, flags) & GAPX_MIXED) ? 1 : 0];
  if (waves_fwd) *waves_fwd = t.waves_fwd;
  if (waves_bwd) *waves_bwd = t.waves_bwd;
  if (gpr_rows) *gpr_rows = t.gpr_rows;
  if (gpr_blocks) *gpr_blocks = t.gpr_blocks;
  if (halves) *halves = t.halves;
  return GAPX_OK;
}

int gapx_ctx_set_tuning(gapx_ctx *c, int flags, int waves_fwd, int waves_bwd, int gpr_rows, int gpr_blocks,
                        int halves) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  auto ok_w = [](int w) { return w == 1 || w == 2 || w == 4; };
  if (!ok_w(waves_fwd) || !ok_w(waves_bwd) || (gpr_rows != 64 && gpr_rows != 128) || gpr_blocks < 1 ||
      gpr_blocks > (1 << 20) || (halves != 0 && halves != 1))
    return fail(GAPX_ERR_ARG,
                "gapx_ctx_set_tuning: waves must be 1, 2 or 4, rows 64 or 128, blocks in [1, 2^20], halves 0 or 1");
  Tuning &t = c->tune[(run_flags(*c->m, flags) & GAPX_MIXED) ? 1 : 0];
  t.waves_fwd = waves_fwd;
  t.waves_bwd = waves_bwd;
  t.gpr_rows = gpr_rows;
  t.gpr_blocks = gpr_blocks;
  t.halves = halves;
  c->invalidate_graph();
  return GAPX_OK;
}

int gapx_ctx_use_graph(gapx_ctx *c, int on) {
  if (!c) return fail(GAPX_ERR_ARG, "null context");
  c->use_graph = on != 0;
  if (!on) c->invalidate_graph();
  return GAPX_OK;
}

int gapx_ctx_counts(gapx_ctx *c, long long *pairs, long long *triplets) {
  if (!c || !c->ran) return fail(GAPX_ERR_ARG, "no evaluation has run on this context");
  GAPX_HIP(hipStreamSynchronize(c->stream));
  int rc = ctx_settle(*c);
  if (rc) return rc;
  if (pairs) *pairs = c->host_stats[0];
  if (triplets) *triplets = c->host_stats[4];
  return GAPX_OK;
}

int gapx_ctx_device_bytes(gapx_ctx *c, long long *bytes) {
  if (!c || !bytes) return fail(GAPX_ERR_ARG, "null argument");
  *bytes = (long long)c->bytes;
  return GAPX_OK;
}

void gapx_ctx_free(gapx_ctx *c) { ctx_free(c); }

}  // extern "C"
