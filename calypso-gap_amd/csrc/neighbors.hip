// Periodic neighbour lists on the GPU: cell-list binning in fractional
// space, then a CSR list of every (j, image) within rcut of each atom.
//
// Replaces the brute-force O(N^2 * images) search of FGAP_CALC
// (gappy/libgap/gap_calc.f90:85-120).  The neighbour set is the same: every
// image R_j + n.L with |R_j + n.L - R_i| <= rcut, (j, n) != (i, 0), self-images
// included (gap_calc.f90:97-101).  Positions are wrapped into the cell first;
// the stored image n refers to the positions as given.
//
// Statistics (kept pairs, largest count, overflow) are per-atom values folded
// by one small kernel at the end of a run: same-address atomics from every
// wavefront serialise in the memory system and cost more than the search.
//
// Bins: nb_k = max(1, floor(2 d_k / rcut)) along each lattice direction (d_k
// the interplanar spacing); a neighbour can be at most S_k = ceil(rcut nb_k /
// d_k) bins away (2 where the cell holds two bins per cutoff), and offsets beyond the cell wrap with an image shift, so cells
// smaller than rcut need no special case.  Atoms inside a bin are kept in
// ascending index order so the lists (and every sum over them) are
// deterministic.
#include <algorithm>
#include <type_traits>

#include "gapx_internal.h"

namespace gapx {

namespace {

// floor(a / b) for b >= 1 and |a| < 2^21 through FP32, given invb within
// 2^-23 (relative) of 1 / b, as inv_small gives it: (a + 1/2) / b is at
// least 1/(2b) from every integer, while the error of invb and the rounding
// of the product move it by at most |a + 1/2| (2^-23 + 2^-24) / b < 3/(8b),
// so the floor is exact.
__device__ __forceinline__ int floordiv_small(int a, float invb) {
  return (int)floorf(((float)a + 0.5f) * invb);
}

// 1 / b by v_rcp_f32 (at most 1 ulp from the exact value)
__device__ __forceinline__ float inv_small(int b) { return __builtin_amdgcn_rcpf((float)b); }

__global__ void k_fill_int(int *p, int n, int v) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void k_zero3(int *a, int na, int *b, int nb, int *c, int nc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < na) a[t] = 0;
  if (t < nb) b[t] = 0;
  if (t < nc) c[t] = 0;
}

__global__ void k_fill_double(double *p, size_t n, double v) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// Wrap, bin and count, then scan, in one launch of 256-thread blocks.  Every block
// wraps and bins its atoms (device-scope atomics on bin_count); the last
// block to finish (a ticket counter, with device fences on both sides) then
// scans the counts into bin_start and leaves the buffers ready for the next
// run: bin_count zero (so no clearing launch is needed), the scatter cursors
// zero, the run's statistics zero and the ticket back at zero.  The host
// clears bin_count and the ticket once whenever they are newly allocated.
__global__ __launch_bounds__(256) void k_wrap_scan(int N, int nbins, const double *__restrict__ pos,
                                                   const int *__restrict__ atom_struct,
                                                   const StructParams *__restrict__ sp, NeighborBufs nb,
                                                   int *__restrict__ ticket, int *__restrict__ stats) {
  const int i = blockIdx.x * 256 + threadIdx.x, t = threadIdx.x;
  if (i < N) {
    const StructParams &P = sp[atom_struct[i]];
    const double r[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int bk3[3], sh[3];
    for (int k = 0; k < 3; ++k) {
      const double f = r[0] * P.Linv[k] + r[1] * P.Linv[3 + k] + r[2] * P.Linv[6 + k];
      const double fl = floor(f);
      sh[k] = (int)fl;
      const int bk = (int)floor((f - fl) * P.nb[k]);
      bk3[k] = bk < 0 ? 0 : (bk >= P.nb[k] ? P.nb[k] - 1 : bk);
    }
    for (int c = 0; c < 3; ++c) {
      nb.wpos[3 * i + c] = r[c] - sh[0] * P.L[c] - sh[1] * P.L[3 + c] - sh[2] * P.L[6 + c];
      nb.shift[3 * i + c] = sh[c];
    }
    const int bin = P.bin_base + (bk3[0] * P.nb[1] + bk3[1]) * P.nb[2] + bk3[2];
    nb.atom_bin[i] = bin;
    atomicAdd(&nb.bin_count[bin], 1);
  }
  __shared__ int s_last;
  __threadfence();
  __syncthreads();
  if (t == 0) s_last = (atomicAdd(ticket, 1) == (int)gridDim.x - 1);
  __syncthreads();
  if (!s_last) return;
  __threadfence();  // every block's count atomics before this block's plain loads
  // Exclusive scan of bin_count in tiles of 2048 bins, eight consecutive bins
  // per thread read as two 16-byte words (the next tile's issued before this
  // one is scanned): a wave prefix sum of the thread totals, the four wave
  // totals through LDS, then bin_start written and bin_count and bin_cursor
  // cleared with 16-byte stores.
  __shared__ int s_wsum[4];
  const int lane = t & 63, wv = t >> 6;
  auto load8 = [&](int k0, int (&v)[8]) {
    if (k0 + 8 <= nbins) {
      const int4 x = *(const int4 *)(nb.bin_count + k0), y = *(const int4 *)(nb.bin_count + k0 + 4);
      v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w, v[4] = y.x, v[5] = y.y, v[6] = y.z, v[7] = y.w;
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (k0 + u < nbins) ? nb.bin_count[k0 + u] : 0;
    }
  };
  auto store8 = [&](int *p, int k0, const int (&v)[8]) {
    if (k0 + 8 <= nbins) {
      *(int4 *)(p + k0) = make_int4(v[0], v[1], v[2], v[3]);
      *(int4 *)(p + k0 + 4) = make_int4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < nbins) p[k0 + u] = v[u];
    }
  };
  const int zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int carry = 0;
  int cur[8], nxt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  load8(8 * t, cur);
  for (int tile0 = 0; tile0 < nbins; tile0 += 2048) {
    const int k0 = tile0 + 8 * t;
    if (tile0 + 2048 < nbins) load8(k0 + 2048, nxt);
    int tot = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) tot += cur[u];
    int x = tot;  // inclusive wave prefix of the thread totals
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[wv] = x;
    __syncthreads();
    int run = carry + x - tot, tile_total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      run += (w < wv) ? s_wsum[w] : 0;
      tile_total += s_wsum[w];
    }
    __syncthreads();  // s_wsum is rewritten by the next tile
    int st[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      st[u] = run;
      run += cur[u];
    }
    store8(nb.bin_start, k0, st);
    store8(nb.bin_count, k0, zero);
    store8(nb.bin_cursor, k0, zero);
    carry += tile_total;
#pragma unroll
    for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
  }
  if (t == 0) nb.bin_start[nbins] = carry;
  if (t < 8) stats[t] = 0;
  if (t == 0) *ticket = 0;
}

__global__ void k_scatter(int N, NeighborBufs nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int b = nb.atom_bin[i];
  int slot = nb.bin_start[b] + atomicAdd(&nb.bin_cursor[b], 1);
  nb.bin_atoms[slot] = i;
}

// Copy atom j's wrapped position and (index, cell shift) to bin slot q, so
// the pair search reads one contiguous record per candidate.
__device__ __forceinline__ void put_sorted(const NeighborBufs &nb, int q, int j) {
  nb.bin_atoms[q] = j;
  double2 *sp = (double2 *)(nb.spos + 4 * (size_t)q);
  sp[0] = make_double2(nb.wpos[3 * j], nb.wpos[3 * j + 1]);
  sp[1] = make_double2(nb.wpos[3 * j + 2], 0.0);
  ((int4 *)nb.sinfo)[q] = make_int4(j, nb.shift[3 * j], nb.shift[3 * j + 1], nb.shift[3 * j + 2]);
}

// Order each bin's atoms by index: one wavefront per bin, every element's
// rank is the number of smaller indices in the bin (bins hold tens of atoms).
__global__ __launch_bounds__(64) void k_rank_bins(int nbins, NeighborBufs nb) {
  const int b = blockIdx.x, lane = threadIdx.x;
  if (b >= nbins) return;
  const int lo = nb.bin_start[b], n = nb.bin_start[b + 1] - lo;
  if (n <= 64) {
    const int v = (lane < n) ? nb.bin_atoms[lo + lane] : 0x7fffffff;
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += (__shfl(v, k, 64) < v) ? 1 : 0;
    if (lane < n) put_sorted(nb, lo + rank, v);
  } else {
    for (int e = lane; e < n; e += 64) {
      const int v = nb.bin_atoms[lo + e];
      int rank = 0;
      for (int k = 0; k < n; ++k) rank += (nb.bin_atoms[lo + k] < v) ? 1 : 0;
      nb.bin_cursor[lo + rank] = v;  // bin_cursor doubles as scratch
    }
    __syncthreads();
    for (int e = lane; e < n; e += 64) put_sorted(nb, lo + e, nb.bin_cursor[lo + e]);
  }
}

// One workgroup per structure does the whole binning in LDS: wrap and bin
// every atom, count per bin, exclusive scan, scatter, then order each bin's
// atoms by index and write the bin-sorted records; workgroup 0 also clears
// the run's statistics.  Same output as k_wrap_scan + k_scatter +
// k_rank_bins, in one launch, for structures of at most kFusedAtoms atoms
// (bins_max bins and atoms_max atoms size the LDS arrays, see
// fused_bin_lds_bytes).  Thread t owns atoms t, t + 1024, .. from the wrap to
// the sorted write and keeps their shifts and bins in registers (the wrapped
// positions it wrote itself are read back once, all together, at the end).
__global__ __launch_bounds__(1024) void k_bin_struct(const double *__restrict__ pos,
                                                     const StructParams *__restrict__ sp, NeighborBufs nb,
                                                     int nbins_total, int bins_max, int *__restrict__ stats) {
  constexpr int PER = kFusedAtoms / 1024;
  extern __shared__ int smem_i[];
  // [bins_max + 1]: counts, then bin starts, then (after the scatter, whose
  // cursors they are) bin ends
  int *s_cnt = smem_i;
  // [max(atoms_max, 1024)]: the scan's partial sums, then the local atom
  // index of every slot, unsorted inside each bin
  int *s_tmp = s_cnt + bins_max + 1;
  int *s_part = s_tmp;
  const int s = blockIdx.x, t = threadIdx.x;
  const StructParams &P = sp[s];
  const int n = P.natoms, base = P.atom_base, nbs = P.nb[0] * P.nb[1] * P.nb[2];
  if (s == 0 && t < 8) stats[t] = 0;
  for (int b = t; b < nbs; b += 1024) s_cnt[b] = 0;
  __syncthreads();
  int sh[PER][3], lb[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int a = t + 1024 * u;
    lb[u] = 0;
    if (a < n) {
      const int i = base + a;
      const double r[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      int bk3[3];
      for (int k = 0; k < 3; ++k) {
        const double f = r[0] * P.Linv[k] + r[1] * P.Linv[3 + k] + r[2] * P.Linv[6 + k];
        const double fl = floor(f);
        sh[u][k] = (int)fl;
        const int bk = (int)floor((f - fl) * P.nb[k]);
        bk3[k] = bk < 0 ? 0 : (bk >= P.nb[k] ? P.nb[k] - 1 : bk);
      }
      for (int c = 0; c < 3; ++c) {
        nb.wpos[3 * i + c] = r[c] - sh[u][0] * P.L[c] - sh[u][1] * P.L[3 + c] - sh[u][2] * P.L[6 + c];
        nb.shift[3 * i + c] = sh[u][c];
      }
      lb[u] = (bk3[0] * P.nb[1] + bk3[1]) * P.nb[2] + bk3[2];
      nb.atom_bin[i] = P.bin_base + lb[u];
      atomicAdd(&s_cnt[lb[u]], 1);
    }
  }
  __syncthreads();
  // exclusive scan of the counts
  const int per = (nbs + 1023) / 1024, lo = t * per, hi = min(nbs, lo + per);
  int sum = 0;
  for (int b = lo; b < hi; ++b) sum += s_cnt[b];
  s_part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = (t >= off) ? s_part[t - off] : 0;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  int run = s_part[t] - sum;
  for (int b = lo; b < hi; ++b) {
    const int v = s_cnt[b];
    s_cnt[b] = run;
    run += v;
  }
  __syncthreads();
  for (int b = t; b < nbs; b += 1024) nb.bin_start[P.bin_base + b] = base + s_cnt[b];
  if (t == 0 && s == (int)gridDim.x - 1) nb.bin_start[nbins_total] = base + n;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int a = t + 1024 * u;
    if (a < n) s_tmp[atomicAdd(&s_cnt[lb[u]], 1)] = a;
  }
  __syncthreads();
  // each atom's rank inside its bin = the number of smaller indices there
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int a = t + 1024 * u;
    if (a < n) {
      const int b0 = lb[u] > 0 ? s_cnt[lb[u] - 1] : 0, b1 = s_cnt[lb[u]];
      int rank = 0;
      for (int k = b0; k < b1; ++k) rank += (s_tmp[k] < a) ? 1 : 0;
      const int q = base + b0 + rank, j = base + a;
      nb.bin_atoms[q] = j;
      double2 *p2 = (double2 *)(nb.spos + 4 * (size_t)q);
      p2[0] = make_double2(nb.wpos[3 * j], nb.wpos[3 * j + 1]);
      p2[1] = make_double2(nb.wpos[3 * j + 2], 0.0);
      ((int4 *)nb.sinfo)[q] = make_int4(j, sh[u][0], sh[u][1], sh[u][2]);
    }
  }
}

__device__ __forceinline__ unsigned long long nb_key(int j, int n0, int n1, int n2) {
  return ((unsigned long long)(unsigned)j << 24) | ((unsigned long long)((n0 + 128) & 255) << 16) |
         ((unsigned long long)((n1 + 128) & 255) << 8) | (unsigned long long)((n2 + 128) & 255);
}

// The compact key: j (below 2^17) and each image component modulo 32.  Two
// images of one atom within a list differ by at most 2 S + 2 <= 32 cells per
// axis when S <= 15, so the key stays unique within a list, and a self image
// (j = i) has |n_k| <= S, so its components are recovered by sign extension.
__device__ __forceinline__ uint32_t nb_key32(int j, int n0, int n1, int n2) {
  return ((uint32_t)j << 15) | ((uint32_t)(n0 & 31) << 10) | ((uint32_t)(n1 & 31) << 5) | (uint32_t)(n2 & 31);
}
__device__ __forceinline__ int sext5(uint32_t v) { return (int)((v & 31u) ^ 16u) - 16; }

// Triplet lists, from the neighbour records the building wave just staged in
// LDS.  Pairs (a, b), a < b, with r_a, r_b, r_ab all within the largest
// angular cutoff are appended (ballot + prefix count) to the atom's list at
// i * tcap, packed as a | b << 16, in diagonal order: the pairs with b - a = d
// form a run of K - d entries.  Runs d and K - d together hold K entries, so
// one wave step takes both (full lanes for K <= 64).  Consecutive entries of
// a run have distinct a and distinct b, so the backward pass's LDS atomics on
// per-neighbour accumulators rarely collide within one instruction.  Entry c
// of a step (c < K) always has neighbour c as one end (a = c in run d, b = c
// in run K - d), so each lane keeps its own records in registers (K <= 256)
// and reads only the partner's, record (c + d) mod K: the first K / 2
// records are repeated after the K staged ones, so the read index is c + d.
// The cutoff test is a squared comparison with a relative margin of 1e-6; a
// triplet admitted by the margin contributes zero, since the descriptor
// kernels apply each line's cutoff to every distance.  Records with r beyond
// the margin are staged with a NaN x (trip_record), so every distance test
// that involves them fails and no separate r test is needed.
__device__ __forceinline__ float trip_rc(double rcA) { return (float)rcA * (1.0f + 1e-6f); }

__device__ __forceinline__ float4 trip_record(double dx, double dy, double dz, double r, float rc) {
  const float rf = (float)r;
  return make_float4(rf <= rc ? (float)dx : __builtin_nanf(""), (float)dy, (float)dz, rf);
}

// LDS records the triplet builder needs for lists of up to kcap neighbours:
// the K staged ones and, for K <= 256, the repeated first half after them
// (at most 384 records in all), so never more than 64 KB for kcap <= 4096
__host__ __device__ constexpr size_t trip_lds_records(int kcap) {
  return (size_t)std::max(kcap, std::min(kcap, 256) + std::min(kcap, 256) / 2);
}

__device__ __forceinline__ void triplets_from_lds(int i, int K, float4 *s_p, double rcA, const TripletBufs &tb,
                                                  int lane) {
  constexpr int kOwn = 4;  // register-held record chunks: K <= 64 kOwn
  K = __builtin_amdgcn_readfirstlane(K);
  const float rc = trip_rc(rcA);
  const float rc2 = rc * rc;
  uint32_t *out = tb.list + (size_t)i * tb.tcap;
  int count = 0;
  // pairs of the step whose lanes pass `ok`, appended after `count` entries
  auto emit = [&](bool ok, uint32_t pair) {
    const unsigned long long m = __ballot(ok);
    const unsigned pre = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    if (ok && (int)pre < tb.tcap - count) out[count + pre] = pair;
    count += __popcll(m);
  };
  if (K <= 64 * kOwn) {
    for (int t = lane; t < K / 2; t += 64) s_p[K + t] = s_p[t];
    __syncthreads();
    const int KX = K + K / 2;  // records readable at c + d
    float4 own[kOwn];
    uint32_t self[kOwn];  // c | c << 16
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
      const int c = 64 * q + lane;
      own[q] = s_p[min(c, max(K - 1, 0))];
      if (c >= K) own[q].x = __builtin_nanf("");  // idle lanes fail every test
      self[q] = (uint32_t)c | ((uint32_t)c << 16);
    }
    // one step: run d1 (K - d1 pairs) and run K - d1 (d1 pairs), or, when the
    // two are the same run (2 d1 = K), its n = K - d1 pairs only
    auto step = [&](int d1, int n, auto whole) {
      const int n1 = K - d1;
      const uint32_t lo = (uint32_t)(d1 - K), hi = (uint32_t)d1 << 16;
#pragma unroll
      for (int q = 0; q < kOwn; ++q) {
        if (64 * q >= n) break;
        const int c = 64 * q + lane;
        const float4 pp = s_p[min(c + d1, KX - 1)];
        const float ex = pp.x - own[q].x, ey = pp.y - own[q].y, ez = pp.z - own[q].z;
        bool ok = ex * ex + ey * ey + ez * ez <= rc2;
        if (!decltype(whole)::value) ok = ok && c < n;
        // run d1: (c, c + d1); run K - d1: (c + d1 - K, c)
        emit(ok, self[q] + (c < n1 ? hi : lo));
      }
    };
    int d1 = 1;
    for (; 2 * d1 < K; ++d1) step(d1, K, std::true_type{});
    if (2 * d1 == K) step(d1, K - d1, std::false_type{});
  } else {
    for (int d1 = 1; 2 * d1 <= K; ++d1) {
      // run d1 (K - d1 pairs), then run d2 = K - d1 (d1 pairs) unless it is d1
      const int d2 = K - d1;
      const int n1 = K - d1, n = (d2 != d1) ? K : n1;
      for (int c0 = 0; c0 < n; c0 += 64) {
        const int c = c0 + lane;
        const bool first = c < n1;
        const int a = first ? c : c - n1, b = a + (first ? d1 : d2);
        // both records read unconditionally (idle lanes read record 0) and
        // the tests combined without short-circuits: one LDS round trip
        const bool live = c < n;
        const float4 pa = s_p[live ? a : 0], pb = s_p[live ? b : 0];
        const float ex = pb.x - pa.x, ey = pb.y - pa.y, ez = pb.z - pa.z;
        emit(live & (ex * ex + ey * ey + ez * ez <= rc2), (uint32_t)a | ((uint32_t)b << 16));
      }
    }
  }
  if (lane == 0) {
    tb.count[i] = min(count, tb.tcap);
    tb.raw[i] = count;
  }
}

// One wavefront per centre atom.  The (2S+1)^3 bin images are taken 64 at a
// time, one per lane (bin range and image shift); a wave prefix sum over the
// bin sizes flattens their atoms into one candidate sequence, which the lanes
// then test 64 at a time, each finding its image by a binary search over the
// prefix sums.  Accepted pairs are appended with ballot + prefix count, so the
// list order is image order, then index order inside the bin (deterministic).
// Slots live at i * kcap + a; counts above kcap are recorded, not written.
// With TRIP the wave also keeps each kept neighbour's (dx, dy, dz, r) in LDS
// (FP32) and builds the atom's triplet list from there (triplets_from_lds).
template <bool TRIP>
__global__ __launch_bounds__(64) void k_build(int N, const int *__restrict__ atom_struct,
                                              const StructParams *__restrict__ sp, double rcut, NeighborBufs nb,
                                              double rcA, TripletBufs tb) {
  extern __shared__ __align__(16) unsigned char smem[];
  float4 *s_p = (float4 *)smem;  // [trip_lds_records(kcap)] with TRIP
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= N) return;
  const StructParams &P = sp[atom_struct[i]];
  const int b = nb.atom_bin[i] - P.bin_base;
  const int bi2 = b % P.nb[2], bi1 = (b / P.nb[2]) % P.nb[1], bi0 = b / (P.nb[1] * P.nb[2]);
  const double xi = nb.wpos[3 * i], yi = nb.wpos[3 * i + 1], zi = nb.wpos[3 * i + 2];
  const int si0 = nb.shift[3 * i], si1 = nb.shift[3 * i + 1], si2 = nb.shift[3 * i + 2];
  const int m1 = 2 * P.S[1] + 1, m2 = 2 * P.S[2] + 1, nimg = (2 * P.S[0] + 1) * m1 * m2;
  // image and bin indices stay below 2^20 (the host keeps S <= 48)
  const float inv_m1 = inv_small(m1), inv_m2 = inv_small(m2);
  const float inv_n0 = inv_small(P.nb[0]), inv_n1 = inv_small(P.nb[1]), inv_n2 = inv_small(P.nb[2]);
  const size_t base = (size_t)i * nb.kcap;
  const double2 *sp2 = (const double2 *)nb.spos;
  const int4 *sinfo = (const int4 *)nb.sinfo;
  // Image bins that lie wholly beyond rcut + 0.01 A of the centre are skipped
  // (no candidate there can pass, so the list is the same).  Image o of the
  // centre's bin holds the points o_k + v_k - u_k bins away along axis k,
  // v_k in [0, 1], u_k the centre's offset in its own bin; with the bin
  // metric factor R (|displacement| = |R f|) each component of R f lies in an
  // interval [c_a + m_a, c_a + M_a], c_a = sum_b R_ab (o_b - u_b), and the sum
  // of the squared distances of those intervals from 0 bounds the squared
  // distance from below.  FP32 suffices: its rounding stays far inside the
  // 0.01 A margin.
  const float R00 = P.R[0], R01 = P.R[1], R02 = P.R[2], R11 = P.R[3], R12 = P.R[4], R22 = P.R[5];
  float cu0, cu1, cu2;  // -R u
  {
    float u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double f = xi * P.Linv[k] + yi * P.Linv[3 + k] + zi * P.Linv[6 + k];
      u[k] = (float)(f * P.nb[k] - (k == 0 ? bi0 : k == 1 ? bi1 : bi2));
    }
    cu0 = -(R00 * u[0] + R01 * u[1] + R02 * u[2]);
    cu1 = -(R11 * u[1] + R12 * u[2]);
    cu2 = -(R22 * u[2]);
  }
  const float m0 = fminf(R00, 0.f) + fminf(R01, 0.f) + fminf(R02, 0.f), m1r = fminf(R11, 0.f) + fminf(R12, 0.f);
  const float M0 = fmaxf(R00, 0.f) + fmaxf(R01, 0.f) + fmaxf(R02, 0.f), M1r = fmaxf(R11, 0.f) + fmaxf(R12, 0.f);
  const float m2r = fminf(R22, 0.f), M2r = fmaxf(R22, 0.f);
  const float rp = (float)rcut + 0.01f, rp2 = rp * rp;
  int count = 0;
  for (int g0 = 0; g0 < nimg; g0 += 64) {
    const int img = g0 + lane;
    int lo = 0, n = 0, t0 = 0, t1 = 0, t2 = 0;
    if (img < nimg) {
      const int q2 = floordiv_small(img, inv_m2), q1 = floordiv_small(q2, inv_m1);
      const int o2 = img - q2 * m2 - P.S[2], o1 = q2 - q1 * m1 - P.S[1], o0 = q1 - P.S[0];
      const int bb0 = bi0 + o0, bb1 = bi1 + o1, bb2 = bi2 + o2;
      t0 = floordiv_small(bb0, inv_n0);
      t1 = floordiv_small(bb1, inv_n1);
      t2 = floordiv_small(bb2, inv_n2);
      const int w0 = bb0 - t0 * P.nb[0], w1 = bb1 - t1 * P.nb[1], w2 = bb2 - t2 * P.nb[2];
      const float c0 = cu0 + R00 * o0 + R01 * o1 + R02 * o2, c1 = cu1 + R11 * o1 + R12 * o2, c2 = cu2 + R22 * o2;
      const float e0 = fmaxf(fmaxf(c0 + m0, -(c0 + M0)), 0.f), e1 = fmaxf(fmaxf(c1 + m1r, -(c1 + M1r)), 0.f),
                  e2 = fmaxf(fmaxf(c2 + m2r, -(c2 + M2r)), 0.f);
      if (e0 * e0 + e1 * e1 + e2 * e2 <= rp2) {
        const int bin = P.bin_base + (w0 * P.nb[1] + w1) * P.nb[2] + w2;
        lo = nb.bin_start[bin];
        n = nb.bin_start[bin + 1] - lo;
      }
    }
    int incl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    const int first = lo - (incl - n);  // candidate c of this image sits at slot first + c
    for (int c0 = 0; c0 < total; c0 += 64) {
      const int c = c0 + lane;
      int src = 0;  // number of images whose prefix ends at or before c
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(incl, src + step - 1, 64) <= c) src += step;
      const int q = __shfl(first, src, 64) + c;
      const int T0 = __shfl(t0, src, 64), T1 = __shfl(t1, src, 64), T2 = __shfl(t2, src, 64);
      bool ok = false;
      int4 inf = make_int4(0, 0, 0, 0);
      double dx = 0, dy = 0, dz = 0, r = 0;
      if (c < total) {
        inf = sinfo[q];
        const double2 pxy = sp2[2 * (size_t)q], pz = sp2[2 * (size_t)q + 1];
        if (!(T0 == 0 && T1 == 0 && T2 == 0 && inf.x == i)) {
          const double sx = T0 * P.L[0] + T1 * P.L[3] + T2 * P.L[6];
          const double sy = T0 * P.L[1] + T1 * P.L[4] + T2 * P.L[7];
          const double sz = T0 * P.L[2] + T1 * P.L[5] + T2 * P.L[8];
          dx = pxy.x + sx - xi;
          dy = pxy.y + sy - yi;
          dz = pz.x + sz - zi;
          r = sqrt(dx * dx + dy * dy + dz * dz);
          ok = !(r > rcut);
        }
      }
      const unsigned long long m = __ballot(ok);
      const int pos = count + __popcll(m & ((1ull << lane) - 1ull));
      if (ok && pos < nb.kcap) {
        const size_t sl = base + pos;
        nb.nb_j[sl] = inf.x;
        nb.nb_rev[sl] = -1;  // k_reverse fills it from whichever side leads the pair
        if (nb.key32)
          ((uint32_t *)nb.nb_key)[sl] = nb_key32(inf.x, T0 - inf.y + si0, T1 - inf.z + si1, T2 - inf.w + si2);
        else
          nb.nb_key[sl] = nb_key(inf.x, T0 - inf.y + si0, T1 - inf.z + si1, T2 - inf.w + si2);
        double2 *d2 = (double2 *)(nb.nb_d + 4 * sl);
        d2[0] = make_double2(dx, dy);
        d2[1] = make_double2(dz, r);
        if (TRIP) s_p[pos] = trip_record(dx, dy, dz, r, trip_rc(rcA));
      }
      count += __popcll(m);
    }
  }
  if (lane == 0) {
    nb.nb_count[i] = min(count, nb.kcap);
    nb.nb_raw[i] = count;
  }
  if (TRIP) {
    __syncthreads();
    triplets_from_lds(i, min(count, nb.kcap), s_p, rcA, tb, lane);
  }
}

// Triplet lists of slots filled by the caller (CAR2ACSF's neighbour table):
// one wavefront per atom stages the slots' (dx, dy, dz, r) in LDS as FP32 and
// builds the list as k_build<true> does.
__global__ __launch_bounds__(64) void k_triplets_from_slots(int N, NeighborBufs nb, double rcA, TripletBufs tb) {
  extern __shared__ __align__(16) unsigned char smem[];
  float4 *s_p = (float4 *)smem;  // [trip_lds_records(kcap)]
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= N) return;
  const int K = min(nb.nb_count[i], nb.kcap);
  const size_t base = (size_t)i * nb.kcap;
  for (int a = lane; a < K; a += 64) {
    const double *d = nb.nb_d + 4 * (base + a);
    s_p[a] = trip_record(d[0], d[1], d[2], d[3], trip_rc(rcA));
  }
  __syncthreads();
  triplets_from_lds(i, K, s_p, rcA, tb, lane);
}

// |pos_i - (pos_j + n1 L1 + n2 L2 + n3 L3)| in the order of operations of
// FGET_BOND (gappy/libgap/get_bond.f90:58-60): each product rounded, the
// sums left to right, no fused multiply-adds.
__device__ __forceinline__ double bond_distance(const double *pi, const double *pj, int n0, int n1, int n2,
                                                const double *L) {
#pragma clang fp contract(off)
  double dr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double xyz = ((pj[c] + (double)n0 * L[c]) + (double)n1 * L[3 + c]) + (double)n2 * L[6 + c];
    dr[c] = pi[c] - xyz;
  }
  return sqrt((dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2]);
}

// FGET_BOND over a cell list (structure 0 of sp): one wavefront per atom
// walks the bin images as k_build does; a candidate within rsel of the
// wrapped positions gets its image index n relative to the positions as given
// and, when |n_k| <= nabc_k (the reference's image range), the reference's
// distance, which enters the minimum when it is at most rcut.  The minimum
// is taken over the wave (shuffles), over the workgroup's four waves (LDS),
// and across workgroups by one atomicMin on the bit pattern (positive
// doubles order like their bits), so the result does not depend on order.
__global__ __launch_bounds__(256) void k_min_bond(int N, const StructParams *__restrict__ sp,
                                                  const double *__restrict__ pos, NeighborBufs nb, double rsel,
                                                  double rcut, int na0, int na1, int na2,
                                                  unsigned long long *__restrict__ out) {
  __shared__ double s_best[4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wv;
  double best = 1e300;
  if (i < N) {
    const StructParams &P = sp[0];
    const int b = nb.atom_bin[i] - P.bin_base;
    const int bi2 = b % P.nb[2], bi1 = (b / P.nb[2]) % P.nb[1], bi0 = b / (P.nb[1] * P.nb[2]);
    const double xi = nb.wpos[3 * i], yi = nb.wpos[3 * i + 1], zi = nb.wpos[3 * i + 2];
    const int si0 = nb.shift[3 * i], si1 = nb.shift[3 * i + 1], si2 = nb.shift[3 * i + 2];
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const int m1 = 2 * P.S[1] + 1, m2 = 2 * P.S[2] + 1, nimg = (2 * P.S[0] + 1) * m1 * m2;
    const float inv_m1 = inv_small(m1), inv_m2 = inv_small(m2);
    const float inv_n0 = inv_small(P.nb[0]), inv_n1 = inv_small(P.nb[1]), inv_n2 = inv_small(P.nb[2]);
    const double2 *sp2 = (const double2 *)nb.spos;
    const int4 *sinfo = (const int4 *)nb.sinfo;
    for (int g0 = 0; g0 < nimg; g0 += 64) {
      const int img = g0 + lane;
      int lo = 0, n = 0, t0 = 0, t1 = 0, t2 = 0;
      if (img < nimg) {
        const int q2 = floordiv_small(img, inv_m2), q1 = floordiv_small(q2, inv_m1);
        const int o2 = img - q2 * m2 - P.S[2], o1 = q2 - q1 * m1 - P.S[1], o0 = q1 - P.S[0];
        const int bb0 = bi0 + o0, bb1 = bi1 + o1, bb2 = bi2 + o2;
        t0 = floordiv_small(bb0, inv_n0);
        t1 = floordiv_small(bb1, inv_n1);
        t2 = floordiv_small(bb2, inv_n2);
        const int w0 = bb0 - t0 * P.nb[0], w1 = bb1 - t1 * P.nb[1], w2 = bb2 - t2 * P.nb[2];
        const int bin = P.bin_base + (w0 * P.nb[1] + w1) * P.nb[2] + w2;
        lo = nb.bin_start[bin];
        n = nb.bin_start[bin + 1] - lo;
      }
      int incl = n;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      const int total = __shfl(incl, 63, 64);
      const int first = lo - (incl - n);
      for (int c0 = 0; c0 < total; c0 += 64) {
        const int c = c0 + lane;
        int src = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
          if (__shfl(incl, src + step - 1, 64) <= c) src += step;
        const int q = __shfl(first, src, 64) + c;
        const int T0 = __shfl(t0, src, 64), T1 = __shfl(t1, src, 64), T2 = __shfl(t2, src, 64);
        if (c < total) {
          const int4 inf = sinfo[q];
          if (!(T0 == 0 && T1 == 0 && T2 == 0 && inf.x == i)) {
            const double2 pxy = sp2[2 * (size_t)q], pz = sp2[2 * (size_t)q + 1];
            const double sx = T0 * P.L[0] + T1 * P.L[3] + T2 * P.L[6];
            const double sy = T0 * P.L[1] + T1 * P.L[4] + T2 * P.L[7];
            const double sz = T0 * P.L[2] + T1 * P.L[5] + T2 * P.L[8];
            const double dx = pxy.x + sx - xi, dy = pxy.y + sy - yi, dz = pz.x + sz - zi;
            const int n0 = T0 - inf.y + si0, n1 = T1 - inf.z + si1, n2 = T2 - inf.w + si2;
            if (dx * dx + dy * dy + dz * dz <= rsel * rsel && abs(n0) <= na0 && abs(n1) <= na1 && abs(n2) <= na2) {
              const double pj[3] = {pos[3 * inf.x], pos[3 * inf.x + 1], pos[3 * inf.x + 2]};
              const double dis = bond_distance(pi, pj, n0, n1, n2, P.L);
              if (!(dis > rcut)) best = fmin(best, dis);
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o, 64));
  if (lane == 0) s_best[wv] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = fmin(fmin(s_best[0], s_best[1]), fmin(s_best[2], s_best[3]));
    if (m < 1e300) atomicMin(out, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// For slot (i -> j, n) find the slot (j -> i, -n), and the other way round;
// the slots start at -1 (k_build writes it), which stays where the partner is
// absent (a pair within one rounding step of rcut can be seen from one side
// only).  Each pair is searched once, from its leading side (i < j, or i == j
// with n lexicographically positive), which writes both entries: no other
// lane writes either.  One wavefront per centre atom, one slot per lane; j's
// keys (64-bit, or 32-bit with K32: half the bytes read) are read kRevBatch
// at a time (independent 16-byte loads) and the first match wins.  nb_key
// carries kRevBatch keys of padding past the last atom's slots, so the
// batched reads stay in bounds; entries at or beyond j's count are never
// compared.
template <bool K32>
__global__ __launch_bounds__(64) void k_reverse(int N, NeighborBufs nb) {
  typedef typename std::conditional<K32, uint32_t, unsigned long long>::type Key;
  constexpr int PER = 16 / (int)sizeof(Key);  // keys per 16-byte load
  const Key *keys = (const Key *)nb.nb_key;
  const int i = blockIdx.x;
  if (i >= N) return;
  const int K = nb.nb_count[i];
  const size_t base = (size_t)i * nb.kcap;
  for (int a = (int)threadIdx.x; a < K; a += 64) {
    const Key key = keys[base + a];
    int j, n0, n1, n2;
    if constexpr (K32) {
      j = (int)(key >> 15);
      n0 = sext5(key >> 10), n1 = sext5(key >> 5), n2 = sext5(key);
    } else {
      j = (int)(key >> 24);
      n0 = (int)((key >> 16) & 255) - 128, n1 = (int)((key >> 8) & 255) - 128, n2 = (int)(key & 255) - 128;
    }
    const bool lead = i < j || (i == j && (n0 > 0 || (n0 == 0 && (n1 > 0 || (n1 == 0 && n2 > 0)))));
    if (!lead) continue;
    Key want;
    if constexpr (K32)
      want = nb_key32(i, -n0, -n1, -n2);
    else
      want = nb_key(i, -n0, -n1, -n2);
    const size_t jb = (size_t)j * nb.kcap;
    const uint4 *kj = (const uint4 *)(keys + jb);  // kcap is a multiple of 8: 16-byte aligned rows
    const int Kj = nb.nb_count[j];
    int found = -1;
    for (int q0 = 0; q0 < Kj && found < 0; q0 += kRevBatch) {
      Key k[kRevBatch];
#pragma unroll
      for (int u = 0; u < kRevBatch / PER; ++u) {
        const uint4 v = kj[q0 / PER + u];
        __builtin_memcpy(k + PER * u, &v, 16);
      }
#pragma unroll
      for (int u = kRevBatch - 1; u >= 0; --u)
        if (q0 + u < Kj && k[u] == want) found = q0 + u;
    }
    if (found >= 0) {
      nb.nb_rev[base + a] = (int)(jb + found);
      nb.nb_rev[jb + found] = (int)(base + a);
    }
  }
}

template <typename V>
__device__ __forceinline__ V wave_reduce_sum(V v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Everything after the backward pass in one launch of 256-thread blocks
// (P = rb.nparts structure parts, T = ceil(N / kReduceAtoms) statistics parts):
//   blocks [0, P)          energy and virial of one part of a structure (at
//                          most kReduceAtoms atoms, four per thread, their
//                          loads in flight together); a structure in one part
//                          writes its energy and stress in GPa directly
//                          (gap_calc.f90:152-154, 189-203), else the last part
//                          to finish adds the parts' sums in part order;
//   blocks [P, P + T)      the list statistics of kReduceAtoms atoms each,
//                          folded with integer atomics; the last block writes
//                          them out and clears the accumulators;
//   blocks from Q = P + T rounded up to a multiple of 8 (the blocks in
//                          between are empty): forces, one wavefront per atom,
//                          four atoms per group g < G, the groups laid out so
//                          that each XCD takes a contiguous range of atoms
//                          (xcd_unit: a structure's reverse slots are read
//                          from one L2): F_m = sum of own slots - sum of
//                          reverse slots in slot order (zeros without grad).
// The sums go first so that they run beside the gather blocks.
template <typename SF>
__global__ __launch_bounds__(256) void k_forces_reduce(int N, int G, int nstruct, NeighborBufs nb,
                                                       const SF *__restrict__ slot_f, double *__restrict__ force,
                                                       const StructParams *__restrict__ sp,
                                                       const double *__restrict__ eatom,
                                                       const double *__restrict__ vir, int grad,
                                                       double *__restrict__ ene, double *__restrict__ stress6,
                                                       TripletBufs tb, ReduceBufs rb, int *stats_out) {
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  constexpr int U = kReduceAtoms / 256;
  const int T = (N + kReduceAtoms - 1) / kReduceAtoms;
  __shared__ int s_last;
  if (b < rb.nparts) {
    const int2 pm = rb.part_map[b];
    const int s = pm.x, p = pm.y;
    if (s >= nstruct) return;
    const StructParams &P = sp[s];
    const int parts = (P.natoms + kReduceAtoms - 1) / kReduceAtoms;
    const int lo = p * kReduceAtoms, hi = min(P.natoms, lo + kReduceAtoms);
    __shared__ double red[4][7];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    double e[U], v[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = P.atom_base + min(lo + t + 256 * u, hi - 1);
      e[u] = eatom[i];
#pragma unroll
      for (int c = 0; c < 6; ++c) v[u][c] = grad ? vir[6 * i + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lo + t + 256 * u < hi) {
        acc[0] += e[u];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[1 + c] += v[u][c];
      }
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      const double w = wave_reduce_sum(acc[c]);
      if (lane == 0) red[wv][c] = w;
    }
    __syncthreads();
    double tot = 0.0;
    if (t < 7) tot = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    if (parts > 1) {
      if (t < 7) rb.part_sum[8 * (size_t)b + t] = tot;
      __threadfence();
      __syncthreads();
      if (t == 0) s_last = (atomicAdd(&rb.ticket[s], 1) == parts - 1);
      __syncthreads();
      if (!s_last) return;
      __threadfence();
      if (t < 7) {
        const size_t first = (size_t)(b - p);  // the structure's first part block
        tot = split_sum(parts, [&](int q) {
          return __hip_atomic_load(&rb.part_sum[8 * (first + q) + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        });
      }
      if (t == 0) rb.ticket[s] = 0;
    }
    if (t < 7) {
      if (t == 0)
        ene[s] = tot;
      else
        stress6[6 * s + t - 1] = grad ? tot * (-(1.0 / kGPa) / P.vol) : 0.0;
    }
    return;
  }
  if (b < rb.nparts + T) {
    const int lo = (b - rb.nparts) * kReduceAtoms;
    __shared__ long long red[4][6];
    long long v[6] = {0, 0, 0, 0, 0, 0};
    int r[U], c[U], tr[U], tc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int a = min(lo + t + 256 * u, N - 1);
      r[u] = nb.nb_raw[a];
      c[u] = nb.nb_count[a];
      tr[u] = tb.raw ? tb.raw[a] : 0;
      tc[u] = tb.raw ? tb.count[a] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lo + t + 256 * u < N) {
        v[0] += c[u];
        v[1] = max(v[1], (long long)r[u]);
        v[2] += (r[u] > nb.kcap) ? 1 : 0;
        v[3] += tc[u];
        v[4] = max(v[4], (long long)tr[u]);
        v[5] += (tr[u] > tb.tcap) ? 1 : 0;
      }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const long long w = __shfl_xor(v[k], o, 64);
        v[k] = (k == 1 || k == 4) ? max(v[k], w) : v[k] + w;
      }
    }
    if (lane == 0)
      for (int k = 0; k < 6; ++k) red[wv][k] = v[k];
    __syncthreads();
    if (t < 6) {
      long long a = 0;
      for (int w = 0; w < 4; ++w) a = (t == 1 || t == 4) ? max(a, red[w][t]) : a + red[w][t];
      int *dst = &rb.stats_acc[t < 3 ? t : t + 1];  // the output layout, [3] and [7] left out
      if (t == 1 || t == 4)
        atomicMax(dst, (int)a);
      else
        atomicAdd(dst, (int)a);
    }
    __threadfence();
    __syncthreads();
    if (t == 0) s_last = (atomicAdd(&rb.stats_acc[7], 1) == T - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    int val = 0;
    if (t < 8 && t != 3 && t != 7) val = __hip_atomic_load(&rb.stats_acc[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t == 3) val = nb.total[3];
    if (t < 8) stats_out[t] = val;
    if (rb.sticky && t < 8) rb.sticky[t] = max(rb.sticky[t], val);
    if (rb.overflow_now) {  // lanes 2, 3 and 6 hold the overflow counts
      const unsigned long long m = __ballot((t == 2 || t == 3 || t == 6) && val != 0);
      if (t == 0) *rb.overflow_now = (m != 0ull) ? 1 : 0;
    }
    __syncthreads();
    if (t < 8) rb.stats_acc[t] = 0;
    return;
  }
  const int Q = xcd_blocks(rb.nparts + T);  // the gather blocks start at an 8-aligned index
  if (b < Q) return;
  const int g = xcd_unit(b - Q, xcd_blocks(G)), m = 4 * g + wv;
  if (g >= G || m >= N) return;
  double f0 = 0, f1 = 0, f2 = 0;
  if (grad) {
    // the atom's own slots and reverse indices are read over its whole
    // capacity, in flight together with its count (slots at or beyond the
    // count hold stale values and are not added), so the gather waits for
    // memory twice: own slots, then the reverse slots they name
    const int K = nb.nb_count[m];
    const size_t base = (size_t)m * nb.kcap;
    for (int a = lane; a < nb.kcap; a += 64) {
      const size_t sl = base + a;
      const double o0 = (double)slot_f[3 * sl], o1 = (double)slot_f[3 * sl + 1], o2 = (double)slot_f[3 * sl + 2];
      const int rv = nb.nb_rev[sl];
      // an empty asm that takes the loaded values: the compiler must issue
      // the loads here, before the comparison with K, instead of moving
      // them under it (no instruction is emitted)
      asm volatile("" ::"v"(o0), "v"(o1), "v"(o2), "v"(rv));
      const bool live = a < K;
      const int r = live ? rv : -1;
      if (live) {
        f0 += o0;
        f1 += o1;
        f2 += o2;
      }
      if (r >= 0) {
        f0 -= (double)slot_f[3 * (size_t)r];
        f1 -= (double)slot_f[3 * (size_t)r + 1];
        f2 -= (double)slot_f[3 * (size_t)r + 2];
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      f0 += __shfl_xor(f0, o, 64);
      f1 += __shfl_xor(f1, o, 64);
      f2 += __shfl_xor(f2, o, 64);
    }
  }
  if (lane == 0) {
    force[3 * m] = f0;
    force[3 * m + 1] = f1;
    force[3 * m + 2] = f2;
  }
}

}  // namespace

hipError_t launch_fill_int(hipStream_t s, int *p, int n, int v) {
  if (n <= 0) return hipSuccess;
  k_fill_int<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_zero_ints(hipStream_t s, int *a, int na, int *b, int nb, int *c, int nc) {
  const int n = std::max(na, std::max(nb, nc));
  if (n <= 0) return hipSuccess;
  k_zero3<<<(n + 255) / 256, 256, 0, s>>>(a, na, b, nb, c, nc);
  return hipGetLastError();
}


hipError_t launch_fill_double(hipStream_t s, double *p, size_t n, double v) {
  if (n == 0) return hipSuccess;
  k_fill_double<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_wrap_scan(hipStream_t s, int N, int nbins, const double *pos, const int *atom_struct,
                            const StructParams *sp, NeighborBufs nb, int *ticket, int *stats) {
  k_wrap_scan<<<(N + 255) / 256, 256, 0, s>>>(N, nbins, pos, atom_struct, sp, nb, ticket, stats);
  return hipGetLastError();
}

hipError_t launch_scatter_bins(hipStream_t s, int N, int nbins, NeighborBufs nb) {
  k_scatter<<<(N + 255) / 256, 256, 0, s>>>(N, nb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_rank_bins<<<nbins, 64, 0, s>>>(nbins, nb);
  return hipGetLastError();
}

size_t fused_bin_lds_bytes(int atoms_max, int bins_max) {
  return sizeof(int) * ((size_t)bins_max + 1 + (size_t)std::max(atoms_max, 1024));
}

hipError_t launch_bin_structures(hipStream_t s, int nstruct, const double *pos, const StructParams *sp,
                                 NeighborBufs nb, int nbins_total, int atoms_max, int bins_max, int *stats) {
  const size_t lds = fused_bin_lds_bytes(atoms_max, bins_max);
  if (atoms_max > kFusedAtoms || lds > kFusedLdsBytes) return hipErrorInvalidValue;
  k_bin_struct<<<nstruct, 1024, lds, s>>>(pos, sp, nb, nbins_total, bins_max, stats);
  return hipGetLastError();
}

hipError_t launch_build_neighbors(hipStream_t s, int N, const int *atom_struct, const StructParams *sp,
                                  double rcut, NeighborBufs nb, double rcA, const TripletBufs *tb) {
  if (tb)
    k_build<true><<<N, 64, sizeof(float4) * trip_lds_records(nb.kcap), s>>>(N, atom_struct, sp, rcut, nb, rcA, *tb);
  else
    k_build<false><<<N, 64, 0, s>>>(N, atom_struct, sp, rcut, nb, 0.0, TripletBufs{});
  return hipGetLastError();
}

hipError_t launch_triplets_from_slots(hipStream_t s, int N, NeighborBufs nb, double rcA, TripletBufs tb) {
  if (N <= 0) return hipSuccess;
  if (nb.kcap < 1 || nb.kcap > kMaxNeighbors) return hipErrorInvalidValue;
  k_triplets_from_slots<<<N, 64, sizeof(float4) * trip_lds_records(nb.kcap), s>>>(N, nb, rcA, tb);
  return hipGetLastError();
}

hipError_t launch_min_bond(hipStream_t s, int N, const StructParams *sp, const double *pos, NeighborBufs nb,
                           double rsel, double rcut, const int nabc[3], unsigned long long *out,
                           unsigned long long init) {
  if (N <= 0) return hipErrorInvalidValue;
  k_set_u64<<<1, 64, 0, s>>>(out, init);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_min_bond<<<(N + 3) / 4, 256, 0, s>>>(N, sp, pos, nb, rsel, rcut, nabc[0], nabc[1], nabc[2], out);
  return hipGetLastError();
}

hipError_t launch_reverse_slots(hipStream_t s, int N, NeighborBufs nb) {
  if (nb.kcap % 8 != 0) return hipErrorInvalidValue;  // 16-byte aligned key rows
  if (nb.key32)
    k_reverse<true><<<N, 64, 0, s>>>(N, nb);
  else
    k_reverse<false><<<N, 64, 0, s>>>(N, nb);
  return hipGetLastError();
}

hipError_t launch_forces_reduce(hipStream_t s, int N, int grad, const NeighborBufs &nb, const void *slot_f,
                                bool slot32, double *force, int nstruct, const StructParams *sp,
                                const double *eatom, const double *atom_vir, double *ene, double *stress6,
                                const TripletBufs &tb, const ReduceBufs &rb, int *stats_out) {
  if (N <= 0 || nstruct <= 0 || rb.nparts < nstruct) return hipErrorInvalidValue;
  const int G = (N + 3) / 4, T = (N + kReduceAtoms - 1) / kReduceAtoms;
  const dim3 grid(xcd_blocks(rb.nparts + T) + xcd_blocks(G));
  if (slot32)
    k_forces_reduce<float><<<grid, 256, 0, s>>>(N, G, nstruct, nb, (const float *)slot_f, force, sp, eatom,
                                                atom_vir, grad, ene, stress6, tb, rb, stats_out);
  else
    k_forces_reduce<double><<<grid, 256, 0, s>>>(N, G, nstruct, nb, (const double *)slot_f, force, sp, eatom,
                                                 atom_vir, grad, ene, stress6, tb, rb, stats_out);
  return hipGetLastError();
}

}  // namespace gapx
