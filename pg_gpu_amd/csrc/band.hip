// Tiled int8-MFMA path for banded LD sums.
//
// Counts for a 64 x 64 tile of SNP pairs are one dense contraction over the
// samples of a population: with 0/1 indicator planes (haplotypes: h, v;
// genotypes: v, e1, e2) stored SNP-major and sample-contiguous, every product
// of an A-side plane (SNPs i) and a B-side plane (SNPs j) is an int8 GEMM with
// exact int32 accumulation (v_mfma_i32_32x32x32_i8).  The joint counts follow
// by inclusion-exclusion (e.g. n10 = h_i.v_j - h_i.h_j), as in the reference's
// own GEMM form _tile_counts (ld_statistics.py:281-298).  Without missing data
// only the allele planes enter the MFMA; the valid-plane products are the
// per-SNP plane sums and the sample count.
//
// Each workgroup of 4 waves owns one tile; wave w owns the 32 x 32 sub-tile
// (rows 32*(w/2), cols 32*(w%2)).  Populations are processed one after the
// other; all but the last park their counts in LDS as int16.  The epilogue then
// evaluates the statistics of each band pair in FP64 from exact counts, keys
// the pair (window of the left SNP, or distance bin), and reduces per key in a
// fixed order: per-lane running sums, wave butterflies at key changes, one LDS
// row per wave, a fixed-order sum over the 4 waves into a per-tile partial
// row, and a fixed-order per-key reduction over tiles.  Results are therefore
// identical from run to run.
#include <algorithm>
#include <array>
#include <utility>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <sys/stat.h>
#include <unistd.h>
#include <string>
#include <type_traits>
#include <vector>

#include "band_core.h"
#include "ld_math.h"

namespace pgld {
#include "epilogue.inc"
}

namespace pgld {

template <int FAMILY, int P, int SET>
struct StatSet;
#define PGLD_STATSETS(P)                                                                           \
  template <> struct StatSet<PGLD_FAMILY_HAP, P, PGLD_SET_MOMENTS> {                               \
    static constexpr int N = EPI_HAP_MOMENTS_##P##_N;                                              \
    __device__ static void eval(const HapP* v, double* out) { epi_hap_moments_##P(v, out); }       \
    __device__ static double scale(int c) { return epi_hap_moments_##P##_scale(c); }               \
  };                                                                                               \
  template <> struct StatSet<PGLD_FAMILY_HAP, P, PGLD_SET_SELF> {                                  \
    static constexpr int N = EPI_HAP_SELF_##P##_N;                                                 \
    __device__ static void eval(const HapP* v, double* out) { epi_hap_self_##P(v, out); }          \
    __device__ static double scale(int c) { return epi_hap_self_##P##_scale(c); }                  \
  };                                                                                               \
  template <> struct StatSet<PGLD_FAMILY_GENO, P, PGLD_SET_MOMENTS> {                              \
    static constexpr int N = EPI_GENO_MOMENTS_##P##_N;                                             \
    __device__ static void eval(const GenoP* v, double* out) { epi_geno_moments_##P(v, out); }     \
    __device__ static double scale(int c) { return epi_geno_moments_##P##_scale(c); }              \
  };                                                                                               \
  template <> struct StatSet<PGLD_FAMILY_GENO, P, PGLD_SET_SELF> {                                 \
    static constexpr int N = EPI_GENO_SELF_##P##_N;                                                \
    __device__ static void eval(const GenoP* v, double* out) { epi_geno_self_##P(v, out); }        \
    __device__ static double scale(int c) { return epi_geno_self_##P##_scale(c); }                 \
  };
PGLD_STATSETS(1)
PGLD_STATSETS(2)
PGLD_STATSETS(3)
PGLD_STATSETS(4)
#undef PGLD_STATSETS

// The fused kernel sums a statistic set in groups of consecutive columns,
// one pass over the sub-tile's pairs per group, so that a group's running
// sums stay in registers: one group except for the 105 four-population
// haplotype moments (three groups of 35).
template <int FAMILY, int P, int SET>
struct StatGroups {
  using SS = StatSet<FAMILY, P, SET>;
  static constexpr int NG = 1;
  template <int G>
  static constexpr int off() { return 0; }
  template <int G>
  static constexpr int count() { return SS::N; }
  template <int G, typename V>
  __device__ static void eval(const V* v, double* out) { SS::eval(v, out); }
};
template <>
struct StatGroups<PGLD_FAMILY_HAP, 4, PGLD_SET_MOMENTS> {
  static constexpr int NG = EPI_HAP_MOMENTS_4_NG;
  static_assert(NG == 3, "three generated groups");
  template <int G>
  static constexpr int off() { return EPI_HAP_MOMENTS_4_GOFF[G]; }
  template <int G>
  static constexpr int count() { return EPI_HAP_MOMENTS_4_GOFF[G + 1] - EPI_HAP_MOMENTS_4_GOFF[G]; }
  template <int G>
  __device__ static void eval(const HapP* v, double* out) {
    if constexpr (G == 0)
      epi_hap_moments_4_g0(v, out);
    else if constexpr (G == 1)
      epi_hap_moments_4_g1(v, out);
    else
      epi_hap_moments_4_g2(v, out);
  }
};

template <>
struct StatGroups<PGLD_FAMILY_GENO, 4, PGLD_SET_MOMENTS> {
  static constexpr int NG = EPI_GENO_MOMENTS_4_NG;
  static_assert(NG == 3, "three generated groups");
  template <int G>
  static constexpr int off() { return EPI_GENO_MOMENTS_4_GOFF[G]; }
  template <int G>
  static constexpr int count() { return EPI_GENO_MOMENTS_4_GOFF[G + 1] - EPI_GENO_MOMENTS_4_GOFF[G]; }
  template <int G>
  __device__ static void eval(const GenoP* v, double* out) {
    if constexpr (G == 0)
      epi_geno_moments_4_g0(v, out);
    else if constexpr (G == 1)
      epi_geno_moments_4_g1(v, out);
    else
      epi_geno_moments_4_g2(v, out);
  }
};

// The joint genotype sets of three or four populations park the nine counts
// of their populations, 72 KiB per population and workgroup: more than LDS
// holds beside the operand rows, so they go to a global workspace (one slot
// per workgroup of a launch; launches of at most GSTASH_TILES tiles).  The
// last population is parked too, which frees its 144 accumulator registers
// for the statistics.
constexpr bool gstash_set(int family, int P, int set) {
  return family == PGLD_FAMILY_GENO && P >= 3 && set == PGLD_SET_MOMENTS;
}
constexpr int64_t GSTASH_TILES = 1024;
constexpr int64_t gstash_elems(int P) { return (int64_t)P * 4 * 9 * 16 * 64; }

template <typename F, int... G>
__device__ __forceinline__ void for_each_group(std::integer_sequence<int, G...>, F&& f) {
  (f(std::integral_constant<int, G>{}), ...);
}

template <int FAMILY>
struct PopVal;
template <>
struct PopVal<PGLD_FAMILY_HAP> {
  using T = HapP;
  __device__ static T make(const int* c) { return hap_pop_int(c[0], c[1], c[2], c[3]); }
};
template <>
struct PopVal<PGLD_FAMILY_GENO> {
  using T = GenoP;
  __device__ static T make(const int* c) { return geno_pop_int(c); }
};

// ---------------------------------------------------------------------------
// Fused band kernel: counts for all populations, statistics, keyed sums.
// ---------------------------------------------------------------------------
// Two workgroups per CU where the register budget allows it without spills, so
// the MFMA phase of one tile overlaps the FP64 epilogue of another.
constexpr int band_min_waves(int family, int P, int set) {
  return ((family == PGLD_FAMILY_HAP && (P <= 2 || set == PGLD_SET_SELF)) || (family == PGLD_FAMILY_GENO && P == 1))
             ? 2
             : 1;
}

// One tile of the fused band pass.  key_base moves the tile's local keys
// (distance bin or window) into the plan's key space: 0 for a single band,
// region * n_keys_local for the tiles of a region batch.  The pair sums are
// folded over the wave by the halving butterfly in groups of SCAT values
// (more values per group: fewer shuffles, more live registers).
#ifndef PGLD_HAP2_GROUP
#define PGLD_HAP2_GROUP 1
#endif
// Kernel variants (KV, bit-identical results; the batch autotuner chooses
// bits 0, 1 and 3): bit 0 set = each population's K loop on its own (no load
// stream across populations), bit 1 set = fast-path rows in a rolled loop,
// bit 2 set = the within-population haplotype sets with missing data read
// 1 / (n)_4 from an LDS table (BandArgs::rtab_*) instead of computing it per
// pair (set by the launcher when the table fits), bit 3 set = lean: each
// population's K loop on its own with one register stage, one fragment set
// and unpadded (swizzled) LDS operand rows, registers and LDS capped for
// three workgroups per CU (the extra waves hide the latencies the stages
// would).
constexpr int KV_PER_POP = 1, KV_ROLLED = 2, KV_RTAB = 4, KV_LEAN = 8;

// Fills the reciprocal table of a KV_RTAB launch; the workgroup barrier that
// follows the tile set-up orders it before the first read.
template <int KV>
__device__ __forceinline__ void fill_recip_table(unsigned char* smem, const BandArgs& ba) {
  if constexpr ((KV & KV_RTAB) != 0) {
    double* tab = reinterpret_cast<double*>(smem + ba.rtab_off);
    for (int x = threadIdx.x; x <= ba.rtab_n; x += BAND_THREADS) tab[x] = hap_self_recip((double)x);
  }
}

template <int FAMILY, bool MISSING, int P, int SET, int SCAT = 8, bool NIB = false, int KV = 0>
__device__ __forceinline__ void band_sums_tile(const BandArgs& ba, const PopArgs& pa, const BandTile& tile,
                                               int32_t key_base, unsigned char* smem) {
  using PL = Planes<FAMILY, MISSING>;
  using SS = StatSet<FAMILY, P, SET>;
  constexpr int NC = PL::NC;
  // haplotype within-population sets evaluate the moments straight from the
  // accumulators (hap_self_stats); their fast-path rows are unrolled
  constexpr bool HAP_SELF = FAMILY == PGLD_FAMILY_HAP && (SET == PGLD_SET_SELF || P == 1);
  // the two-population haplotype moments set evaluates from centred moments
  // (hap_two_pop_stats) rather than from the joint counts
  constexpr bool HAP2 = FAMILY == PGLD_FAMILY_HAP && P == 2 && SET == PGLD_SET_MOMENTS;
  constexpr bool FAST_UNROLL = (HAP_SELF || HAP2) && !(KV & KV_ROLLED);
  // rows per scheduling group of the unrolled fast path
  constexpr int FAST_GROUP = HAP2 ? PGLD_HAP2_GROUP : 2;
  // statistics of every pair are finite: integer numerators times reciprocals
  // that are 0 where a falling factorial is 0 (haplotype sets), or divided by
  // (n)_4 only where n >= 4 (one genotype population)
  constexpr bool FINITE_STATS = HAP_SELF || HAP2 || (FAMILY == PGLD_FAMILY_GENO && P == 1);
  static_assert(!(KV & KV_RTAB) || (HAP_SELF && MISSING), "reciprocal table: within-population sets, missing data");
  // 1 / (n)_4 of a pair with n valid samples (n <= the population's size <= rtab_n)
  const double* rtab = reinterpret_cast<const double*>(smem + ba.rtab_off);
  auto self_recip = [&](double n) -> double {
    if constexpr ((KV & KV_RTAB) != 0)
      return rtab[(int)n];
    else
      return hap_self_recip(n);
  };
  constexpr int NS = SS::N + 1;  // statistics + pair count
  // smem layout: [reduction rows: 4 * n_slots * NS doubles][stash: (P-1) * 4 * NC * 16 * 64 int16][operands]
  // (lean: [operands][reduction rows]; then the reciprocal table at rtab_off)
  static_assert(!(KV & KV_LEAN) || (SET == PGLD_SET_SELF && P > 1), "lean: within-population sets");
  __shared__ double s_pos_i[BT], s_pos_j[BT], s_kc_i[BT], s_kc_j[BT];
  __shared__ int s_wkey[BT];
  __shared__ double s_edges[MAX_LDS_EDGES];

  // the wave index (and so the sub-tile's offsets) is wave-uniform: scalar
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // The lean variant puts the operand buffer first, at a fixed LDS address
  // (its swizzled addresses then need no per-tile base), the reduction rows
  // after it.
  constexpr bool OP_FIRST = (KV & KV_LEAN) != 0;
  constexpr int OP_FIRST_BYTES = OP_FIRST ? opbuf_bytes_swz<PL::N>() : 0;
  double* red = reinterpret_cast<double*>(smem + OP_FIRST_BYTES);
  const int red_n = 4 * tile.n_slots * NS;
  constexpr bool GSTASH = gstash_set(FAMILY, P, SET);
  unsigned char* after_red = smem + OP_FIRST_BYTES + ((red_n * 8 + 15) / 16) * 16;
  int16_t* stash = GSTASH ? ba.gstash + (int64_t)blockIdx.x * gstash_elems(P) : reinterpret_cast<int16_t*>(after_red);

  // The tile's positions, key coordinates and bin edges are loaded into
  // registers, the population stream's first operand chunks right behind
  // them (both latencies overlap), and then copied to LDS.
  double v_pi = 0.0, v_pj = 0.0, v_ki = 0.0, v_kj = 0.0, v_e = 0.0;
  if (tid < BT) {
    const int64_t gi = min((int64_t)tile.i0 + tid, ba.n_var - 1), gj = min((int64_t)tile.j0 + tid, ba.n_var - 1);
    v_pi = ba.pos[gi];
    v_pj = ba.pos[gj];
    v_ki = ba.kc[gi];
    v_kj = ba.kc[gj];
  }
  static_assert(MAX_LDS_EDGES <= BAND_THREADS, "one edge per thread");
  const bool edges_in_lds = ba.key_mode != PGLD_KEY_WINDOW && ba.n_keys < MAX_LDS_EDGES;
  if (edges_in_lds && tid <= ba.n_keys) v_e = ba.edges[tid];
  // Every band pair of the tile has its bin in the tile's key range [klo, khi]
  // (the plan bounds it by the tile's smallest and largest distance) or none:
  // only those bins are searched, and a one-bin tile needs two compares.
  const int klo = tile.key_lo - key_base, khi = klo + tile.n_slots - 1;
  double e_lo = 0.0, e_hi = 0.0;
  if (ba.key_mode != PGLD_KEY_WINDOW) {
    e_lo = ba.edges[klo];
    e_hi = ba.edges[khi + 1];
  }
  // sets with several populations run their K loops as one load stream
  // across the populations (PopStream): every within-population set (except
  // the per-population kernel variant) and the haplotype moments sets
  using PS = PopStream<FAMILY, MISSING, NIB>;
  constexpr bool STREAM =
      P > 1 && (SET == PGLD_SET_SELF ? !(KV & (KV_PER_POP | KV_LEAN)) : FAMILY == PGLD_FAMILY_HAP);
  v4i pre0[STREAM ? PS::ITEMS : 1], pre1[STREAM ? PS::ITEMS : 1];
  bool pre1_ok = false;
  if constexpr (STREAM) PS::start(pa, ba.n_var, tile.i0, tile.j0, pre0, pre1, pre1_ok);
  for (int t = tid; t < red_n; t += BAND_THREADS) red[t] = 0.0;
  if (tid < BT) {
    s_pos_i[tid] = v_pi;
    s_pos_j[tid] = v_pj;
    s_kc_i[tid] = v_ki;
    s_kc_j[tid] = v_kj;
    if (ba.key_mode == PGLD_KEY_WINDOW) {
      const double wk = floor((v_ki - ba.origin) / ba.width);
      // wk in [0, n_keys): compared as an int once it is known to fit one
      s_wkey[tid] = (wk >= 0.0 && wk < 2147483648.0 && (int)wk < ba.n_keys) ? (int)wk : -1;
    }
  }
  if (edges_in_lds && tid <= ba.n_keys) s_edges[tid] = v_e;
  __syncthreads();

  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int ri = tile.i0 + wr, rj = tile.j0 + wc;
  // wave-uniform skip: sub-tile entirely on/below the diagonal, past the end,
  // or farther apart than max_dist
  const int64_t last_i = min((int64_t)ri + 31, ba.n_var - 1);
  bool wave_live = (int64_t)rj < ba.n_var && rj + 31 > ri && (int64_t)ri < ba.n_var &&
                   s_pos_j[wc] - s_pos_i[last_i - tile.i0] <= ba.max_dist;

  const int jl = wc + (lane & 31);
  const int64_t j = (int64_t)tile.j0 + jl;
  double* wred = red + w * tile.n_slots * NS;

  // Walks the wave's 32 x 32 sub-tile in accumulator order, evaluates M
  // statistics of every band pair with eval(r, i, out), keys the pair and keeps
  // per-lane running sums that are folded into this wave's LDS row at key
  // changes: statistic s goes to column col0 + s * cstride, the pair count (if
  // `count`) to the last column.  Pair keys can be kept between passes over
  // the same sub-tile: kmode 1 records them in kcode (one byte per row: 0 = no
  // key, else key - klo + 1), kmode 2 reads them back, kmode 0 neither.
  v4u kcode = {0u, 0u, 0u, 0u};
  // A sub-tile whose pairs are all valid (j > i everywhere, inside the panel,
  // the farthest pair within max_dist) and share one key (keys are monotone
  // in the left SNP's window and in the distance, so the extreme pairs decide)
  // needs no per-pair key or validity work.
  // A sub-tile whose pairs are all valid and fall in two consecutive keys
  // (duo_key, duo_key + 1) is summed into two running sets and flushed once
  // (sets with finite statistics; the others take the per-pair path):
  // a pair goes to the second key when its left SNP is in window duo_key + 1
  // (windows) or its key distance is at least edges[duo_key + 1], the edge
  // between the two bins (the bin search's own comparison).
  // Both cases are kept in one wave-uniform scalar through the K loops:
  // key_sel = 2 * key (one key) or 2 * key + 1 (two keys), -1 otherwise.
  int key_sel = -1;
  if (wave_live && rj > ri + 31 && (int64_t)rj + 31 < ba.n_var && s_pos_j[wc + 31] - s_pos_i[wr] <= ba.max_dist) {
    if (ba.key_mode == PGLD_KEY_WINDOW) {
      const int k0 = s_wkey[wr], k1 = s_wkey[wr + 31];
      if (k0 >= 0 && (k1 == k0 || k1 == k0 + 1)) key_sel = 2 * k0 + (k1 - k0);
    } else {
      const double d_lo = s_kc_j[wc] - s_kc_i[wr + 31], d_hi = s_kc_j[wc + 31] - s_kc_i[wr];
      if (d_lo >= e_lo && d_hi < e_hi) {
        const int b_lo = klo == khi ? klo
                                    : (edges_in_lds ? key_bin_in(d_lo, s_edges, klo, khi + 1)
                                                    : key_bin_in(d_lo, ba.edges, klo, khi + 1));
        const int b_hi = klo == khi ? klo
                                    : (edges_in_lds ? key_bin_in(d_hi, s_edges, klo, khi + 1)
                                                    : key_bin_in(d_hi, ba.edges, klo, khi + 1));
        if (b_hi - b_lo <= 1) key_sel = 2 * b_lo + (b_hi - b_lo);
      }
    }
  }
  key_sel = __builtin_amdgcn_readfirstlane(key_sel);  // every input above is wave-uniform
  auto epilogue = [&](auto m_tag, int col0, int cstride, bool count, int kmode, auto&& eval) {
    constexpr int M = decltype(m_tag)::value;
    double lacc[M + 1];
#pragma unroll
    for (int s = 0; s <= M; ++s) lacc[s] = 0.0;
    int ck = -1;
    auto flush = [&](bool flag) {
      unsigned long long part = __ballot(flag);
      while (part) {
        const int leader = __ffsll((long long)part) - 1;
        const int kk = __builtin_amdgcn_readlane(ck, leader);  // leader is wave-uniform
        const bool mine = flag && ck == kk;
        part &= ~__ballot(mine);
        const int slot = kk - tile.key_lo;
        const bool slot_ok = slot >= 0 && slot < tile.n_slots;
        double* row = wred + slot * NS;
        if (slot_ok) {
          double vals[M + 1];
#pragma unroll
          for (int s = 0; s <= M; ++s) vals[s] = mine ? lacc[s] : 0.0;
          if (count)
            wave_sums_scatter<M + 1, SCAT>(vals, lane, [&](int s, double x) {
              row[s < M ? col0 + s * cstride : NS - 1] += x;
            });
          else
            wave_sums_scatter<M, SCAT>(vals, lane, [&](int s, double x) { row[col0 + s * cstride] += x; });
        }
      }
      if (flag) {
#pragma unroll
        for (int s = 0; s <= M; ++s) lacc[s] = 0.0;
        ck = -1;
      }
    };
    if (key_sel >= 0 && !(key_sel & 1)) {
      // every lane holds 16 valid pairs of one key
      auto row = [&](int r) {
        const int64_t i = (int64_t)tile.i0 + wr + acc_row(r, lane);
        double out[M];
        eval(r, i, out);
#pragma unroll
        for (int s = 0; s < M; ++s) lacc[s] += out[s];
      };
      if constexpr (FAST_UNROLL) {
        // straight-line: accumulator elements addressed statically; rows are
        // scheduled two at a time (two independent FP64 chains interleave,
        // without the register growth of scheduling all 16 together)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          row(r);
          if ((r + 1) % FAST_GROUP == 0) {
            // the running sums are complete here (an empty asm operand keeps
            // the compiler from sinking the rows' last products to the end)
#pragma unroll
            for (int q = 0; q < M; ++q) asm volatile("" : "+v"(lacc[q]));
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      } else {
#pragma unroll 1
        for (int r = 0; r < 16; ++r) row(r);
      }
      // This wave's row for that key takes exactly one flush per epilogue
      // call and the calls write disjoint columns, so the totals are stored
      // without reading the row back (x + 0.0 is what adding x into the
      // cleared row gives).
      lacc[M] = 16.0;
      double* frow = wred + ((key_sel >> 1) - klo) * NS;
      if (count)
        wave_sums_scatter<M + 1, SCAT>(lacc, lane, [&](int s, double x) {
          frow[s < M ? col0 + s * cstride : NS - 1] = x + 0.0;
        });
      else
        wave_sums_scatter<M, SCAT>(lacc, lane, [&](int s, double x) { frow[col0 + s * cstride] = x + 0.0; });
      return;
    }
    if (FINITE_STATS && key_sel >= 0) {
      // every lane holds 16 valid pairs, each of key duo_key (running sums
      // lacc) or duo_key + 1 (lsec).  Keys are monotone in the left SNP, and
      // the lane's 16 rows ascend with r, so the second key's pairs are the
      // lane's last n_sec rows (windows) or its first n_sec rows (bins, the
      // larger distances): n_sec is counted once per sub-tile, with the
      // per-pair tests of the bin search and the window key, and kept in
      // kcode between passes.  The statistics are finite, so each pair is
      // weighted by 1 / 0 into both sets: fma(1, x, acc) = acc + x and
      // fma(0, x, acc) = acc.  (Sets whose statistics can be NaN take the
      // per-pair path.)
      const int duo_key = key_sel >> 1;
      const bool win = ba.key_mode == PGLD_KEY_WINDOW;
      int n_sec = 0;
      if (kmode == 2) {
        n_sec = (int)kcode[0];
      } else {
        if (win) {
#pragma unroll
          for (int r = 0; r < 16; ++r) n_sec += s_wkey[wr + acc_row(r, lane)] != duo_key ? 1 : 0;
        } else {
          const double edge = edges_in_lds ? s_edges[duo_key + 1] : ba.edges[duo_key + 1];
          const double kc_j = s_kc_j[jl];
#pragma unroll
          for (int r = 0; r < 16; ++r) n_sec += edge <= kc_j - s_kc_i[wr + acc_row(r, lane)] ? 1 : 0;
        }
        if (kmode == 1) kcode[0] = (uint32_t)n_sec;
      }
      // the second key's rows: r in [sec_lo, sec_lo + n_sec)
      const int sec_lo = win ? 16 - n_sec : 0;
      double lsec[M];
#pragma unroll
      for (int s = 0; s < M; ++s) lsec[s] = 0.0;
      auto row = [&](int r) {
        const int64_t i = (int64_t)tile.i0 + wr + acc_row(r, lane);
        const bool sec = (unsigned)(r - sec_lo) < (unsigned)n_sec;
        double out[M];
        eval(r, i, out);
        const double wb = sec ? 1.0 : 0.0, wa = sec ? 0.0 : 1.0;
#pragma unroll
        for (int s = 0; s < M; ++s) {
          lacc[s] = fma(wa, out[s], lacc[s]);
          lsec[s] = fma(wb, out[s], lsec[s]);
        }
      };
      if constexpr (FAST_UNROLL) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          row(r);
          if ((r + 1) % FAST_GROUP == 0) {
#pragma unroll
            for (int q = 0; q < M; ++q) asm volatile("" : "+v"(lacc[q]), "+v"(lsec[q]));
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      } else {
#pragma unroll 1
        for (int r = 0; r < 16; ++r) row(r);
      }
      // both key rows take exactly one flush from this wave per epilogue
      // call (as in the one-key case above), one row after the other
      double* frow = wred + (duo_key - klo) * NS;
      auto put = [&](const double* vals, double* krow) {
        if (count)
          wave_sums_scatter<M + 1, SCAT>(vals, lane, [&](int s, double x) {
            krow[s < M ? col0 + s * cstride : NS - 1] = x + 0.0;
          });
        else
          wave_sums_scatter<M, SCAT>(vals, lane, [&](int s, double x) { krow[col0 + s * cstride] = x + 0.0; });
      };
      lacc[M] = (double)(16 - n_sec);
      put(lacc, frow);
      double vs[M + 1];
#pragma unroll
      for (int s = 0; s < M; ++s) vs[s] = lsec[s];
      vs[M] = (double)n_sec;
      put(vs, frow + NS);
      return;
    }
#pragma unroll 1
    for (int r = 0; r < 16; ++r) {
      const int il = wr + acc_row(r, lane);
      const int64_t i = (int64_t)tile.i0 + il;
      int key = -1;
      if (kmode == 2) {
        const int code = (int)((kcode[r >> 2] >> (8 * (r & 3))) & 0xffu);
        if (code) key = key_base + klo + code - 1;
      } else {
        const bool valid = i < j && j < ba.n_var && (s_pos_j[jl] - s_pos_i[il]) <= ba.max_dist;
        if (valid) {
          if (ba.key_mode == PGLD_KEY_WINDOW) {
            key = s_wkey[il];
          } else {
            double d = s_kc_j[jl] - s_kc_i[il];
            if (ba.key_mode == PGLD_KEY_BIN_ABS) d = fabs(d);
            if (!(d >= e_lo) || !(d < e_hi))
              key = -1;
            else if (klo == khi)
              key = klo;
            else
              key = edges_in_lds ? key_bin_in(d, s_edges, klo, khi + 1) : key_bin_in(d, ba.edges, klo, khi + 1);
          }
          if (key >= 0) key += key_base;
        }
        if (kmode == 1 && key >= 0)
          kcode[r >> 2] |= (uint32_t)(key - key_base - klo + 1) << (8 * (r & 3));
      }
      const bool has = key >= 0;
      const bool need = has && ck >= 0 && key != ck;
      if (__ballot(need)) flush(need);
      if (has) {
        ck = key;
        double out[M];
        eval(r, i, out);
#pragma unroll
        for (int s = 0; s < M; ++s) lacc[s] += out[s];
        lacc[M] += 1.0;
      }
    }
    flush(ck >= 0);
  };

  WaveAcc<PL::N, NIB> acc;
  if constexpr (SET == PGLD_SET_SELF && P > 1) {
    // Independent per-population statistics (column s * P + p): each
    // population's products are followed at once by its own epilogue, so no
    // counts are parked in LDS and more workgroups fit on a CU.
    // The K chunks of all populations form one load stream (PopStream): the
    // next population's first operands are in flight during this one's
    // epilogue.
    using S1 = StatSet<FAMILY, 1, PGLD_SET_SELF>;
    unsigned char* opbuf = OP_FIRST ? smem : reinterpret_cast<unsigned char*>(stash);
#pragma unroll 1
    for (int p = 0; p < P; ++p) {
      if constexpr ((KV & KV_LEAN) != 0)
        block_products<FAMILY, MISSING, false, NIB, 1, false, true>(pa.planes[p], ba.n_var, pa.k_pad[p], tile.i0,
                                                                    tile.j0, wr, wc, wave_live, opbuf, acc,
                                                                    pa.n_samples[p]);
      else if constexpr ((KV & KV_PER_POP) != 0)
        block_products<FAMILY, MISSING, false, NIB>(pa.planes[p], ba.n_var, pa.k_pad[p], tile.i0, tile.j0, wr, wc,
                                                    wave_live, opbuf, acc);
      else
        PS::template products<P>(pa, p, ba.n_var, tile.i0, tile.j0, wr, wc, wave_live, opbuf, acc, pre0, pre1,
                                 pre1_ok);
      // the first population's pass records the pair keys, the others reuse them
      const int kmode = tile.n_slots > 255 ? 0 : (p == 0 ? 1 : 2);
      if (wave_live)
        epilogue(std::integral_constant<int, S1::N>{}, p, P, p == 0, kmode, [&](int r, int64_t i, double* out) {
          if constexpr (HAP_SELF) {
            double n, a, b, s;
            hap_moments<MISSING>(acc, r, pa.marg[p], i, j, pa.n_samples[p], n, a, b, s);
            hap_self_stats(n, a, b, s, self_recip(n), out);
          } else {
            int c[NC];
            elem_counts<FAMILY, MISSING>(acc, r, pa.marg[p], ba.n_var, i, j, pa.n_samples[p], c);
            typename PopVal<FAMILY>::T V = PopVal<FAMILY>::make(c);
            S1::eval(&V, out);
          }
        });
    }
  } else {
    unsigned char* opbuf = GSTASH ? after_red : after_red + (size_t)(P - 1) * 4 * NC * 16 * 64 * 2;
    // with the load stream the next population's first operand chunks are in
    // flight while this one's counts are parked
#pragma unroll 1
    for (int p = 0; p < P; ++p) {
      if constexpr (STREAM)
        PS::template products<P>(pa, p, ba.n_var, tile.i0, tile.j0, wr, wc, wave_live, opbuf, acc, pre0, pre1,
                                 pre1_ok);
      else
        block_products<FAMILY, MISSING, false, NIB>(pa.planes[p], ba.n_var, pa.k_pad[p], tile.i0, tile.j0, wr, wc,
                                                    wave_live, opbuf, acc);
      if ((p == P - 1 && !GSTASH) || !wave_live) continue;
      // all populations but the last park their counts in LDS as int16 (the
      // two-population haplotype set parks the centred moments n, Sa, Sb, Sab);
      // with the global workspace the last one too, so that its accumulators
      // are free during the statistics
      const int64_t jj = rj + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = ri + acc_row(r, lane);
        int c[NC];
        if constexpr (HAP2) {
          double n, a, b, s;
          hap_moments<MISSING>(acc, r, pa.marg[p], min(i, ba.n_var - 1), min(jj, ba.n_var - 1), pa.n_samples[p], n,
                               a, b, s);
          c[0] = (int)n;
          c[1] = (int)a;
          c[2] = (int)b;
          c[3] = (int)s;
        } else {
          elem_counts<FAMILY, MISSING>(acc, r, pa.marg[p], ba.n_var, min(i, ba.n_var - 1), min(jj, ba.n_var - 1),
                                       pa.n_samples[p], c);
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) stash[((((p * 4 + w) * NC + q) * 16 + r) << 6) + lane] = (int16_t)c[q];
      }
    }
    using SG = StatGroups<FAMILY, P, SET>;
    // with several groups the first group's pass records the pair keys and
    // the others reuse them
    if (wave_live)
      for_each_group(std::make_integer_sequence<int, SG::NG>{}, [&](auto g_tag) {
        constexpr int G = decltype(g_tag)::value;
        const int kmode = SG::NG == 1 || tile.n_slots > 255 ? 0 : (G == 0 ? 1 : 2);
        epilogue(std::integral_constant<int, SG::template count<G>()>{}, SG::template off<G>(), 1, G == 0, kmode,
                 [&](int r, int64_t i, double* out) {
          if constexpr (HAP_SELF && P == 1) {
            double n, a, b, s;
            hap_moments<MISSING>(acc, r, pa.marg[0], i, j, pa.n_samples[0], n, a, b, s);
            hap_self_stats(n, a, b, s, self_recip(n), out);
          } else if constexpr (HAP2) {
            double n[2], a[2], b[2], s[2];
            const int16_t* st = stash + (((w * NC) * 16 + r) << 6) + lane;
            n[0] = st[0];
            a[0] = st[16 * 64];
            b[0] = st[2 * 16 * 64];
            s[0] = st[3 * 16 * 64];
            hap_moments<MISSING>(acc, r, pa.marg[1], i, j, pa.n_samples[1], n[1], a[1], b[1], s[1]);
            hap_two_pop_stats(n, a, b, s, out);
          } else {
            typename PopVal<FAMILY>::T V[P];
#pragma unroll
            for (int p = 0; p < (GSTASH ? P : P - 1); ++p) {
              int c[NC];
#pragma unroll
              for (int q = 0; q < NC; ++q) c[q] = stash[((((p * 4 + w) * NC + q) * 16 + r) << 6) + lane];
              V[p] = PopVal<FAMILY>::make(c);
            }
            if constexpr (!GSTASH) {
              int c[NC];
              elem_counts<FAMILY, MISSING>(acc, r, pa.marg[P - 1], ba.n_var, i, j, pa.n_samples[P - 1], c);
              V[P - 1] = PopVal<FAMILY>::make(c);
            }
            SG::template eval<G>(V, out);
          }
        });
      });
  }
  __syncthreads();
  // fixed-order sum over the four wave rows into this tile's partial row,
  // times the statistic's power-of-two factor (left out of the per-pair values)
  const int row_n = tile.n_slots * NS;
  double* dst = ba.partials + tile.slot_base * NS;
  for (int t = tid; t < row_n; t += BAND_THREADS) {
    const int c = t % NS;
    double f = 1.0;
    if (c < NS - 1) {
      if constexpr (SET == PGLD_SET_SELF && P > 1)
        f = StatSet<FAMILY, 1, PGLD_SET_SELF>::scale(c / P);  // column s * P + p
      else
        f = SS::scale(c);
    }
    dst[t] = (((red[t] + red[row_n + t]) + red[2 * row_n + t]) + red[3 * row_n + t]) * f;
  }
}

template <int FAMILY, bool MISSING, int P, int SET, bool NIB, int KV = 0>
__global__ __launch_bounds__(BAND_THREADS, band_min_waves(FAMILY, P, SET)) void k_band_sums(BandArgs ba,
                                                                                            PopArgs pa) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int64_t t = ba.tile0 + xcd_tile(blockIdx.x, gridDim.x);
  const BandTile tile = ba.tiles[ba.order ? ba.order[t] : t];
  if (tile.n_slots <= 0) return;
  fill_recip_table<KV>(smem, ba);
  band_sums_tile<FAMILY, MISSING, P, SET, 8, NIB, KV>(ba, pa, tile, 0, smem);
}

// Region batch, persistent: workgroup b takes tiles order[b], order[b + grid],
// ... .  Each tile's region supplies its populations' planes, its SNP range
// (positions are concatenated over regions) and its key offset.  The per-key
// reduction runs over slots in schedule-independent order, so the result does
// not depend on the grid or the tile order.
// The persistent loop adds live state; genotype and 4-population joint
// (moments) instances take the whole register file (their LDS use allows one
// workgroup per CU).
constexpr int batch_min_waves(int family, int P, int set) {
  return (family == PGLD_FAMILY_HAP && (P <= 3 || set == PGLD_SET_SELF)) ? band_min_waves(family, P, set) : 1;
}

template <int FAMILY, bool MISSING, int P, int SET, bool NIB, int KV = 0>
__global__ __launch_bounds__(BAND_THREADS, (KV & KV_LEAN) ? 3 : batch_min_waves(FAMILY, P, SET)) void k_batch_sums(
    BandArgs ba, const RegionDesc* __restrict__ regions, const int32_t* __restrict__ order, int64_t n_tiles,
    int32_t xcd_split) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  // xcd_split: workgroup b runs on XCD b % 8 (round-robin dispatch); each XCD
  // walks its own contiguous eighth of the order so that neighbouring tiles
  // (same region, shared plane rows) meet in one L2.
  int64_t t0 = blockIdx.x, step = gridDim.x, t_end = n_tiles;
  if (xcd_split && gridDim.x % N_XCD == 0) {
    const int64_t x = blockIdx.x % N_XCD;
    t0 = n_tiles * x / N_XCD + blockIdx.x / N_XCD;
    step = gridDim.x / N_XCD;
    t_end = n_tiles * (x + 1) / N_XCD;
  }
  fill_recip_table<KV>(smem, ba);  // once per workgroup: the table outlives the tiles
  for (int64_t t = t0; t < t_end; t += step) {
    const int32_t ot = order[t];
    if (ot < 0) continue;  // padding of the XCD-blocked order
    const BandTile tile = ba.tiles[ot];
    if (tile.n_slots <= 0) continue;
    const RegionDesc& rd = regions[tile.region];
    BandArgs lb = ba;
    lb.pos = ba.pos + rd.snp_base;
    lb.kc = ba.kc + rd.snp_base;
    lb.n_var = rd.n_var;
    const int32_t key_base = rd.key_base;
    __syncthreads();  // the previous tile's LDS is no longer read
    // rd.pa stays in global memory: the population loop indexes it at run time
    // one value per butterfly for the statistic sets that would otherwise spill
    band_sums_tile<FAMILY, MISSING, P, SET, (SET == PGLD_SET_SELF ? 8 : 1), NIB, KV>(lb, rd.pa, tile, key_base,
                                                                                      smem);
  }
}

// Counts mode: one population, per-pair counts in the band's (i, j) order.
template <int FAMILY, bool MISSING, bool POPC>
__global__ __launch_bounds__(BAND_THREADS) void k_band_counts(BandArgs ba, PopArgs pa) {
  using PL = Planes<FAMILY, MISSING>;
  constexpr int NC = PL::NC;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const BandTile tile = ba.tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int ri = tile.i0 + wr, rj = tile.j0 + wc;
  const bool wave_live = (int64_t)rj < ba.n_var && (int64_t)ri < ba.n_var && rj + 31 > ri;
  WaveAcc<PL::N> acc;
  block_products<FAMILY, MISSING, POPC>(pa.planes[0], ba.n_var, pa.k_pad[0], tile.i0, tile.j0, wr, wc, wave_live,
                                        smem, acc);
  if (!wave_live) return;
  const int64_t j = rj + (lane & 31);
  if (j >= ba.n_var) return;
  const double pj = ba.pos[j];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t i = ri + acc_row(r, lane);
    if (i < j && pj - ba.pos[i] <= ba.max_dist) {
      int c[NC];
      elem_counts<FAMILY, MISSING>(acc, r, pa.marg[0], ba.n_var, i, j, pa.n_samples[0], c);
      int32_t* row = ba.counts_out + (ba.pair_base[i] + (j - i - 1)) * NC;
#pragma unroll
      for (int q = 0; q < NC; ++q) row[q] = c[q];
    }
  }
}

// Per-key reduction over tile partial rows, in ascending tile order.
// Fixed-order sums of the partial rows listed in key_list[lo, hi) (the rows
// lo .. hi - 1 themselves when key_list is null), every column of a row at
// once: thread t takes column t % cols of the slot group t / cols (groups
// stride over the rows, so the cols threads of a group read one row
// contiguously), and the group totals are added in group order.  out(s, v)
// receives column s.  The order depends only on the row list and ns_total.
template <class F>
__device__ __forceinline__ void fixed_order_rows(const double* __restrict__ partials,
                                                 const int64_t* __restrict__ key_list, int64_t lo, int64_t hi,
                                                 int32_t ns_total, double* buf, F&& out) {
  const int t = threadIdx.x;
  for (int c0 = 0; c0 < ns_total; c0 += 256) {
    const int cols = min(ns_total - c0, 256), groups = 256 / cols;
    const int g = t / cols, s = t - g * cols;
    double x = 0.0;
    if (g < groups) {
#pragma unroll 4
      for (int64_t e = lo + g; e < hi; e += groups) x += partials[(key_list ? key_list[e] : e) * ns_total + c0 + s];
    }
    __syncthreads();  // the previous column block has read buf
    if (g < groups) buf[g * cols + s] = x;
    __syncthreads();
    if (t < cols) {
      double v = 0.0;
      for (int q = 0; q < groups; ++q) v += buf[q * cols + t];
      out(c0 + t, v);
    }
  }
}

// First level of the two-level key reduction: chunk c adds its slots (entries
// [lohi[2c], lohi[2c + 1]) of the key list) into part2[c][s].
__global__ __launch_bounds__(256) void k_reduce_chunks(const double* __restrict__ partials,
                                                       const int64_t* __restrict__ key_list,
                                                       const int64_t* __restrict__ lohi, int32_t ns_total,
                                                       double* __restrict__ part2) {
  __shared__ double buf[256];
  const int64_t c = blockIdx.x;
  fixed_order_rows(partials, key_list, lohi[2 * c], lohi[2 * c + 1], ns_total, buf,
                   [&](int s, double v) { part2[c * ns_total + s] = v; });
}

// Per-key sums over the key's slots (key_list entries key_off[k] ..
// key_off[k + 1]; with key_list null the entries are the slots themselves).
// A key of one first-level chunk sums to the same value either way: its
// chunk total plus the zero totals of the empty groups.
__global__ __launch_bounds__(256) void k_band_reduce(const double* __restrict__ partials,
                                                     const int32_t* __restrict__ key_off,
                                                     const int64_t* __restrict__ key_list, int32_t ns_total,
                                                     int32_t n_stats, double* __restrict__ sums,
                                                     double* __restrict__ key_pairs) {
  __shared__ double buf[256];
  const int k = blockIdx.x;
  fixed_order_rows(partials, key_list, key_off[k], key_off[k + 1], ns_total, buf, [&](int s, double v) {
    if (s < n_stats) sums[(int64_t)k * n_stats + s] = v;
    else if (key_pairs) key_pairs[k] = v;
  });
}

// ---------------------------------------------------------------------------
// Plane packing: (rows x variants) int8 -> [plane][variant][k_pad] 0/1 bytes
// ---------------------------------------------------------------------------
// Tile of 128 samples x 128 variants per block.  Phase 1: every row segment is
// read as one full 128-byte line (32 lanes x dword) into LDS.  Phase 2: each
// thread takes 16 samples x 4 variants, turns four 4 x 4 byte blocks into
// variant-major dwords with byte permutes, derives the plane bytes and the
// domain / missing flags four samples at a time (SWAR on dwords) and writes
// 16 sample bytes per (plane, variant): 8 lanes fill one 128-byte line.  The
// per-variant plane counts are summed over those 8 lanes with shuffles (four
// variants' counts packed as bytes) and added to the margins by one lane.
constexpr int PK_S = 128, PK_V = 128;
constexpr int PK_Q = PK_V / 4;  // dword columns (variant quads) per sample row

// LDS column swizzle: sample row s keeps variant quad q at column q ^ pk_swz(s).
// Phase 2's lanes read rows 16 apart; the swizzle puts them on distinct banks.
__device__ __forceinline__ int pk_swz(int s) { return ((s >> 4) & 3) << 3; }

// 4 x 4 byte transpose: r[i] holds bytes (v0..v3) of sample i; on return
// c[j] holds bytes (s0..s3) of variant j.
__device__ __forceinline__ void transpose4x4(const uint32_t r[4], uint32_t c[4]) {
  const uint32_t t0 = __builtin_amdgcn_perm(r[1], r[0], 0x05010400u);
  const uint32_t t1 = __builtin_amdgcn_perm(r[1], r[0], 0x07030602u);
  const uint32_t t2 = __builtin_amdgcn_perm(r[3], r[2], 0x05010400u);
  const uint32_t t3 = __builtin_amdgcn_perm(r[3], r[2], 0x07030602u);
  c[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
  c[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
  c[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
  c[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// Sets flags[f] unless another wave already has: one relaxed load instead of
// thousands of same-address atomics serialised at one L2 channel.
__device__ __forceinline__ void set_flag_once(int32_t* flags, int f) {
  if (__hip_atomic_load(&flags[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) atomicOr(&flags[f], 1);
}

// NIB: FP4 planes (two samples per byte as e2m1 0 / 1.0 codes, rows of
// k_pad / 2 bytes) instead of int8 0/1 bytes.
// One 128-sample x 128-variant tile t of a population (variant blocks
// fastest: nx = ceil(n_var / 128) tiles per sample block).
template <int FAMILY, bool WORDS, bool NIB>
__device__ __forceinline__ void pack_tile(const int8_t* __restrict__ data, int64_t n_var,
                                          const int32_t* __restrict__ rows, int64_t n_sel, int64_t k_pad,
                                          int8_t* __restrict__ planes, int32_t* __restrict__ marg,
                                          int32_t* __restrict__ flags, int64_t t, int64_t nx, uint32_t* tile,
                                          int64_t* srow) {
  constexpr int NPL = FAMILY == PGLD_FAMILY_HAP ? 2 : 3;
  const int tid = threadIdx.x;
  static_assert(PK_V == 128 && PK_S == 128, "index math below assumes 128 x 128 tiles");
  const int64_t v0 = (t % nx) * PK_V, k0 = (t / nx) * PK_S;
  if (tid < PK_S) srow[tid] = k0 + tid < n_sel ? (int64_t)(rows ? rows[k0 + tid] : k0 + tid) * n_var : -1;
  __syncthreads();
  // phase 1: 32 threads per row (one dword = 4 variants each), 8 rows per
  // pass; all 16 loads of a thread are independent and issued back to back
  {
    const int q = tid & 31, rr = tid >> 5;
    uint32_t words_in[PK_S / 8];
#pragma unroll
    for (int u = 0; u < PK_S / 8; ++u) {
      const int64_t base = srow[rr + 8 * u];
      uint32_t word = 0xffffffffu;  // padding samples and variants read as missing
      if (base >= 0) {
        const int64_t at = base + v0 + 4 * q;
        if (WORDS && v0 + 4 * q + 3 < n_var) {
          word = *reinterpret_cast<const uint32_t*>(data + at);
        } else if (v0 + 4 * q + 3 < n_var) {
          // a row that is not 4-byte aligned: the two aligned dwords that
          // hold bytes at .. at + 3, joined with v_alignbyte (each holds a
          // byte of the row, so neither read leaves the row's pages)
          const uintptr_t a = reinterpret_cast<uintptr_t>(data + at);
          const uint32_t* pa = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
          const uint32_t o = (uint32_t)(a & 3);
          const uint32_t lo = pa[0];
          word = o ? __builtin_amdgcn_alignbyte(pa[1], lo, o) : lo;  // bytes o .. o + 3 of {pa[1], lo}
        } else {
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const uint32_t byte = v0 + 4 * q + b < n_var ? (uint8_t)data[at + b] : 0xffu;
            word = (word & ~(0xffu << (8 * b))) | (byte << (8 * b));
          }
        }
      }
      words_in[u] = word;
    }
#pragma unroll
    for (int u = 0; u < PK_S / 8; ++u) {
      const int s = rr + 8 * u;
      tile[s * PK_Q + (q ^ pk_swz(s))] = words_in[u];
    }
  }
  __syncthreads();
  // phase 2: thread -> (16-sample group sg, variant quad vq); the 8 lanes of
  // one variant quad cover its 128 samples
  const int sg = tid & 7, vq = tid >> 3;
  const int qcol = vq ^ ((sg & 3) << 3);  // == vq ^ pk_swz(16 sg + anything below 16)
  uint32_t cols[4][4];  // [variant j][sample dword i]: bytes = samples 16 sg + 4 i .. + 3
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t r[4], c[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) r[e] = tile[(16 * sg + 4 * i + e) * PK_Q + qcol];
    transpose4x4(r, c);
#pragma unroll
    for (int j = 0; j < 4; ++j) cols[j][i] = c[j];
  }
  // samples of this group that exist, as a byte mask per sample dword
  const int64_t s_first = k0 + 16 * sg;
  uint32_t smask[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) m |= (s_first + 4 * i + b < n_sel ? 0xffu : 0u) << (8 * b);
    smask[i] = m;
  }
  constexpr uint32_t ONES = 0x01010101u;
  int any_neg = 0, any_bad = 0;
  uint32_t cnt4[NPL];  // byte j = this lane's count of variant 4 vq + j (<= 16)
#pragma unroll
  for (int p = 0; p < NPL; ++p) cnt4[p] = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int v = 4 * vq + j;
    if (v0 + v >= n_var) continue;
    uint32_t pw[NPL][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t w = cols[j][i];
      const uint32_t neg = (w >> 7) & ONES;    // 1 per negative byte
      const uint32_t val = neg ^ ONES;         // 1 per byte >= 0
      const uint32_t nonneg_m = val * 0xffu, neg_m = neg * 0xffu;
      uint32_t bad = (~w & 0x7f7f7f7fu) & neg_m;  // negative but not -1
      if constexpr (FAMILY == PGLD_FAMILY_HAP) {
        bad |= (w & 0x7e7e7e7eu) & nonneg_m;     // > 1
        pw[0][i] = w & val;                      // x == 1 (bit 0 of a valid 0/1 byte)
        pw[1][i] = val;                          // x >= 0
      } else {
        bad |= ((w & 0x7c7c7c7cu) | (w & (w >> 1) & ONES)) & nonneg_m;  // > 2
        pw[0][i] = val;                          // x >= 0
        pw[1][i] = w & ~(w >> 1) & val;          // x == 1
        pw[2][i] = (w >> 1) & ~w & val;          // x == 2
      }
      any_neg |= (neg_m & smask[i]) != 0;
      any_bad |= (bad & smask[i]) != 0;
#pragma unroll
      for (int p = 0; p < NPL; ++p) cnt4[p] += (uint32_t)__popc(pw[p][i]) << (8 * j);
    }
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      if constexpr (NIB) {
        // four 0/1 bytes of a dword -> four 0/1 nibbles (byte b -> bits 4b),
        // times the e2m1 code of 1.0; the haplotype allele plane is centred:
        // +1 (0x2) where x == 1, -1 (0xA) where x == 0, 0 where missing
        auto nib = [](uint32_t x) {
          const uint32_t t = (x | (x >> 4)) & 0x00ff00ffu;
          return (t | (t >> 8)) & 0xffffu;
        };
        uint32_t h[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          // per nibble: valid -> 0xA, minus 0x8 where x == 1 (no borrows: x == 1 implies valid)
          if (FAMILY == PGLD_FAMILY_HAP && p == 0)
            h[i] = nib(pw[1][i]) * (uint32_t)FP4_MINUS_ONE - nib(pw[0][i]) * (uint32_t)(FP4_MINUS_ONE - FP4_ONE);
          else
            h[i] = nib(pw[p][i]) * (uint32_t)FP4_ONE;
        }
        uint2 val2;
        val2.x = h[0] | (h[1] << 16);
        val2.y = h[2] | (h[3] << 16);
        *reinterpret_cast<uint2*>(planes + ((int64_t)p * n_var + v0 + v) * (k_pad / 2) + (k0 + 16 * sg) / 2) = val2;
      } else {
        int4 val4;
        val4.x = (int)pw[p][0];
        val4.y = (int)pw[p][1];
        val4.z = (int)pw[p][2];
        val4.w = (int)pw[p][3];
        *reinterpret_cast<int4*>(planes + ((int64_t)p * n_var + v0 + v) * k_pad + k0 + 16 * sg) = val4;
      }
    }
  }
  // sum the byte-packed counts over the 8 lanes of the quad (<= 128 per byte)
#pragma unroll
  for (int p = 0; p < NPL; ++p) {
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) cnt4[p] += (uint32_t)__shfl_xor((int)cnt4[p], o);
  }
  if (sg == 0) {
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = (cnt4[p] >> (8 * j)) & 0xff;
        if (c) atomicAdd(&marg[(int64_t)p * n_var + v0 + 4 * vq + j], c);
      }
    }
  }
  unsigned long long mneg = __ballot(any_neg), mbad = __ballot(any_bad);
  if ((tid & 63) == 0) {
    if (mneg) set_flag_once(flags, 0);
    if (mbad) set_flag_once(flags, 1);
  }
}

template <int FAMILY, bool WORDS, bool NIB = false>
__global__ __launch_bounds__(256) void k_pack(const int8_t* __restrict__ data, int64_t n_var,
                                              const int32_t* __restrict__ rows, int64_t n_sel, int64_t k_pad,
                                              int8_t* __restrict__ planes, int32_t* __restrict__ marg,
                                              int32_t* __restrict__ flags) {
  __shared__ uint32_t tile[PK_S * PK_Q];  // [sample][variant quad ^ pk_swz(sample)]
  __shared__ int64_t srow[PK_S];
  // Block (x, y) of the launch grid packs tile t = xcd_tile(linear id),
  // variant blocks fastest: each XCD takes a contiguous run of the tiles, so
  // neighbouring variant blocks of the same sample rows run on one XCD -- the
  // 128-byte lines that straddle two of them (rows of n_var bytes are rarely
  // line-aligned) are fetched into one L2, and reads stay in the same DRAM rows.
  const int64_t nx = gridDim.x;
  const int64_t t = xcd_tile((int64_t)blockIdx.x + (int64_t)blockIdx.y * nx, nx * gridDim.y);
  pack_tile<FAMILY, WORDS, NIB>(data, n_var, rows, n_sel, k_pad, planes, marg, flags, t, nx, tile, srow);
}

// Many populations (of many regions) in one launch: job j owns the global
// tiles [job.tile0, job.tile0 + nx ny); the block's job is found by binary
// search over the table (one workgroup-uniform search per block).
struct PackJob {
  const int8_t* data;
  const int32_t* rows;
  int8_t* planes;
  int32_t* marg;
  int32_t* flags;
  int64_t n_var, n_sel, k_pad, tile0, nx;
  int32_t words, pad_;
};

__device__ __forceinline__ int find_job(const PackJob* __restrict__ jobs, int n_jobs, int64_t t) {
  int lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (jobs[mid].tile0 <= t) lo = mid; else hi = mid;
  }
  return lo;
}

template <int FAMILY, bool NIB>
__global__ __launch_bounds__(256) void k_pack_batch(const PackJob* __restrict__ jobs, int n_jobs, int64_t n_tiles) {
  __shared__ uint32_t tile[PK_S * PK_Q];
  __shared__ int64_t srow[PK_S];
  const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= n_tiles) return;
  const PackJob& j = jobs[find_job(jobs, n_jobs, t)];
  const int64_t lt = t - j.tile0;
  if (j.words)
    pack_tile<FAMILY, true, NIB>(j.data, j.n_var, j.rows, j.n_sel, j.k_pad, j.planes, j.marg, j.flags, lt, j.nx,
                                 tile, srow);
  else
    pack_tile<FAMILY, false, NIB>(j.data, j.n_var, j.rows, j.n_sel, j.k_pad, j.planes, j.marg, j.flags, lt, j.nx,
                                  tile, srow);
}

// Clears every job's margins and flags (the packing adds into them): one
// grid-stride pass over the concatenated (planes x n_var + 2) words.
template <int NPL>
__global__ __launch_bounds__(256) void k_pack_batch_clear(const PackJob* __restrict__ jobs, int n_jobs) {
  for (int j = blockIdx.x; j < n_jobs; j += gridDim.x) {
    const PackJob& jb = jobs[j];
    const int64_t n = (int64_t)NPL * jb.n_var;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) jb.marg[e] = 0;
    if (threadIdx.x < 2) jb.flags[threadIdx.x] = 0;
  }
}

// int8 0/1 planes -> bit planes: word w of a row holds samples 32w .. 32w+31
// (bit b = sample 32w + b); rows padded with zero words to bit_row_bytes.
__global__ void k_pack_bits(const int8_t* __restrict__ planes, int64_t n_rows, int64_t k_pad,
                            uint32_t* __restrict__ bits, int64_t words) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_rows * words) return;
  const int64_t row = t / words, w = t - row * words;
  uint32_t out = 0;
  if (32 * w < k_pad) {
    const v4i* src = reinterpret_cast<const v4i*>(planes + row * k_pad + 32 * w);
    const v4i lo = src[0], hi = src[1];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const uint32_t x = (uint32_t)(d < 4 ? lo[d] : hi[d - 4]);  // four 0/1 bytes
      const uint32_t nib = (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
      out |= nib << (4 * d);
    }
  }
  bits[t] = out;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int64_t upper_index(const double* pos, int64_t n, int64_t from, double ub) {
  return std::upper_bound(pos + from, pos + n, ub) - pos;
}

static int bin_host(double d, const std::vector<double>& e) {
  int n = (int)e.size() - 1;
  if (!(d >= e[0]) || !(d < e[n])) return -1;
  return (int)(std::upper_bound(e.begin(), e.end(), d) - e.begin()) - 1;
}

// ---------------------------------------------------------------------------
// Band schedule construction (host)
// ---------------------------------------------------------------------------
struct TileBuild {
  std::vector<BandTile> tiles;
  std::vector<BandTileHost> host;
  std::vector<std::pair<int32_t, int64_t>> key_slot;  // (key, slot) in tile order
  int64_t n_slots = 0;
  int max_ts = 0;
};

static int check_plan_args(const double* pos, const double* key_coord, int64_t n_var, double max_dist,
                           int32_t key_mode, const double* edges, int32_t n_keys, double width,
                           std::vector<double>& e) {
  PGLD_REQUIRE(max_dist >= 0.0, "band plan: negative max_dist");
  PGLD_REQUIRE(key_mode == PGLD_KEY_BIN_ABS || key_mode == PGLD_KEY_BIN || key_mode == PGLD_KEY_WINDOW,
               "band plan: unknown key mode %d", key_mode);
  PGLD_REQUIRE(n_keys >= 1, "band plan: n_keys must be >= 1");
  PGLD_REQUIRE(key_mode == PGLD_KEY_WINDOW ? width > 0.0 : edges != nullptr,
               "band plan: windows need width > 0, bins need edges");
  const double* kc = key_coord ? key_coord : pos;
  for (int64_t v = 1; v < n_var; ++v) {
    PGLD_REQUIRE(pos[v] >= pos[v - 1], "band plan: positions not sorted at %lld", (long long)v);
    PGLD_REQUIRE(kc[v] >= kc[v - 1], "band plan: key coordinate decreases at %lld", (long long)v);
  }
  e.clear();
  if (key_mode != PGLD_KEY_WINDOW) {
    e.assign(edges, edges + n_keys + 1);
    for (int k = 0; k < n_keys; ++k) PGLD_REQUIRE(e[k] < e[k + 1], "band plan: edges must increase");
  }
  return PGLD_OK;
}

// Tiles of one region (SNPs [0, n_var) of pos / kc): 64 x 64 tiles from each
// 64-SNP row to the last column any of its SNPs reaches within max_dist, each
// with the key range its pairs can fall in (keys shifted by key_base).
// Returns the region's pair count; base (if given) receives the exclusive
// prefix of per-SNP pair counts.
static int64_t build_region(const double* pos, const double* kc, int64_t n_var, double max_dist, int32_t key_mode,
                            const std::vector<double>& e, int32_t n_keys, double origin, double width, int32_t region,
                            int32_t key_base, std::vector<int64_t>* base_out, TileBuild& tb) {
  std::vector<int64_t> per(n_var), base(n_var + 1, 0);
  for (int64_t v = 0; v < n_var; ++v) {
    int64_t u = upper_index(pos, n_var, v, pos[v] + max_dist);
    per[v] = std::max<int64_t>(0, u - v - 1);
    base[v + 1] = base[v] + per[v];
  }
  auto key_i = [&](int64_t i) -> int64_t {  // window key of the left SNP
    return (int64_t)std::floor((kc[i] - origin) / width);
  };
  const int64_t n_rows = ceil_div(n_var, BT);
  for (int64_t r = 0; r < n_rows; ++r) {
    const int64_t i0 = r * BT, il = std::min(i0 + BT - 1, n_var - 1);
    if (base[il + 1] - base[i0] == 0) continue;
    const int64_t jhi = std::max(il, il + per[il]);
    for (int64_t tc = r; tc * BT <= jhi; ++tc) {
      const int64_t j0 = tc * BT, jl = std::min(j0 + BT - 1, n_var - 1);
      int64_t klo, khi;
      if (key_mode == PGLD_KEY_WINDOW) {
        klo = std::max<int64_t>(0, key_i(i0));
        khi = std::min<int64_t>(n_keys - 1, key_i(il));
      } else {
        double dmin = std::max(0.0, kc[j0] - kc[il]);
        double dmax = kc[jl] - kc[i0];
        if (dmax < e[0] || dmin >= e[n_keys]) {
          klo = 1;
          khi = 0;
        } else {
          klo = dmin < e[0] ? 0 : bin_host(dmin, e);
          khi = dmax >= e[n_keys] ? n_keys - 1 : bin_host(dmax, e);
        }
      }
      BandTile t{};
      t.i0 = (int32_t)i0;
      t.j0 = (int32_t)j0;
      t.key_lo = (int32_t)(klo + key_base);
      t.n_slots = khi >= klo ? (int32_t)(khi - klo + 1) : 0;
      t.slot_base = tb.n_slots;
      t.region = region;
      for (int q = 0; q < t.n_slots; ++q) tb.key_slot.push_back({(int32_t)(t.key_lo + q), tb.n_slots + q});
      tb.n_slots += t.n_slots;
      tb.max_ts = std::max(tb.max_ts, t.n_slots);
      tb.tiles.push_back(t);
      tb.host.push_back(BandTileHost{region, (int32_t)(il - i0 + 1), (int32_t)(jl - j0 + 1), (int32_t)(tc == r),
                                     (int32_t)r, (int32_t)tc});
    }
  }
  if (base_out) base_out->assign(base.begin(), base.end() - 1);
  return base[n_var];
}

static int upload_plan(pgld_band_plan* plan, const double* pos, const double* kc, int64_t n_var,
                       const std::vector<double>& e, TileBuild& tb, const std::vector<int64_t>& base) {
  plan->n_tiles = (int64_t)tb.tiles.size();
  plan->n_slots = tb.n_slots;
  plan->max_tile_slots = tb.max_ts;
  const int32_t n_keys = plan->n_keys;
  std::vector<int32_t> key_off(n_keys + 1, 0);
  for (auto& ks : tb.key_slot) key_off[ks.first + 1]++;
  for (int k = 0; k < n_keys; ++k) key_off[k + 1] += key_off[k];
  std::vector<int64_t> key_list(tb.key_slot.size());
  {
    std::vector<int32_t> fill(key_off.begin(), key_off.end() - 1);
    for (auto& ks : tb.key_slot) key_list[fill[ks.first]++] = ks.second;
  }
  auto up = [&](void** dst, const void* src, size_t bytes) -> int {
    hipError_t err = hipMalloc(dst, bytes ? bytes : 8);
    if (err == hipSuccess && src && bytes) err = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    return check_hip(err, "band plan upload");
  };
  int rc = PGLD_OK;
  if (!rc) rc = up((void**)&plan->d_pos, pos, n_var * sizeof(double));
  if (!rc) rc = up((void**)&plan->d_kc, kc, n_var * sizeof(double));
  if (!rc) rc = up((void**)&plan->d_edges, e.empty() ? nullptr : e.data(), e.size() * sizeof(double));
  if (!rc) rc = up(&plan->d_tiles, tb.tiles.data(), tb.tiles.size() * sizeof(BandTile));
  if (!rc) rc = up((void**)&plan->d_pair_base, base.empty() ? nullptr : base.data(), base.size() * sizeof(int64_t));
  if (!rc) rc = up((void**)&plan->d_key_off, key_off.data(), key_off.size() * sizeof(int32_t));
  {
    // keys whose slot lists are long (e.g. few distance bins over many tiles)
    // are summed in two fixed-order levels: chunks of RED_CHUNK slots, then the
    // chunks of each key.  A key of at most RED_CHUNK slots sums to the same
    // value either way, so every plan holding a given key (a region batch or
    // the region's own band) reduces it identically.
    int64_t longest = 0;
    for (int k = 0; k < n_keys; ++k) longest = std::max<int64_t>(longest, key_off[k + 1] - key_off[k]);
    if (!rc && longest > RED_CHUNK) {
      std::vector<int64_t> lo;
      std::vector<int32_t> off(n_keys + 1, 0);
      for (int k = 0; k < n_keys; ++k) {
        off[k] = (int32_t)lo.size();
        for (int64_t e = key_off[k]; e < key_off[k + 1]; e += RED_CHUNK) lo.push_back(e);
      }
      off[n_keys] = (int32_t)lo.size();
      std::vector<int64_t> hi(lo.size());
      for (int k = 0; k < n_keys; ++k)
        for (int32_t c = off[k]; c < off[k + 1]; ++c) hi[c] = std::min<int64_t>(lo[c] + RED_CHUNK, key_off[k + 1]);
      std::vector<int64_t> lohi(2 * lo.size());
      for (size_t c = 0; c < lo.size(); ++c) {
        lohi[2 * c] = lo[c];
        lohi[2 * c + 1] = hi[c];
      }
      plan->n_red_chunks = (int64_t)lo.size();
      rc = up((void**)&plan->d_red_lo, lohi.data(), lohi.size() * sizeof(int64_t));
      if (!rc) rc = up((void**)&plan->d_red_off, off.data(), off.size() * sizeof(int32_t));
    }
  }
  if (!rc) rc = up((void**)&plan->d_key_list, key_list.empty() ? nullptr : key_list.data(),
                   key_list.size() * sizeof(int64_t));
  if (!rc && plan->n_regions == 0) {
    // 2 x 2 groups of tiles (super-tiles), ordered by 8 x 8 blocks of super-tiles
    // and row-major inside a block.  The kernel hands each XCD a contiguous run
    // of this list, so the 64 super-tiles an XCD holds at once (32 CUs x 2)
    // share 8 row and 8 column strips of operands in its L2; in plain
    // row-major order they share one row strip and re-fetch every column strip.
    constexpr int64_t SG = 8;
    std::vector<std::pair<int64_t, int32_t>> keyed(tb.tiles.size());
    const int64_t n_sc = ceil_div(std::max<int64_t>(n_var, 1), 2 * BT) + 1;
    const int64_t n_gc = ceil_div(n_sc, SG);
    for (size_t t = 0; t < tb.tiles.size(); ++t) {
      const int64_t si = (tb.tiles[t].i0 / BT) >> 1, sj = (tb.tiles[t].j0 / BT) >> 1;
      const int64_t blk = (si / SG) * n_gc + sj / SG;
      keyed[t] = {(blk * SG + si % SG) * SG + sj % SG, (int32_t)t};
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& b) {
                       return a.first < b.first;
                     });
    std::vector<SuperTile> sup;
    for (size_t k = 0; k < keyed.size();) {
      SuperTile st{};
      const BandTile& first = tb.tiles[keyed[k].second];
      st.i0 = (int32_t)(((first.i0 / BT) >> 1) * 2 * BT);
      st.j0 = (int32_t)(((first.j0 / BT) >> 1) * 2 * BT);
      for (int q = 0; q < 4; ++q) st.t[q] = -1;
      const int64_t key = keyed[k].first;
      for (; k < keyed.size() && keyed[k].first == key; ++k) {
        const BandTile& bt = tb.tiles[keyed[k].second];
        st.t[((bt.i0 - st.i0) / BT) * 2 + (bt.j0 - st.j0) / BT] = keyed[k].second;
      }
      sup.push_back(st);
    }
    plan->n_super = (int64_t)sup.size();
    rc = up(&plan->d_super, sup.empty() ? nullptr : sup.data(), sup.size() * sizeof(SuperTile));
  }
  if (!rc && plan->n_regions == 0) {
    // Launch order of the fused sums: tiles by 8 x 8 blocks (block-row major,
    // row-major inside a block).  Each XCD walks a contiguous run of it, so the
    // ~64 tiles it holds at once read 8 row and 8 column strips of operands
    // from its L2 instead of one row strip and ~20 column strips (a row-major
    // walk along a band, whose strips overflow the L2 once several populations
    // or planes are read).  Slots, keys and the reduction order belong to the
    // tiles, so the order changes no result.
    constexpr int64_t SB = 8;
    const int64_t n_bc = ceil_div(std::max<int64_t>(n_var, 1), BT * SB) + 1;
    std::vector<std::pair<int64_t, int32_t>> keyed(tb.tiles.size());
    for (size_t t = 0; t < tb.tiles.size(); ++t) {
      const int64_t ti = tb.tiles[t].i0 / BT, tj = tb.tiles[t].j0 / BT;
      keyed[t] = {(((ti / SB) * n_bc + tj / SB) * SB + ti % SB) * SB + tj % SB, (int32_t)t};
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& b) {
                       return a.first < b.first;
                     });
    std::vector<int32_t> order(keyed.size());
    for (size_t k = 0; k < keyed.size(); ++k) order[k] = keyed[k].second;
    rc = up((void**)&plan->d_sums_order, order.empty() ? nullptr : order.data(), order.size() * sizeof(int32_t));
  }
  if (!rc && plan->n_regions == 0) {
    // tile lists by tile-row (contiguous in the schedule) and tile-column
    const int64_t n_tr = ceil_div(std::max<int64_t>(n_var, 1), BT);
    std::vector<int64_t> row_off(n_tr + 1, 0), col_off(n_tr + 1, 0);
    for (const BandTile& t : tb.tiles) {
      row_off[t.i0 / BT + 1]++;
      col_off[t.j0 / BT + 1]++;
    }
    for (int64_t r = 0; r < n_tr; ++r) {
      row_off[r + 1] += row_off[r];
      col_off[r + 1] += col_off[r];
    }
    std::vector<int32_t> col_list(tb.tiles.size());
    std::vector<int64_t> fill(col_off.begin(), col_off.end() - 1);
    for (size_t t = 0; t < tb.tiles.size(); ++t) col_list[fill[tb.tiles[t].j0 / BT]++] = (int32_t)t;
    rc = up((void**)&plan->d_row_off, row_off.data(), row_off.size() * sizeof(int64_t));
    if (!rc) rc = up((void**)&plan->d_col_off, col_off.data(), col_off.size() * sizeof(int64_t));
    if (!rc) rc = up((void**)&plan->d_col_list, col_list.empty() ? nullptr : col_list.data(),
                     col_list.size() * sizeof(int32_t));
  }
  if (plan->n_regions > 0) plan->h_tiles.swap(tb.host);
  return rc;
}

// The lean kernels keep three workgroups per CU: whether the table still
// fits beside lds bytes of each (allocations rounded up to 1 KiB).
static bool lean_table_fits(size_t lds, size_t tab) {
  const size_t with = (lds + STATIC_LDS_BYTES + tab + 1023) / 1024 * 1024;
  return 3 * with <= 160 * 1024;
}

// Whether the reciprocal table (tab bytes) fits beside a workgroup's lds
// bytes without lowering the workgroups per CU below min(current, 2).
static bool recip_table_fits(size_t lds, size_t tab) {
  const size_t cu = 160 * 1024, base = lds + STATIC_LDS_BYTES, with = base + tab;
  return with <= cu && cu / with >= std::min<size_t>(cu / base, 2);
}

template <int FAMILY, bool MISSING, int P, int SET, bool NIB>
static void launch_sums(const pgld_band_plan* plan, const BandArgs& ba_in, const PopArgs& pa, size_t lds,
                        hipStream_t s) {
  BandArgs ba = ba_in;
  ba.rtab_n = -1;
  ba.rtab_off = 0;
  // the within-population haplotype sets with missing data take 1 / (n)_4
  // from an LDS table of n = 0 .. the largest population when it fits
  constexpr bool TAB = FAMILY == PGLD_FAMILY_HAP && MISSING && (SET == PGLD_SET_SELF || P == 1);
  // the autotuned variants (plan->kvar) exist for the FP4 within-population
  // haplotype sets of region batches; the lean one (kvar 3) leaves the table
  // out so that three workgroups fit a CU's LDS
  constexpr bool VARIANTS = FAMILY == PGLD_FAMILY_HAP && SET == PGLD_SET_SELF && P > 1 && NIB;
  const bool lean = VARIANTS && plan->n_regions > 0 && plan->kvar == 3;
  if (lean) lds -= (size_t)opbuf_bytes<Planes<FAMILY, MISSING>::N>() - opbuf_bytes_swz<Planes<FAMILY, MISSING>::N>();
  bool tab = false;
  if constexpr (TAB) {
    int smax = 0;
    if (plan->n_regions > 0)
      smax = plan->bound_max_samples;
    else
      for (int p = 0; p < P; ++p) smax = std::max(smax, pa.n_samples[p]);
    const size_t bytes = (((size_t)smax + 1) * 8 + 15) / 16 * 16;
    if (lean ? lean_table_fits(lds, bytes) : recip_table_fits(lds, bytes)) {
      tab = true;
      ba.rtab_n = smax;
      ba.rtab_off = (int32_t)lds;
      lds += bytes;
    }
  }
  if (plan->n_regions > 0) {
    if constexpr (gstash_set(FAMILY, P, SET)) {
      return;  // run_sums refuses these sets for region batches
    } else {
      const int m = plan->order_mode;
      const int64_t len = plan->order_len[m];
      const int64_t grid = plan->grid > 0 ? std::min<int64_t>(plan->grid, len) : len;
      const auto* rd = reinterpret_cast<const RegionDesc*>(plan->d_regions);
      const int32_t* ord = plan->d_order + plan->order_off[m];
      const int32_t xs = (int32_t)plan->xcd_split;
      const dim3 g((unsigned)grid), b(BAND_THREADS);
      auto go = [&](auto kv_tag) {
        constexpr int KV = decltype(kv_tag)::value;
        hipLaunchKernelGGL((k_batch_sums<FAMILY, MISSING, P, SET, NIB, KV>), g, b, lds, s, ba, rd, ord, len, xs);
      };
      if constexpr (VARIANTS) {
        if (lean) {
          if constexpr (TAB) {
            if (tab) return go(std::integral_constant<int, KV_LEAN | KV_RTAB>{});
          }
          return go(std::integral_constant<int, KV_LEAN>{});
        }
      }
      if constexpr (TAB) {
        if (tab) {
          if constexpr (VARIANTS) {
            if (plan->kvar == KV_PER_POP) return go(std::integral_constant<int, KV_RTAB | KV_PER_POP>{});
            if (plan->kvar == KV_ROLLED) return go(std::integral_constant<int, KV_RTAB | KV_ROLLED>{});
          }
          return go(std::integral_constant<int, KV_RTAB>{});
        }
        return go(std::integral_constant<int, 0>{});
      } else {
        if constexpr (VARIANTS) {
          if (plan->kvar == KV_PER_POP) return go(std::integral_constant<int, KV_PER_POP>{});
          if (plan->kvar == KV_ROLLED) return go(std::integral_constant<int, KV_ROLLED>{});
        }
        return go(std::integral_constant<int, 0>{});
      }
    }
  } else {
    const dim3 g((unsigned)plan->n_tiles), b(BAND_THREADS);
    if constexpr (gstash_set(FAMILY, P, SET)) {
      // one workspace slot per workgroup of a launch of at most GSTASH_TILES
      for (int64_t t0 = 0; t0 < plan->n_tiles; t0 += GSTASH_TILES) {
        ba.tile0 = t0;
        const dim3 gc((unsigned)std::min<int64_t>(GSTASH_TILES, plan->n_tiles - t0));
        hipLaunchKernelGGL((k_band_sums<FAMILY, MISSING, P, SET, NIB>), gc, b, lds, s, ba, pa);
      }
      return;
    }
    if constexpr (TAB) {
      if (tab) {
        hipLaunchKernelGGL((k_band_sums<FAMILY, MISSING, P, SET, NIB, KV_RTAB>), g, b, lds, s, ba, pa);
        return;
      }
    }
    hipLaunchKernelGGL((k_band_sums<FAMILY, MISSING, P, SET, NIB>), g, b, lds, s, ba, pa);
  }
}

typedef void (*SumsLauncher)(const pgld_band_plan*, const BandArgs&, const PopArgs&, size_t, hipStream_t);

template <int FAMILY, bool MISSING, bool NIB>
static SumsLauncher pick_launcher(int P, int set) {
  if constexpr (FAMILY == PGLD_FAMILY_HAP) {
    if (set == PGLD_SET_MOMENTS) {
      if (P == 1) return launch_sums<FAMILY, MISSING, 1, PGLD_SET_MOMENTS, NIB>;
      if (P == 2) return launch_sums<FAMILY, MISSING, 2, PGLD_SET_MOMENTS, NIB>;
      if (P == 3) return launch_sums<FAMILY, MISSING, 3, PGLD_SET_MOMENTS, NIB>;
      if (P == 4) return launch_sums<FAMILY, MISSING, 4, PGLD_SET_MOMENTS, NIB>;
    } else {
      if (P == 1) return launch_sums<FAMILY, MISSING, 1, PGLD_SET_MOMENTS, NIB>;
      if (P == 2) return launch_sums<FAMILY, MISSING, 2, PGLD_SET_SELF, NIB>;
      if (P == 3) return launch_sums<FAMILY, MISSING, 3, PGLD_SET_SELF, NIB>;
      if (P == 4) return launch_sums<FAMILY, MISSING, 4, PGLD_SET_SELF, NIB>;
    }
  } else {
    if (P == 1) return launch_sums<FAMILY, MISSING, 1, PGLD_SET_MOMENTS, NIB>;
    if (set == PGLD_SET_MOMENTS) {
      if (P == 2) return launch_sums<FAMILY, MISSING, 2, PGLD_SET_MOMENTS, NIB>;
      if (P == 3) return launch_sums<FAMILY, MISSING, 3, PGLD_SET_MOMENTS, NIB>;
      if (P == 4) return launch_sums<FAMILY, MISSING, 4, PGLD_SET_MOMENTS, NIB>;
    } else {
      if (P == 2) return launch_sums<FAMILY, MISSING, 2, PGLD_SET_SELF, NIB>;
      if (P == 3) return launch_sums<FAMILY, MISSING, 3, PGLD_SET_SELF, NIB>;
      if (P == 4) return launch_sums<FAMILY, MISSING, 4, PGLD_SET_SELF, NIB>;
    }
  }
  return nullptr;
}

static int n_stats_of(int family, int P, int set) {
  if (set == PGLD_SET_SELF && P > 1) return 3 * P;
  return P == 1 ? 3 : (P == 2 ? 15 : (P == 3 ? 45 : 105));
}

}  // namespace pgld


using namespace pgld;

// L2-blocked tile order for a region batch.  Workgroup b of a launch runs on
// XCD b % 8, and each XCD holds about 64 workgroups at once; so the tiles are
// cut into 8 x 8 blocks of one region (8 A and 8 B operand panels serve 64
// tiles from one L2), the blocks are dealt to the XCD with the least work so
// far (regions by descending work per tile, largest first), and the eight
// per-XCD streams are interleaved position by position.  Streams that run out
// are padded with -1 so that position p stays on XCD p % 8; the padding is
// skipped by the kernel.  Persistent grids that are a multiple of 8 keep the
// same mapping (workgroup b walks positions b, b + grid, ...).
static std::vector<int32_t> blocked_order(const std::vector<BandTileHost>& h, const std::vector<double>& wsum,
                                          const std::vector<double>& cost) {
  constexpr int SB = 8;
  const int64_t T = (int64_t)h.size();
  const int R = (int)wsum.size();
  std::vector<std::vector<int32_t>> by_region(R);
  for (int64_t t = 0; t < T; ++t) by_region[h[t].region].push_back((int32_t)t);
  std::vector<int32_t> rorder(R);
  for (int r = 0; r < R; ++r) rorder[r] = r;
  std::stable_sort(rorder.begin(), rorder.end(), [&](int32_t a, int32_t b) { return wsum[a] > wsum[b]; });
  std::vector<std::vector<int32_t>> stream(N_XCD);
  std::vector<double> load(N_XCD, 0.0);
  for (int32_t r : rorder) {
    std::vector<int32_t>& ts = by_region[r];
    // region tiles are row-major; a stable sort by block keeps rows within a block
    std::stable_sort(ts.begin(), ts.end(), [&](int32_t a, int32_t b) {
      const int32_t ba = h[a].ti / SB, bb = h[b].ti / SB;
      if (ba != bb) return ba < bb;
      return h[a].tj / SB < h[b].tj / SB;
    });
    for (size_t s = 0; s < ts.size();) {
      size_t e = s;
      double c = 0.0;
      while (e < ts.size() && h[ts[e]].ti / SB == h[ts[s]].ti / SB && h[ts[e]].tj / SB == h[ts[s]].tj / SB)
        c += cost[ts[e++]];
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      stream[x].insert(stream[x].end(), ts.begin() + s, ts.begin() + e);
      load[x] += c;
      s = e;
    }
  }
  size_t longest = 0;
  for (auto& st : stream) longest = std::max(longest, st.size());
  std::vector<int32_t> out;
  out.reserve(longest * N_XCD);
  for (size_t q = 0; q < longest; ++q)
    for (int x = 0; x < N_XCD; ++x) out.push_back(q < stream[x].size() ? stream[x][q] : -1);
  return out;
}

extern "C" {

static int pack_planes(int32_t family, const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows,
                       int64_t n_sel, int64_t k_pad, int8_t* planes, int32_t* marg, int32_t* flags, void* stream,
                       bool nib) {
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_pack_planes: bad family");
  if (!rows) n_sel = n_rows;
  PGLD_REQUIRE(n_var >= 0 && n_sel >= 0, "pgld_pack_planes: negative size");
  PGLD_REQUIRE(k_pad % (nib ? 256 : 128) == 0 && k_pad >= n_sel,
               "pgld_pack_planes: k_pad must be a multiple of %d >= n_sel", nib ? 256 : 128);
  PGLD_REQUIRE(n_sel <= 32767, "pgld_pack_planes: at most 32767 samples per population");
  PGLD_REQUIRE(data && planes && marg && flags, "pgld_pack_planes: null pointer");
  const int npl = family == PGLD_FAMILY_HAP ? 2 : 3;
  hipStream_t s = as_stream(stream);
  int rc = fill_i32(flags, 2, 0, s);
  if (rc) return rc;
  if (n_var == 0) return PGLD_OK;
  rc = fill_i32(marg, (int64_t)npl * n_var, 0, s);
  if (rc) return rc;
  if (k_pad == 0) return PGLD_OK;
  dim3 g((unsigned)ceil_div(n_var, PK_V), (unsigned)(k_pad / PK_S));
  const bool words = (n_var % 4 == 0) && (reinterpret_cast<uintptr_t>(data) % 4 == 0);
#define PGLD_PACK(F, W, N) \
  hipLaunchKernelGGL((k_pack<F, W, N>), g, dim3(256), 0, s, data, n_var, rows, n_sel, k_pad, planes, marg, flags)
#define PGLD_PACK_W(F, N) \
  if (words) PGLD_PACK(F, true, N); else PGLD_PACK(F, false, N)
  if (family == PGLD_FAMILY_HAP) {
    if (nib) { PGLD_PACK_W(PGLD_FAMILY_HAP, true); } else { PGLD_PACK_W(PGLD_FAMILY_HAP, false); }
  } else {
    if (nib) { PGLD_PACK_W(PGLD_FAMILY_GENO, true); } else { PGLD_PACK_W(PGLD_FAMILY_GENO, false); }
  }
#undef PGLD_PACK_W
#undef PGLD_PACK
  return check_launch("k_pack");
}

int pgld_pack_planes(int32_t family, const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows,
                     int64_t n_sel, int64_t k_pad, int8_t* planes, int32_t* marg, int32_t* flags, void* stream) {
  return pack_planes(family, data, n_rows, n_var, rows, n_sel, k_pad, planes, marg, flags, stream, false);
}

int pgld_pack_planes_fp4(int32_t family, const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows,
                         int64_t n_sel, int64_t k_pad, int8_t* planes, int32_t* marg, int32_t* flags, void* stream) {
  return pack_planes(family, data, n_rows, n_var, rows, n_sel, k_pad, planes, marg, flags, stream, true);
}

}  // extern "C"

struct pgld_pack_batch {
  int32_t family = 0;
  bool fp4 = false;
  int32_t n_jobs = 0;
  int64_t n_tiles = 0;
  pgld::PackJob* d_jobs = nullptr;
};

extern "C" {

int pgld_pack_batch_create(int32_t family, int32_t fp4, int32_t n_jobs, const int8_t* const* data,
                           const int64_t* n_rows, const int64_t* n_var, const int32_t* const* rows,
                           const int64_t* n_sel, const int64_t* k_pad, int8_t* const* planes, int32_t* const* marg,
                           int32_t* const* flags, pgld_pack_batch** out) {
  PGLD_REQUIRE(out, "pgld_pack_batch_create: null output");
  *out = nullptr;
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_pack_batch_create: bad family");
  PGLD_REQUIRE(n_jobs >= 1 && data && n_rows && n_var && rows && n_sel && k_pad && planes && marg && flags,
               "pgld_pack_batch_create: null pointer or no jobs");
  std::vector<PackJob> jobs(n_jobs);
  int64_t t0 = 0;
  for (int j = 0; j < n_jobs; ++j) {
    const int64_t ns = rows[j] ? n_sel[j] : n_rows[j];
    PGLD_REQUIRE(n_var[j] >= 0 && ns >= 0, "pgld_pack_batch_create: negative size (job %d)", j);
    PGLD_REQUIRE(k_pad[j] % (fp4 ? 256 : 128) == 0 && k_pad[j] >= ns && k_pad[j] > 0,
                 "pgld_pack_batch_create: k_pad must be a positive multiple of %d >= n_sel (job %d)", fp4 ? 256 : 128,
                 j);
    PGLD_REQUIRE(ns <= 32767, "pgld_pack_batch_create: at most 32767 samples per population (job %d)", j);
    PGLD_REQUIRE(data[j] && planes[j] && marg[j] && flags[j], "pgld_pack_batch_create: null pointer (job %d)", j);
    PackJob& b = jobs[j];
    b.data = data[j];
    b.rows = rows[j];
    b.planes = planes[j];
    b.marg = marg[j];
    b.flags = flags[j];
    b.n_var = n_var[j];
    b.n_sel = ns;
    b.k_pad = k_pad[j];
    b.nx = ceil_div(n_var[j], PK_V);
    b.tile0 = t0;
    b.words = (n_var[j] % 4 == 0) && (reinterpret_cast<uintptr_t>(data[j]) % 4 == 0);
    b.pad_ = 0;
    t0 += b.nx * (k_pad[j] / PK_S);
  }
  PGLD_REQUIRE(t0 < 0x7fffffff, "pgld_pack_batch_create: too many tiles");
  auto* pb = new pgld_pack_batch();
  pb->family = family;
  pb->fp4 = fp4 != 0;
  pb->n_jobs = n_jobs;
  pb->n_tiles = t0;
  hipError_t e = hipMalloc((void**)&pb->d_jobs, jobs.size() * sizeof(PackJob));
  if (e == hipSuccess)
    e = hipMemcpy(pb->d_jobs, jobs.data(), jobs.size() * sizeof(PackJob), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (pb->d_jobs) (void)hipFree(pb->d_jobs);
    delete pb;
    set_error("pgld_pack_batch_create: %s", hipGetErrorString(e));
    return (int)e;
  }
  *out = pb;
  return PGLD_OK;
}

int pgld_pack_batch_run(const pgld_pack_batch* pb, void* stream) {
  PGLD_REQUIRE(pb && pb->d_jobs, "pgld_pack_batch_run: null batch");
  hipStream_t s = as_stream(stream);
  const unsigned clear_blocks = (unsigned)std::min<int64_t>(pb->n_jobs, 4096);
  if (pb->family == PGLD_FAMILY_HAP)
    hipLaunchKernelGGL(k_pack_batch_clear<2>, dim3(clear_blocks), dim3(256), 0, s, pb->d_jobs, pb->n_jobs);
  else
    hipLaunchKernelGGL(k_pack_batch_clear<3>, dim3(clear_blocks), dim3(256), 0, s, pb->d_jobs, pb->n_jobs);
  int rc = check_launch("k_pack_batch_clear");
  if (rc || pb->n_tiles == 0) return rc;
  const dim3 g((unsigned)pb->n_tiles), b(256);
  if (pb->family == PGLD_FAMILY_HAP) {
    if (pb->fp4) hipLaunchKernelGGL((k_pack_batch<PGLD_FAMILY_HAP, true>), g, b, 0, s, pb->d_jobs, pb->n_jobs, pb->n_tiles);
    else hipLaunchKernelGGL((k_pack_batch<PGLD_FAMILY_HAP, false>), g, b, 0, s, pb->d_jobs, pb->n_jobs, pb->n_tiles);
  } else {
    if (pb->fp4) hipLaunchKernelGGL((k_pack_batch<PGLD_FAMILY_GENO, true>), g, b, 0, s, pb->d_jobs, pb->n_jobs, pb->n_tiles);
    else hipLaunchKernelGGL((k_pack_batch<PGLD_FAMILY_GENO, false>), g, b, 0, s, pb->d_jobs, pb->n_jobs, pb->n_tiles);
  }
  return check_launch("k_pack_batch");
}

int pgld_pack_batch_destroy(pgld_pack_batch* pb) {
  if (!pb) return PGLD_OK;
  if (pb->d_jobs) (void)hipFree(pb->d_jobs);
  delete pb;
  return PGLD_OK;
}

int pgld_band_plan_create(const double* pos, int64_t n_var, double max_dist, int32_t key_mode,
                          const double* key_coord, const double* edges, int32_t n_keys, double origin, double width,
                          pgld_band_plan** out) {
  PGLD_REQUIRE(out, "pgld_band_plan_create: null output");
  *out = nullptr;
  PGLD_REQUIRE(n_var >= 0 && n_var < 0x7fffff00, "pgld_band_plan_create: bad n_var");
  PGLD_REQUIRE(n_var == 0 || pos, "pgld_band_plan_create: null pos");
  std::vector<double> e;
  int rc = check_plan_args(pos, key_coord, n_var, max_dist, key_mode, edges, n_keys, width, e);
  if (rc) return rc;
  const double* kc = key_coord ? key_coord : pos;
  auto* plan = new pgld_band_plan();
  plan->n_var = n_var;
  plan->max_dist = max_dist;
  plan->key_mode = key_mode;
  plan->n_keys = n_keys;
  plan->n_keys_local = n_keys;
  plan->origin = origin;
  plan->width = width;
  TileBuild tb;
  std::vector<int64_t> base;
  plan->n_pairs = build_region(pos, kc, n_var, max_dist, key_mode, e, n_keys, origin, width, 0, 0, &base, tb);
  rc = upload_plan(plan, pos, kc, n_var, e, tb, base);
  if (rc) {
    pgld_band_plan_destroy(plan);
    return rc;
  }
  *out = plan;
  return PGLD_OK;
}

int pgld_batch_plan_create(int32_t n_regions, const int64_t* n_var, const double* pos, double max_dist,
                           int32_t key_mode, const double* key_coord, const double* edges, int32_t n_keys,
                           double origin, double width, pgld_band_plan** out) {
  PGLD_REQUIRE(out, "pgld_batch_plan_create: null output");
  *out = nullptr;
  PGLD_REQUIRE(n_regions >= 1 && n_var, "pgld_batch_plan_create: need n_regions >= 1 and n_var");
  PGLD_REQUIRE((int64_t)n_regions * n_keys < 0x7fffffff, "pgld_batch_plan_create: too many keys");
  std::vector<int64_t> rbase(n_regions + 1, 0);
  for (int r = 0; r < n_regions; ++r) {
    PGLD_REQUIRE(n_var[r] >= 0 && n_var[r] < 0x7fffff00, "pgld_batch_plan_create: bad n_var of region %d", r);
    rbase[r + 1] = rbase[r] + n_var[r];
  }
  const int64_t total = rbase[n_regions];
  PGLD_REQUIRE(total == 0 || pos, "pgld_batch_plan_create: null pos");
  std::vector<double> e;
  for (int r = 0; r < n_regions; ++r) {
    int rc = check_plan_args(pos + rbase[r], key_coord ? key_coord + rbase[r] : nullptr, n_var[r], max_dist,
                             key_mode, edges, n_keys, width, e);
    if (rc) return rc;
  }
  const double* kc = key_coord ? key_coord : pos;
  auto* plan = new pgld_band_plan();
  plan->n_var = total;
  plan->max_dist = max_dist;
  plan->key_mode = key_mode;
  plan->n_keys = n_regions * n_keys;
  plan->n_keys_local = n_keys;
  plan->origin = origin;
  plan->width = width;
  plan->n_regions = n_regions;
  plan->h_region_nvar.assign(n_var, n_var + n_regions);
  plan->h_region_base.assign(rbase.begin(), rbase.end() - 1);
  TileBuild tb;
  int64_t n_pairs = 0;
  for (int r = 0; r < n_regions; ++r)
    n_pairs += build_region(pos + rbase[r], kc + rbase[r], n_var[r], max_dist, key_mode, e, n_keys, origin, width, r,
                            r * n_keys, nullptr, tb);
  plan->n_pairs = n_pairs;
  std::vector<int64_t> no_base;
  int rc = upload_plan(plan, pos, kc, total, e, tb, no_base);
  if (rc) {
    pgld_band_plan_destroy(plan);
    return rc;
  }
  *out = plan;
  return PGLD_OK;
}

int pgld_band_plan_info(const pgld_band_plan* plan, int64_t* n_tiles, int64_t* n_pairs, int64_t* n_slots) {
  PGLD_REQUIRE(plan, "pgld_band_plan_info: null plan");
  if (n_tiles) *n_tiles = plan->n_tiles;
  if (n_pairs) *n_pairs = plan->n_pairs;
  if (n_slots) *n_slots = plan->n_slots;
  return PGLD_OK;
}

int pgld_band_plan_destroy(pgld_band_plan* plan) {
  if (!plan) return PGLD_OK;
  void* ptrs[] = {plan->d_pos,      plan->d_kc,       plan->d_edges,     plan->d_tiles,    plan->d_pair_base,
                  plan->d_key_off,  plan->d_key_list, plan->d_partials,  plan->d_regions,  plan->d_order,
                  plan->d_row_off,  plan->d_col_off,  plan->d_col_list,  plan->d_snp_part, plan->d_snp_cnt,
                  plan->d_super,    plan->d_red_lo,   plan->d_red_off,  plan->d_sums_order, plan->d_gstash};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : plan->ev)
    if (e) (void)hipEventDestroy(e);
  delete plan;
  return PGLD_OK;
}

int pgld_band_plan_set_timing(pgld_band_plan* plan, int32_t enable) {
  PGLD_REQUIRE(plan, "pgld_band_plan_set_timing: null plan");
  if (enable && !plan->ev[0]) {
    PGLD_TRY_HIP(hipEventCreate(&plan->ev[0]), "pgld_band_plan_set_timing event");
    PGLD_TRY_HIP(hipEventCreate(&plan->ev[1]), "pgld_band_plan_set_timing event");
  }
  plan->timing = enable != 0;
  return PGLD_OK;
}

int pgld_band_plan_kernel_ms(const pgld_band_plan* plan, float* ms) {
  PGLD_REQUIRE(plan && ms, "pgld_band_plan_kernel_ms: null pointer");
  PGLD_REQUIRE(plan->ev[0], "pgld_band_plan_kernel_ms: timing was never enabled on this plan");
  PGLD_TRY_HIP(hipEventSynchronize(plan->ev[1]), "pgld_band_plan_kernel_ms sync");
  PGLD_TRY_HIP(hipEventElapsedTime(ms, plan->ev[0], plan->ev[1]), "pgld_band_plan_kernel_ms elapsed");
  return PGLD_OK;
}

static int fill_pop_args(int32_t n_pops, const int8_t* const* planes, const int32_t* const* marg,
                         const int64_t* k_pad, const int32_t* n_samples, PopArgs& pa) {
  pa = PopArgs{};
  for (int p = 0; p < n_pops; ++p) {
    PGLD_REQUIRE(planes[p] && marg[p], "band: null planes/marg for pop %d", p);
    PGLD_REQUIRE(k_pad[p] % KSTEP == 0 && k_pad[p] >= KSTEP && k_pad[p] >= n_samples[p] && n_samples[p] >= 0,
                 "band: k_pad must be a positive multiple of 128 >= n_samples (pop %d)", p);
    PGLD_REQUIRE(n_samples[p] <= 32767, "band: at most 32767 samples per population");
    PGLD_REQUIRE(k_pad[p] <= MAX_K_PAD, "band: k_pad must be at most 32768 (pop %d)", p);
    pa.planes[p] = planes[p];
    pa.marg[p] = marg[p];
    pa.k_pad[p] = k_pad[p];
    pa.n_samples[p] = n_samples[p];
  }
  return PGLD_OK;
}

}  // extern "C"

namespace pgld {

static SumsLauncher launcher_for(int32_t family, int32_t n_pops, bool miss, int32_t stat_set, bool fp4 = false) {
  if (fp4) {
    if (family == PGLD_FAMILY_HAP)
      return miss ? pick_launcher<PGLD_FAMILY_HAP, true, true>(n_pops, stat_set)
                  : pick_launcher<PGLD_FAMILY_HAP, false, true>(n_pops, stat_set);
    return miss ? pick_launcher<PGLD_FAMILY_GENO, true, true>(n_pops, stat_set)
                : pick_launcher<PGLD_FAMILY_GENO, false, true>(n_pops, stat_set);
  }
  if (family == PGLD_FAMILY_HAP)
    return miss ? pick_launcher<PGLD_FAMILY_HAP, true, false>(n_pops, stat_set)
                : pick_launcher<PGLD_FAMILY_HAP, false, false>(n_pops, stat_set);
  return miss ? pick_launcher<PGLD_FAMILY_GENO, true, false>(n_pops, stat_set)
              : pick_launcher<PGLD_FAMILY_GENO, false, false>(n_pops, stat_set);
}

// LDS bytes of one fused workgroup, or 0 when the tile's keys x statistics do
// not fit (160 KiB of LDS per workgroup on gfx950, minus the static arrays).
static size_t sums_lds_bytes(const pgld_band_plan* plan, int32_t family, int32_t n_pops, int32_t stat_set, bool miss,
                             int n_stats) {
  const int ns_total = n_stats + 1;
  const int nc = family == PGLD_FAMILY_HAP ? 4 : 9;
  const size_t red_bytes = (((size_t)4 * plan->max_tile_slots * ns_total * 8 + 15) / 16) * 16;
  // per-population statistics need no parked counts (each population's
  // epilogue follows its own products)
  const size_t stash_bytes = stat_set == PGLD_SET_SELF || gstash_set(family, n_pops, stat_set)
                                 ? 0
                                 : (size_t)(n_pops - 1) * 4 * nc * 16 * 64 * 2;
  const int n_planes_used = family == PGLD_FAMILY_HAP ? (miss ? 2 : 1) : (miss ? 3 : 2);
  const size_t op_bytes = (size_t)2 * n_planes_used * BT * OP_ROW;
  if (red_bytes > (size_t)RED_LDS_BYTES || red_bytes + stash_bytes + op_bytes + STATIC_LDS_BYTES > 160 * 1024)
    return 0;
  return red_bytes + stash_bytes + op_bytes;
}

// Workspace: one partial row per (tile, key) slot, then one per reduction
// chunk when the plan reduces in two levels.
static int ensure_partials(pgld_band_plan* pl, int64_t ns_total) {
  const size_t need = (size_t)(std::max<int64_t>(1, pl->n_slots) + pl->n_red_chunks) * ns_total;
  if (pl->partials_cap >= need) return PGLD_OK;
  if (pl->d_partials) PGLD_TRY_HIP(hipFree(pl->d_partials), "band workspace free");
  pl->d_partials = nullptr;
  pl->partials_cap = 0;
  PGLD_TRY_HIP(hipMalloc((void**)&pl->d_partials, need * sizeof(double)), "band workspace");
  pl->partials_cap = need;
  return PGLD_OK;
}

// The fused pass + per-key reduction.  pa is used by single-band plans;
// region batches read their populations from the bound region table.  With
// allow_alloc false (region batches, which may be captured into a graph) the
// workspace must already be large enough.
// A caller-supplied workspace (work, work_cap doubles) replaces the plan's.
static int run_sums(pgld_band_plan* pl, int32_t family, int32_t n_pops, const PopArgs& pa, int32_t with_missing,
                    int32_t stat_set, double* sums, double* key_pairs, hipStream_t s, bool allow_alloc,
                    bool fp4 = false, double* work = nullptr, size_t work_cap = 0) {
  const bool miss = with_missing != 0;
  SumsLauncher L = launcher_for(family, n_pops, miss, stat_set, fp4);
  if (!L) {
    set_error("band sums: no fused kernel for family %d with %d populations and stat set %d", family, n_pops,
              stat_set);
    return PGLD_EUNSUPPORTED;
  }
  const int n_stats = n_stats_of(family, n_pops, stat_set);
  const int ns_total = n_stats + 1;
  const size_t lds = sums_lds_bytes(pl, family, n_pops, stat_set, miss, n_stats);
  if (!lds) {
    set_error("band sums: a tile touches %d keys; %d statistics x that many keys exceed LDS", pl->max_tile_slots,
              n_stats);
    return PGLD_EUNSUPPORTED;
  }
  int16_t* gstash = nullptr;
  if (gstash_set(family, n_pops, stat_set) && pl->n_tiles > 0) {
    if (pl->n_regions > 0 || work || !allow_alloc) {
      set_error("band sums: the %d-population genotype moments need the plan's own workspace (single bands)",
                n_pops);
      return PGLD_EUNSUPPORTED;
    }
    const size_t elems = (size_t)gstash_elems(n_pops) * (size_t)std::min<int64_t>(GSTASH_TILES, pl->n_tiles);
    if (pl->gstash_cap < elems) {
      if (pl->d_gstash) PGLD_TRY_HIP(hipFree(pl->d_gstash), "band count workspace free");
      pl->d_gstash = nullptr;
      pl->gstash_cap = 0;
      PGLD_TRY_HIP(hipMalloc((void**)&pl->d_gstash, elems * sizeof(int16_t)), "band count workspace");
      pl->gstash_cap = elems;
    }
    gstash = pl->d_gstash;
  }
  const size_t need = (size_t)(std::max<int64_t>(1, pl->n_slots) + pl->n_red_chunks) * ns_total;
  double* partials = pl->d_partials;
  if (work) {
    if (work_cap < need) {
      set_error("band sums: the plan needs %zu workspace doubles, the fixed workspace holds %zu", need, work_cap);
      return PGLD_ECAPACITY;
    }
    partials = work;
  } else if (allow_alloc) {
    int rc = ensure_partials(pl, ns_total);
    if (rc) return rc;
    partials = pl->d_partials;
  } else if (pl->partials_cap < need) {
    set_error("band sums: workspace not bound for this statistic set");
    return PGLD_EINVAL;
  }
  // k_band_reduce writes every key (zero where no tile touches it), so the
  // outputs need no clearing and the call records no memset node when
  // captured into a graph.
  BandArgs ba{};
  ba.tiles = reinterpret_cast<const BandTile*>(pl->d_tiles);
  ba.pos = pl->d_pos;
  ba.kc = pl->d_kc;
  ba.edges = pl->d_edges;
  ba.n_var = pl->n_var;
  ba.max_dist = pl->max_dist;
  ba.origin = pl->origin;
  ba.width = pl->width;
  ba.key_mode = pl->key_mode;
  ba.n_keys = pl->n_keys_local;
  ba.partials = partials;
  ba.order = pl->n_regions == 0 ? pl->d_sums_order : nullptr;
  ba.gstash = gstash;
  if (pl->n_tiles > 0) {
    if (pl->timing) PGLD_TRY_HIP(hipEventRecord(pl->ev[0], s), "band sums event");
    L(pl, ba, pa, lds, s);
    int rc = check_launch("k_band_sums");
    if (rc) return rc;
    if (pl->timing) PGLD_TRY_HIP(hipEventRecord(pl->ev[1], s), "band sums event");
  }
  if (pl->n_red_chunks > 0) {
    double* part2 = partials + (size_t)std::max<int64_t>(1, pl->n_slots) * ns_total;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)pl->n_red_chunks), dim3(256), 0, s, partials,
                       pl->d_key_list, pl->d_red_lo, ns_total, part2);
    int rc = check_launch("k_reduce_chunks");
    if (rc) return rc;
    hipLaunchKernelGGL(k_band_reduce, dim3((unsigned)pl->n_keys), dim3(256), 0, s, part2, pl->d_red_off,
                       (const int64_t*)nullptr, ns_total, n_stats, sums, key_pairs);
    return check_launch("k_band_reduce");
  }
  hipLaunchKernelGGL(k_band_reduce, dim3((unsigned)pl->n_keys), dim3(256), 0, s, partials, pl->d_key_off,
                     pl->d_key_list, ns_total, n_stats, sums, key_pairs);
  return check_launch("k_band_reduce");
}

}  // namespace pgld

extern "C" {

int pgld_band_sums(const pgld_band_plan* plan, int32_t family, int32_t n_pops, const int8_t* const* planes,
                   const int32_t* const* marg, const int64_t* k_pad, const int32_t* n_samples, int32_t with_missing,
                   int32_t stat_set, double* sums, double* key_pairs, void* stream) {
  PGLD_REQUIRE(plan && planes && marg && k_pad && n_samples && sums, "pgld_band_sums: null pointer");
  PGLD_REQUIRE(plan->n_regions == 0, "pgld_band_sums: region batches run through pgld_batch_sums");
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_band_sums: bad family");
  PGLD_REQUIRE(n_pops >= 1 && n_pops <= 4, "pgld_band_sums: n_pops must be 1..4");
  PGLD_REQUIRE(stat_set == PGLD_SET_MOMENTS || stat_set == PGLD_SET_SELF, "pgld_band_sums: bad stat_set");
  PopArgs pa;
  int rc = fill_pop_args(n_pops, planes, marg, k_pad, n_samples, pa);
  if (rc) return rc;
  return run_sums(const_cast<pgld_band_plan*>(plan), family, n_pops, pa, with_missing, stat_set, sums, key_pairs,
                  as_stream(stream), true);
}

int pgld_band_sums_fp4(const pgld_band_plan* plan, int32_t family, int32_t n_pops, const int8_t* const* planes,
                       const int32_t* const* marg, const int64_t* k_pad, const int32_t* n_samples,
                       int32_t with_missing, int32_t stat_set, double* sums, double* key_pairs, void* stream) {
  PGLD_REQUIRE(plan && planes && marg && k_pad && n_samples && sums, "pgld_band_sums_fp4: null pointer");
  PGLD_REQUIRE(plan->n_regions == 0, "pgld_band_sums_fp4: region batches take int8 planes");
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_band_sums_fp4: bad family");
  PGLD_REQUIRE(n_pops >= 1 && n_pops <= 4, "pgld_band_sums_fp4: n_pops must be 1..4");
  PGLD_REQUIRE(stat_set == PGLD_SET_MOMENTS || stat_set == PGLD_SET_SELF, "pgld_band_sums_fp4: bad stat_set");
  for (int p = 0; p < n_pops; ++p)
    PGLD_REQUIRE(k_pad[p] % 256 == 0, "pgld_band_sums_fp4: k_pad must be a multiple of 256 (FP4 planes)");
  PopArgs pa;
  int rc = fill_pop_args(n_pops, planes, marg, k_pad, n_samples, pa);
  if (rc) return rc;
  return run_sums(const_cast<pgld_band_plan*>(plan), family, n_pops, pa, with_missing, stat_set, sums, key_pairs,
                  as_stream(stream), true, true);
}

// ---------------------------------------------------------------------------
// Allocation-free fused sums: the partial-row workspace is a fixed device
// array (one copy per GPU, resident until the process ends) instead of the
// plan's hipMalloc'd buffer.  Calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace pgld {
constexpr size_t NOALLOC_WORK_DOUBLES = (size_t)1 << 22;  // 32 MiB
static __device__ __align__(256) double scratch[NOALLOC_WORK_DOUBLES];
constexpr int MAX_DEVICES = 64;
static double* g_scratch_ptr[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Device address of this GPU's copy of the array, looked up once per device
// with that device current (the array's name is a host-side stand-in).
static int scratch_on_current_device(double** out) {
  int dev = 0;
  PGLD_TRY_HIP(hipGetDevice(&dev), "pgld_band_sums_fp4_noalloc: hipGetDevice");
  PGLD_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "pgld_band_sums_fp4_noalloc: device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    PGLD_TRY_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "pgld_band_sums_fp4_noalloc: workspace address");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return PGLD_OK;
}
}  // namespace pgld

extern "C" {

int pgld_band_sums_fp4_noalloc(const pgld_band_plan* plan, int32_t family, int32_t n_pops,
                               const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                               const int32_t* n_samples, int32_t with_missing, int32_t stat_set, double* sums,
                               double* key_pairs, void* stream) {
  PGLD_REQUIRE(plan && planes && marg && k_pad && n_samples && sums, "pgld_band_sums_fp4_noalloc: null pointer");
  PGLD_REQUIRE(plan->n_regions == 0, "pgld_band_sums_fp4_noalloc: single bands only");
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_band_sums_fp4_noalloc: bad family");
  PGLD_REQUIRE(n_pops >= 1 && n_pops <= 4, "pgld_band_sums_fp4_noalloc: n_pops must be 1..4");
  PGLD_REQUIRE(stat_set == PGLD_SET_MOMENTS || stat_set == PGLD_SET_SELF, "pgld_band_sums_fp4_noalloc: bad stat_set");
  for (int p = 0; p < n_pops; ++p)
    PGLD_REQUIRE(k_pad[p] % 256 == 0, "pgld_band_sums_fp4_noalloc: k_pad must be a multiple of 256 (FP4 planes)");
  PopArgs pa;
  int rc = fill_pop_args(n_pops, planes, marg, k_pad, n_samples, pa);
  if (rc) return rc;
  const int64_t ns_total = n_stats_of(family, n_pops, stat_set) + 1;
  const size_t need = (size_t)(std::max<int64_t>(1, plan->n_slots) + plan->n_red_chunks) * ns_total;
  if (need > NOALLOC_WORK_DOUBLES) {
    set_error("pgld_band_sums_fp4_noalloc: the plan needs %zu workspace doubles, more than the fixed %zu "
              "(use pgld_band_sums_fp4)", need, NOALLOC_WORK_DOUBLES);
    return PGLD_ECAPACITY;
  }
  double* work = nullptr;
  rc = scratch_on_current_device(&work);
  if (rc) return rc;
  return run_sums(const_cast<pgld_band_plan*>(plan), family, n_pops, pa, with_missing, stat_set, sums, key_pairs,
                  as_stream(stream), false, true, work, NOALLOC_WORK_DOUBLES);
}

static int batch_bind(pgld_band_plan* plan, int32_t family, int32_t n_pops, int32_t stat_set,
                      const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                      const int32_t* n_samples, bool fp4) {
  PGLD_REQUIRE(plan && planes && marg && k_pad && n_samples, "pgld_batch_bind: null pointer");
  PGLD_REQUIRE(plan->n_regions > 0, "pgld_batch_bind: not a region batch plan");
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_batch_bind: bad family");
  PGLD_REQUIRE(n_pops >= 1 && n_pops <= 4, "pgld_batch_bind: n_pops must be 1..4");
  PGLD_REQUIRE(stat_set == PGLD_SET_MOMENTS || stat_set == PGLD_SET_SELF, "pgld_batch_bind: bad stat_set");
  PGLD_REQUIRE(launcher_for(family, n_pops, true, stat_set), "pgld_batch_bind: no fused kernel for this set");
  if (fp4)
    for (int64_t e = 0; e < (int64_t)plan->n_regions * n_pops; ++e)
      PGLD_REQUIRE(k_pad[e] % 256 == 0, "pgld_batch_bind_fp4: k_pad must be a multiple of 256 (FP4 planes)");
  const int R = plan->n_regions;
  std::vector<RegionDesc> rd(R);
  std::vector<double> wsum(R, 0.0);  // MFMA work per tile pair of the region (sum of k_pad)
  for (int r = 0; r < R; ++r) {
    int rc = fill_pop_args(n_pops, planes + (size_t)r * n_pops, marg + (size_t)r * n_pops, k_pad + (size_t)r * n_pops,
                           n_samples + (size_t)r * n_pops, rd[r].pa);
    if (rc) return rc;
    rd[r].n_var = plan->h_region_nvar[r];
    rd[r].snp_base = plan->h_region_base[r];
    rd[r].key_base = r * plan->n_keys_local;
    for (int p = 0; p < n_pops; ++p) wsum[r] += (double)k_pad[(size_t)r * n_pops + p];
  }
  // tile orders: cost-descending (longest first over the persistent grid),
  // natural, and L2-blocked per XCD
  const int64_t T = plan->n_tiles;
  std::vector<double> cost(T);
  std::vector<int32_t> order;
  order.reserve(3 * T + 64);
  for (int64_t t = 0; t < T; ++t) {
    const BandTileHost& h = plan->h_tiles[t];
    const double pairs = h.diag ? 0.5 * h.rows * h.cols : (double)h.rows * h.cols;
    cost[t] = (double)h.rows * h.cols * wsum[h.region] + 64.0 * pairs;  // K loop + epilogue
    order.push_back((int32_t)t);
  }
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
  for (int64_t t = 0; t < T; ++t) order.push_back((int32_t)t);
  const std::vector<int32_t> blocked = blocked_order(plan->h_tiles, wsum, cost);
  order.insert(order.end(), blocked.begin(), blocked.end());
  plan->order_off[0] = 0;
  plan->order_off[1] = T;
  plan->order_off[2] = 2 * T;
  plan->order_len[0] = plan->order_len[1] = T;
  plan->order_len[2] = (int64_t)blocked.size();
  if (order.empty()) order.push_back(-1);
  if (plan->d_order) {
    (void)hipFree(plan->d_order);
    plan->d_order = nullptr;
  }
  PGLD_TRY_HIP(hipMalloc((void**)&plan->d_order, order.size() * sizeof(int32_t)), "batch order");
  PGLD_TRY_HIP(hipMemcpy(plan->d_order, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice),
               "batch order upload");
  if (!plan->d_regions) PGLD_TRY_HIP(hipMalloc(&plan->d_regions, (size_t)R * sizeof(RegionDesc)), "batch regions");
  PGLD_TRY_HIP(hipMemcpy(plan->d_regions, rd.data(), (size_t)R * sizeof(RegionDesc), hipMemcpyHostToDevice),
               "batch regions upload");
  int rc = ensure_partials(plan, n_stats_of(family, n_pops, stat_set) + 1);
  if (rc) return rc;
  plan->bound_family = family;
  plan->bound_fp4 = fp4;
  plan->bound_pops = n_pops;
  plan->bound_set = stat_set;
  double s_sum = 0.0;
  for (int64_t e = 0; e < (int64_t)R * n_pops; ++e) s_sum += (double)n_samples[e];
  plan->bound_mean_samples = s_sum / ((double)R * n_pops);
  plan->bound_max_samples = *std::max_element(n_samples, n_samples + (size_t)R * n_pops);
  return PGLD_OK;
}

int pgld_batch_bind(pgld_band_plan* plan, int32_t family, int32_t n_pops, int32_t stat_set,
                    const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                    const int32_t* n_samples) {
  return batch_bind(plan, family, n_pops, stat_set, planes, marg, k_pad, n_samples, false);
}

int pgld_batch_bind_fp4(pgld_band_plan* plan, int32_t family, int32_t n_pops, int32_t stat_set,
                        const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                        const int32_t* n_samples) {
  return batch_bind(plan, family, n_pops, stat_set, planes, marg, k_pad, n_samples, true);
}

int pgld_batch_sums(pgld_band_plan* plan, int32_t with_missing, double* sums, double* key_pairs, void* stream) {
  PGLD_REQUIRE(plan && sums, "pgld_batch_sums: null pointer");
  PGLD_REQUIRE(plan->n_regions > 0 && plan->bound_family >= 0, "pgld_batch_sums: plan is not a bound region batch");
  PopArgs none{};
  return run_sums(plan, plan->bound_family, plan->bound_pops, none, with_missing, plan->bound_set, sums, key_pairs,
                  as_stream(stream), false, plan->bound_fp4);
}

int pgld_batch_schedule(pgld_band_plan* plan, int32_t grid, int32_t order, int32_t xcd_split,
                        int32_t kernel_variant) {
  PGLD_REQUIRE(plan && plan->n_regions > 0, "pgld_batch_schedule: not a region batch plan");
  PGLD_REQUIRE(grid >= 0 && order >= 0 && order <= 2, "pgld_batch_schedule: bad grid or order");
  PGLD_REQUIRE(kernel_variant >= 0 && kernel_variant <= 3, "pgld_batch_schedule: kernel variant must be 0 .. 3");
  plan->grid = grid;
  plan->order_mode = order;
  plan->xcd_split = xcd_split != 0;
  plan->kvar = kernel_variant;
  return PGLD_OK;
}

}  // extern "C"

namespace pgld {

// Autotuner results, kept for the process and in a text file (one line per
// shape class: "key grid order xcd_split kernel_variant ms") so that later
// processes skip the timing.  The file is PGLD_TUNE_CACHE if set (empty: no
// file), else $XDG_CACHE_HOME/pgld/batch_tune.txt or ~/.cache/pgld/batch_tune.txt.
struct TuneEntry {
  int32_t grid, order, xcd, kvar;
  float ms;
};
static std::mutex g_tune_mu;
static std::map<std::string, TuneEntry> g_tune;
static bool g_tune_loaded = false;

static std::string tune_file() {
  const char* e = getenv("PGLD_TUNE_CACHE");
  if (e) return std::string(e);
  const char* x = getenv("XDG_CACHE_HOME");
  std::string dir;
  if (x && *x) {
    dir = x;
  } else {
    const char* h = getenv("HOME");
    if (!h || !*h) return "";
    dir = std::string(h) + "/.cache";
  }
  return dir + "/pgld/batch_tune.txt";
}

static void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const std::string path = tune_file();
  if (path.empty()) return;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  char key[256];
  TuneEntry t;
  while (fscanf(f, "%255s %d %d %d %d %f", key, &t.grid, &t.order, &t.xcd, &t.kvar, &t.ms) == 6) {
    if (t.grid >= 0 && t.order >= 0 && t.order <= 2 && t.kvar >= 0 && t.kvar <= 3) g_tune[key] = t;
  }
  fclose(f);
}

// Rewrites the whole file (write to a temporary, then rename); failures to
// write only lose the persistence.
static void tune_store_locked() {
  const std::string path = tune_file();
  if (path.empty()) return;
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos && slash > 0) {
    const std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  }
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return;
  for (const auto& kv : g_tune)
    fprintf(f, "%s %d %d %d %d %.4f\n", kv.first.c_str(), kv.second.grid, kv.second.order, kv.second.xcd,
            kv.second.kvar, kv.second.ms);
  const bool ok = fclose(f) == 0;
  if (ok) (void)rename(tmp.c_str(), path.c_str());
  else (void)remove(tmp.c_str());
}

}  // namespace pgld

extern "C" {

// Autotuner: times every schedule candidate of the persistent launch on this
// plan (one run each after one warm-up, HIP events) and keeps the fastest.
// Candidates: grid (one workgroup per tile, 1x / 2x / 4x the CU count) x tile
// order (cost-descending, natural, natural with an XCD split, L2-blocked), and
// for the kernels that have them the kernel variants (K loop streamed across
// populations or per population, fast-path rows unrolled or rolled), all with
// bit-identical sums.  The choice is cached per shape class -- GPU
// architecture and CU count, family, planes, populations, statistic set,
// missing data, log2 of the tile count, of the mean region size, of the pairs
// per SNP (the band width) and of the mean population size -- in the process
// and on disk.
int pgld_batch_autotune(pgld_band_plan* plan, int32_t with_missing, double* sums, double* key_pairs, void* stream,
                        int32_t* grid_out, int32_t* order_out, int32_t* xcd_out, int32_t* kvar_out,
                        float* best_ms) {
  PGLD_REQUIRE(plan && sums, "pgld_batch_autotune: null pointer");
  PGLD_REQUIRE(plan->n_regions > 0 && plan->bound_family >= 0, "pgld_batch_autotune: plan is not a bound batch");
  hipStream_t s = as_stream(stream);
  int dev = 0, cus = 256;
  PGLD_TRY_HIP(hipGetDevice(&dev), "pgld_batch_autotune device");
  PGLD_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "pgld_batch_autotune CUs");
  hipDeviceProp_t prop;
  PGLD_TRY_HIP(hipGetDeviceProperties(&prop, dev), "pgld_batch_autotune properties");
  std::string arch(prop.gcnArchName);
  for (char& c : arch)
    if (c == ' ' || c == ':') c = '_';
  double mean_k = 0.0;
  for (int64_t r = 0; r < plan->n_regions; ++r) mean_k += (double)plan->h_region_nvar[r];
  mean_k /= (double)plan->n_regions;
  const double pairs_per_snp = plan->n_var > 0 ? (double)plan->n_pairs / (double)plan->n_var : 0.0;
  const bool variants = plan->bound_family == PGLD_FAMILY_HAP && plan->bound_set == PGLD_SET_SELF &&
                        plan->bound_pops > 1 && plan->bound_fp4;
  char key[256];
  snprintf(key, sizeof(key), "%s/cu%d/f%d/fp4%d/P%d/set%d/miss%d/t%d/k%d/w%d/s%d", arch.c_str(), cus,
           plan->bound_family, (int)plan->bound_fp4, plan->bound_pops, plan->bound_set, with_missing != 0,
           (int)std::log2((double)std::max<int64_t>(1, plan->n_tiles)), (int)std::log2(1.0 + mean_k),
           (int)std::log2(1.0 + pairs_per_snp), (int)std::log2(1.0 + plan->bound_mean_samples));
  auto publish = [&](const TuneEntry& t, float ms) {
    plan->grid = t.grid;
    plan->order_mode = t.order;
    plan->xcd_split = t.xcd != 0;
    plan->kvar = t.kvar;
    if (grid_out) *grid_out = t.grid;
    if (order_out) *order_out = t.order;
    if (xcd_out) *xcd_out = t.xcd;
    if (kvar_out) *kvar_out = t.kvar;
    if (best_ms) *best_ms = ms;
  };
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      publish(it->second, -1.0f);
      return PGLD_OK;
    }
  }
  std::vector<TuneEntry> cands;
  for (int32_t g : {0, cus, 2 * cus, 4 * cus}) {
    cands.push_back({g, 0, 0, 0, 0.f});  // cost-descending, strided
    cands.push_back({g, 1, 0, 0, 0.f});  // natural, strided
    cands.push_back({g, 1, 1, 0, 0.f});  // natural, one contiguous eighth per XCD
    cands.push_back({g, 2, 0, 0, 0.f});  // 8 x 8 tile blocks dealt to the XCDs
  }
  if (variants)
    for (int32_t kv : {1, 2, 3}) {
      cands.push_back({0, 2, 0, kv, 0.f});
      cands.push_back({0, 0, 0, kv, 0.f});
    }
  hipEvent_t e0, e1;
  PGLD_TRY_HIP(hipEventCreate(&e0), "autotune event");
  PGLD_TRY_HIP(hipEventCreate(&e1), "autotune event");
  float best = 1e30f;
  TuneEntry win = cands[0];
  int rc = PGLD_OK;
  for (size_t c = 0; c < cands.size() && !rc; ++c) {
    plan->grid = cands[c].grid;
    plan->order_mode = cands[c].order;
    plan->xcd_split = cands[c].xcd != 0;
    plan->kvar = cands[c].kvar;
    if (c == 0) rc = pgld_batch_sums(plan, with_missing, sums, key_pairs, stream);  // warm-up
    if (!rc) rc = check_hip(hipEventRecord(e0, s), "autotune event");
    if (!rc) rc = pgld_batch_sums(plan, with_missing, sums, key_pairs, stream);
    if (!rc) rc = check_hip(hipEventRecord(e1, s), "autotune event");
    if (!rc) rc = check_hip(hipEventSynchronize(e1), "autotune sync");
    float ms = 0.0f;
    if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "autotune elapsed");
    if (!rc && ms < best) {
      best = ms;
      win = cands[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  win.ms = best;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = win;
    tune_store_locked();
  }
  publish(win, best);
  return PGLD_OK;
}

// ---- HIP graphs --------------------------------------------------------------
struct pgld_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int pgld_graph_begin(void* stream) {
  PGLD_REQUIRE(stream, "pgld_graph_begin: capture needs a created stream, not the null stream");
  PGLD_TRY_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal), "pgld_graph_begin");
  return PGLD_OK;
}

int pgld_graph_end(void* stream, pgld_graph** out) {
  PGLD_REQUIRE(stream && out, "pgld_graph_end: null argument");
  *out = nullptr;
  auto* g = new pgld_graph();
  int rc = check_hip(hipStreamEndCapture(as_stream(stream), &g->graph), "pgld_graph_end capture");
  if (!rc) rc = check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "pgld_graph_end instantiate");
  if (rc) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return rc;
  }
  *out = g;
  return PGLD_OK;
}

int pgld_graph_launch(pgld_graph* g, void* stream) {
  PGLD_REQUIRE(g && g->exec, "pgld_graph_launch: null graph");
  PGLD_TRY_HIP(hipGraphLaunch(g->exec, as_stream(stream)), "pgld_graph_launch");
  return PGLD_OK;
}

int pgld_graph_destroy(pgld_graph* g) {
  if (!g) return PGLD_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return PGLD_OK;
}

static int band_counts(const pgld_band_plan* plan, int32_t family, const int8_t* planes, bool popc, const int32_t* marg,
                     int64_t k_pad, int32_t n_samples, int32_t with_missing, const int64_t* pair_base,
                     int32_t* counts, void* stream) {
  PGLD_REQUIRE(plan && planes && marg && counts, "pgld_band_counts: null pointer");
  PGLD_REQUIRE(plan->n_regions == 0, "pgld_band_counts: not defined for region batches");
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_band_counts: bad family");
  PopArgs pa;
  int rc = fill_pop_args(1, &planes, &marg, &k_pad, &n_samples, pa);
  if (rc) return rc;
  if (plan->n_tiles == 0) return PGLD_OK;
  BandArgs ba{};
  ba.tiles = reinterpret_cast<const BandTile*>(plan->d_tiles);
  ba.pos = plan->d_pos;
  ba.n_var = plan->n_var;
  ba.max_dist = plan->max_dist;
  ba.counts_out = counts;
  ba.pair_base = pair_base ? pair_base : plan->d_pair_base;
  dim3 g((unsigned)plan->n_tiles), b(BAND_THREADS);
  hipStream_t s = as_stream(stream);
  const bool miss = with_missing != 0;
  const int npl = family == PGLD_FAMILY_HAP ? (miss ? 2 : 1) : (miss ? 3 : 2);
  const size_t lds = (size_t)2 * npl * BT * OP_ROW;
#define PGLD_COUNTS(F, M, B) hipLaunchKernelGGL((k_band_counts<F, M, B>), g, b, lds, s, ba, pa)
  if (family == PGLD_FAMILY_HAP) {
    if (popc) { if (miss) PGLD_COUNTS(PGLD_FAMILY_HAP, true, true); else PGLD_COUNTS(PGLD_FAMILY_HAP, false, true); }
    else { if (miss) PGLD_COUNTS(PGLD_FAMILY_HAP, true, false); else PGLD_COUNTS(PGLD_FAMILY_HAP, false, false); }
  } else {
    if (popc) { if (miss) PGLD_COUNTS(PGLD_FAMILY_GENO, true, true); else PGLD_COUNTS(PGLD_FAMILY_GENO, false, true); }
    else { if (miss) PGLD_COUNTS(PGLD_FAMILY_GENO, true, false); else PGLD_COUNTS(PGLD_FAMILY_GENO, false, false); }
  }
#undef PGLD_COUNTS
  return check_launch("k_band_counts");
}

int pgld_band_counts(const pgld_band_plan* plan, int32_t family, const int8_t* planes, const int32_t* marg,
                     int64_t k_pad, int32_t n_samples, int32_t with_missing, const int64_t* pair_base,
                     int32_t* counts, void* stream) {
  return band_counts(plan, family, planes, false, marg, k_pad, n_samples, with_missing, pair_base, counts, stream);
}

int pgld_band_counts_bits(const pgld_band_plan* plan, int32_t family, const uint32_t* bits, const int32_t* marg,
                          int64_t k_pad, int32_t n_samples, int32_t with_missing, const int64_t* pair_base,
                          int32_t* counts, void* stream) {
  return band_counts(plan, family, reinterpret_cast<const int8_t*>(bits), true, marg, k_pad, n_samples, with_missing,
                     pair_base, counts, stream);
}

int pgld_pack_bits(const int8_t* planes, int64_t n_planes, int64_t n_var, int64_t k_pad, uint32_t* bits,
                   void* stream) {
  PGLD_REQUIRE(planes && bits, "pgld_pack_bits: null pointer");
  PGLD_REQUIRE(n_planes >= 1 && n_var >= 0 && k_pad % 128 == 0, "pgld_pack_bits: bad shape");
  const int64_t words = bit_row_bytes(k_pad) / 4, n = n_planes * n_var * words;
  if (n == 0) return PGLD_OK;
  hipLaunchKernelGGL(k_pack_bits, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, as_stream(stream), planes,
                     n_planes * n_var, k_pad, bits, words);
  return check_launch("k_pack_bits");
}

}  // extern "C"
