// Dense / banded r^2 through the int8-MFMA tile path.
//
// For every pair (i < j) of a band plan the haplotype counts come from the same
// LDS-staged MFMA products as the fused LD sums (block_products), and r^2 is
// evaluated from the exact integer counts:
//
//   n11 = h_i.h_j, nA = h_i.v_j, nB = v_i.h_j, n = v_i.v_j   (with missing data)
//   n11 = h_i.h_j, nA = sum h_i, nB = sum h_j, n = S          (without)
//   diff = n * n11 - nA * nB          (= n^2 D, an exact int64)
//   r^2  = (diff * diff) / ((nA (n - nA)) * (nB (n - nB)))   in FP64,
//
// NaN when the denominator is 0 (a site monomorphic among the pair's valid
// samples), the convention of ld_statistics.r_squared (ld_statistics.py:
// 194-217).  Outputs, each optional: a per-key histogram of r^2 over [0, 1]
// with n_hist equal bins plus one column counting NaN pairs (integer, so the
// result is exact and order-free), a sparse list of the pairs with r^2 >
// thresh (FP32 values), and r^2 of every band pair in the plan's (i, j) order
// (FP64).
#include "band_core.h"

namespace pgld {

struct R2Args {
  const BandTile* tiles;
  const double* pos;
  const double* kc;
  const double* edges;
  const int64_t* pair_base;
  int64_t n_var;
  double max_dist, origin, width;
  int32_t key_mode, n_keys, n_hist;
  double thresh;
  unsigned long long* hist;  // (n_keys, n_hist + 1) or null
  int32_t* sp_i;             // sparse list or null
  int32_t* sp_j;
  float* sp_r2;
  int64_t sp_cap;
  unsigned long long* sp_count;
  double* r2_out;        // per band pair or null
  double* snp_part;      // per tile: 64 row sums then 64 column sums, or null
  int32_t* snp_cnt;      // per tile: 64 row counts then 64 column counts
};

// Per-pair estimators of the r^2 family
enum R2Est : int32_t {
  EST_R2 = 0,        // r^2 from the pair's own counts (ld_statistics.r_squared); NaN where undefined
  EST_R2_NAIVE = 1,  // frequency r^2 with per-site p over all valid samples (_tile_r2_naive); 0 where undefined
  EST_SIGMA_D2 = 2,  // unbiased D^2 / pi_2 (_tile_sigma_d2); 0 unless n >= 4 and pi_2 != 0
  EST_RH_R = 3,      // genotypes: signed Rogers-Huff r of 0/1/2 dosages; NaN where a site is constant
  EST_DIP_R2 = 4,    // genotypes: dosage r^2 over valid samples, site means (_r2_matrix_diploid); 0 where undefined
  EST_D_NAIVE = 5,   // signed D = p_AB - p_i p_j with the frequencies of EST_R2_NAIVE (pairwise_LD_v)
};

constexpr int R2_MAX_HIST_LDS = 8192;  // ints of per-tile histogram kept in LDS

// r^2-family value of a haplotype pair from its counts (n11 = h_i.h_j, nA, nB
// the pair-valid allele counts, n the pair-valid sample count); is_nan marks
// an undefined r^2, sval / scnt the pair's contribution to per-SNP sums.
// Counts are below 2^15 (at most 32767 samples per population), so the
// integer products below are exact in 24-bit multiplies and 32-bit results.
template <bool MISSING, int EST>
__device__ __forceinline__ double hap_pair_value(int32_t n11, int32_t nA, int32_t nB, int32_t n,
                                                 const int32_t* __restrict__ marg, int64_t n_var, int64_t i,
                                                 int64_t j, int32_t S, bool& is_nan, double& sval, int& scnt) {
  double r2;
  is_nan = false;
  if constexpr (EST == EST_R2) {
    const int32_t diff = __mul24(n, n11) - __mul24(nA, nB);
    const double dd = (double)diff;
    const double num = dd * dd;
    const double den = (double)__mul24(nA, n - nA) * (double)__mul24(nB, n - nB);
    is_nan = !(den > 0.0);
    r2 = den > 0.0 ? num / den : __builtin_nan("");
    sval = is_nan ? 0.0 : r2;
    scnt = !is_nan;
  } else if constexpr (EST == EST_R2_NAIVE || EST == EST_D_NAIVE) {
#pragma clang fp contract(off)
    // per-site frequencies over every valid sample of the site, joint
    // frequency over the pair's valid samples (the reference's order of
    // operations: p_AB - p_i p_j, then D^2 / (p_i q_i p_j q_j))
    const double vi = MISSING ? (double)marg[n_var + i] : (double)S;
    const double vj = MISSING ? (double)marg[n_var + j] : (double)S;
    const double pi = vi > 0.0 ? (double)marg[i] / vi : 0.0;
    const double pj = vj > 0.0 ? (double)marg[j] / vj : 0.0;
    const double pab = n > 0 ? (double)n11 / (double)n : 0.0;
    const double D = pab - pi * pj;
    if constexpr (EST == EST_D_NAIVE) {
      r2 = D;
    } else {
      const double den = (pi * (1.0 - pi)) * (pj * (1.0 - pj));
      r2 = den > 0.0 ? (D * D) / den : 0.0;
    }
    sval = r2;
    scnt = 1;
  } else {
    const double c1 = (double)n11, c2 = (double)(nA - n11), c3 = (double)(nB - n11);
    const double c4 = (double)(n - nA - nB + n11);
    const double dd_num = c1 * (c1 - 1.0) * c4 * (c4 - 1.0) + c2 * (c2 - 1.0) * c3 * (c3 - 1.0) -
                          2.0 * c1 * c2 * c3 * c4;
    const double pi2_num = (c1 + c2) * (c1 + c3) * (c2 + c4) * (c3 + c4) -
                           c1 * c4 * (-1.0 + c1 + 3.0 * c2 + 3.0 * c3 + c4) -
                           c2 * c3 * (-1.0 + 3.0 * c1 + c2 + c3 + 3.0 * c4);
    const bool ok = n >= 4;
    r2 = (ok && pi2_num != 0.0) ? dd_num / pi2_num : 0.0;
    sval = r2;
    scnt = ok;
  }
  return r2;
}


template <int FAMILY, bool MISSING, bool POPC, int EST, bool NIB = false>
__global__ __launch_bounds__(BAND_THREADS, 2) void k_band_r2(R2Args ra, const int8_t* __restrict__ planes,
                                                             const int32_t* __restrict__ marg, int64_t k_pad,
                                                             int32_t S) {
  using PL = Planes<FAMILY, MISSING>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ double s_pos_i[BT], s_pos_j[BT], s_kc_i[BT], s_kc_j[BT];
  __shared__ int s_wkey[BT];
  __shared__ int s_hist[R2_MAX_HIST_LDS];
  __shared__ double s_rsum[2][BT], s_csum[2][BT];  // [column half][row], [row half][column]
  __shared__ int s_rcnt[2][BT], s_ccnt[2][BT];

  const BandTile tile = ra.tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int hw = ra.n_hist + 1;
  const int h_n = ra.hist ? tile.n_slots * hw : 0;
  for (int t = tid; t < h_n; t += BAND_THREADS) s_hist[t] = 0;
  if (ra.snp_part && tid < 2 * BT) {
    (&s_rsum[0][0])[tid] = 0.0;
    (&s_csum[0][0])[tid] = 0.0;
    (&s_rcnt[0][0])[tid] = 0;
    (&s_ccnt[0][0])[tid] = 0;
  }
  if (tid < BT) {
    int64_t gi = min((int64_t)tile.i0 + tid, ra.n_var - 1), gj = min((int64_t)tile.j0 + tid, ra.n_var - 1);
    s_pos_i[tid] = ra.pos[gi];
    s_pos_j[tid] = ra.pos[gj];
    s_kc_i[tid] = ra.kc[gi];
    s_kc_j[tid] = ra.kc[gj];
    if (ra.key_mode == PGLD_KEY_WINDOW) {
      const double wk = floor((ra.kc[gi] - ra.origin) / ra.width);
      s_wkey[tid] = (wk >= 0.0 && wk < (double)ra.n_keys) ? (int)wk : -1;
    }
  }
  __syncthreads();

  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int ri = tile.i0 + wr, rj = tile.j0 + wc;
  const int64_t last_i = min((int64_t)ri + 31, ra.n_var - 1);
  const bool wave_live = (int64_t)rj < ra.n_var && rj + 31 > ri && (int64_t)ri < ra.n_var &&
                         s_pos_j[wc] - s_pos_i[last_i - tile.i0] <= ra.max_dist;
  WaveAcc<PL::N, NIB> acc;  // float accumulators (FP4 planes) hold exact integer counts
  block_products<FAMILY, MISSING, POPC, NIB>(planes, ra.n_var, k_pad, tile.i0, tile.j0, wr, wc, wave_live, smem,
                                                 acc);

  // pairs of histogram cell 0 (first key, r^2 below 1 / n_hist) are counted
  // in a register and added once per wave
  int cell0 = 0;
  if (wave_live) {
    const int jl = wc + (lane & 31);
    const int64_t j = (int64_t)tile.j0 + jl;
    double csum = 0.0;
    int ccnt = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int il = wr + acc_row(r, lane);
      const int64_t i = (int64_t)tile.i0 + il;
      const bool valid = i < j && j < ra.n_var && s_pos_j[jl] - s_pos_i[il] <= ra.max_dist;
      float r2f = 0.0f;
      bool hit = false;
      double sval = 0.0;
      int scnt = 0;
      if (valid && FAMILY == PGLD_FAMILY_GENO) {
        // dosage g = e1 + 2 e2; pair-valid sums from the plane products,
        // site sums from the per-SNP plane sums [v, e1, e2]
        int64_t gg, gib, gjb, nb;
        if constexpr (MISSING) {
          gg = (int64_t)acc.c[1][1][r] + 2 * (int64_t)acc.c[1][2][r] + 2 * (int64_t)acc.c[2][1][r] +
               4 * (int64_t)acc.c[2][2][r];
          gib = (int64_t)acc.c[1][0][r] + 2 * (int64_t)acc.c[2][0][r];  // g_i over samples valid at j
          gjb = (int64_t)acc.c[0][1][r] + 2 * (int64_t)acc.c[0][2][r];
          nb = acc.c[0][0][r];
        } else {
          gg = (int64_t)acc.c[0][0][r] + 2 * (int64_t)acc.c[0][1][r] + 2 * (int64_t)acc.c[1][0][r] +
               4 * (int64_t)acc.c[1][1][r];
          gib = 0;
          gjb = 0;
          nb = S;
        }
        const int64_t si = (int64_t)marg[ra.n_var + i] + 2 * (int64_t)marg[2 * ra.n_var + i];
        const int64_t sj = (int64_t)marg[ra.n_var + j] + 2 * (int64_t)marg[2 * ra.n_var + j];
        const int64_t qi = (int64_t)marg[ra.n_var + i] + 4 * (int64_t)marg[2 * ra.n_var + i];
        const int64_t qj = (int64_t)marg[ra.n_var + j] + 4 * (int64_t)marg[2 * ra.n_var + j];
        const int64_t ni = MISSING ? (int64_t)marg[i] : (int64_t)S, nj = MISSING ? (int64_t)marg[j] : (int64_t)S;
        if constexpr (!MISSING) {
          gib = si;
          gjb = sj;
        }
        double val;
        bool is_nan = false;
        if constexpr (EST == EST_RH_R) {
          // n cov = n sum(g_i g_j) - sum(g_i) sum(g_j); n ssd = n sum(g^2) - sum(g)^2 (exact)
          const double cn = (double)(nb * gg - si * sj);
          const double den = (double)(nb * qi - si * si) * (double)(nb * qj - sj * sj);
          is_nan = !(den > 0.0);
          val = den > 0.0 ? cn / sqrt(den) : __builtin_nan("");
        } else {
          // (cov n_i n_j)^2 / ((var_i n_i)(var_j n_j) n_i n_j), cov and var about the site means
          const int64_t c = gg * ni * nj - sj * gib * ni - si * gjb * nj + si * sj * nb;
          const int64_t vi = ni * qi - si * si, vj = nj * qj - sj * sj;
          const double den = (double)vi * (double)vj * (double)ni * (double)nj;
          const double cd = (double)c;
          val = (vi > 0 && vj > 0 && ni > 0 && nj > 0) ? (cd * cd) / den : 0.0;
        }
        r2f = (float)val;
        sval = is_nan ? 0.0 : val;
        scnt = !is_nan;
        if (ra.r2_out) ra.r2_out[ra.pair_base[i] + (j - i - 1)] = val;
        hit = !is_nan && val > ra.thresh;
        if (ra.hist) {
          int key;
          if (ra.key_mode == PGLD_KEY_WINDOW) {
            key = s_wkey[il];
          } else {
            double d = s_kc_j[jl] - s_kc_i[il];
            if (ra.key_mode == PGLD_KEY_BIN_ABS) d = fabs(d);
            key = key_bin(d, ra.edges, ra.n_keys);
          }
          const int slot = key - tile.key_lo;
          if (key >= 0 && slot >= 0 && slot < tile.n_slots) {
            int b = ra.n_hist;
            if (!is_nan) b = max(0, min((int)(val * (double)ra.n_hist), ra.n_hist - 1));
            if (slot * hw + b == 0)
              ++cell0;
            else
              atomicAdd(&s_hist[slot * hw + b], 1);
          }
        }
      } else if (valid && FAMILY == PGLD_FAMILY_HAP) {
        int n11, nA, nB, n;
        hap_plane_counts<MISSING>(acc, r, (const GLOBAL_AS int32_t*)marg, i, j, S, n11, nA, nB, n);
        bool is_nan = false;
        const double r2 = hap_pair_value<MISSING, EST>(n11, nA, nB, n, marg, ra.n_var, i, j, S, is_nan, sval, scnt);
        r2f = (float)r2;
        if (ra.hist) {
          int key;
          if (ra.key_mode == PGLD_KEY_WINDOW) {
            key = s_wkey[il];
          } else {
            double d = s_kc_j[jl] - s_kc_i[il];
            if (ra.key_mode == PGLD_KEY_BIN_ABS) d = fabs(d);
            key = key_bin(d, ra.edges, ra.n_keys);
          }
          const int slot = key - tile.key_lo;
          if (key >= 0 && slot >= 0 && slot < tile.n_slots) {
            int b = ra.n_hist;
            if (!is_nan) b = max(0, min((int)(r2 * (double)ra.n_hist), ra.n_hist - 1));
            if (slot * hw + b == 0)
              ++cell0;
            else
              atomicAdd(&s_hist[slot * hw + b], 1);
          }
        }
        hit = !is_nan && r2 > ra.thresh;
        if (ra.r2_out) ra.r2_out[ra.pair_base[i] + (j - i - 1)] = r2;
      }
      if (ra.snp_part) {
        // row sum over this half-wave's 32 columns (fixed butterfly order)
        double rs = sval;
        int rc = scnt;
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) {
          rs += __shfl_xor(rs, o);
          rc += __shfl_xor(rc, o);
        }
        if ((lane & 31) == 0) {
          s_rsum[wc >> 5][il] = rs;
          s_rcnt[wc >> 5][il] = rc;
        }
        csum += sval;
        ccnt += scnt;
      }
      if (ra.sp_count) {
        const unsigned long long m = __ballot(hit);
        if (m) {
          const int leader = __ffsll((long long)m) - 1;
          unsigned long long base = 0;
          if (lane == leader) base = atomicAdd(ra.sp_count, (unsigned long long)__popcll(m));
          base = __shfl(base, leader);
          if (hit) {
            const unsigned long long slot = base + __popcll(m & ((1ull << lane) - 1ull));
            if ((int64_t)slot < ra.sp_cap) {
              ra.sp_i[slot] = (int32_t)i;
              ra.sp_j[slot] = (int32_t)j;
              ra.sp_r2[slot] = r2f;
            }
          }
        }
      }
    }
    if (ra.snp_part) {
      // column sums: the 16 rows of each half, then both halves
      csum += __shfl_xor(csum, 32);
      ccnt += __shfl_xor(ccnt, 32);
      if (lane < 32) {
        s_csum[wr >> 5][jl] = csum;
        s_ccnt[wr >> 5][jl] = ccnt;
      }
    }
  }
  if (ra.snp_part) {
    __syncthreads();
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (tid < BT) {
      ra.snp_part[t * 2 * BT + tid] = s_rsum[0][tid] + s_rsum[1][tid];
      ra.snp_cnt[t * 2 * BT + tid] = s_rcnt[0][tid] + s_rcnt[1][tid];
    } else if (tid < 2 * BT) {
      ra.snp_part[t * 2 * BT + tid] = s_csum[0][tid - BT] + s_csum[1][tid - BT];
      ra.snp_cnt[t * 2 * BT + tid] = s_ccnt[0][tid - BT] + s_ccnt[1][tid - BT];
    }
  }
  if (ra.hist) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cell0 += __shfl_xor(cell0, o);
    if (lane == 0 && cell0) atomicAdd(&s_hist[0], cell0);
    __syncthreads();
    for (int t = tid; t < h_n; t += BAND_THREADS) {
      const int v = s_hist[t];
      if (v) atomicAdd(&ra.hist[(int64_t)(tile.key_lo + t / hw) * hw + t % hw], (unsigned long long)v);
    }
  }
}

// ---------------------------------------------------------------------------
// Wide r^2 kernel: haplotypes without missing data (one int8 plane per side).
// One workgroup per 128 x 128 super-tile; per 128-byte K chunk 128 A rows and
// 128 B rows are staged in LDS and wave w accumulates the 64 x 64 plan tile
// q = w of the super-tile as 2 x 2 MFMA blocks, so each operand fragment read
// from LDS feeds two MFMAs (half the LDS traffic per MFMA of the 64 x 64
// kernel).  The per-pair values, keys and outputs are those of k_band_r2 (the
// same function evaluates r^2); each wave keeps its plan tile's histogram in
// its own quarter of the LDS histogram.  Per-SNP sums are left to k_band_r2.
// ---------------------------------------------------------------------------
constexpr int R2W_ROWS = 2 * BT;
constexpr int R2W_HIST_PER_WAVE = R2_MAX_HIST_LDS / 4;
constexpr int R2W_OP_BYTES = 2 * R2W_ROWS * OP_ROW;
constexpr int R2W_RESIDENT = 64;  // workgroups resident per XCD (32 CUs x 2)

// NIB: FP4 planes (rows of k_pad / 2 bytes, 32 samples per 16-byte
// fragment) on v_mfma_f32_32x32x64_f8f6f4 with float accumulators.
template <int EST, bool NIB = false>
__global__ __launch_bounds__(BAND_THREADS, 2) void k_band_r2_wide(R2Args ra, const SuperTile* __restrict__ sts,
                                                                  const int8_t* __restrict__ planes,
                                                                  const int32_t* __restrict__ marg, int64_t k_pad,
                                                                  int32_t S) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ double s_pos_i[R2W_ROWS], s_pos_j[R2W_ROWS], s_kc_i[R2W_ROWS], s_kc_j[R2W_ROWS];
  __shared__ int s_wkey[R2W_ROWS];
  __shared__ int s_hist[R2_MAX_HIST_LDS];

  const SuperTile st = sts[xcd_tile(blockIdx.x, gridDim.x)];
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int hw = ra.n_hist + 1;
  const int own = w == 0 ? st.t[0] : w == 1 ? st.t[1] : w == 2 ? st.t[2] : st.t[3];
  const BandTile tile = own >= 0 ? ra.tiles[own] : BandTile{};
  int* wh = s_hist + w * R2W_HIST_PER_WAVE;
  const int h_n = (ra.hist && own >= 0) ? tile.n_slots * hw : 0;
  for (int t = lane; t < h_n; t += 64) wh[t] = 0;
  if (tid < R2W_ROWS) {
    const int64_t gi = min((int64_t)st.i0 + tid, ra.n_var - 1), gj = min((int64_t)st.j0 + tid, ra.n_var - 1);
    s_pos_i[tid] = ra.pos[gi];
    s_pos_j[tid] = ra.pos[gj];
    s_kc_i[tid] = ra.kc[gi];
    s_kc_j[tid] = ra.kc[gj];
    if (ra.key_mode == PGLD_KEY_WINDOW) {
      const double wk = floor((ra.kc[gi] - ra.origin) / ra.width);
      s_wkey[tid] = (wk >= 0.0 && wk < (double)ra.n_keys) ? (int)wk : -1;
    }
  }
  __syncthreads();

  // this wave's plan tile: rows wr .. wr + 63, columns wc .. wc + 63 of the super-tile
  const int wr = (w >> 1) * BT, wc = (w & 1) * BT;
  const int64_t ri = (int64_t)st.i0 + wr, rj = (int64_t)st.j0 + wc;
  const int64_t last_i = min(ri + BT - 1, ra.n_var - 1);
  const bool live = own >= 0 && rj < ra.n_var && rj + BT - 1 > ri && ri < ra.n_var &&
                    s_pos_j[wc] - s_pos_i[last_i - st.i0] <= ra.max_dist;

  // ---- K loop: LDS-staged 128 + 128 rows per chunk, two register stages ----
  constexpr int ITEMS = 2 * R2W_ROWS * 8 / BAND_THREADS;
  const GLOBAL_AS int8_t* src[ITEMS];
  int dst[ITEMS];
#pragma unroll
  for (int m = 0; m < ITEMS; ++m) {
    const int e = tid + BAND_THREADS * m;
    const int q = e & 7, row = (e >> 3) & (R2W_ROWS - 1), side = e >> 10;
    const int64_t snp = min((int64_t)(side ? st.j0 : st.i0) + row, ra.n_var - 1);
    src[m] = (const GLOBAL_AS int8_t*)planes + snp * (NIB ? k_pad / 2 : k_pad) + 16 * q;  // plane 0 (h)
    dst[m] = (side * R2W_ROWS + row) * OP_ROW + 16 * q;
  }
  if constexpr (NIB) k_pad /= 2;  // from here on: bytes per operand row
  typename std::conditional<NIB, v16f, v16i>::type acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;
  v4i pre0[ITEMS], pre1[ITEMS];
  // K chunks are visited in a rotated order (integer sums do not depend on
  // it): the rotation grows with the workgroup's position in its XCD's run,
  // by one full turn per R2W_RESIDENT workgroups.  Workgroups start about
  // evenly spaced in time once the XCD is full, so those resident together
  // work on about the same chunk at the same time and share the operand
  // strips of their neighbourhood in L2 instead of streaming them at random
  // offsets.  Loads past the last step read a valid (unused) chunk.
  const int nch = (int)(k_pad / KSTEP);
  const int rot = (int)(((int64_t)(blockIdx.x / N_XCD) * nch / R2W_RESIDENT) % nch);
  auto load = [&](v4i* pre, int64_t k) {
    int c = (int)(k / KSTEP) + rot;
    if (c >= nch) c -= nch;
    if (c >= nch) c -= nch;
    const int64_t kk = (int64_t)c * KSTEP;
#pragma unroll
    for (int m = 0; m < ITEMS; ++m) pre[m] = *(const GLOBAL_AS v4i*)(src[m] + kk);
  };
  load(pre0, 0);
  load(pre1, KSTEP);
  const int kh = 16 * (lane >> 5);
  auto step = [&](v4i* pre, int64_t k0) {
    __syncthreads();  // the previous chunk has been read by every wave
#pragma unroll
    for (int m = 0; m < ITEMS; ++m) *reinterpret_cast<v4i*>(smem + dst[m]) = pre[m];
    __syncthreads();
    load(pre, k0 + 2 * KSTEP);
    if (live) {
      v4i A[2][2], B[2][2];
      auto frag = [&](int sub, v4i* a_, v4i* b_) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          a_[h] = *reinterpret_cast<const v4i*>(smem + (wr + 32 * h + (lane & 31)) * OP_ROW + 32 * sub + kh);
          b_[h] = *reinterpret_cast<const v4i*>(smem + (R2W_ROWS + wc + 32 * h + (lane & 31)) * OP_ROW +
                                                32 * sub + kh);
        }
      };
      frag(0, A[0], B[0]);
#pragma unroll
      for (int sub = 0; sub < 4; ++sub) {
        if (sub < 3) frag(sub + 1, A[(sub + 1) & 1], B[(sub + 1) & 1]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            if constexpr (NIB) {
              const v4i x = A[sub & 1][a], y = B[sub & 1][b];
              const v8i x8 = {x[0], x[1], x[2], x[3], 0, 0, 0, 0}, y8 = {y[0], y[1], y[2], y[3], 0, 0, 0, 0};
              acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x8, y8, acc[a][b], 4, 4, 0, FP4_UNIT_SCALE,
                                                                          0, FP4_UNIT_SCALE);
            } else {
              acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[sub & 1][a], B[sub & 1][b], acc[a][b], 0, 0, 0);
            }
          }
      }
    }
  };
  {
    int64_t k0 = 0;
    for (; k0 + 2 * KSTEP <= k_pad; k0 += 2 * KSTEP) {
      step(pre0, k0);
      step(pre1, k0 + KSTEP);
    }
    if (k0 < k_pad) step(pre0, k0);
  }

  // ---- per-pair r^2, histogram, sparse list, per-pair output ----
  // Most pairs land in histogram cell 0 (first key, r^2 below 1 / n_hist):
  // each lane counts those in a register and the wave adds the total once,
  // instead of 64 lanes' atomics on one LDS address per row.
  int cell0 = 0;
  if (live) {
    // one 32 x 32 block at a time (a runtime loop keeps the unrolled row body
    // count at 16)
#pragma unroll 1
    for (int blk = 0; blk < 4; ++blk) {
      const int a = blk >> 1, b = blk & 1;
      v16i cur;
      auto take = [&](const auto& x) {  // float accumulators (FP4) hold exact integer counts
        if constexpr (NIB) {
#pragma unroll
          for (int r = 0; r < 16; ++r) cur[r] = (int)x[r];
        } else {
          cur = x;
        }
      };
      if (blk == 0)
        take(acc[0][0]);
      else if (blk == 1)
        take(acc[0][1]);
      else if (blk == 2)
        take(acc[1][0]);
      else
        take(acc[1][1]);
      const int jl = wc + 32 * b + (lane & 31);
      const int64_t j = (int64_t)st.j0 + jl;
      // one row of the block; FAST: every pair valid and in key slot
      // slot_in (no per-pair validity or key work), else both are tested
      auto row = [&](int r, auto fast_tag, int slot_in) {
        constexpr bool FAST = decltype(fast_tag)::value;
        const int il = wr + 32 * a + acc_row(r, lane);
        const int64_t i = (int64_t)st.i0 + il;
        const bool valid = FAST || (i < j && j < ra.n_var && s_pos_j[jl] - s_pos_i[il] <= ra.max_dist);
        bool hit = false;
        float r2f = 0.0f;
        if (valid) {
          bool is_nan = false;
          double sval = 0.0;
          int scnt = 0;
          // FP4 planes: cur is the centred product Sab; n11 = (S + Sa + Sb + Sab) / 4
          // with Sa = 2 marg_i - S, Sb = 2 marg_j - S
          const int n11 = NIB ? (cur[r] + 2 * (marg[i] + marg[j]) - S) >> 2 : cur[r];
          const double r2 = hap_pair_value<false, EST>(n11, marg[i], marg[j], S, marg, ra.n_var, i, j, S, is_nan,
                                                       sval, scnt);
          r2f = (float)r2;
          if (ra.hist) {
            int slot = slot_in;
            if (!FAST) {
              int key;
              if (ra.key_mode == PGLD_KEY_WINDOW) {
                key = s_wkey[il];
              } else {
                double d = s_kc_j[jl] - s_kc_i[il];
                if (ra.key_mode == PGLD_KEY_BIN_ABS) d = fabs(d);
                key = key_bin(d, ra.edges, ra.n_keys);
              }
              slot = key >= 0 ? key - tile.key_lo : -1;
              if (slot >= tile.n_slots) slot = -1;
            }
            if (slot >= 0) {
              int hb = ra.n_hist;
              if (!is_nan) hb = max(0, min((int)(r2 * (double)ra.n_hist), ra.n_hist - 1));
              const int cell = slot * hw + hb;
              if (cell == 0)
                ++cell0;
              else
                atomicAdd(&wh[cell], 1);
            }
          }
          hit = !is_nan && r2 > ra.thresh;
          if (ra.r2_out) ra.r2_out[ra.pair_base[i] + (j - i - 1)] = r2;
        }
        if (ra.sp_count) {
          const unsigned long long m = __ballot(hit);
          if (m) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(ra.sp_count, (unsigned long long)__popcll(m));
            base = __shfl(base, leader);
            if (hit) {
              const unsigned long long slot = base + __popcll(m & ((1ull << lane) - 1ull));
              if ((int64_t)slot < ra.sp_cap) {
                ra.sp_i[slot] = (int32_t)i;
                ra.sp_j[slot] = (int32_t)j;
                ra.sp_r2[slot] = r2f;
              }
            }
          }
        }
      };
      // the block's pairs are all valid when its first column lies right of
      // its last row, its last column exists and its farthest pair is within
      // max_dist; window keys are monotone in the row, so equal keys at the
      // first and last row mean one key for the block (histograms with
      // distance keys take the per-pair path)
      const int ri0 = wr + 32 * a, cj0 = wc + 32 * b;
      int fast_slot = -1;
      if ((int64_t)st.j0 + cj0 > (int64_t)st.i0 + ri0 + 31 && (int64_t)st.j0 + cj0 + 31 < ra.n_var &&
          s_pos_j[cj0 + 31] - s_pos_i[ri0] <= ra.max_dist) {
        if (!ra.hist) {
          fast_slot = 0;
        } else if (ra.key_mode == PGLD_KEY_WINDOW) {
          const int k0 = s_wkey[ri0], k1 = s_wkey[ri0 + 31];
          if (k0 >= 0 && k0 == k1 && k0 - tile.key_lo < tile.n_slots) fast_slot = k0 - tile.key_lo;
        }
      }
      if (fast_slot >= 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) row(r, std::true_type{}, fast_slot);
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) row(r, std::false_type{}, -1);
      }
    }
  }
  if (h_n > 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cell0 += __shfl_xor(cell0, o);
    if (lane == 0 && cell0) wh[0] += cell0;
  }
  // the wave's histogram quarter is written by this wave only
  for (int t = lane; t < h_n; t += 64) {
    const int v = wh[t];
    if (v) atomicAdd(&ra.hist[(int64_t)(tile.key_lo + t / hw) * hw + t % hw], (unsigned long long)v);
  }
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

// Per-SNP totals from the tiles' partial rows: sums[v] / counts[v] over the
// pairs (v, j > v), sums[n_var + v] / counts[n_var + v] over the pairs
// (i < v, v); tiles are visited in schedule order (deterministic).
__global__ void k_snp_reduce(const double* __restrict__ part, const int32_t* __restrict__ cnt,
                             const int64_t* __restrict__ row_off, const int64_t* __restrict__ col_off,
                             const int32_t* __restrict__ col_list, int64_t n_var, double* __restrict__ sums,
                             int64_t* __restrict__ counts) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_var) return;
  const int64_t tr = v / BT, q = v % BT;
  double s = 0.0;
  int64_t c = 0;
  for (int64_t t = row_off[tr]; t < row_off[tr + 1]; ++t) {
    s += part[t * 2 * BT + q];
    c += cnt[t * 2 * BT + q];
  }
  sums[v] = s;
  if (counts) counts[v] = c;
  s = 0.0;
  c = 0;
  for (int64_t e = col_off[tr]; e < col_off[tr + 1]; ++e) {
    const int64_t t = col_list[e];
    s += part[t * 2 * BT + BT + q];
    c += cnt[t * 2 * BT + BT + q];
  }
  sums[n_var + v] = s;
  if (counts) counts[n_var + v] = c;
}

}  // namespace pgld

using namespace pgld;

extern "C" {

static int band_r2(const pgld_band_plan* plan, const int8_t* planes, bool popc, const int32_t* marg, int64_t k_pad,
                   int32_t n_samples, int32_t with_missing, int32_t estimator, int32_t n_hist, double thresh,
                   uint64_t* hist, int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap,
                   uint64_t* n_sparse, double* r2_out, double* snp_sums, int64_t* snp_counts, void* stream,
                   bool fp4 = false) {
  if (fp4) PGLD_REQUIRE(k_pad % 256 == 0, "pgld_band_ld_pairs_fp4: k_pad must be a multiple of 256 (FP4 planes)");
  PGLD_REQUIRE(estimator >= EST_R2 && estimator <= EST_D_NAIVE, "pgld_band_ld_pairs: unknown estimator %d",
               estimator);
  PGLD_REQUIRE(!(estimator == EST_RH_R && with_missing),
               "pgld_band_ld_pairs: Rogers-Huff r needs genotypes without missing values");
  PGLD_REQUIRE(plan && planes && marg, "pgld_band_r2: null pointer");
  PGLD_REQUIRE(plan->n_regions == 0, "pgld_band_r2: not defined for region batches");
  PGLD_REQUIRE(k_pad % KSTEP == 0 && k_pad >= KSTEP && k_pad >= n_samples && n_samples >= 0,
               "pgld_band_r2: k_pad must be a positive multiple of 128 >= n_samples");
  // the 24-bit count products of hap_pair_value need counts below 2^15
  PGLD_REQUIRE(n_samples <= 32767, "pgld_band_r2: at most 32767 samples per population");
  PGLD_REQUIRE(k_pad <= MAX_K_PAD, "pgld_band_r2: k_pad must be at most 32768");
  PGLD_REQUIRE(!hist || n_hist >= 1, "pgld_band_r2: n_hist must be >= 1");
  PGLD_REQUIRE(!hist || (int64_t)plan->max_tile_slots * (n_hist + 1) <= R2_MAX_HIST_LDS,
               "pgld_band_r2: %d keys per tile x %d histogram bins exceed LDS", plan->max_tile_slots, n_hist + 1);
  PGLD_REQUIRE(!n_sparse || (sparse_i && sparse_j && sparse_r2) || sparse_cap == 0,
               "pgld_band_r2: sparse outputs need i, j and r2 buffers");
  hipStream_t s = as_stream(stream);
  // zero the outputs with a kernel (no memset node when captured into a graph)
  if (hist) {
    const int64_t n = (int64_t)plan->n_keys * (n_hist + 1);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(hist), n);
  }
  if (n_sparse)
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned long long*>(n_sparse),
                       (int64_t)1);
  int rc = check_launch("k_fill_u64");
  if (rc) return rc;
  if (plan->n_tiles == 0) return PGLD_OK;
  R2Args ra{};
  ra.tiles = reinterpret_cast<const BandTile*>(plan->d_tiles);
  ra.pos = plan->d_pos;
  ra.kc = plan->d_kc;
  ra.edges = plan->d_edges;
  ra.pair_base = plan->d_pair_base;
  ra.n_var = plan->n_var;
  ra.max_dist = plan->max_dist;
  ra.origin = plan->origin;
  ra.width = plan->width;
  ra.key_mode = plan->key_mode;
  ra.n_keys = plan->n_keys;
  ra.n_hist = n_hist;
  ra.thresh = thresh;
  ra.hist = reinterpret_cast<unsigned long long*>(hist);
  ra.sp_i = sparse_i;
  ra.sp_j = sparse_j;
  ra.sp_r2 = sparse_r2;
  ra.sp_cap = n_sparse ? sparse_cap : 0;
  ra.sp_count = reinterpret_cast<unsigned long long*>(n_sparse);
  ra.r2_out = r2_out;
  auto* pl = const_cast<pgld_band_plan*>(plan);
  if (snp_sums) {
    if (!pl->d_snp_part) {
      PGLD_TRY_HIP(hipMalloc((void**)&pl->d_snp_part, (size_t)plan->n_tiles * 2 * BT * sizeof(double)),
                   "pgld_band_ld_pairs workspace");
      PGLD_TRY_HIP(hipMalloc((void**)&pl->d_snp_cnt, (size_t)plan->n_tiles * 2 * BT * sizeof(int32_t)),
                   "pgld_band_ld_pairs workspace");
    }
    ra.snp_part = pl->d_snp_part;
    ra.snp_cnt = pl->d_snp_cnt;
  }
  const bool miss = with_missing != 0;
  const bool geno = estimator == EST_RH_R || estimator == EST_DIP_R2;
  const int npl = geno ? (miss ? 3 : 2) : (miss ? 2 : 1);
  const size_t lds = (size_t)2 * npl * BT * OP_ROW;
  dim3 g((unsigned)plan->n_tiles), b(BAND_THREADS);
  if (plan->timing) PGLD_TRY_HIP(hipEventRecord(plan->ev[0], s), "pgld_band_r2 event");
  // haplotypes without missing data and no per-SNP sums: the 128 x 128 kernel
  const bool wide = !geno && !miss && !popc && !snp_sums && plan->n_super > 0 && estimator != EST_D_NAIVE &&
                    (!hist || (int64_t)plan->max_tile_slots * (n_hist + 1) <= R2W_HIST_PER_WAVE);
  if (wide && fp4) {
    dim3 gw((unsigned)plan->n_super);
    const auto* sts = reinterpret_cast<const SuperTile*>(plan->d_super);
    if (estimator == EST_R2)
      hipLaunchKernelGGL((k_band_r2_wide<EST_R2, true>), gw, b, R2W_OP_BYTES, s, ra, sts, planes, marg, k_pad,
                         n_samples);
    else if (estimator == EST_R2_NAIVE)
      hipLaunchKernelGGL((k_band_r2_wide<EST_R2_NAIVE, true>), gw, b, R2W_OP_BYTES, s, ra, sts, planes, marg, k_pad,
                         n_samples);
    else
      hipLaunchKernelGGL((k_band_r2_wide<EST_SIGMA_D2, true>), gw, b, R2W_OP_BYTES, s, ra, sts, planes, marg, k_pad,
                         n_samples);
    rc = check_launch("k_band_r2_wide");
    if (rc) return rc;
    if (plan->timing) PGLD_TRY_HIP(hipEventRecord(plan->ev[1], s), "pgld_band_r2 event");
    return PGLD_OK;
  }
  if (wide) {
    dim3 gw((unsigned)plan->n_super);
    const auto* sts = reinterpret_cast<const SuperTile*>(plan->d_super);
    if (estimator == EST_R2)
      hipLaunchKernelGGL((k_band_r2_wide<EST_R2>), gw, b, R2W_OP_BYTES, s, ra, sts, planes, marg, k_pad, n_samples);
    else if (estimator == EST_R2_NAIVE)
      hipLaunchKernelGGL((k_band_r2_wide<EST_R2_NAIVE>), gw, b, R2W_OP_BYTES, s, ra, sts, planes, marg, k_pad,
                         n_samples);
    else
      hipLaunchKernelGGL((k_band_r2_wide<EST_SIGMA_D2>), gw, b, R2W_OP_BYTES, s, ra, sts, planes, marg, k_pad,
                         n_samples);
    rc = check_launch("k_band_r2_wide");
    if (rc) return rc;
    if (plan->timing) PGLD_TRY_HIP(hipEventRecord(plan->ev[1], s), "pgld_band_r2 event");
    return PGLD_OK;
  }
  // FP4 planes take the NIB instances (never with bit planes: popc and fp4 come from separate entry points)
#define PGLD_R2(F, M, B, E)                                                                                  \
  do {                                                                                                       \
    if (fp4)                                                                                                 \
      hipLaunchKernelGGL((k_band_r2<F, M, B, E, !B>), g, b, lds, s, ra, planes, marg, k_pad, n_samples);     \
    else                                                                                                     \
      hipLaunchKernelGGL((k_band_r2<F, M, B, E, false>), g, b, lds, s, ra, planes, marg, k_pad, n_samples);  \
  } while (0)
#define PGLD_R2_EST(M, B)                                                        \
  do {                                                                           \
    if (estimator == EST_R2) PGLD_R2(PGLD_FAMILY_HAP, M, B, EST_R2);             \
    else if (estimator == EST_R2_NAIVE) PGLD_R2(PGLD_FAMILY_HAP, M, B, EST_R2_NAIVE); \
    else PGLD_R2(PGLD_FAMILY_HAP, M, B, EST_SIGMA_D2);                           \
  } while (0)
  if (estimator == EST_RH_R) {
    PGLD_R2(PGLD_FAMILY_GENO, false, false, EST_RH_R);
  } else if (estimator == EST_DIP_R2) {
    if (miss) PGLD_R2(PGLD_FAMILY_GENO, true, false, EST_DIP_R2);
    else PGLD_R2(PGLD_FAMILY_GENO, false, false, EST_DIP_R2);
  } else if (popc) {
    if (miss) PGLD_R2_EST(true, true); else PGLD_R2_EST(false, true);
  } else if (estimator == EST_D_NAIVE) {
    if (miss) PGLD_R2(PGLD_FAMILY_HAP, true, false, EST_D_NAIVE);
    else PGLD_R2(PGLD_FAMILY_HAP, false, false, EST_D_NAIVE);
  } else {
    if (miss) PGLD_R2_EST(true, false); else PGLD_R2_EST(false, false);
  }
#undef PGLD_R2_EST
#undef PGLD_R2
  rc = check_launch("k_band_r2");
  if (rc) return rc;
  if (plan->timing) PGLD_TRY_HIP(hipEventRecord(plan->ev[1], s), "pgld_band_r2 event");
  if (snp_sums) {
    hipLaunchKernelGGL(k_snp_reduce, dim3((unsigned)ceil_div(plan->n_var, 256)), dim3(256), 0, s, pl->d_snp_part,
                       pl->d_snp_cnt, plan->d_row_off, plan->d_col_off, plan->d_col_list, plan->n_var, snp_sums,
                       snp_counts);
    rc = check_launch("k_snp_reduce");
  }
  return rc;
}

int pgld_band_r2(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                 int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                 int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap, uint64_t* n_sparse,
                 double* r2_out, void* stream) {
  return band_r2(plan, planes, false, marg, k_pad, n_samples, with_missing, EST_R2, n_hist, thresh, hist, sparse_i,
                 sparse_j, sparse_r2, sparse_cap, n_sparse, r2_out, nullptr, nullptr, stream);
}

int pgld_band_ld_pairs(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                       int32_t n_samples, int32_t with_missing, int32_t estimator, int32_t n_hist, double thresh,
                       uint64_t* hist, int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap,
                       uint64_t* n_sparse, double* values, double* snp_sums, int64_t* snp_counts, void* stream) {
  return band_r2(plan, planes, false, marg, k_pad, n_samples, with_missing, estimator, n_hist, thresh, hist,
                 sparse_i, sparse_j, sparse_r2, sparse_cap, n_sparse, values, snp_sums, snp_counts, stream);
}

int pgld_band_ld_pairs_fp4(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                           int32_t n_samples, int32_t with_missing, int32_t estimator, int32_t n_hist, double thresh,
                           uint64_t* hist, int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap,
                           uint64_t* n_sparse, double* values, double* snp_sums, int64_t* snp_counts, void* stream) {
  return band_r2(plan, planes, false, marg, k_pad, n_samples, with_missing, estimator, n_hist, thresh, hist,
                 sparse_i, sparse_j, sparse_r2, sparse_cap, n_sparse, values, snp_sums, snp_counts, stream, true);
}

int pgld_band_r2_bits(const pgld_band_plan* plan, const uint32_t* bits, const int32_t* marg, int64_t k_pad,
                      int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                      int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap, uint64_t* n_sparse,
                      double* r2_out, void* stream) {
  return band_r2(plan, reinterpret_cast<const int8_t*>(bits), true, marg, k_pad, n_samples, with_missing, EST_R2,
                 n_hist, thresh, hist, sparse_i, sparse_j, sparse_r2, sparse_cap, n_sparse, r2_out, nullptr, nullptr,
                 stream);
}

}  // extern "C"
