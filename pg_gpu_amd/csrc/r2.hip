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
  double* r2_out;  // per band pair or null
};

constexpr int R2_MAX_HIST_LDS = 8192;  // ints of per-tile histogram kept in LDS

template <bool MISSING, bool POPC>
__global__ __launch_bounds__(BAND_THREADS, 2) void k_band_r2(R2Args ra, const int8_t* __restrict__ planes,
                                                             const int32_t* __restrict__ marg, int64_t k_pad,
                                                             int32_t S) {
  using PL = Planes<PGLD_FAMILY_HAP, MISSING>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ double s_pos_i[BT], s_pos_j[BT], s_kc_i[BT], s_kc_j[BT];
  __shared__ int s_wkey[BT];
  __shared__ int s_hist[R2_MAX_HIST_LDS];

  const BandTile tile = ra.tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int hw = ra.n_hist + 1;
  const int h_n = ra.hist ? tile.n_slots * hw : 0;
  for (int t = tid; t < h_n; t += BAND_THREADS) s_hist[t] = 0;
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
  WaveAcc<PL::N> acc;
  block_products<PGLD_FAMILY_HAP, MISSING, POPC>(planes, ra.n_var, k_pad, tile.i0, tile.j0, wr, wc, wave_live, smem,
                                                 acc);

  if (wave_live) {
    const int jl = wc + (lane & 31);
    const int64_t j = (int64_t)tile.j0 + jl;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int il = wr + acc_row(r, lane);
      const int64_t i = (int64_t)tile.i0 + il;
      const bool valid = i < j && j < ra.n_var && s_pos_j[jl] - s_pos_i[il] <= ra.max_dist;
      float r2f = 0.0f;
      bool hit = false;
      if (valid) {
        int64_t n11 = acc.c[0][0][r], nA, nB, n;
        if constexpr (MISSING) {
          nA = acc.c[0][1][r];
          nB = acc.c[1][0][r];
          n = acc.c[1][1][r];
        } else {
          nA = marg[i];
          nB = marg[j];
          n = S;
        }
        const int64_t diff = n * n11 - nA * nB;
        const double dd = (double)diff;
        const double num = dd * dd;
        const double den = (double)(nA * (n - nA)) * (double)(nB * (n - nB));
        const double r2 = den > 0.0 ? num / den : __builtin_nan("");
        const bool is_nan = !(den > 0.0);
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
            if (!is_nan) b = min((int)(r2 * (double)ra.n_hist), ra.n_hist - 1);
            atomicAdd(&s_hist[slot * hw + b], 1);
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
    }
  }
  if (ra.hist) {
    __syncthreads();
    for (int t = tid; t < h_n; t += BAND_THREADS) {
      const int v = s_hist[t];
      if (v) atomicAdd(&ra.hist[(int64_t)(tile.key_lo + t / hw) * hw + t % hw], (unsigned long long)v);
    }
  }
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

}  // namespace pgld

using namespace pgld;

extern "C" {

static int band_r2(const pgld_band_plan* plan, const int8_t* planes, bool popc, const int32_t* marg, int64_t k_pad,
                   int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                   int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap, uint64_t* n_sparse,
                   double* r2_out, void* stream) {
  PGLD_REQUIRE(plan && planes && marg, "pgld_band_r2: null pointer");
  PGLD_REQUIRE(plan->n_regions == 0, "pgld_band_r2: not defined for region batches");
  PGLD_REQUIRE(k_pad % KSTEP == 0 && k_pad >= n_samples && n_samples >= 0,
               "pgld_band_r2: k_pad must be a multiple of 128 >= n_samples");
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
  const bool miss = with_missing != 0;
  const size_t lds = miss ? (size_t)opbuf_bytes<2>() : (size_t)opbuf_bytes<1>();
  dim3 g((unsigned)plan->n_tiles), b(BAND_THREADS);
  if (plan->timing) PGLD_TRY_HIP(hipEventRecord(plan->ev[0], s), "pgld_band_r2 event");
  if (popc) {
    if (miss) hipLaunchKernelGGL((k_band_r2<true, true>), g, b, lds, s, ra, planes, marg, k_pad, n_samples);
    else hipLaunchKernelGGL((k_band_r2<false, true>), g, b, lds, s, ra, planes, marg, k_pad, n_samples);
  } else {
    if (miss) hipLaunchKernelGGL((k_band_r2<true, false>), g, b, lds, s, ra, planes, marg, k_pad, n_samples);
    else hipLaunchKernelGGL((k_band_r2<false, false>), g, b, lds, s, ra, planes, marg, k_pad, n_samples);
  }
  rc = check_launch("k_band_r2");
  if (rc) return rc;
  if (plan->timing) PGLD_TRY_HIP(hipEventRecord(plan->ev[1], s), "pgld_band_r2 event");
  return PGLD_OK;
}

int pgld_band_r2(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                 int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                 int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap, uint64_t* n_sparse,
                 double* r2_out, void* stream) {
  return band_r2(plan, planes, false, marg, k_pad, n_samples, with_missing, n_hist, thresh, hist, sparse_i, sparse_j,
                 sparse_r2, sparse_cap, n_sparse, r2_out, stream);
}

int pgld_band_r2_bits(const pgld_band_plan* plan, const uint32_t* bits, const int32_t* marg, int64_t k_pad,
                      int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                      int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap, uint64_t* n_sparse,
                      double* r2_out, void* stream) {
  return band_r2(plan, reinterpret_cast<const int8_t*>(bits), true, marg, k_pad, n_samples, with_missing, n_hist,
                 thresh, hist, sparse_i, sparse_j, sparse_r2, sparse_cap, n_sparse, r2_out, stream);
}

}  // extern "C"
