// Shared pieces of the tiled int8-MFMA kernels: the band plan, the tile
// schedule, indicator-plane selection, the LDS-staged MFMA product loop and
// the count reconstruction.  Included by band.hip and r2.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <vector>

#include "common.h"
#include "../../include/pgld.h"

struct BandTileHost {  // tile geometry kept on the host for scheduling
  int32_t region, rows, cols, diag;
  int32_t ti, tj;  // tile row / column (64-SNP blocks) within the region
};

struct pgld_band_plan {
  int64_t n_var = 0, n_pairs = 0, n_tiles = 0, n_slots = 0;
  double max_dist = 0.0, origin = 0.0, width = 1.0;
  int32_t key_mode = 0, n_keys = 0, max_tile_slots = 0;
  double* d_pos = nullptr;       // band coordinate
  double* d_kc = nullptr;        // key coordinate
  double* d_edges = nullptr;     // n_keys + 1 (bin modes)
  void* d_tiles = nullptr;       // BandTile[n_tiles]
  int64_t* d_pair_base = nullptr;  // exclusive prefix of per-variant pair counts
  int32_t* d_key_off = nullptr;    // n_keys + 1
  int64_t* d_key_list = nullptr;   // slots grouped by key, ascending tile order
  // two-level key reduction (keys with many slots): chunk c sums key_list
  // entries [d_red_lo[c], d_red_lo[c + 1]) of one key; key k owns chunks
  // [d_red_off[k], d_red_off[k + 1])
  int64_t n_red_chunks = 0;
  int64_t* d_red_lo = nullptr;
  int32_t* d_red_off = nullptr;
  double* d_partials = nullptr;    // workspace
  size_t partials_cap = 0;         // doubles
  bool timing = false;             // record events around each k_band_sums launch
  hipEvent_t ev[2] = {nullptr, nullptr};
  // region batches (n_regions > 0): keys are region * n_keys_local + key
  int32_t n_regions = 0, n_keys_local = 0;
  std::vector<int64_t> h_region_nvar, h_region_base;
  std::vector<BandTileHost> h_tiles;   // host copy of the schedule, for ordering
  void* d_regions = nullptr;           // RegionDesc[n_regions], written by bind
  int32_t* d_order = nullptr;          // tile orders, see order_mode (-1 entries are skipped)
  int32_t bound_family = -1, bound_pops = 0, bound_set = -1;
  bool bound_fp4 = false;              // the bound planes are FP4 (pgld_batch_bind_fp4)
  int32_t grid = 0;                    // persistent grid (0 = one workgroup per tile)
  int32_t order_mode = 0;              // 0 cost-descending, 1 natural, 2 L2-blocked per XCD
  int64_t order_off[3] = {0, 0, 0};    // start of each order in d_order
  int64_t order_len[3] = {0, 0, 0};    // its length (the blocked order holds -1 padding)
  bool xcd_split = false;              // each XCD walks a contiguous eighth of the order
  int32_t kvar = 0;                    // kernel variant of the fused batch pass (see KV_* in band.hip)
  double bound_mean_samples = 0.0;     // mean population size of the bound batch (autotuner key)
  int32_t bound_max_samples = 0;       // largest population of the bound batch
  int64_t partials_ns = 0;             // statistics + 1 the workspace was sized for
  // per-SNP sums of pair values (single bands): tiles by tile-row / tile-column
  int64_t* d_row_off = nullptr;        // tile-row r -> tiles [row_off[r], row_off[r + 1])
  int64_t* d_col_off = nullptr;        // tile-column c -> col_list[col_off[c] .. col_off[c + 1])
  int32_t* d_col_list = nullptr;
  int32_t* d_sums_order = nullptr;     // single plans: launch order of the fused sums tiles (8 x 8 blocks)
  void* d_super = nullptr;             // SuperTile[n_super] (single plans): 2 x 2 groups of tiles
  int64_t n_super = 0;
  double* d_snp_part = nullptr;        // workspace: per tile 64 row + 64 column sums
  int32_t* d_snp_cnt = nullptr;
  int16_t* d_gstash = nullptr;         // workspace: parked counts of the 3-4 population genotype sets
  size_t gstash_cap = 0;               // int16 elements
};

// Address space 1 (global memory) on AMDGPU.
#define GLOBAL_AS __attribute__((address_space(1)))

namespace pgld {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

constexpr int BT = 64;       // tile edge (SNPs)
constexpr int KSTEP = 128;   // K bytes per loop iteration: 4 MFMAs of K = 32
// Largest K padding of a population's planes: 32,767 samples padded to a
// multiple of 128 (int8) or 256 (FP4).  Operand offsets within a 64-row block
// (OpBlock) are 24-bit products that rely on it.
constexpr int64_t MAX_K_PAD = 32768;
constexpr int BAND_THREADS = 256;
constexpr int RED_LDS_BYTES = 64 * 1024;
constexpr int RED_CHUNK = 256;      // slots per first-level chunk of a long key reduction
constexpr int MAX_LDS_EDGES = 130;  // bin edges kept in LDS when n_keys < 130
// static LDS of k_band_sums: positions, key coordinates, window keys, edges
constexpr int STATIC_LDS_BYTES = 4 * 64 * 8 + 64 * 4 + MAX_LDS_EDGES * 8;

struct BandTile {
  int32_t i0, j0, key_lo, n_slots;  // i0 / j0 are SNP indices within the tile's region
  int64_t slot_base;
  int32_t region, pad_;
};

struct PopArgs {
  const int8_t* planes[4];
  const int32_t* marg[4];
  int64_t k_pad[4];
  int32_t n_samples[4];
};

// One region of a region batch: its populations' planes, its SNP range in the
// plan's concatenated coordinates, and its first key.
struct RegionDesc {
  PopArgs pa;
  int64_t n_var;
  int64_t snp_base;
  int32_t key_base, pad_;
};

// ---------------------------------------------------------------------------
// Plane selection per family / missing-data mode
// ---------------------------------------------------------------------------
template <int FAMILY, bool MISSING>
struct Planes;
template <>
struct Planes<PGLD_FAMILY_HAP, true> {  // planes [h, v]; products {h, v} x {h, v}
  static constexpr int N = 2, FIRST = 0, NPL = 2, NC = 4;
};
template <>
struct Planes<PGLD_FAMILY_HAP, false> {  // h only
  static constexpr int N = 1, FIRST = 0, NPL = 2, NC = 4;
};
template <>
struct Planes<PGLD_FAMILY_GENO, true> {  // planes [v, e1, e2]; all 9 products
  static constexpr int N = 3, FIRST = 0, NPL = 3, NC = 9;
};
template <>
struct Planes<PGLD_FAMILY_GENO, false> {  // e1, e2 only
  static constexpr int N = 2, FIRST = 1, NPL = 3, NC = 9;
};

// Accumulators of the N x N plane products: int32 for the int8 MFMA, float
// for the FP4 MFMA (exact: counts stay far below 2^24).
template <int N, bool F = false>
struct WaveAcc {
  using T = typename std::conditional<F, v16f, v16i>::type;
  T c[N][N];
};

// FP4 planes: the indicators as e2m1 codes (0 -> 0x0, 1 -> 0x2 = 1.0, and
// -1 -> 0xA for the centred haplotype allele plane below), two samples per byte, the even sample in the low nibble; a row holds
// k_pad / 2 bytes, k_pad a multiple of 256.  v_mfma_f32_32x32x64_f8f6f4 with
// e2m1 operands (cbsz = blgp = 4) and unit scales (E8M0 127) takes the 16
// bytes a lane reads for one int8 MFMA as 32 samples of row (lane mod 32),
// K half (lane div 32), nibble k low first -- the same lane layout on both
// operands, so one FP4 MFMA covers 64 samples where the int8 one covers 32.
constexpr int FP4_ONE = 0x2;
// Haplotype FP4 planes hold the allele plane centred: a' = +1 (0x2) for the
// derived allele, -1 (0xA) for the ancestral one, 0 where missing; the valid
// plane keeps 0 / 1.  The products then are the centred moments the
// statistics use (hap_plane_counts, hap_moments).  Genotype FP4 planes hold
// 0 / 1 indicators on every plane.
constexpr int FP4_MINUS_ONE = 0xA;
constexpr int FP4_UNIT_SCALE = 0x7f;

// LDS-staged products for the whole workgroup (all four waves must call it:
// it contains barriers).  Per 128-byte K chunk the 64 A rows and 64 B rows of
// every plane are copied with full-line loads (8 lanes per 128-byte line) into
// LDS rows padded to 144 bytes, which makes the MFMA-layout ds_read_b128 reads
// (lane l -> row l % 32) bank-conflict free.
constexpr int OP_ROW = 144;  // padded LDS row stride (bytes)

template <int N>
__host__ __device__ constexpr int opbuf_bytes() { return 2 * N * BT * OP_ROW; }

// Bit-plane rows for the popcount variant: 32 samples per uint32 word, rows
// padded to a multiple of 128 bytes (1,024 samples) with zero words.
__host__ __device__ inline int64_t bit_row_bytes(int64_t k_pad) { return 128 * ((k_pad + 1023) / 1024); }

// Wave-uniform value in scalar registers.
__device__ __forceinline__ int64_t uniform_i64(int64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Operand rows of one population for a tile's K loop.  The byte address of
// a copy is a wave-uniform part -- the side's block base (first plane used,
// first SNP of the tile's 64-row block on that side), the plane and the K
// offset -- plus this thread's 32-bit offset within the block: (row - first
// row) x row bytes + 16 (t % 8), below 64 x 32,768 bytes.  The loads then
// take their base from scalar registers and need no 64-bit address
// arithmetic on the VALU.  Copy m of thread t covers bytes 16 (t % 8) .. of
// row 32 (m % 2) + t / 8 of side (m / 2) / N, plane (m / 2) % N, and lands at
// LDS offset (t / 8) OP_ROW + 16 (t % 8) + 32 OP_ROW m.
struct OpBlock {
  const GLOBAL_AS int8_t* side[2];
  int64_t plane;
  uint32_t off[2][2];  // side, row half
};
__device__ __forceinline__ OpBlock op_block(const int8_t* planes, int first_plane, int64_t row, int64_t n_var,
                                            int i0, int j0) {
  OpBlock s;
  row = uniform_i64(row);
  s.plane = n_var * row;
  const GLOBAL_AS int8_t* base = (const GLOBAL_AS int8_t*)uniform_i64((int64_t)planes) + first_plane * s.plane;
  s.side[0] = (const GLOBAL_AS int8_t*)uniform_i64((int64_t)(base + (int64_t)i0 * row));
  s.side[1] = (const GLOBAL_AS int8_t*)uniform_i64((int64_t)(base + (int64_t)j0 * row));
  const int t = threadIdx.x >> 3;
  const uint32_t q = 16u * (threadIdx.x & 7);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    // rows past the panel's end read its last row (never used)
    const int ri = (int)(min((int64_t)i0 + t + 32 * h, n_var - 1) - i0);
    const int rj = (int)(min((int64_t)j0 + t + 32 * h, n_var - 1) - j0);
    s.off[0][h] = __umul24((uint32_t)ri, (uint32_t)row) + q;
    s.off[1][h] = __umul24((uint32_t)rj, (uint32_t)row) + q;
  }
  return s;
}
// The same copies as buffer loads: one resource per side and plane (base in
// scalar registers), the thread's 32-bit offset, the K offset as the scalar
// offset -- no per-chunk address arithmetic on the VALU and no 64-bit
// per-thread pointers kept live.  Every offset stays inside its 64-row block.
template <int N>
struct OpRsrc {
  __amdgpu_buffer_rsrc_t r[2][N];
};
constexpr int BUF_RSRC_WORD3 = 0x00020000;  // raw buffer, 32-bit data format
template <int N>
__device__ __forceinline__ OpRsrc<N> op_rsrc(const OpBlock& s) {
  OpRsrc<N> q;
#pragma unroll
  for (int sd = 0; sd < 2; ++sd)
#pragma unroll
    for (int a = 0; a < N; ++a)
      q.r[sd][a] = __builtin_amdgcn_make_buffer_rsrc((void*)(s.side[sd] + (int64_t)a * s.plane), (short)0,
                                                     0x7fffffff, BUF_RSRC_WORD3);
  return q;
}
template <int N>
__device__ __forceinline__ void op_load_buf(v4i* pre, const OpRsrc<N>& q, const OpBlock& s, int k) {
#pragma unroll
  for (int m = 0; m < 4 * N; ++m) {
    const int a = (m >> 1) % N, sd = (m >> 1) / N, h = m & 1;
    pre[m] = __builtin_amdgcn_raw_buffer_load_b128(q.r[sd][a], s.off[sd][h], k, 0);
  }
}

// The 2 N BT x 16-byte copies of the K chunk at byte offset k (uniform).
template <int N>
__device__ __forceinline__ void op_load(v4i* pre, const OpBlock& s, int64_t k) {
#pragma unroll
  for (int m = 0; m < 4 * N; ++m) {
    const int a = (m >> 1) % N, sd = (m >> 1) / N, h = m & 1;
    const GLOBAL_AS int8_t* b = s.side[sd] + ((int64_t)a * s.plane + k);
    pre[m] = *(const GLOBAL_AS v4i*)(b + s.off[sd][h]);
  }
}

// POPC = false: int8 0/1 planes (row_bytes = k_pad), products on the int8
// MFMA.  POPC = true: bit planes (row_bytes = bit_row_bytes(k_pad)), products
// as AND + popcount on the VALU, accumulated into the same accumulator
// layout as the MFMA (lane l -> column l % 32, rows acc_row(r, l)), so every
// epilogue works with either.
// PFS: register stages of operand loads in flight (0 = two for up to two
// planes per side, else one); FDB: the MFMA operand fragments of the next
// sub-step are read while the current one's MFMAs run (two fragment sets).
// One stage and one fragment set keep fewer registers live, for kernels that
// hide the latency with more waves per SIMD instead.
// SWZ: operand rows unpadded (OP_ROW_SWZ bytes) with their 16-byte units
// permuted by row (unit u of row r at u ^ (r mod 8)), which keeps the
// MFMA-layout reads bank-conflict free in 1/9 less LDS at one XOR per
// sub-step.
constexpr int OP_ROW_SWZ = 128;
template <int N>
__host__ __device__ constexpr int opbuf_bytes_swz() { return 2 * N * BT * OP_ROW_SWZ; }

template <int FAMILY, bool MISSING, bool POPC = false, bool NIB = false, int PFS = 0, bool FDB = true,
          bool SWZ = false>
__device__ __forceinline__ void block_products(const int8_t* __restrict__ planes, int64_t n_var, int64_t k_pad,
                                               int i0, int j0, int wr, int wc, bool wave_live,
                                               unsigned char* __restrict__ opbuf,
                                               WaveAcc<Planes<FAMILY, MISSING>::N, NIB>& acc, int n_samples = -1) {
  static_assert(!(POPC && NIB), "bit planes and FP4 planes are separate layouts");
  static_assert(!(POPC && SWZ), "the popcount form reads the padded layout");
  constexpr int RS = SWZ ? OP_ROW_SWZ : OP_ROW;  // LDS row stride
  if constexpr (POPC) k_pad = bit_row_bytes(k_pad);  // from here on: bytes per operand row
  if constexpr (NIB) k_pad /= 2;
  using PL = Planes<FAMILY, MISSING>;
  constexpr int N = PL::N;
  constexpr int ITEMS = 2 * N * BT * 8 / BAND_THREADS;  // 16-byte copies per thread per chunk
  static_assert(ITEMS == 4 * N, "copy layout");
  const int tid = threadIdx.x, lane = tid & 63;
  // plane pointers may come from a region table in memory; the explicit
  // global address space keeps the loads off the flat path (whose counter
  // would make every LDS wait also wait for the prefetch)
  const OpBlock ob = op_block(planes, PL::FIRST, k_pad, n_var, i0, j0);
  // copy m of thread t: row t / 8 + 32 (m % 2) of block m / 2, unit t % 8
  // (swizzled by the row, whose low three bits are those of t / 8)
  const int dbase = (tid >> 3) * RS + 16 * (SWZ ? ((tid & 7) ^ ((tid >> 3) & 7)) : (tid & 7));
  // the MFMAs of the first sub-step start from zero accumulators; waves
  // without MFMAs (and the popcount form) clear them here
  if (POPC || !wave_live) {
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc.c[a][b][r] = 0;
  }
  // register stages: the loads of chunk c + PF are in flight while the MFMAs
  // of chunk c run (PF = 2 where registers allow: up to 2 planes per side)
  constexpr int PF = PFS > 0 ? PFS : (N <= 2 ? 2 : 1);
  v4i pre0[ITEMS], pre1[PF == 2 ? ITEMS : 1];
  // Loads are issued unconditionally (a chunk past the end re-reads the last
  // chunk and is never used) so that every path keeps the same number of
  // loads in flight and the wait before a chunk's LDS store covers only that
  // chunk, not the younger prefetch.  k_pad is a multiple of KSTEP.
  [[maybe_unused]] const OpRsrc<N> rs = SWZ ? op_rsrc<N>(ob) : OpRsrc<N>{};
  // k_pad <= MAX_K_PAD: K offsets fit an int
  const int kp = (int)k_pad;
  auto load = [&](v4i* pre, int k) {
    if constexpr (SWZ)
      op_load_buf<N>(pre, rs, ob, k < kp ? k : kp - KSTEP);
    else
      op_load<N>(pre, ob, k < kp ? k : kp - KSTEP);
  };
  load(pre0, 0);
  if constexpr (PF == 2) load(pre1, KSTEP);
  // The two workgroups of a CU alternate between this loop (MFMA, LDS) and
  // the FP64 epilogue; the raised wave priority lets the K loop's
  // instructions issue first when both contend, the epilogue fills the gaps.
  __builtin_amdgcn_s_setprio(2);
  const int ra = wr + (lane & 31), rb = wc + (lane & 31), kh = 16 * (lane >> 5);
  const int c16 = ((lane & 7) << 4) ^ kh;
  const int swz_a = (ra * RS) ^ c16, swz_b = ((N * BT + rb) * RS) ^ c16;
  // FP4 rows without the register-stage double buffer of the fragments: a
  // sub-step whose 64 samples all lie past the population's last sample
  // (given n_samples >= 0) multiplies padding nibbles, which are 0, so its
  // fragment reads and MFMAs are skipped -- except the first chunk's first
  // sub-step, which starts the accumulators
  const int kvalid = (NIB && !FDB && n_samples >= 0) ? 32 * ((n_samples + 63) / 64) : kp;
  auto step = [&](auto first_tag, v4i* pre, int k0) {
    constexpr bool FIRST = decltype(first_tag)::value;
    __syncthreads();  // the previous chunk has been read by every wave
#pragma unroll
    for (int m = 0; m < ITEMS; ++m) *reinterpret_cast<v4i*>(opbuf + dbase + 32 * RS * m) = pre[m];
    __syncthreads();
    load(pre, k0 + PF * KSTEP);
    if (wave_live && POPC) {
      // 8 groups of four 32-sample words per 128-byte chunk; the A rows of a
      // half-wave are shared (LDS broadcast), B is this lane's column
      const int half = lane >> 5;
#pragma unroll 2
      for (int q = 0; q < 8; ++q) {
        v4i B[N];
#pragma unroll
        for (int b = 0; b < N; ++b)
          B[b] = *reinterpret_cast<const v4i*>(opbuf + ((1 * N + b) * BT + rb) * OP_ROW + 16 * q);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = wr + 8 * (r >> 2) + 4 * half + (r & 3);
          v4i A[N];
#pragma unroll
          for (int a = 0; a < N; ++a)
            A[a] = *reinterpret_cast<const v4i*>(opbuf + ((0 * N + a) * BT + row) * OP_ROW + 16 * q);
#pragma unroll
          for (int a = 0; a < N; ++a)
#pragma unroll
            for (int b = 0; b < N; ++b) {
              int c = acc.c[a][b][r];
#pragma unroll
              for (int w = 0; w < 4; ++w) c += __popc((unsigned)(A[a][w] & B[b][w]));
              acc.c[a][b][r] = c;
            }
        }
      }
    } else if (wave_live) {
      // with FDB, operand fragments of substep st + 1 are read while the
      // MFMAs of substep st run
      constexpr int NF = FDB ? 2 : 1;
      v4i A[NF][N], B[NF][N];
      auto frag = [&](int st, v4i* a_, v4i* b_) {
        if constexpr (SWZ) {
          // unit 2 st + lane / 32 of rows ra, rb (both = lane mod 8, mod 8)
          // sits at unit (2 st + lane / 32) ^ (lane mod 8): row bases are
          // multiples of 128, so the address is base ^ c16 ^ 32 st with
          // c16 = 16 (lane mod 8) ^ kh -- one XOR per sub-step and side (the
          // empty asm operands keep the compiler from hoisting all of them
          // into live registers)
          int s32 = 32 * st;
          asm volatile("" : "+s"(s32));
          s32 &= 0x60;  // known bits again, so the LDS base folds into the offset field
          const int xa = swz_a ^ s32, xb = swz_b ^ s32;
#pragma unroll
          for (int a = 0; a < N; ++a) {
            a_[a] = *reinterpret_cast<const v4i*>(opbuf + xa + a * BT * RS);
            b_[a] = *reinterpret_cast<const v4i*>(opbuf + xb + a * BT * RS);
          }
        } else {
#pragma unroll
          for (int a = 0; a < N; ++a) {
            a_[a] = *reinterpret_cast<const v4i*>(opbuf + ((0 * N + a) * BT + ra) * RS + 32 * st + kh);
            b_[a] = *reinterpret_cast<const v4i*>(opbuf + ((1 * N + a) * BT + rb) * RS + 32 * st + kh);
          }
        }
      };
      if constexpr (FDB) frag(0, A[0], B[0]);
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        if constexpr (!FDB) {
          if (!(FIRST && st == 0) && k0 + 32 * st >= kvalid) break;  // wave-uniform
        }
        if constexpr (FDB) {
          if (st < 3) frag(st + 1, A[(st + 1) % NF], B[(st + 1) % NF]);
        } else {
          frag(st, A[0], B[0]);
        }
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int b = 0; b < N; ++b) {
            const bool zero = FIRST && st == 0;
            if constexpr (NIB) {
              const v4i x = A[st % NF][a], y = B[st % NF][b];
              const v8i x8 = {x[0], x[1], x[2], x[3], 0, 0, 0, 0}, y8 = {y[0], y[1], y[2], y[3], 0, 0, 0, 0};
              acc.c[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                  x8, y8, zero ? v16f{} : acc.c[a][b], 4, 4, 0, FP4_UNIT_SCALE, 0, FP4_UNIT_SCALE);
            } else {
              acc.c[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[st % NF][a], B[st % NF][b],
                                                                  zero ? v16i{} : acc.c[a][b], 0, 0, 0);
            }
          }
      }
    }
  };
  // chunk c is staged from pre0 (c even) or pre1 (c odd); k_pad >= KSTEP
  const std::true_type first;
  const std::false_type later;
  step(first, pre0, 0);
  if constexpr (PF == 2) {
    int k0 = KSTEP;
    if (k0 < kp) {
      step(later, pre1, k0);
      k0 += KSTEP;
    }
    for (; k0 + 2 * KSTEP <= kp; k0 += 2 * KSTEP) {
      step(later, pre0, k0);
      step(later, pre1, k0 + KSTEP);
    }
    if (k0 < kp) step(later, pre0, k0);
  } else {
    for (int k0 = KSTEP; k0 < kp; k0 += KSTEP) step(later, pre0, k0);
  }
  __builtin_amdgcn_s_setprio(0);
}

// K loop of population p of a tile whose populations are evaluated one
// after another (the within-population sets), with the operand loads run
// across population boundaries: chunk c of population p is staged in
// register set pre[c % 2] (pre0 / pre1, owned by the caller and live across
// calls), consuming a chunk issues the load two chunks ahead, and the loads
// issued by the last two chunks are population p + 1's chunks 0 and 1.  They
// stay in flight through p's epilogue (which issues no vector-memory loads),
// so only the tile's first chunks wait for memory.  On entry pre0 holds (or
// is loading) chunk 0 of p and pre1 chunk 1 (anything when p has one chunk).
// All four waves must call it (it contains barriers).
template <int FAMILY, bool MISSING, bool NIB>
struct PopStream {
  using PL = Planes<FAMILY, MISSING>;
  static constexpr int N = PL::N;
  static constexpr int ITEMS = 2 * N * BT * 8 / BAND_THREADS;
  static_assert(ITEMS == 4 * N, "copy layout");

  __device__ __forceinline__ static int64_t row_bytes(int64_t k_pad) { return NIB ? k_pad / 2 : k_pad; }
  // One population's operand rows (OpBlock) and its K chunks.
  struct Src {
    OpBlock ob;
    int nc;
  };
  __device__ __forceinline__ static Src source(const PopArgs& pa, int p, int64_t n_var, int i0, int j0) {
    Src s;
    const int64_t row = row_bytes(pa.k_pad[p]);
    s.ob = op_block(pa.planes[p], PL::FIRST, row, n_var, i0, j0);
    s.nc = (int)(uniform_i64(row) / KSTEP);
    return s;
  }
  __device__ __forceinline__ static void load(v4i (&pre)[ITEMS], const Src& s, int chunk) {
    op_load<N>(pre, s.ob, (int64_t)chunk * KSTEP);
  }

  // K loop of population p of a tile whose populations are evaluated one
  // after another (the within-population sets), with the operand loads run
  // across population boundaries: chunk c of p is staged in register set
  // pre0 (c even) or pre1 (c odd), owned by the caller and live across
  // calls; consuming a chunk issues the load two chunks ahead, and the loads
  // issued by p's last two chunks are p + 1's chunks 0 and 1.  They stay in
  // flight through p's epilogue (which issues no vector-memory loads), so
  // only the tile's first chunks wait for memory.  On entry pre0 holds chunk
  // 0 of p and, if pre1_ok, pre1 chunk 1 (pre1_ok is false after a one-chunk
  // population, whose K loop refills pre0 only); on return pre1_ok says the
  // same of p + 1.  All four waves must call it (it contains barriers).
  template <int P>
  __device__ __forceinline__ static void products(const PopArgs& pa, int p, int64_t n_var, int i0, int j0, int wr,
                                                  int wc, bool wave_live, unsigned char* __restrict__ opbuf,
                                                  WaveAcc<N, NIB>& acc, v4i (&pre0)[ITEMS], v4i (&pre1)[ITEMS],
                                                  bool& pre1_ok) {
    const Src cur = source(pa, p, n_var, i0, j0);
    const bool has_next = p + 1 < P;
    const Src nxt = source(pa, has_next ? p + 1 : p, n_var, i0, j0);
    const int nc = cur.nc;
    const int lane = threadIdx.x & 63;
    const int dbase = (threadIdx.x >> 3) * OP_ROW + 16 * (threadIdx.x & 7);
    // the first chunk's MFMAs start from zero accumulators; idle waves clear them
    if (!wave_live) {
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = 0; b < N; ++b)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc.c[a][b][r] = 0;
    }
    __builtin_amdgcn_s_setprio(2);
    const int ra = wr + (lane & 31), rb = wc + (lane & 31), kh = 16 * (lane >> 5);
    auto mfmas = [&](auto first_tag) {
      constexpr bool FIRST = decltype(first_tag)::value;
      v4i A[2][N], B[2][N];
      auto frag = [&](int st, v4i* a_, v4i* b_) {
#pragma unroll
        for (int a = 0; a < N; ++a) {
          a_[a] = *reinterpret_cast<const v4i*>(opbuf + ((0 * N + a) * BT + ra) * OP_ROW + 32 * st + kh);
          b_[a] = *reinterpret_cast<const v4i*>(opbuf + ((1 * N + a) * BT + rb) * OP_ROW + 32 * st + kh);
        }
      };
      frag(0, A[0], B[0]);
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        if (st < 3) frag(st + 1, A[(st + 1) & 1], B[(st + 1) & 1]);
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int b = 0; b < N; ++b) {
            const bool zero = FIRST && st == 0;
            if constexpr (NIB) {
              const v4i x = A[st & 1][a], y = B[st & 1][b];
              const v8i x8 = {x[0], x[1], x[2], x[3], 0, 0, 0, 0}, y8 = {y[0], y[1], y[2], y[3], 0, 0, 0, 0};
              acc.c[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                  x8, y8, zero ? v16f{} : acc.c[a][b], 4, 4, 0, FP4_UNIT_SCALE, 0, FP4_UNIT_SCALE);
            } else {
              acc.c[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[st & 1][a], B[st & 1][b],
                                                                  zero ? v16i{} : acc.c[a][b], 0, 0, 0);
            }
          }
      }
    };
    auto stage = [&](v4i (&pre)[ITEMS]) {
      __syncthreads();  // the previous chunk has been read by every wave
#pragma unroll
      for (int m = 0; m < ITEMS; ++m) *reinterpret_cast<v4i*>(opbuf + dbase + 32 * OP_ROW * m) = pre[m];
      __syncthreads();
    };
    // refill pre after consuming chunk c: chunk c + 2 of p, else chunk c % 2
    // of p + 1 (chunk 0 when p + 1 has a single chunk, or of p again past the
    // last population: read, never used).  One load sequence on every path,
    // so the older register set's wait leaves the younger set's loads in flight.
    auto refill = [&](v4i (&pre)[ITEMS], int c) {
      const bool same = c + 2 < nc;
      const int par = c & 1;
      const int chunk = same ? c + 2 : ((has_next && nxt.nc > par) ? par : 0);
      Src s;
      s.ob.side[0] = same ? cur.ob.side[0] : nxt.ob.side[0];
      s.ob.side[1] = same ? cur.ob.side[1] : nxt.ob.side[1];
      s.ob.plane = same ? cur.ob.plane : nxt.ob.plane;
#pragma unroll
      for (int sd = 0; sd < 2; ++sd)
#pragma unroll
        for (int h = 0; h < 2; ++h) s.ob.off[sd][h] = same ? cur.ob.off[sd][h] : nxt.ob.off[sd][h];
      load(pre, s, chunk);
    };
    // chunk 0 outside the loop; pre1's late load (after a one-chunk
    // population) is issued once pre0 is in LDS
    const std::true_type first;
    const std::false_type later;
    stage(pre0);
    if (!pre1_ok && nc > 1) load(pre1, cur, 1);
    refill(pre0, 0);
    if (wave_live) mfmas(first);
    int c = 1;
#pragma unroll 1
    for (; c + 2 <= nc; c += 2) {
      stage(pre1);
      refill(pre1, c);
      if (wave_live) mfmas(later);
      stage(pre0);
      refill(pre0, c + 1);
      if (wave_live) mfmas(later);
    }
    if (c < nc) {
      stage(pre1);
      refill(pre1, c);
      if (wave_live) mfmas(later);
    }
    pre1_ok = nc > 1;
    __builtin_amdgcn_s_setprio(0);
  }

  // first two chunks of a tile's population 0
  __device__ __forceinline__ static void start(const PopArgs& pa, int64_t n_var, int i0, int j0,
                                               v4i (&pre0)[ITEMS], v4i (&pre1)[ITEMS], bool& pre1_ok) {
    const Src s = source(pa, 0, n_var, i0, j0);
    load(pre0, s, 0);
    load(pre1, s, s.nc > 1 ? 1 : 0);
    pre1_ok = s.nc > 1;
  }
};

// Haplotype plane products of accumulator element r (row i, col j) as the
// four 0/1 counts hh = #(a = 1, b = 1), hv = #(a = 1, b valid), vh, vv over
// the pair-valid samples.  int8 planes hold the 0/1 indicators, so the
// products are these counts.  FP4 planes hold the allele plane centred
// (a' = 2a - 1 on valid samples, 0 where missing; see FP4_MINUS_ONE), so
// their products are the centred moments Sab = sum a'b', Sa = sum a' v_j,
// Sb = sum v_i b' and n = vv, and hh = (n + Sa + Sb + Sab) / 4, hv = (n + Sa)
// / 2, vh = (n + Sb) / 2 (float accumulators hold these integers exactly).
// Without missing data only the allele-plane product is formed; the rest
// follows from the per-SNP allele sums (marg) and the sample count.
template <bool MISSING, class ACC>
__device__ __forceinline__ void hap_plane_counts(const ACC& acc, int r, const GLOBAL_AS int32_t* marg, int64_t i,
                                                 int64_t j, int32_t S, int& hh, int& hv, int& vh, int& vv) {
  if constexpr (std::is_same<typename ACC::T, v16f>::value) {
    const int sab = (int)acc.c[0][0][r];
    int sa, sb;
    if constexpr (MISSING) {
      sa = (int)acc.c[0][1][r];
      sb = (int)acc.c[1][0][r];
      vv = (int)acc.c[1][1][r];
      hv = (vv + sa) >> 1;
      vh = (vv + sb) >> 1;
    } else {
      hv = marg[i];
      vh = marg[j];
      vv = S;
      sa = 2 * hv - vv;
      sb = 2 * vh - vv;
    }
    hh = (vv + sa + sb + sab) >> 2;
  } else {
    hh = acc.c[0][0][r];
    if constexpr (MISSING) {
      hv = acc.c[0][1][r];
      vh = acc.c[1][0][r];
      vv = acc.c[1][1][r];
    } else {
      hv = marg[i];  // plane 0 (h) sums
      vh = marg[j];
      vv = S;
    }
  }
}

// Joint counts of accumulator element r (row i, col j) in the reference's
// column order: haplotypes [n11, n10, n01, n00]; genotypes n00 .. n22.
template <int FAMILY, bool MISSING, class ACC>
__device__ __forceinline__ void elem_counts(const ACC& acc, int r,
                                            const int32_t* __restrict__ marg_, int64_t n_var, int64_t i, int64_t j,
                                            int32_t S, int* out) {
  const GLOBAL_AS int32_t* marg = (const GLOBAL_AS int32_t*)marg_;
  if constexpr (FAMILY == PGLD_FAMILY_HAP) {
    int hh, hv, vh, vv;
    hap_plane_counts<MISSING>(acc, r, marg, i, j, S, hh, hv, vh, vv);
    out[0] = hh;
    out[1] = hv - hh;
    out[2] = vh - hh;
    out[3] = vv - hv - vh + hh;
  } else {
    int N[9];  // N[3a+b], a/b: 0 = valid, 1 = e1, 2 = e2
    if constexpr (MISSING) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) N[3 * a + b] = (int)acc.c[a][b][r];
    } else {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) N[3 * (a + 1) + (b + 1)] = (int)acc.c[a][b][r];
      N[0] = S;
      N[1] = marg[n_var + j];      // e1 sums
      N[2] = marg[2 * n_var + j];  // e2 sums
      N[3] = marg[n_var + i];
      N[6] = marg[2 * n_var + i];
    }
    int n11 = N[4], n12 = N[5], n21 = N[7], n22 = N[8];
    int n10 = N[3] - n11 - n12;
    int n20 = N[6] - n21 - n22;
    int n01 = N[1] - n11 - n21;
    int n02 = N[2] - n12 - n22;
    out[0] = N[0] - n10 - n20 - n01 - n02 - n11 - n12 - n21 - n22;
    out[1] = n01; out[2] = n02;
    out[3] = n10; out[4] = n11; out[5] = n12;
    out[6] = n20; out[7] = n21; out[8] = n22;
  }
}

// Centred haplotype moments of accumulator element r: n = vv, Sa = 2 hv - vv,
// Sb = 2 vh - vv, Sab = 4 hh - 2 hv - 2 vh + vv (a' = 2a - 1 over the
// pair-valid samples).  FP4 planes give them straight from the products
// (centred allele plane, see hap_plane_counts); int8 accumulators combine
// the counts in int32.
template <bool MISSING, class ACC>
__device__ __forceinline__ void hap_moments(const ACC& acc, int r, const int32_t* __restrict__ marg_, int64_t i,
                                            int64_t j, int32_t S, double& n, double& a, double& b, double& s) {
  const GLOBAL_AS int32_t* marg = (const GLOBAL_AS int32_t*)marg_;
  if constexpr (std::is_same<typename ACC::T, v16f>::value) {
    s = acc.c[0][0][r];
    if constexpr (MISSING) {
      a = acc.c[0][1][r];
      b = acc.c[1][0][r];
      n = acc.c[1][1][r];
    } else {
      a = 2 * marg[i] - S;
      b = 2 * marg[j] - S;
      n = S;
    }
  } else {
    const int hh = acc.c[0][0][r];
    int hv, vh, vv;
    if constexpr (MISSING) {
      hv = acc.c[0][1][r];
      vh = acc.c[1][0][r];
      vv = acc.c[1][1][r];
    } else {
      hv = marg[i];
      vh = marg[j];
      vv = S;
    }
    n = vv;
    a = 2 * hv - vv;
    b = 2 * vh - vv;
    s = 4 * hh - 2 * (hv + vh) + vv;
  }
}

__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// Lane exchanges of doubles without the LDS crossbar (ds_bpermute would
// queue behind the operand traffic of the K loops).  DPP moves within 16-lane
// rows: row_mirror (lane i <- 15 - i), row_half_mirror (i <- 7 - i within
// eight), quad_perm [2,3,0,1] and [1,0,3,2].
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_ROW_MIRROR = 0x140, DPP_ROW_HALF_MIRROR = 0x141, DPP_QUAD_2301 = 0x4E, DPP_QUAD_1032 = 0xB1;

// Halving exchange across lane bit O = 32 or 16 (v_permlane32_swap /
// v_permlane16_swap swap the odd 32- or 16-lane rows of the first operand
// with the even rows of the second): returns x(l) + x(l ^ O) in lanes with
// bit O clear and y(l ^ O) + y(l) in the others.  With x = y every lane gets
// the pair total, the same bits in both lanes of a pair.
template <int O>
__device__ __forceinline__ double swap_sum(double x, double y) {
  static_assert(O == 32 || O == 16, "permlane swap width");
  const unsigned xl = (unsigned)__double2loint(x), xh = (unsigned)__double2hiint(x);
  const unsigned yl = (unsigned)__double2loint(y), yh = (unsigned)__double2hiint(y);
  if constexpr (O == 32) {
    const auto l = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  } else {
    const auto l = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  }
}

// Wave-wide totals of NV per-lane values, in groups of CH = min(CHMAX, NV
// rounded up to a power of two).  Each group goes through a halving
// butterfly over lane bits 32, 16, 8, 4, 2, 1: at each of the first log2 CH
// levels a lane keeps one half of its values, adds the partner lane's copy
// of that half and hands over the other half (the partner differs in that
// lane bit: l ^ 32, l ^ 16, then 15 - l within the 16-lane row), so a group
// costs CH - 1 + (6 - log2 CH) exchanges instead of 6 CH; the remaining
// levels add full totals (partners 15 - l, 7 - l within eight, l ^ 2,
// l ^ 1).  Afterwards lane l holds the total of value base + ((l >> (6 -
// log2 CH)) & (CH - 1)), identical in the 64 / CH lanes that differ in the
// low bits; the lane whose low bits are zero calls put(index, total).  The
// summation tree is fixed, so the totals are deterministic.
template <int NV, int CHMAX = 8, typename F>
__device__ __forceinline__ void wave_sums_scatter(const double* vals, int lane, F&& put) {
  static_assert(CHMAX == 8 || CHMAX == 4 || CHMAX == 2 || CHMAX == 1, "group size");
  constexpr int CH = NV > 4 && CHMAX >= 8 ? 8 : NV > 2 && CHMAX >= 4 ? 4 : NV > 1 && CHMAX >= 2 ? 2 : 1;
  constexpr int LG = CH == 8 ? 3 : CH == 4 ? 2 : CH == 2 ? 1 : 0;
#pragma unroll
  for (int base = 0; base < NV; base += CH) {
    double v[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) v[k] = base + k < NV ? vals[base + k] : 0.0;
    // level 0: lane bit 32
    if constexpr (LG > 0) {
#pragma unroll
      for (int k = 0; k < CH / 2; ++k) v[k] = swap_sum<32>(v[k], v[k + CH / 2]);
    } else {
      v[0] = swap_sum<32>(v[0], v[0]);
    }
    // level 1: lane bit 16
    if constexpr (LG > 1) {
#pragma unroll
      for (int k = 0; k < CH / 4; ++k) v[k] = swap_sum<16>(v[k], v[k + CH / 4]);
    } else {
      v[0] = swap_sum<16>(v[0], v[0]);
    }
    // level 2: lane bit 8 (partner 15 - l within the row)
    if constexpr (LG > 2) {
      const bool up = (lane & 8) != 0;
      const double send = up ? v[0] : v[1];
      const double keep = up ? v[1] : v[0];
      v[0] = keep + dpp_f64<DPP_ROW_MIRROR>(send);
    } else {
      v[0] += dpp_f64<DPP_ROW_MIRROR>(v[0]);
    }
    v[0] += dpp_f64<DPP_ROW_HALF_MIRROR>(v[0]);
    v[0] += dpp_f64<DPP_QUAD_2301>(v[0]);
    v[0] += dpp_f64<DPP_QUAD_1032>(v[0]);
    const int idx = base + ((lane >> (6 - LG)) & (CH - 1));
    if ((lane & ((64 >> LG) - 1)) == 0 && idx < NV) put(idx, v[0]);
  }
}

__device__ __forceinline__ int32_t key_bin(double d, const double* edges, int32_t n_keys) {
  if (!(d >= edges[0]) || !(d < edges[n_keys])) return -1;
  int lo = 0, hi = n_keys;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (edges[mid] <= d) lo = mid; else hi = mid;
  }
  return lo;
}

// Bin of d among bins [lo, hi) given edges[lo] <= d < edges[hi].
__device__ __forceinline__ int32_t key_bin_in(double d, const double* edges, int lo, int hi) {
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (edges[mid] <= d) lo = mid; else hi = mid;
  }
  return lo;
}

// Workgroups are dealt round-robin to the 8 XCDs (workgroup b runs on XCD
// b % 8).  Tile t of the band schedule is given to workgroup b such that each
// XCD walks one contiguous range of tiles: consecutive tiles share their A rows
// and most B columns, so they hit the same L2.
constexpr int N_XCD = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
  const int64_t x = b % N_XCD, slot = b / N_XCD;
  const int64_t per = n / N_XCD, rem = n % N_XCD;
  return x * per + (x < rem ? x : rem) + slot;
}

// 128 x 128 block of the pair matrix: the (up to) four 64 x 64 plan tiles
// (i0 + 64 (q >> 1), j0 + 64 (q & 1)), q = 0..3, as indices into the plan's
// tile list (-1 where the band has no such tile).
struct SuperTile {
  int32_t i0, j0;
  int32_t t[4];
};

struct BandArgs {
  const BandTile* tiles;
  const double* pos;
  const double* kc;
  const double* edges;
  int64_t n_var;
  double max_dist, origin, width;
  int32_t key_mode, n_keys;
  double* partials;
  int32_t* counts_out;       // counts mode
  const int64_t* pair_base;  // counts mode
  const int32_t* order;      // fused sums of a single band: tile launch order (null: list order)
  // reciprocal table of the within-population haplotype sets with missing
  // data: 1 / (n)_4 for n = 0 .. rtab_n as doubles at byte rtab_off of the
  // dynamic LDS (kernel variants with KV_RTAB only)
  int32_t rtab_n, rtab_off;
  // fused sums of a single band: the launch covers tiles tile0 .. tile0 +
  // grid - 1 of the launch order; gstash (3-4 population genotype moments)
  // holds each workgroup's parked counts, GSTASH_ELEMS int16 per workgroup
  int64_t tile0;
  int16_t* gstash;
};

}  // namespace pgld
