// Plain per-pair counting kernels: one thread per SNP pair, counting sample
// by sample.  A 256-thread block takes 256 consecutive pairs; when the union
// of their SNP columns fits in LDS, each round stages 16 samples x those
// columns (coalesced byte loads along the variant axis of the row-major
// matrix, packed 4 samples per dword), and every thread then reads its two
// columns from LDS.  Otherwise the thread reads its two columns directly.
//
// Counting works on 4 packed int8 samples at a time: the per-byte indicator
// bits v = (x >= 0), h = (x == 1) (haplotypes) or e1 = (x == 1), e2 = (x == 2)
// (genotypes) sit in bit 0 of each byte, and popcounts of their ANDs give the
// joint counts.
#include "common.h"
#include "../../include/pgld.h"

namespace pgld {

constexpr int PL_THREADS = 256;
constexpr int PL_KS = 16;        // samples per staging round
constexpr int PL_KW = PL_KS / 4; // dwords per staged column
constexpr int PL_MAXC = 2048;    // staged columns (2048 * 16 B = 32 KB)

struct Acc4 {
  int hh, hv, vh, vv;
};
struct Acc9 {
  int n[9];  // {v, e1, e2} x {v, e1, e2}
};

__device__ __forceinline__ uint32_t valid_bits(uint32_t x) { return (~x >> 7) & 0x01010101u; }

__device__ __forceinline__ void acc_hap(Acc4& a, uint32_t xi, uint32_t xj) {
  uint32_t vi = valid_bits(xi), vj = valid_bits(xj);
  uint32_t hi = xi & vi, hj = xj & vj;
  a.hh += __popc(hi & hj);
  a.hv += __popc(hi & vj);
  a.vh += __popc(vi & hj);
  a.vv += __popc(vi & vj);
}

__device__ __forceinline__ void acc_geno(Acc9& a, uint32_t xi, uint32_t xj) {
  uint32_t pi[3], pj[3];
  pi[0] = valid_bits(xi);
  pj[0] = valid_bits(xj);
  pi[1] = xi & pi[0];
  pj[1] = xj & pj[0];
  pi[2] = (xi >> 1) & pi[0];
  pj[2] = (xj >> 1) & pj[0];
#pragma unroll
  for (int a_ = 0; a_ < 3; ++a_)
#pragma unroll
    for (int b_ = 0; b_ < 3; ++b_) a.n[3 * a_ + b_] += __popc(pi[a_] & pj[b_]);
}

template <int FAMILY>
struct PlainAcc;
template <>
struct PlainAcc<PGLD_FAMILY_HAP> {
  Acc4 a{0, 0, 0, 0};
  __device__ void add(uint32_t xi, uint32_t xj) { acc_hap(a, xi, xj); }
  __device__ void store(int64_t p, int32_t* counts, int32_t* nvalid) const {
    int4 c;
    c.x = a.hh;
    c.y = a.hv - a.hh;
    c.z = a.vh - a.hh;
    c.w = a.vv - a.hv - a.vh + a.hh;
    reinterpret_cast<int4*>(counts)[p] = c;
    if (nvalid) nvalid[p] = a.vv;
  }
};
template <>
struct PlainAcc<PGLD_FAMILY_GENO> {
  Acc9 a{{0, 0, 0, 0, 0, 0, 0, 0, 0}};
  __device__ void add(uint32_t xi, uint32_t xj) { acc_geno(a, xi, xj); }
  __device__ void store(int64_t p, int32_t* counts, int32_t* nvalid) const {
    const int* N = a.n;  // N[3*a+b], a/b: 0 = valid, 1 = e1, 2 = e2
    int n11 = N[4], n12 = N[5], n21 = N[7], n22 = N[8];
    int n10 = N[3] - n11 - n12;
    int n20 = N[6] - n21 - n22;
    int n01 = N[1] - n11 - n21;
    int n02 = N[2] - n12 - n22;
    int n00 = N[0] - n10 - n20 - n01 - n02 - n11 - n12 - n21 - n22;
    int32_t* row = counts + p * 9;
    row[0] = n00; row[1] = n01; row[2] = n02;
    row[3] = n10; row[4] = n11; row[5] = n12;
    row[6] = n20; row[7] = n21; row[8] = n22;
    if (nvalid) nvalid[p] = N[0];
  }
};

__device__ __forceinline__ int block_min(int v, int* red) {
  for (int o = 32; o > 0; o >>= 1) {
    int w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = red[0];
  for (int k = 1; k < PL_THREADS / 64; ++k) r = red[k] < r ? red[k] : r;
  return r;
}

template <int FAMILY>
__global__ __launch_bounds__(PL_THREADS) void k_counts_plain(const int8_t* __restrict__ data, int64_t n_var,
                                                             const int32_t* __restrict__ rows, int64_t n_sel,
                                                             const int32_t* __restrict__ ii,
                                                             const int32_t* __restrict__ jj, int64_t n_pairs,
                                                             int32_t* __restrict__ counts,
                                                             int32_t* __restrict__ nvalid) {
  __shared__ uint32_t stage[PL_MAXC * PL_KW];
  __shared__ int64_t srow[PL_KS];
  __shared__ int red[PL_THREADS / 64];

  const int tid = threadIdx.x;
  const int64_t p = (int64_t)blockIdx.x * PL_THREADS + tid;
  const bool in_range = p < n_pairs;
  int i = in_range ? ii[p] : 0;
  int j = in_range ? jj[p] : 0;
  // Out-of-range SNP indices never index memory: such pairs get -1 counts.
  const bool active = in_range && i >= 0 && j >= 0 && (int64_t)i < n_var && (int64_t)j < n_var;
  if (!active) i = j = 0x7fffffff;
  const int lo = block_min(i < j ? i : j, red);
  const int hi = -block_min(active ? -(i > j ? i : j) : 0x7fffffff, red);
  const int ncols = lo == 0x7fffffff ? PL_MAXC + 1 : hi - lo + 1;
  PlainAcc<FAMILY> acc;

  if (ncols <= PL_MAXC) {
    const int ci = (i - lo) * PL_KW, cj = (j - lo) * PL_KW;
    for (int64_t k0 = 0; k0 < n_sel; k0 += PL_KS) {
      const int ks = (int)((n_sel - k0) < PL_KS ? (n_sel - k0) : PL_KS);
      __syncthreads();
      if (tid < PL_KS) srow[tid] = tid < ks ? (int64_t)(rows ? rows[k0 + tid] : k0 + tid) * n_var : -1;
      __syncthreads();
      for (int e = tid; e < ncols * PL_KW; e += PL_THREADS) {
        const int c = e % ncols;  // consecutive threads -> consecutive columns (coalesced)
        const int w = e / ncols;
        uint32_t word = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int64_t base = srow[4 * w + b];
          const uint32_t byte = base >= 0 ? (uint8_t)data[base + lo + c] : 0xffu;  // pad = missing
          word |= byte << (8 * b);
        }
        stage[c * PL_KW + w] = word;
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int w = 0; w < PL_KW; ++w) acc.add(stage[ci + w], stage[cj + w]);
      }
    }
  } else if (active) {
    for (int64_t k0 = 0; k0 < n_sel; k0 += 4) {
      uint32_t xi = 0, xj = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        uint32_t bi = 0xffu, bj = 0xffu;
        if (k0 + b < n_sel) {
          const int64_t base = (int64_t)(rows ? rows[k0 + b] : k0 + b) * n_var;
          bi = (uint8_t)data[base + i];
          bj = (uint8_t)data[base + j];
        }
        xi |= bi << (8 * b);
        xj |= bj << (8 * b);
      }
      acc.add(xi, xj);
    }
  }
  if (active) {
    acc.store(p, counts, nvalid);
  } else if (in_range) {
    for (int c = 0; c < (FAMILY == PGLD_FAMILY_HAP ? 4 : 9); ++c)
      counts[p * (FAMILY == PGLD_FAMILY_HAP ? 4 : 9) + c] = -1;
    if (nvalid) nvalid[p] = -1;
  }
}

template <int FAMILY>
static int launch_plain(const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows, int64_t n_sel,
                        const int32_t* ii, const int32_t* jj, int64_t n_pairs, int32_t* counts, int32_t* nvalid,
                        void* stream, const char* name) {
  PGLD_REQUIRE(n_rows >= 0 && n_var >= 0 && n_pairs >= 0, "%s: negative size", name);
  PGLD_REQUIRE(n_var < 0x7fffffff, "%s: n_var must fit int32", name);
  if (!rows) n_sel = n_rows;
  PGLD_REQUIRE(n_sel >= 0, "%s: negative n_sel", name);
  if (n_pairs == 0) return PGLD_OK;
  PGLD_REQUIRE(data && ii && jj && counts, "%s: null pointer", name);
  hipLaunchKernelGGL(k_counts_plain<FAMILY>, dim3((unsigned)ceil_div(n_pairs, PL_THREADS)), dim3(PL_THREADS), 0,
                     as_stream(stream), data, n_var, rows, n_sel, ii, jj, n_pairs, counts, nvalid);
  return check_launch(name);
}

}  // namespace pgld

using namespace pgld;

extern "C" {

int pgld_counts_pairs_hap(const int8_t* hap, int64_t n_rows, int64_t n_var, const int32_t* rows, int64_t n_sel,
                          const int32_t* idx_i, const int32_t* idx_j, int64_t n_pairs, int32_t* counts,
                          int32_t* n_valid, void* stream) {
  return launch_plain<PGLD_FAMILY_HAP>(hap, n_rows, n_var, rows, n_sel, idx_i, idx_j, n_pairs, counts, n_valid,
                                       stream, "pgld_counts_pairs_hap");
}

int pgld_counts_pairs_geno(const int8_t* geno, int64_t n_rows, int64_t n_var, const int32_t* rows, int64_t n_sel,
                           const int32_t* idx_i, const int32_t* idx_j, int64_t n_pairs, int32_t* counts,
                           int32_t* n_valid, void* stream) {
  return launch_plain<PGLD_FAMILY_GENO>(geno, n_rows, n_var, rows, n_sel, idx_i, idx_j, n_pairs, counts, n_valid,
                                        stream, "pgld_counts_pairs_geno");
}

}  // extern "C"
