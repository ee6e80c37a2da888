// Error handling, value scans, pair enumeration, pair keys, keyed sums and
// the r / r^2 / D' kernels of libpgld.
#include <algorithm>

#include "common.h"
#include "ld_math.h"
#include "../../include/pgld.h"

namespace pgld {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

const char* get_error() { return g_err.c_str(); }

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return PGLD_OK;
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return check_launch("k_fill_i32");
}

// ---------------------------------------------------------------------------
// Value scan (has_missing / domain check)
// ---------------------------------------------------------------------------
__global__ void k_scan_rows(const int8_t* __restrict__ data, int64_t n_var, const int32_t* __restrict__ rows,
                            int64_t n_sel, int32_t max_value, int32_t* __restrict__ flags) {
  int64_t total = n_sel * n_var;
  int any_neg = 0, any_bad = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = t / n_var, c = t - r * n_var;
    int64_t row = rows ? rows[r] : r;
    int v = data[row * n_var + c];
    any_neg |= (v < 0);
    any_bad |= (v < -1) | (v > max_value);
  }
  // wave-level OR, then one atomic per wave
  unsigned long long mneg = __ballot(any_neg);
  unsigned long long mbad = __ballot(any_bad);
  if ((threadIdx.x & 63) == 0) {
    if (mneg) atomicOr(&flags[0], 1);
    if (mbad) atomicOr(&flags[1], 1);
  }
}

// Per-variant allele counts over selected rows: alt[v] = sum of values > 0
// (haplotypes: number of 1s; genotypes: dosage sum), valid[v] = number of
// values >= 0.  One thread per variant, rows walked in order.
__global__ void k_allele_counts(const int8_t* __restrict__ data, int64_t n_var, const int32_t* __restrict__ rows,
                                int64_t n_sel, int32_t* __restrict__ alt, int32_t* __restrict__ valid) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_var) return;
  int a = 0, n = 0;
  for (int64_t k = 0; k < n_sel; ++k) {
    const int x = data[(int64_t)(rows ? rows[k] : k) * n_var + v];
    a += x > 0 ? x : 0;
    n += x >= 0;
  }
  alt[v] = a;
  valid[v] = n;
}

// ---------------------------------------------------------------------------
// Pair enumeration
// ---------------------------------------------------------------------------
__global__ void k_pairs_per_variant(const double* __restrict__ pos, int64_t n, double max_dist,
                                    int64_t* __restrict__ per) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double ub = pos[v] + max_dist;
  // first index with pos > ub (searchsorted side='right'), searching (v, n)
  int64_t lo = v + 1, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (pos[mid] <= ub) lo = mid + 1; else hi = mid;
  }
  int64_t c = lo - v - 1;
  per[v] = c > 0 ? c : 0;
}

__global__ void k_pairs_fill(const int64_t* __restrict__ excl, int64_t v_lo, int64_t n_anchor, int64_t n_pairs,
                             int32_t* __restrict__ idx_i, int32_t* __restrict__ idx_j) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_pairs) return;
  // last anchor k with excl[k] <= t
  int64_t lo = 0, hi = n_anchor;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (excl[mid] <= t) lo = mid; else hi = mid;
  }
  int64_t i = v_lo + lo;
  idx_i[t] = (int32_t)i;
  idx_j[t] = (int32_t)(i + 1 + (t - excl[lo]));
}

// ---------------------------------------------------------------------------
// Keys and keyed sums
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t bin_of(double d, const double* __restrict__ edges, int32_t n_keys) {
  // digitize(d, edges) - 1 for ascending edges: k with e_k <= d < e_{k+1}
  if (!(d >= edges[0]) || !(d < edges[n_keys])) return -1;
  int lo = 0, hi = n_keys;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (edges[mid] <= d) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void k_pair_keys(const double* __restrict__ coord, const int32_t* __restrict__ ii,
                            const int32_t* __restrict__ jj, int64_t n, int32_t mode,
                            const double* __restrict__ edges, int32_t n_keys, double origin, double width,
                            int32_t* __restrict__ keys) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double ci = coord[ii[t]], cj = coord[jj[t]];
  int32_t k;
  if (mode == PGLD_KEY_WINDOW) {
    double w = floor((ci - origin) / width);
    k = (w >= 0.0 && w < (double)n_keys) ? (int32_t)w : -1;
  } else {
    double d = cj - ci;
    if (mode == PGLD_KEY_BIN_ABS) d = fabs(d);
    k = bin_of(d, edges, n_keys);
  }
  keys[t] = k;
}

// One block handles a contiguous slice of pairs; per-block sums are kept in
// LDS (float64 LDS atomics) and flushed with one global atomic per nonzero
// (key, column).
__global__ void k_keyed_sums(const double* __restrict__ vals, const int32_t* __restrict__ keys, int64_t n,
                             int32_t n_cols, int32_t n_keys, double* __restrict__ sums) {
  extern __shared__ double acc[];
  const int32_t nk = n_keys * n_cols;
  for (int32_t t = threadIdx.x; t < nk; t += blockDim.x) acc[t] = 0.0;
  __syncthreads();
  int64_t per_block = (n + gridDim.x - 1) / gridDim.x;
  int64_t lo = (int64_t)blockIdx.x * per_block;
  int64_t hi = lo + per_block < n ? lo + per_block : n;
  for (int64_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
    int32_t k = keys[p];
    if (k < 0 || k >= n_keys) continue;
    const double* row = vals + p * n_cols;
    for (int32_t c = 0; c < n_cols; ++c) atomicAdd(&acc[k * n_cols + c], row[c]);
  }
  __syncthreads();
  for (int32_t t = threadIdx.x; t < nk; t += blockDim.x) {
    double v = acc[t];
    if (v != 0.0) atomicAdd(&sums[t], v);
  }
}

// ---------------------------------------------------------------------------
// r family
// ---------------------------------------------------------------------------
__global__ void k_r_family(const int32_t* __restrict__ counts, const int32_t* __restrict__ nvalid, int64_t n,
                           int32_t which, double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int4 c = reinterpret_cast<const int4*>(counts)[t];
  double a = c.x, b = c.y, cc = c.z, d = c.w;
  double nn = nvalid ? (double)nvalid[t] : a + b + cc + d;
  double r, r2, dp;
  hap_r_family(a, b, cc, d, nn, r, r2, dp);
  out[t] = which == 0 ? r : (which == 1 ? r2 : dp);
}

}  // namespace pgld

using namespace pgld;

extern "C" {

int pgld_version(void) { return 10000; }

const char* pgld_last_error(void) { return get_error(); }

int pgld_scan_rows(const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows, int64_t n_sel,
                   int32_t max_value, int32_t* flags, void* stream) {
  PGLD_REQUIRE(data && flags, "pgld_scan_rows: null pointer");
  PGLD_REQUIRE(n_rows >= 0 && n_var >= 0, "pgld_scan_rows: negative size");
  if (!rows) n_sel = n_rows;
  hipStream_t s = as_stream(stream);
  int rc = fill_i32(flags, 2, 0, s);
  if (rc) return rc;
  int64_t total = n_sel * n_var;
  if (total == 0) return PGLD_OK;
  int64_t blocks = ceil_div(total, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_scan_rows, dim3((unsigned)blocks), dim3(256), 0, s, data, n_var, rows, n_sel, max_value,
                     flags);
  return check_launch("k_scan_rows");
}

int pgld_allele_counts(const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows, int64_t n_sel,
                       int32_t* alt, int32_t* valid, void* stream) {
  PGLD_REQUIRE(n_rows >= 0 && n_var >= 0, "pgld_allele_counts: negative size");
  if (!rows) n_sel = n_rows;
  if (n_var == 0) return PGLD_OK;
  PGLD_REQUIRE(data && alt && valid, "pgld_allele_counts: null pointer");
  hipLaunchKernelGGL(k_allele_counts, dim3((unsigned)ceil_div(n_var, 256)), dim3(256), 0, as_stream(stream), data,
                     n_var, rows, n_sel, alt, valid);
  return check_launch("k_allele_counts");
}

int pgld_pairs_per_variant(const double* pos, int64_t n, double max_dist, int64_t* per, void* stream) {
  PGLD_REQUIRE(n >= 0, "pgld_pairs_per_variant: negative n");
  if (n == 0) return PGLD_OK;
  PGLD_REQUIRE(pos && per, "pgld_pairs_per_variant: null pointer");
  hipLaunchKernelGGL(k_pairs_per_variant, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, as_stream(stream), pos,
                     n, max_dist, per);
  return check_launch("k_pairs_per_variant");
}

int pgld_pairs_fill(const int64_t* excl, int64_t v_lo, int64_t n_anchor, int64_t n_pairs, int32_t* idx_i,
                    int32_t* idx_j, void* stream) {
  PGLD_REQUIRE(n_pairs >= 0 && n_anchor >= 0, "pgld_pairs_fill: negative size");
  if (n_pairs == 0) return PGLD_OK;
  PGLD_REQUIRE(excl && idx_i && idx_j && n_anchor > 0, "pgld_pairs_fill: null pointer / no anchors");
  hipLaunchKernelGGL(k_pairs_fill, dim3((unsigned)ceil_div(n_pairs, 256)), dim3(256), 0, as_stream(stream), excl,
                     v_lo, n_anchor, n_pairs, idx_i, idx_j);
  return check_launch("k_pairs_fill");
}

int pgld_pair_keys(const double* coord, const int32_t* idx_i, const int32_t* idx_j, int64_t n_pairs, int32_t mode,
                   const double* edges, int32_t n_keys, double origin, double width, int32_t* keys, void* stream) {
  PGLD_REQUIRE(n_pairs >= 0 && n_keys > 0, "pgld_pair_keys: bad size");
  PGLD_REQUIRE(mode == PGLD_KEY_BIN_ABS || mode == PGLD_KEY_BIN || mode == PGLD_KEY_WINDOW,
               "pgld_pair_keys: unknown mode %d", mode);
  PGLD_REQUIRE(mode == PGLD_KEY_WINDOW ? width > 0.0 : edges != nullptr, "pgld_pair_keys: missing edges/width");
  if (n_pairs == 0) return PGLD_OK;
  PGLD_REQUIRE(coord && idx_i && idx_j && keys, "pgld_pair_keys: null pointer");
  hipLaunchKernelGGL(k_pair_keys, dim3((unsigned)ceil_div(n_pairs, 256)), dim3(256), 0, as_stream(stream), coord,
                     idx_i, idx_j, n_pairs, mode, edges, n_keys, origin, width, keys);
  return check_launch("k_pair_keys");
}

int pgld_keyed_sums(const double* values, const int32_t* keys, int64_t n_pairs, int32_t n_cols, int32_t n_keys,
                    double* sums, void* stream) {
  PGLD_REQUIRE(n_pairs >= 0 && n_cols > 0 && n_keys > 0, "pgld_keyed_sums: bad size");
  PGLD_REQUIRE((int64_t)n_keys * n_cols * 8 <= 64 * 1024, "pgld_keyed_sums: n_keys*n_cols above 8192");
  if (n_pairs == 0) return PGLD_OK;
  PGLD_REQUIRE(values && keys && sums, "pgld_keyed_sums: null pointer");
  int64_t blocks = ceil_div(n_pairs, 4096);
  if (blocks > 2048) blocks = 2048;
  size_t lds = (size_t)n_keys * n_cols * sizeof(double);
  hipLaunchKernelGGL(k_keyed_sums, dim3((unsigned)blocks), dim3(256), lds, as_stream(stream), values, keys, n_pairs,
                     n_cols, n_keys, sums);
  return check_launch("k_keyed_sums");
}

int pgld_r_family(const int32_t* counts, const int32_t* n_valid, int64_t n_pairs, int32_t which, double* out,
                  void* stream) {
  PGLD_REQUIRE(which >= 0 && which <= 2, "pgld_r_family: which must be 0, 1 or 2");
  PGLD_REQUIRE(n_pairs >= 0, "pgld_r_family: negative size");
  if (n_pairs == 0) return PGLD_OK;
  PGLD_REQUIRE(counts && out, "pgld_r_family: null pointer");
  hipLaunchKernelGGL(k_r_family, dim3((unsigned)ceil_div(n_pairs, 256)), dim3(256), 0, as_stream(stream), counts,
                     n_valid, n_pairs, which, out);
  return check_launch("k_r_family");
}

}  // extern "C"
