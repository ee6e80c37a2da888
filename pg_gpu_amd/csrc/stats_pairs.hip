// Per-pair statistic evaluation from per-population counts, driven by a term
// table (one weighted term per (statistic, population tuple), as moments'
// averaging specs list them).  One thread per pair; the term table is read
// with uniform (scalar) loads, so every lane takes the same branch.
#include "common.h"
#include "ld_math.h"
#include "../../include/pgld.h"

namespace pgld {

struct CountPtrs {
  const int32_t* c[4];
  const int32_t* nv[4];
};

template <int NC, int P>
struct PairCounts {
  double v[P][NC];
  __device__ const double* pop(int q) const {
    // q is wave-uniform: a uniform branch chain, no private-memory indexing
    if (P > 3 && q == 3) return v[P > 3 ? 3 : 0];
    if (P > 2 && q == 2) return v[P > 2 ? 2 : 0];
    if (P > 1 && q == 1) return v[P > 1 ? 1 : 0];
    return v[0];
  }
};

// c[4] carries the population's sample size n (n_valid or the sum of counts)
__device__ __forceinline__ HapP hp_of(const double* c) {
  HapP p = hap_pop(c[0], c[1], c[2], c[3]);
  p.n = c[4];
  return p;
}

// Haplotype term (haplotype_kernels.py:335-581 case dispatch)
template <int P>
__device__ double hap_term(const PairCounts<5, P>& pc, int type, int a, int b, int c, int d) {
  if (type == ST_DD) {
    HapP A = hp_of(pc.pop(a));
    if (a == b) return hap_dd_single(A);
    return hap_dd_between(A, hp_of(pc.pop(b)));
  }
  if (type == ST_DZ) {
    if (a == b && b == c) return hap_dz_iii(hp_of(pc.pop(a)));
    if (a == b) return hap_dz_iij(hp_of(pc.pop(a)), hp_of(pc.pop(c)));
    if (a == c) return hap_dz_iji(hp_of(pc.pop(a)), hp_of(pc.pop(b)));
    if (b == c) return hap_dz_ijj(hp_of(pc.pop(a)), hp_of(pc.pop(b)));
    return hap_dz_diff(hp_of(pc.pop(a)), hp_of(pc.pop(b)), hp_of(pc.pop(c)));
  }
  // pi2 case classification (haplotype_kernels.py:460-486)
  const int q[4] = {a, b, c, d};
  int nu = 0, uniq[4], mult[4];
  for (int t = 0; t < 4; ++t) {
    int k = 0;
    while (k < nu && uniq[k] != q[t]) ++k;
    if (k == nu) {
      uniq[nu] = q[t];
      mult[nu] = 0;
      ++nu;
    }
    ++mult[k];
  }
  int mc = 0, tp = -1, sp = -1;
  for (int k = 0; k < nu; ++k) {
    if (mult[k] > mc) mc = mult[k];
    if (mult[k] == 3) tp = uniq[k];
    if (mult[k] == 1) sp = uniq[k];
  }
  if (nu == 1) return hap_pi2_iiii(hp_of(pc.pop(a)));
  if (mc == 3) return hap_pi2_triple(hp_of(pc.pop(tp)), hp_of(pc.pop(sp)));
  if (a == b && c == d) {
    HapP A = hp_of(pc.pop(a)), C = hp_of(pc.pop(c));
    return type == ST_PI2_SYM ? hap_pi2_iikk_sym(A, C) : hap_pi2_iikk(A, C);
  }
  if (a == b) return hap_pi2_iikl(hp_of(pc.pop(a)), hp_of(pc.pop(c)), hp_of(pc.pop(d)));
  if (c == d) return hap_pi2_ijkk(hp_of(pc.pop(a)), hp_of(pc.pop(b)), hp_of(pc.pop(c)));
  if ((a == c && b == d) || (a == d && b == c)) return hap_pi2_ijij(hp_of(pc.pop(a)), hp_of(pc.pop(b)));
  if (nu == 3) {
    int s, x, y;
    if (a == c) { s = a; x = b; y = d; }
    else if (a == d) { s = a; x = b; y = c; }
    else if (b == c) { s = b; x = a; y = d; }
    else { s = b; x = a; y = c; }
    return hap_pi2_shared(hp_of(pc.pop(s)), hp_of(pc.pop(x)), hp_of(pc.pop(y)));
  }
  return hap_pi2_alldiff(hp_of(pc.pop(a)), hp_of(pc.pop(b)), hp_of(pc.pop(c)), hp_of(pc.pop(d)));
}

template <int P>
__device__ double geno_term(const PairCounts<9, P>& pc, int type, int a, int b, int c, int d) {
  if (type == ST_SIGMA_D2) return geno_sigma_d2(geno_pop(pc.pop(a)));
  if (type == ST_DD) {
    GenoP A = geno_pop(pc.pop(a));
    if (a == b) return geno_dd_single(A);
    return geno_dd_between(A, geno_pop(pc.pop(b)));
  }
  if (type == ST_DZ) {
    if (a == b && b == c) return geno_dz_iii(geno_pop(pc.pop(a)));
    if (a == b) return geno_dz_iij(geno_pop(pc.pop(a)), geno_pop(pc.pop(c)));
    if (a == c) return geno_dz_iji(geno_pop(pc.pop(a)), geno_pop(pc.pop(b)));
    if (b == c) return geno_dz_ijj(geno_pop(pc.pop(a)), geno_pop(pc.pop(b)));
    return geno_dz_diff(geno_pop(pc.pop(a)), geno_pop(pc.pop(b)), geno_pop(pc.pop(c)));
  }
  // pi2 case classification (genotype_kernels.py:683-711)
  const int q[4] = {a, b, c, d};
  int nu = 0, uniq[4], mult[4];
  for (int t = 0; t < 4; ++t) {
    int k = 0;
    while (k < nu && uniq[k] != q[t]) ++k;
    if (k == nu) {
      uniq[nu] = q[t];
      mult[nu] = 0;
      ++nu;
    }
    ++mult[k];
  }
  int mc = 0, tp = -1, sp = -1;
  for (int k = 0; k < nu; ++k) {
    if (mult[k] > mc) mc = mult[k];
    if (mult[k] == 3) tp = uniq[k];
    if (mult[k] == 1) sp = uniq[k];
  }
  if (nu == 1) return geno_pi2_iiii(geno_pop(pc.pop(a)));
  if (mc == 3) {
    GenoP T = geno_pop(pc.pop(tp)), S = geno_pop(pc.pop(sp));
    return (a == tp && b == tp) ? geno_pi2_iiij(T, S) : geno_pi2_ijii(T, S);
  }
  if (a == b && c == d) return geno_pi2_iikk(geno_pop(pc.pop(a)), geno_pop(pc.pop(c)));
  if (a == b) return geno_pi2_iikl(geno_pop(pc.pop(a)), geno_pop(pc.pop(c)), geno_pop(pc.pop(d)));
  if (c == d) return geno_pi2_ijkk(geno_pop(pc.pop(a)), geno_pop(pc.pop(b)), geno_pop(pc.pop(c)));
  if ((a == c && b == d) || (a == d && b == c)) return geno_pi2_ijij(geno_pop(pc.pop(a)), geno_pop(pc.pop(b)));
  if (nu == 3) {
    int s, x, y;
    if (a == c) { s = a; x = b; y = d; }
    else if (a == d) { s = a; x = b; y = c; }
    else if (b == c) { s = b; x = a; y = d; }
    else { s = b; x = a; y = c; }
    return geno_pi2_shared(geno_pop(pc.pop(s)), geno_pop(pc.pop(x)), geno_pop(pc.pop(y)));
  }
  return geno_pi2_alldiff(geno_pop(pc.pop(a)), geno_pop(pc.pop(b)), geno_pop(pc.pop(c)), geno_pop(pc.pop(d)));
}

template <int FAMILY, int P>
__global__ __launch_bounds__(256) void k_stats_pairs(CountPtrs cp, int64_t n, const pgld_term* __restrict__ terms,
                                                     const int32_t* __restrict__ off, int32_t n_stats,
                                                     double* __restrict__ out) {
  constexpr int NC = FAMILY == PGLD_FAMILY_HAP ? 4 : 9;
  constexpr int NV = FAMILY == PGLD_FAMILY_HAP ? 5 : 9;  // haplotypes keep n in slot 4
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  PairCounts<NV, P> pc;
#pragma unroll
  for (int q = 0; q < P; ++q) {
#pragma unroll
    for (int k = 0; k < NC; ++k) pc.v[q][k] = (double)cp.c[q][t * NC + k];
    if constexpr (FAMILY == PGLD_FAMILY_HAP)
      pc.v[q][4] = cp.nv[q] ? (double)cp.nv[q][t] : pc.v[q][0] + pc.v[q][1] + pc.v[q][2] + pc.v[q][3];
  }
  for (int s = 0; s < n_stats; ++s) {
    double acc = 0.0;
    for (int e = off[s]; e < off[s + 1]; ++e) {
      const pgld_term tm = terms[e];
      double v;
      if constexpr (FAMILY == PGLD_FAMILY_HAP)
        v = hap_term<P>(pc, tm.type, tm.pops[0], tm.pops[1], tm.pops[2], tm.pops[3]);
      else
        v = geno_term<P>(pc, tm.type, tm.pops[0], tm.pops[1], tm.pops[2], tm.pops[3]);
      acc = (e == off[s] && tm.weight == 1.0) ? v : acc + tm.weight * v;
    }
    out[t * n_stats + s] = acc;
  }
}

template <int FAMILY>
static int launch_stats(int32_t n_pops, CountPtrs cp, int64_t n, const pgld_term* terms, const int32_t* off,
                        int32_t n_stats, double* out, hipStream_t s) {
  dim3 g((unsigned)ceil_div(n, 256)), b(256);
  switch (n_pops) {
    case 1: hipLaunchKernelGGL((k_stats_pairs<FAMILY, 1>), g, b, 0, s, cp, n, terms, off, n_stats, out); break;
    case 2: hipLaunchKernelGGL((k_stats_pairs<FAMILY, 2>), g, b, 0, s, cp, n, terms, off, n_stats, out); break;
    case 3: hipLaunchKernelGGL((k_stats_pairs<FAMILY, 3>), g, b, 0, s, cp, n, terms, off, n_stats, out); break;
    default: hipLaunchKernelGGL((k_stats_pairs<FAMILY, 4>), g, b, 0, s, cp, n, terms, off, n_stats, out); break;
  }
  return check_launch("k_stats_pairs");
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_stats_pairs(int32_t family, int32_t n_pops, const int32_t* const* counts,
                                const int32_t* const* n_valid, int64_t n_pairs, const pgld_term* terms,
                                const int32_t* stat_offsets, int32_t n_stats, double* out, void* stream) {
  PGLD_REQUIRE(family == PGLD_FAMILY_HAP || family == PGLD_FAMILY_GENO, "pgld_stats_pairs: bad family %d", family);
  PGLD_REQUIRE(n_pops >= 1 && n_pops <= 4, "pgld_stats_pairs: n_pops must be 1..4");
  PGLD_REQUIRE(n_pairs >= 0 && n_stats >= 0, "pgld_stats_pairs: negative size");
  if (n_pairs == 0 || n_stats == 0) return PGLD_OK;
  PGLD_REQUIRE(counts && terms && stat_offsets && out, "pgld_stats_pairs: null pointer");
  CountPtrs cp{};
  for (int q = 0; q < n_pops; ++q) {
    PGLD_REQUIRE(counts[q], "pgld_stats_pairs: null counts pointer for pop %d", q);
    cp.c[q] = counts[q];
    cp.nv[q] = n_valid ? n_valid[q] : nullptr;
  }
  hipStream_t s = as_stream(stream);
  if (family == PGLD_FAMILY_HAP) return launch_stats<PGLD_FAMILY_HAP>(n_pops, cp, n_pairs, terms, stat_offsets,
                                                                       n_stats, out, s);
  return launch_stats<PGLD_FAMILY_GENO>(n_pops, cp, n_pairs, terms, stat_offsets, n_stats, out, s);
}
