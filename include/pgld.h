/* libpgld -- MI355X (gfx950) LD pair-statistics library, C ABI.
 *
 * Conventions
 *  - Every buffer argument is a device pointer owned by the caller unless the
 *    comment says "host".  The library owns only band-plan memory.
 *  - Every call is asynchronous on the given hipStream_t (passed as void*);
 *    NULL means the default stream.
 *  - Every call returns an int status: 0 = OK, negative = invalid argument
 *    (-1), input outside the value domain (-2), capacity exceeded (-3) or
 *    unsupported (-4); positive = the hipError_t of a failed HIP call.  The
 *    message of the last failure on the calling thread is pgld_last_error().
 *    Nothing aborts and no C++ exception crosses this boundary.
 *  - Sizes and offsets are int64; variant and row indices are int32, as in
 *    the reference's CuPy arrays.
 *  - Matrices are row-major int8: haplotypes (n_rows = haplotypes, n_var),
 *    values {-1 missing, 0, 1}; genotypes (n_rows = individuals, n_var),
 *    values {-1, 0, 1, 2}.
 *  - Not thread-safe per plan: use one band plan per host thread or stream.
 *
 * The reference (kr-colab/pg_gpu) has no C ABI: its operator boundary is the
 * CuPy-level Python functions named in each comment, which launch NVRTC
 * kernels through _launch (haplotype_kernels.py:11-13,
 * genotype_kernels.py:18-20).  Each entry point below replaces one of them.
 */
#ifndef PGLD_H
#define PGLD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGLD_FAMILY_HAP 0
#define PGLD_FAMILY_GENO 1

#define PGLD_STAT_DD 0
#define PGLD_STAT_DZ 1
#define PGLD_STAT_PI2 2      /* fused-kernel pi2 form (haplotype_kernels.py:167-304) */
#define PGLD_STAT_PI2_SYM 3  /* CuPy pi2 form; differs only for pi2(i,i,k,k) */
#define PGLD_STAT_SIGMA_D2 4 /* genotype sigma_D^2 (genotype_kernels.py:91-169) */

#define PGLD_KEY_BIN_ABS 0 /* bin of |coord[j]-coord[i]| against edges  */
#define PGLD_KEY_BIN 1     /* bin of coord[j]-coord[i] against edges       */
#define PGLD_KEY_WINDOW 2  /* floor((coord[i]-origin)/width), left SNP     */

/* One weighted term of a named statistic: weight * type(pops...).  Mirrors
 * one (weight, stat_type, pop_indices) entry of generate_stat_specs
 * (ld_pipeline.py:388-457).  32 bytes. */
typedef struct pgld_term {
  int32_t type;    /* PGLD_STAT_* */
  int32_t pops[4]; /* population indices; unused slots ignored */
  int32_t pad_;
  double weight;
} pgld_term;

int pgld_version(void);
const char* pgld_last_error(void);

/* Value scan over selected rows: flags[0] = 1 if any value < 0, flags[1] = 1
 * if any value lies outside [-1, max_value].  rows == NULL selects all rows.
 * Replaces the has_missing syncs of ld_pipeline.py:158 and :253. */
int pgld_scan_rows(const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows,
                   int64_t n_sel, int32_t max_value, int32_t* flags, void* stream);

/* Per-variant allele counts over the selected rows: alt[v] = sum of the
 * values > 0 (haplotypes: number of 1s; genotypes: dosage sum) and valid[v]
 * = number of values >= 0 (int32, n_var each).  Feeds the biallelic filters
 * (haplotype_matrix.py:724-789, genotype_matrix.py:506-539, moments_ld.py:
 * 256-285) and the heterozygosity statistics (moments_ld.py:346-388). */
int pgld_allele_counts(const int8_t* data, int64_t n_rows, int64_t n_var, const int32_t* rows, int64_t n_sel,
                       int32_t* alt, int32_t* valid, void* stream);

/* Pair enumeration of iter_pairs_within_distance (ld_pipeline.py:57-122).
 * per[v] = #{j > v : pos[j] - pos[v] <= max_dist}; positions sorted. */
int pgld_pairs_per_variant(const double* pos, int64_t n, double max_dist, int64_t* per, void* stream);
/* Expand the pairs of anchors v_lo .. v_lo+n_anchor-1.  excl[k] is the
 * exclusive prefix sum of per[] over that anchor range (length n_anchor),
 * n_pairs the total.  Pairs are ordered by (i, j). */
int pgld_pairs_fill(const int64_t* excl, int64_t v_lo, int64_t n_anchor, int64_t n_pairs,
                    int32_t* idx_i, int32_t* idx_j, void* stream);

/* Plain per-pair 4-way haplotype counts [n11, n10, n01, n00] (int32, n x 4)
 * and n_valid (int32, n; may be NULL) over the selected rows.  One thread per
 * pair, sample columns staged through LDS.  Replaces compute_counts_for_pairs
 * (ld_pipeline.py:130-175). */
int pgld_counts_pairs_hap(const int8_t* hap, int64_t n_rows, int64_t n_var, const int32_t* rows,
                          int64_t n_sel, const int32_t* idx_i, const int32_t* idx_j, int64_t n_pairs,
                          int32_t* counts, int32_t* n_valid, void* stream);

/* Plain per-pair 9-way genotype counts n00, n01, ..., n22 (int32, n x 9) and
 * n_valid.  Replaces _GENO_COUNTS_KERN / compute_genotype_counts_for_pairs
 * (ld_pipeline.py:183-263). */
int pgld_counts_pairs_geno(const int8_t* geno, int64_t n_rows, int64_t n_var, const int32_t* rows,
                           int64_t n_sel, const int32_t* idx_i, const int32_t* idx_j, int64_t n_pairs,
                           int32_t* counts, int32_t* n_valid, void* stream);

/* Per-pair statistics (float64, n_pairs x n_stats) from per-population
 * counts.  counts is a HOST array of n_pops device pointers, each (n_pairs x
 * 4) for haplotypes or (n_pairs x 9) for genotypes.  n_valid is NULL or a
 * HOST array of n_pops device pointers (entries may be NULL): for haplotypes a
 * non-NULL entry replaces that population's sample size n (the reference's
 * n_valid argument, ld_statistics.py:1082-1113); genotype statistics always
 * use n = sum of the 9 counts, as the single-population genotype kernels do
 * (genotype_kernels.py:35).  terms / stat_offsets
 * (device) list, for statistic s, terms[stat_offsets[s] .. stat_offsets[s+1]).
 * Replaces ld_statistics.dd/dz/pi2 (ld_statistics.py:13-129),
 * compute_multi_pop_statistics_batch_hap (haplotype_kernels.py:584-640),
 * compute_multi_pop_statistics_batch_geno (genotype_kernels.py:881-936) and
 * the sigma_D^2 kernel launch of sigma_d2_geno (ld_statistics_genotype.py:
 * 1593-1597).  n_pops <= 4. */
int pgld_stats_pairs(int32_t family, int32_t n_pops, const int32_t* const* counts,
                     const int32_t* const* n_valid, int64_t n_pairs, const pgld_term* terms,
                     const int32_t* stat_offsets, int32_t n_stats, double* out, void* stream);

/* r (which=0), r^2 (1) or D' (2) per pair from haplotype counts; n_valid may
 * be NULL (n = sum of counts).  Replaces ld_statistics.r/r_squared/d_prime
 * (ld_statistics.py:167-246). */
int pgld_r_family(const int32_t* counts, const int32_t* n_valid, int64_t n_pairs, int32_t which,
                  double* out, void* stream);

/* Key of every pair: PGLD_KEY_BIN_ABS / PGLD_KEY_BIN use edges (device,
 * n_keys + 1, ascending) with [e_k, e_{k+1}) semantics of digitize - 1
 * (moments_ld.py:310-314); PGLD_KEY_WINDOW uses origin / width.  Keys
 * outside [0, n_keys) are written as -1. */
int pgld_pair_keys(const double* coord, const int32_t* idx_i, const int32_t* idx_j, int64_t n_pairs,
                   int32_t mode, const double* edges, int32_t n_keys, double origin, double width,
                   int32_t* keys, void* stream);

/* sums[key, c] += sum of values[p, c] over pairs with keys[p] == key (float64
 * atomics; keys < 0 skipped).  Replaces the cp.add.at scatter of
 * moments_ld.py:335-339 and haplotype_matrix.py:1683-1689. */
int pgld_keyed_sums(const double* values, const int32_t* keys, int64_t n_pairs, int32_t n_cols,
                    int32_t n_keys, double* sums, void* stream);

/* ---- tiled int8-MFMA path ---------------------------------------------- */

/* Indicator planes for the MFMA path: SNP-major, sample-contiguous, K padded
 * with zeros to k_pad (a multiple of 128).  Haplotypes: 2 planes [h, v] with
 * h = (x == 1), v = (x >= 0); genotypes: 3 planes [v, e1, e2] with
 * e1 = (g == 1), e2 = (g == 2).  planes has n_planes * n_var * k_pad bytes;
 * marg (int32, n_planes * n_var) receives each plane's per-SNP sum; flags as
 * in pgld_scan_rows.  rows selects one population. */
int pgld_pack_planes(int32_t family, const int8_t* data, int64_t n_rows, int64_t n_var,
                     const int32_t* rows, int64_t n_sel, int64_t k_pad, int8_t* planes, int32_t* marg,
                     int32_t* flags, void* stream);
/* The same planes in FP4 (no counterpart in the reference: an operand layout
 * for the gfx950 FP4 MFMA): each indicator as an e2m1 code (0x0 or
 * 0x2 = 1.0), two samples per byte, the even sample in the low nibble; the
 * haplotype allele plane is centred instead: 0x2 (+1) where x == 1, 0xA (-1)
 * where x == 0, 0x0 where missing, so that its products are the centred
 * moments the statistics use.  k_pad must be a multiple of 256 and planes
 * has n_planes * n_var * k_pad / 2 bytes.  marg (the 0/1 indicator sums) and
 * flags as above.  Consumed by pgld_band_sums_fp4, pgld_batch_bind_fp4 and
 * pgld_band_ld_pairs_fp4. */
int pgld_pack_planes_fp4(int32_t family, const int8_t* data, int64_t n_rows, int64_t n_var,
                         const int32_t* rows, int64_t n_sel, int64_t k_pad, int8_t* planes, int32_t* marg,
                         int32_t* flags, void* stream);

/* Many pgld_pack_planes(_fp4) calls as one launch (no counterpart in the
 * reference: the populations of a batch of regions are packed once per step,
 * and thousands of small launches would cost more than the packing).  Job j
 * packs data[j] (n_rows[j] x n_var[j]), rows[j] (n_sel[j] entries, or NULL =
 * all rows) into planes[j] / marg[j] / flags[j] with k_pad[j], exactly as the
 * single call would (same bytes, margins and flags); all jobs share family
 * and layout (fp4 != 0: FP4 planes).  Arguments are HOST arrays of n_jobs
 * entries (device pointers for the buffers); create validates them and
 * uploads the job table once; run (capturable into a HIP graph: two kernel
 * launches, no allocation, no memset) may be repeated while the buffers stay
 * valid. */
typedef struct pgld_pack_batch pgld_pack_batch;
int pgld_pack_batch_create(int32_t family, int32_t fp4, int32_t n_jobs, const int8_t* const* data,
                           const int64_t* n_rows, const int64_t* n_var, const int32_t* const* rows,
                           const int64_t* n_sel, const int64_t* k_pad, int8_t* const* planes,
                           int32_t* const* marg, int32_t* const* flags, pgld_pack_batch** out);
int pgld_pack_batch_run(const pgld_pack_batch* pb, void* stream);
int pgld_pack_batch_destroy(pgld_pack_batch* pb);

typedef struct pgld_band_plan pgld_band_plan;

/* Build the tile schedule of the distance band {(i, j): i < j,
 * pos[j] - pos[i] <= max_dist} over SNP tiles of 64, and the key slots each
 * tile can touch.  pos, key_coord and edges are HOST arrays: band membership
 * uses pos; keys use key_coord (NULL = pos; must be nondecreasing, e.g. a
 * genetic map interpolated at pos, moments_ld.py:167-174); edges holds
 * n_keys + 1 values for the bin modes (NULL for windows).  The plan keeps
 * device copies and its own workspace. */
int pgld_band_plan_create(const double* pos, int64_t n_var, double max_dist, int32_t key_mode,
                          const double* key_coord, const double* edges, int32_t n_keys, double origin,
                          double width, pgld_band_plan** plan);
int pgld_band_plan_info(const pgld_band_plan* plan, int64_t* n_tiles, int64_t* n_pairs, int64_t* n_slots);
int pgld_band_plan_destroy(pgld_band_plan* plan);
/* Kernel timing for measurement (no counterpart in the reference, whose
 * kernels are timed from the host): with timing enabled, pgld_band_sums
 * records a pair of HIP events on its stream around the fused kernel, and
 * pgld_band_plan_kernel_ms waits for the last pair and returns the elapsed
 * milliseconds of that launch. */
int pgld_band_plan_set_timing(pgld_band_plan* plan, int32_t enable);
int pgld_band_plan_kernel_ms(const pgld_band_plan* plan, float* ms);

#define PGLD_SET_MOMENTS 0 /* ld_names(n_pops): 3 / 15 / 45 / 105 statistics for 1 / 2 / 3 / 4 pops */
#define PGLD_SET_SELF 1    /* DD_i_i, Dz_i_i_i, pi2_i_i_i_i per population (3 * n_pops) */

/* Fused band pass: int8 MFMA counts for every pair of the band, every
 * population, then the statistics of stat_set (PGLD_SET_MOMENTS: ld_names(n_pops), spec-averaged
 * as generate_stat_specs, ld_pipeline.py:388-457; haplotypes with two
 * populations use the fast-path forms of ld_pipeline.py:271-322), then
 * float64 sums per key, reduced in a fixed order (bit-reproducible run to
 * run).  planes / marg / k_pad / n_samples are HOST arrays of n_pops entries
 * (device pointers for planes and marg); n_samples <= 32767 and k_pad <= 32768
 * per population (as packed).  sums (device, n_keys x n_stats) is
 * overwritten; key_pairs (device, n_keys, may be NULL) receives the number of
 * pairs per key.  n_pops <= 4 (the 105 four-population moments are summed
 * in three column groups; the 3-4 population genotype moments park their
 * counts in a plan-owned global workspace, so they run on single bands only,
 * not on region batches or caller-supplied workspaces).  Replaces the
 * chunk loop of moments_ld._compute_ld_sums (moments_ld.py:248-343) and
 * HaplotypeMatrix.compute_ld_statistics_gpu_* (haplotype_matrix.py:
 * 1592-1866). */
int pgld_band_sums(const pgld_band_plan* plan, int32_t family, int32_t n_pops,
                   const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                   const int32_t* n_samples, int32_t with_missing, int32_t stat_set, double* sums,
                   double* key_pairs, void* stream);
/* pgld_band_sums over FP4 planes (pgld_pack_planes_fp4; no counterpart in the
 * reference): the plane products run on v_mfma_f32_32x32x64_f8f6f4, 64
 * samples per instruction, with float accumulators that hold the counts
 * exactly; the sums equal pgld_band_sums' bit for bit.  Single bands only;
 * k_pad of every population a multiple of 256. */
int pgld_band_sums_fp4(const pgld_band_plan* plan, int32_t family, int32_t n_pops,
                   const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                   const int32_t* n_samples, int32_t with_missing, int32_t stat_set, double* sums,
                   double* key_pairs, void* stream);

/* pgld_band_sums_fp4 without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial-row workspace is a fixed 32 MiB device array (2^22 doubles), one
 * copy per GPU held until the process ends, instead of the plan's own buffer;
 * plans needing more ((n_slots + reduction chunks) x (n_stats + 1) doubles)
 * return PGLD_ECAPACITY, and the 3-4 population genotype moments (which
 * park counts in the plan's own workspace) PGLD_EUNSUPPORTED.  Every call on
 * one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race.  Nothing in the
 * array needs clearing between calls (every used element is written before
 * it is read).  Same arguments and results as pgld_band_sums_fp4. */
int pgld_band_sums_fp4_noalloc(const pgld_band_plan* plan, int32_t family, int32_t n_pops,
                               const int8_t* const* planes, const int32_t* const* marg,
                               const int64_t* k_pad, const int32_t* n_samples, int32_t with_missing,
                               int32_t stat_set, double* sums, double* key_pairs, void* stream);

/* Counts of every band pair through the MFMA path, written per pair in the
 * plan's pair order (i ascending, then j): counts (int32, n_pairs x 4 or x 9)
 * for ONE population.  Used to check the tiled counts bit-exactly against
 * the plain kernel. */
int pgld_band_counts(const pgld_band_plan* plan, int32_t family, const int8_t* planes,
                     const int32_t* marg, int64_t k_pad, int32_t n_samples, int32_t with_missing,
                     const int64_t* pair_base, int32_t* counts, void* stream);

/* r^2 of every pair of the band from one haplotype population's planes,
 * counts through the int8-MFMA tiles: r^2 = diff^2 / (nA (n - nA) nB (n - nB))
 * with diff = n n11 - nA nB computed exactly (int64) and one FP64 division; NaN
 * where the denominator is 0 (ld_statistics.r_squared, ld_statistics.py:
 * 194-217; pairs with missing data use their pair-complete counts).  Outputs,
 * each optional (NULL to skip):
 *   hist      uint64 (n_keys, n_hist + 1): per plan key, pairs with r^2 in
 *             [b / n_hist, (b + 1) / n_hist) (r^2 = 1 in the last bin); the
 *             last column counts NaN pairs (the histogram analogue of
 *             windowed_r_squared, haplotype_matrix.py:1233-1317);
 *   sparse_i / sparse_j / sparse_r2 (FP32), capacity sparse_cap, with
 *             *n_sparse (device uint64) = number of pairs with r^2 > thresh
 *             (may exceed the capacity: only the first sparse_cap are
 *             written; order unspecified);
 *   r2_out    FP64 per band pair in the plan's (i, j) order (the condensed
 *             upper triangle for an all-pairs plan, cf. pairwise_r2,
 *             haplotype_matrix.py:1092-1170).
 * Replaces tally_gpu_haplotypes + r_squared + the host-side binning of
 * windowed_r_squared for dense all-pairs r^2. */
int pgld_band_r2(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                 int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                 int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap, uint64_t* n_sparse,
                 double* r2_out, void* stream);

/* Pair estimators of pgld_band_ld_pairs */
#define PGLD_EST_R2 0       /* r^2 from the pair's counts (ld_statistics.r_squared); NaN where undefined */
#define PGLD_EST_R2_NAIVE 1 /* frequency r^2, per-site p over all valid samples (pairwise_r2,
                               _tile_r2_naive, ld_statistics.py:301-308); 0 where undefined */
#define PGLD_EST_SIGMA_D2 2 /* unbiased D^2 / pi_2 (_tile_sigma_d2, ld_statistics.py:311-341);
                               0 unless n >= 4 and pi_2 != 0 */
#define PGLD_EST_RH_R 3     /* genotype planes, no missing data: signed Rogers-Huff r of 0/1/2 dosages
                               (_tile_rogers_huff_r, ld_statistics.py:410-443); NaN for a constant site */
#define PGLD_EST_DIP_R2 4   /* genotype planes: dosage r^2 about site means over valid samples
                               (_r2_matrix_diploid, ld_statistics.py:975-1021); 0 where undefined */
#define PGLD_EST_D_NAIVE 5  /* haplotype planes: signed D = p_AB - p_i p_j with the frequencies of
                               PGLD_EST_R2_NAIVE (pairwise_LD_v, haplotype_matrix.py:1092-1136) */

/* pgld_band_r2 with a choice of estimator and two more optional outputs
 * (planes / marg are haplotype planes for estimators 0-2 and 5, genotype
 * planes for 3-4):
 * values     FP64 per band pair in (i, j) order (as r2_out);
 * snp_sums   FP64 (2 * n_var): [v] = sum over pairs (v, j > v), [n_var + v] =
 *            sum over pairs (i < v, v) of the estimator (NaN pairs skipped);
 * snp_counts int64 (2 * n_var): the number of pairs counted in each sum (for
 *            PGLD_EST_SIGMA_D2 the pairs with n >= 4).
 * The per-SNP sums are what ZnS and Omega need (ld_statistics.py:534-862)
 * without an m x m matrix.  Histogram bins clamp values outside [0, 1]. */
int pgld_band_ld_pairs(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                       int32_t n_samples, int32_t with_missing, int32_t estimator, int32_t n_hist, double thresh,
                       uint64_t* hist, int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap,
                       uint64_t* n_sparse, double* values, double* snp_sums, int64_t* snp_counts, void* stream);
/* pgld_band_ld_pairs over FP4 planes (pgld_pack_planes_fp4, k_pad a multiple
 * of 256; no counterpart in the reference): every estimator, with or without
 * missing data and per-SNP sums, with its products on
 * v_mfma_f32_32x32x64_f8f6f4 (64 samples per instruction, exact float
 * counts); outputs equal pgld_band_ld_pairs' bit for bit. */
int pgld_band_ld_pairs_fp4(const pgld_band_plan* plan, const int8_t* planes, const int32_t* marg, int64_t k_pad,
                       int32_t n_samples, int32_t with_missing, int32_t estimator, int32_t n_hist, double thresh,
                       uint64_t* hist, int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap,
                       uint64_t* n_sparse, double* values, double* snp_sums, int64_t* snp_counts, void* stream);

/* ---- bit-packed popcount variant of the tile products ------------------ */

/* Bit planes from int8 0/1 planes (n_planes x n_var x k_pad, as written by
 * pgld_pack_planes): row r of bits holds 32 * ceil(k_pad / 1024) uint32 words,
 * word w = samples 32w .. 32w + 31 (bit b = sample 32w + b), zero padded. */
int pgld_pack_bits(const int8_t* planes, int64_t n_planes, int64_t n_var, int64_t k_pad, uint32_t* bits,
                   void* stream);
/* pgld_band_counts / pgld_band_r2 with the products formed by AND + popcount
 * on the VALU from bit planes instead of int8 MFMA (same outputs, bit for
 * bit); k_pad is the int8 planes' padding the bit rows derive from.  Kept for
 * the MFMA-versus-popcount comparison of the tile counts. */
int pgld_band_counts_bits(const pgld_band_plan* plan, int32_t family, const uint32_t* bits, const int32_t* marg,
                          int64_t k_pad, int32_t n_samples, int32_t with_missing, const int64_t* pair_base,
                          int32_t* counts, void* stream);
int pgld_band_r2_bits(const pgld_band_plan* plan, const uint32_t* bits, const int32_t* marg, int64_t k_pad,
                      int32_t n_samples, int32_t with_missing, int32_t n_hist, double thresh, uint64_t* hist,
                      int32_t* sparse_i, int32_t* sparse_j, float* sparse_r2, int64_t sparse_cap,
                      uint64_t* n_sparse, double* r2_out, void* stream);

/* ---- ragged region batches: many independent regions, one launch ------ */

/* A band plan over n_regions independent regions (no pair crosses regions).
 * pos (host) holds the regions' positions one after the other (n_var[r]
 * values for region r, each sorted); key_coord likewise or NULL; edges /
 * n_keys / origin / width describe each region's keys as in
 * pgld_band_plan_create, and the plan's keys are region * n_keys + key.
 * The reference loops over regions from Python (one moments_ld call each,
 * moments_ld.py:40-203); this runs them as one persistent launch. */
int pgld_batch_plan_create(int32_t n_regions, const int64_t* n_var, const double* pos, double max_dist,
                           int32_t key_mode, const double* key_coord, const double* edges, int32_t n_keys,
                           double origin, double width, pgld_band_plan** plan);
/* Bind the packed planes of every region's populations (host arrays of
 * n_regions * n_pops entries, region-major: entry r * n_pops + p), uploads the
 * region table, orders the tiles for the persistent launch and sizes the
 * workspace for the statistic set.  Synchronous; call before any capture. */
int pgld_batch_bind(pgld_band_plan* plan, int32_t family, int32_t n_pops, int32_t stat_set,
                    const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                    const int32_t* n_samples);
/* pgld_batch_bind for FP4 planes (pgld_pack_planes_fp4, every k_pad a
 * multiple of 256; no counterpart in the reference): the bound batch runs its
 * products on the FP4 MFMA, sums bit-identical to the int8 binding. */
int pgld_batch_bind_fp4(pgld_band_plan* plan, int32_t family, int32_t n_pops, int32_t stat_set,
                    const int8_t* const* planes, const int32_t* const* marg, const int64_t* k_pad,
                    const int32_t* n_samples);
/* Fused counts -> statistics -> per-key sums for every region: sums is
 * (n_regions * n_keys, n_stats) FP64, key_pairs (optional) the pair count
 * per key.  Allocates nothing, records no memset: may be captured. */
int pgld_batch_sums(pgld_band_plan* plan, int32_t with_missing, double* sums, double* key_pairs, void* stream);
/* Schedule of the persistent launch: grid workgroups (0 = one per tile),
 * order 0 = cost-descending tiles, 1 = natural, 2 = L2-blocked (8 x 8 tile
 * blocks of one region dealt to the 8 XCDs by work and interleaved so that
 * position p runs on XCD p % 8; use grid 0 or a multiple of 8); xcd_split =
 * each XCD walks a contiguous eighth of the order. */
int pgld_batch_schedule(pgld_band_plan* plan, int32_t grid, int32_t order, int32_t xcd_split,
                        int32_t kernel_variant);
/* kernel_variant (bit-identical sums; used by the FP4 within-population
 * haplotype sets, ignored by the others): 0 = K loop streamed across the
 * populations of a tile, fast-path rows unrolled; 1 = one K loop per
 * population; 2 = streamed, rows in a rolled loop; 3 = lean: one K loop per
 * population with one load stage, registers and LDS capped for three
 * workgroups per CU. */
/* Autotuner: time the schedule and kernel-variant candidates on this plan,
 * keep the fastest.  The choice is cached per shape class (GPU architecture,
 * CUs, family, planes, populations, set, missing data, log2 of tiles, mean
 * region size, pairs per SNP and mean population size) in the process and in
 * a text file -- $PGLD_TUNE_CACHE, else $XDG_CACHE_HOME/pgld/batch_tune.txt or
 * ~/.cache/pgld/batch_tune.txt; PGLD_TUNE_CACHE="" keeps it in the process --
 * so later processes skip the timing.  Outputs may be NULL; *best_ms = -1
 * when the cache answered. */
int pgld_batch_autotune(pgld_band_plan* plan, int32_t with_missing, double* sums, double* key_pairs, void* stream,
                        int32_t* grid, int32_t* order, int32_t* xcd_split, int32_t* kernel_variant,
                        float* best_ms);

/* HIP graph capture of a sequence of libpgld calls on one created stream:
 * pgld_graph_begin(stream); ... calls on that stream ...; pgld_graph_end. The
 * captured calls must not allocate or synchronize (pack with its buffers,
 * pgld_batch_sums).  pgld_graph_launch replays the whole sequence as one
 * submission.  (No counterpart in the reference.) */
typedef struct pgld_graph pgld_graph;
int pgld_graph_begin(void* stream);
int pgld_graph_end(void* stream, pgld_graph** graph);
int pgld_graph_launch(pgld_graph* graph, void* stream);
int pgld_graph_destroy(pgld_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* PGLD_H */
