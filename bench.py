"""Benchmark of the LD pair-statistics path on one MI355X.

Default workload (the headline, the largest single-GPU configuration):
ragged_regions -- 512 independent regions of 500-20,000 SNPs x 50-5,000
haplotypes in 4 populations, 3% MCAR missing, positions uniform over 1 Mb,
all pairs of each region, per-region 16 log-bin DD/Dz/pi2 sums of every
population (FP64).  One step = a HIP graph that repacks every region's
population planes from the resident int8 matrices (one batched launch) and
runs one autotuned persistent fused launch + the per-key reduction.  The line also carries the
error of every statistic against the FP64 oracle on the regions of least
work (all their pairs).

Other workloads (--workload): windowed_hap (100,000 SNPs x 2,000 haplotypes
in two populations, 2% missing, pairs within 100 kb, the 15 two-population
moments statistics per 100 kb window), genotype_bins (50,000 SNPs x 1,000
diploids, 5% missing, all pairs in 1 Mb, 16 log distance bins), plain_hap
(2,000 SNPs x 200 haplotypes, per-pair r^2/DD/Dz/pi2 within 20 kb through the
plain per-pair kernels), dense_r2 (40,000 SNPs x 10,000 haplotypes, r^2 of
all 7.9998e8 pairs, 100-bin histogram + sparse list of pairs with
r^2 > 0.2).

Prints one JSON line: value is Gpair*samples/s for the whole job; roofline is
for the dominant kernel; cpu_baseline is the NumPy oracle on a bounded sample
of the same pairs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT8_OPS = 5.0e15   # MI355X dense int8 MFMA, ops/s
PEAK_FP4_OPS = 10.1e15   # MI355X dense FP4 MFMA, ops/s
PEAK_HBM = 8.0e12        # bytes/s
# HBM bytes per k_band_sums launch on the default workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same counters'
# reading of a 1 GiB device copy (--copy-calibration: FETCH x 2.000, WRITE x
# 1.000): FP4 planes 158,541 + 3,908 KiB, int8 planes 769,840 + 3,908 KiB.
# Summary and commands: profiles/k_band_sums_hap2pop_windows_100k_snps_2000_haps.md
TRAFFIC_BYTES_PER_LAUNCH = {"fp4": 0.3287e9, "int8": 1.581e9}
# Issue counters of the same launch (SQ_INSTS_VALU, which counts the MFMAs
# too, and SQ_INSTS_MFMA per dispatch): the VALU-issue bound of the kernel.
WINDOWED_SQ = {"fp4": {"valu_insts": 3.01273e8, "mfma_insts": 1.44901e7}}
# The same for k_batch_sums on the 512-region ragged batch (FP4 planes, the
# autotuner's lean variant): FETCH_SIZE 2.3243e8 KiB x 2.000 + WRITE_SIZE
# 1.23857e6 KiB x 1.000 per dispatch, and its issue counters.
# Summary and commands: profiles/k_batch_sums_ragged_512_regions.md
RAGGED_TRAFFIC_BYTES = {"fp4": 477.3e9}
RAGGED_SQ = {"fp4": {"valu_insts": 9.83314e10, "mfma_insts": 6.01159e9}}
CLOCK_HZ = 2.4e9         # MI355X peak engine clock
N_SIMD = 256 * 4         # 256 CUs x 4 SIMDs
VALU_ISSUE_CYCLES = 4    # a wave64 VALU instruction (FP64 included) holds its SIMD's issue for 4 cycles
MFMA_FP4_CYCLES = 32     # v_mfma_f32_32x32x64_f8f6f4 with e2m1 operands


def valu_issue_bound(sq, kernel_ms):
    """Issue time of the kernel's non-MFMA VALU instructions spread over every
    SIMD at the peak clock, against the measured kernel time (the MFMA pipe's
    own busy time beside it)."""
    valu = sq["valu_insts"] - sq["mfma_insts"]
    issue_ms = valu * VALU_ISSUE_CYCLES / N_SIMD / CLOCK_HZ * 1e3
    mfma_ms = sq["mfma_insts"] * MFMA_FP4_CYCLES / N_SIMD / CLOCK_HZ * 1e3
    return {"bound": "valu_issue", "valu_insts": valu, "issue_ms": issue_ms, "frac": issue_ms / kernel_ms,
            "mfma_insts": sq["mfma_insts"], "mfma_busy_ms": mfma_ms, "mfma_busy_frac": mfma_ms / kernel_ms,
            "source": "SQ_INSTS_VALU and SQ_INSTS_MFMA per dispatch (rocprofv3 --pmc), kernel_ms live"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ragged_regions",
                    choices=["windowed_hap", "genotype_bins", "plain_hap", "dense_r2", "ragged_regions", "banded_r2"])
    ap.add_argument("--n-var", type=int, default=None)
    ap.add_argument("--n-samples", "--n-hap", dest="n_samples", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-pairs", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-oracle", action="store_true", help="skip the FP64-oracle error check of the ragged batch")
    ap.add_argument("--oracle-regions", type=int, default=4)
    ap.add_argument("--planes", default="fp4", choices=["int8", "fp4"],
                    help="operand planes of the fused sums (windowed and genotype workloads): int8 MFMA, or FP4 "
                         "(e2m1 0 / 1) on the FP4 MFMA")
    ap.add_argument("--method", default="mfma", choices=["mfma", "popcount"],
                    help="tile products of the r^2 workloads: MFMA (int8 or FP4 planes) or AND + popcount on bit "
                         "planes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 device copies of 1 GiB (the HBM-counter reference of the traffic figure)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(torch, step, steps, warmup):
    """Average ms of `steps` calls after `warmup`, HIP events on the current
    stream, synchronized on both sides."""
    out = None
    for _ in range(warmup):
        out = step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        out = step()
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps, out


def band_kernel_ms(plan, packed, with_missing, reps):
    """k_band_sums alone: HIP events recorded by the library around each launch
    on the stream it is launched on, averaged over launches."""
    plan.set_timing(True)
    kms = []
    for _ in range(reps):
        plan.sums(packed, with_missing=with_missing)
        kms.append(plan.kernel_ms())
    plan.set_timing(False)
    return float(np.mean(kms)), float(np.min(kms))


def band_pair_sample(pos, max_dist, want, seed=7):
    """`want` pairs drawn uniformly (with replacement; every pair once when
    want covers the band) from {(i, j): i < j, pos[j] - pos[i] <= max_dist},
    without listing the band: the
    left SNP with probability proportional to its pair count, then the right
    SNP uniformly among its partners.  Sorted by (i, j)."""
    rng = np.random.default_rng(seed)
    pos = np.asarray(pos)
    per = np.searchsorted(pos, pos + max_dist, side="right") - np.arange(len(pos)) - 1
    per = np.maximum(per, 0)
    cum = np.cumsum(per)
    total = int(cum[-1])
    r = np.arange(total) if want >= total else rng.integers(0, total, size=want)
    ii = np.searchsorted(cum, r, side="right")
    jj = ii + 1 + (r - (cum[ii] - per[ii]))
    order = np.lexsort((jj, ii))
    return ii[order], jj[order]


def base_line(args, value, ms_step, workload, config, dtype="int8"):
    return {
        "metric": "pair-sample counts/sec (Gpair*samples/s)",
        "value": value,
        "unit": "Gpair*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": dict({"workload": workload}, **config),
    }


def mfma_roofline(ops, kernel_ms, kernel, macs, traffic=None, fp4=False):
    """MFMA roofline of a kernel: algorithmic ops over its time against the
    dense peak of the MFMA it runs on (int8, or FP4 for FP4 planes; the
    int8-peak fraction is kept beside the FP4 one)."""
    achieved = ops / (kernel_ms * 1e-3)
    peak = PEAK_FP4_OPS if fp4 else PEAK_INT8_OPS
    out = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
           "frac": achieved / peak, "traffic": traffic, "kernel": kernel, "kernel_ms": kernel_ms,
           "ops": f"{'FP4' if fp4 else 'int8'} MFMA ops (1 MAC = 2 ops) of {macs} MACs per pair*sample"}
    if fp4:
        out["frac_of_int8_peak"] = achieved / PEAK_INT8_OPS
    return out


def _plain_compare(torch, plain_step, fused_out, steps):
    ref = plain_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        ref = plain_step()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t1) * 1e3 / steps
    a, b = fused_out.cpu().numpy(), ref.cpu().numpy()
    check = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
    log(f"[bench] plain path {plain_ms:.1f} ms, max rel diff fused vs plain {check:.2e}")
    return plain_ms, check


# ---------------------------------------------------------------------------
def run_windowed_hap(args, torch):
    from pg_gpu_amd import band, synthetic, ld_pipeline as LP
    from pg_gpu_amd import _lib as L

    n_var = args.n_var or 100_000
    n_hap = args.n_samples or 2000
    macs = 4  # haplotypes with missing data: hh, hv, vh, vv
    log("[bench] generating workload")
    hap, pos, pops = synthetic.windowed_haplotype_panel(n_hap=n_hap, n_var=n_var)
    max_dist, window = 100_000.0, 100_000.0
    n_keys = 100
    dev = L.device()
    hap_d = torch.as_tensor(hap, device=dev)
    rows_d = [torch.as_tensor(np.asarray(r, dtype=np.int32), device=dev) for r in pops]
    plan = band.BandPlan(pos, max_dist, window=window, n_keys=n_keys, origin=0.0)
    S = [len(r) for r in pops]
    pair_samples = plan.n_pairs * sum(S)
    log(f"[bench] pairs={plan.n_pairs} tiles={plan.n_tiles} pair_samples={pair_samples:.4e}")

    # the input is validated once (domain, missing data); a timed step then
    # repacks both populations into the same buffers and runs the fused pass
    fp4 = args.planes == "fp4"
    packed = [band.pack_population(hap_d, r, "hap", fp4=fp4) for r in rows_d]
    with_missing = any(p.has_missing for p in packed)

    def step():
        for k, r in enumerate(rows_d):
            band.pack_population(hap_d, r, "hap", validate=False, out=packed[k], fp4=fp4)
        return plan.sums(packed, with_missing=with_missing)

    ms_step, out = timed_steps(torch, step, args.steps, args.warmup)
    value = pair_samples / (ms_step * 1e-3) / 1e9
    log(f"[bench] fused step {ms_step:.3f} ms")
    band_ms, band_min = band_kernel_ms(plan, packed, with_missing, max(3, args.steps))
    default_shape = (n_var, n_hap) == (100_000, 2000)
    roof = mfma_roofline(2.0 * macs * pair_samples, band_ms, "k_band_sums", macs,
                         TRAFFIC_BYTES_PER_LAUNCH[args.planes] if default_shape else None, fp4)
    valu = (valu_issue_bound(WINDOWED_SQ[args.planes], band_ms)
            if default_shape and args.planes in WINDOWED_SQ else None)
    log(f"[bench] k_band_sums {band_ms:.3f} ms (min {band_min:.3f}) -> {roof['achieved']:.1f} int8 TOP/s")

    # plain per-pair path on identical inputs and outputs (speedup vs plain)
    plain_ms = check = None
    if not args.no_plain:
        pos_d = torch.as_tensor(pos.astype(np.float64), device=dev)

        def plain_step():
            sums = torch.zeros((n_keys, 15), dtype=torch.float64, device=dev)
            chunk = LP.estimate_ld_chunk_size(max(S), num_pops=2)
            for ii, jj in LP.iter_pairs_within_distance(pos, max_dist, chunk):
                cl = [LP.compute_counts_for_pairs(hap_d, ii, jj, r) for r in rows_d]
                st = LP.compute_two_pop_statistics_batch(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
                keys = torch.empty(ii.shape[0], dtype=torch.int32, device=dev)
                L.call("pgld_pair_keys", L.ptr(pos_d), L.ptr(ii), L.ptr(jj), ii.shape[0], L.KEY_WINDOW, None,
                       n_keys, 0.0, window, L.ptr(keys), L.stream_handle())
                L.call("pgld_keyed_sums", L.ptr(st), L.ptr(keys), ii.shape[0], 15, n_keys, L.ptr(sums),
                       L.stream_handle())
            return sums
        plain_ms, check = _plain_compare(torch, plain_step, out, args.plain_steps)

    cpu = None
    if not args.no_cpu:
        from oracle import ld_oracle as O
        ii, jj = band_pair_sample(pos, max_dist, args.cpu_pairs or 1_500_000)
        t2 = time.perf_counter()
        acc = np.zeros((n_keys, 15))
        for c0 in range(0, len(ii), 250_000):
            a, b = ii[c0:c0 + 250_000], jj[c0:c0 + 250_000]
            cl = [O.hap_counts(hap, a, b, r) for r in pops]
            st = O.two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
            acc += O.keyed_sums(O.window_key(pos[a], 0.0, window), st, n_keys)
        cpu_s = time.perf_counter() - t2
        cpu = {"value": len(ii) * sum(S) / cpu_s / 1e9, "unit": "Gpair*samples/s", "cores": 1, "kind": "port",
               "sample": f"{len(ii)} random pairs of the same band (of {plan.n_pairs}), both populations, "
                         f"counts + 15 statistics + window sums, NumPy oracle (single thread), {cpu_s:.1f} s"}
        log(f"[bench] cpu oracle {cpu['value']:.4f} Gpair*samples/s over {cpu_s:.1f} s")
        # the BLAS-tiled exact CPU variant (float32 indicator GEMMs on every
        # host thread BLAS uses, FP64 statistics) on the first left SNPs
        i_hi = min(n_var, 20000)
        per = np.searchsorted(pos, pos[:i_hi] + max_dist, side="right") - np.arange(i_hi) - 1
        n_tiled = int(np.maximum(per, 0).sum())
        t2 = time.perf_counter()
        O.tiled_window_sums_hap(hap, pos, pops, max_dist, window, n_keys, 0, i_hi)
        tiled_s = time.perf_counter() - t2
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu_tiled = {"value": n_tiled * sum(S) / tiled_s / 1e9, "unit": "Gpair*samples/s", "cores": threads,
                     "kind": "port",
                     "sample": f"the {n_tiled} band pairs of the first {i_hi} left SNPs, float32 indicator GEMMs "
                               f"(BLAS, {threads} threads) + 15 FP64 statistics (NumPy) + window sums, "
                               f"{tiled_s:.1f} s"}
        log(f"[bench] cpu tiled {cpu_tiled['value']:.4f} Gpair*samples/s over {tiled_s:.1f} s")

    line = base_line(args, value, ms_step,
                     f"batched windowed haplotype LD: {n_var:,} SNPs x {n_hap:,} haplotypes (2 pops x "
                     f"{n_hap // 2:,}), 2% MCAR missing, gamma(0.5) heterogeneous positions over 10 Mb, pairs "
                     "within 100 kb, 15 two-pop stats summed per 100 kb window (FP64)",
                     {"n_var": n_var, "n_hap": n_hap, "pairs": plan.n_pairs, "pair_samples": pair_samples,
                      "planes": args.planes}, dtype="fp4" if fp4 else "int8")
    line.update({"roofline": roof, "valu_roofline": valu, "plain_ms_per_step": plain_ms,
                 "speedup_vs_plain": (plain_ms / ms_step) if plain_ms else None,
                 "fused_vs_plain_max_rel": check, "cpu_baseline": cpu,
                 "cpu_baseline_tiled": cpu_tiled if cpu else None})
    return line


# ---------------------------------------------------------------------------
def run_genotype_bins(args, torch):
    from pg_gpu_amd import band, synthetic, ld_pipeline as LP
    from pg_gpu_amd import _lib as L

    n_var = args.n_var or 50_000
    n_ind = args.n_samples or 1000
    macs = 9  # genotypes with missing data: {v, e1, e2} x {v, e1, e2}
    log("[bench] generating workload")
    g, pos = synthetic.genotype_panel(n_ind=n_ind, n_var=n_var)
    edges = np.logspace(0, 6, 17)
    max_dist = 1e6
    dev = L.device()
    g_d = torch.as_tensor(g, device=dev)
    plan = band.BandPlan(pos, max_dist, edges=edges)
    pair_samples = plan.n_pairs * n_ind
    log(f"[bench] pairs={plan.n_pairs} tiles={plan.n_tiles} pair_samples={pair_samples:.4e}")
    fp4 = args.planes == "fp4"
    packed = [band.pack_population(g_d, None, "geno", fp4=fp4)]
    with_missing = bool(packed[0].has_missing)

    def step():
        band.pack_population(g_d, None, "geno", validate=False, out=packed[0], fp4=fp4)
        return plan.sums(packed, with_missing=with_missing)

    ms_step, out = timed_steps(torch, step, args.steps, args.warmup)
    value = pair_samples / (ms_step * 1e-3) / 1e9
    log(f"[bench] fused step {ms_step:.3f} ms")
    band_ms, band_min = band_kernel_ms(plan, packed, with_missing, max(3, args.steps))
    roof = mfma_roofline(2.0 * macs * pair_samples, band_ms, "k_band_sums", macs, fp4=fp4)
    log(f"[bench] k_band_sums {band_ms:.3f} ms (min {band_min:.3f}) -> {roof['achieved']:.1f} TOP/s")

    plain_ms = check = None
    if not args.no_plain:
        pos_d = torch.as_tensor(pos.astype(np.float64), device=dev)
        edges_d = torch.as_tensor(edges, device=dev)
        specs = LP.generate_stat_specs(1)

        def plain_step():
            sums = torch.zeros((16, 3), dtype=torch.float64, device=dev)
            chunk = LP.estimate_ld_chunk_size(n_ind, num_pops=1)
            for ii, jj in LP.iter_pairs_within_distance(pos, max_dist, chunk):
                c, nv = LP.compute_genotype_counts_for_pairs(g_d, ii, jj)
                st = LP.stats_from_counts("geno", [c], [nv], specs)
                keys = torch.empty(ii.shape[0], dtype=torch.int32, device=dev)
                L.call("pgld_pair_keys", L.ptr(pos_d), L.ptr(ii), L.ptr(jj), ii.shape[0], L.KEY_BIN,
                       L.ptr(edges_d), 16, 0.0, 1.0, L.ptr(keys), L.stream_handle())
                L.call("pgld_keyed_sums", L.ptr(st), L.ptr(keys), ii.shape[0], 3, 16, L.ptr(sums),
                       L.stream_handle())
            return sums
        plain_ms, check = _plain_compare(torch, plain_step, out, args.plain_steps)

    cpu = None
    if not args.no_cpu:
        from oracle import ld_oracle as O
        ii, jj = band_pair_sample(pos, max_dist, args.cpu_pairs or 1_000_000)
        t2 = time.perf_counter()
        acc = np.zeros((16, 3))
        for c0 in range(0, len(ii), 200_000):
            a, b = ii[c0:c0 + 200_000], jj[c0:c0 + 200_000]
            c, nv = O.geno_counts(g, a, b, None)
            st = O.multi_pop_stats_geno([c], [nv])
            acc += O.keyed_sums(O.bin_index(pos[b] - pos[a], edges), st, 16)
        cpu_s = time.perf_counter() - t2
        cpu = {"value": len(ii) * n_ind / cpu_s / 1e9, "unit": "Gpair*samples/s", "cores": 1, "kind": "port",
               "sample": f"{len(ii)} random pairs of the {plan.n_pairs} in the span, 9-way counts + DD/Dz/pi2 + "
                         f"bin sums, NumPy oracle (single thread), {cpu_s:.1f} s"}
        log(f"[bench] cpu oracle {cpu['value']:.4f} Gpair*samples/s over {cpu_s:.1f} s")

    line = base_line(args, value, ms_step,
                     f"genotype LD: {n_var:,} SNPs x {n_ind:,} diploids, 5% missing, HWE Beta(0.5,0.5) dosages, "
                     "all pairs in a 1 Mb span, 16 log-spaced distance bins (1 bp - 1 Mb), single-population "
                     "DD/Dz/pi2 FP64 bin sums (no biallelic filter)",
                     {"n_var": n_var, "n_ind": n_ind, "pairs": plan.n_pairs, "pair_samples": pair_samples,
                      "planes": args.planes}, dtype="fp4" if fp4 else "int8")
    line.update({"roofline": roof, "plain_ms_per_step": plain_ms,
                 "speedup_vs_plain": (plain_ms / ms_step) if plain_ms else None,
                 "fused_vs_plain_max_rel": check, "cpu_baseline": cpu})
    return line


# ---------------------------------------------------------------------------
def run_plain_hap(args, torch):
    """Per-pair outputs through the plain drop-in path: pair enumeration,
    plain counts, DD/Dz/pi2 and r^2 per pair."""
    from pg_gpu_amd import synthetic, ld_pipeline as LP, ld_statistics as LS
    from pg_gpu_amd import _lib as L

    n_var = args.n_var or 2000
    n_hap = args.n_samples or 200
    hap, pos = synthetic.plain_haplotype_panel(n_hap=n_hap, n_var=n_var)
    max_dist = 20_000
    dev = L.device()
    hap_d = torch.as_tensor(hap, device=dev)
    ii, jj = next(iter(LP.iter_pairs_within_distance(pos, max_dist, 10**7)))
    n_pairs = int(ii.shape[0])
    pair_samples = n_pairs * n_hap

    def step():
        c, nv = LP.compute_counts_for_pairs(hap_d, ii, jj)
        st = LP.compute_single_pop_statistics_batch(c, nv)
        r2 = LS.r_squared(c, nv)
        return st, r2

    ms_step, _ = timed_steps(torch, step, args.steps, args.warmup)
    value = pair_samples / (ms_step * 1e-3) / 1e9
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps)
    e0.record()
    for _ in range(reps):
        LP.compute_counts_for_pairs(hap_d, ii, jj)
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / reps
    alg_bytes = 2.0 * pair_samples  # two sample columns per pair, no reuse
    roof = {"bound": "hbm", "achieved": alg_bytes / (k_ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": alg_bytes / (k_ms * 1e-3) / PEAK_HBM, "traffic": None, "kernel": "k_counts_plain",
            "kernel_ms": k_ms, "bytes": "2 B per pair*sample (both SNP columns, no reuse)"}
    log(f"[bench] plain step {ms_step:.3f} ms, counts kernel {k_ms:.3f} ms")
    cpu = None
    if not args.no_cpu:
        from oracle import ld_oracle as O
        ii_h, jj_h = L.to_numpy(ii).astype(np.int64), L.to_numpy(jj).astype(np.int64)
        t2 = time.perf_counter()
        reps_cpu = 0
        while True:
            c, nv = O.hap_counts(hap, ii_h, jj_h, None)
            O.multi_pop_stats_hap([c], [nv])
            O.r_squared(c, nv)
            reps_cpu += 1
            if time.perf_counter() - t2 > 10.0:
                break
        cpu_s = (time.perf_counter() - t2) / reps_cpu
        cpu = {"value": pair_samples / cpu_s / 1e9, "unit": "Gpair*samples/s", "cores": 1, "kind": "port",
               "sample": f"all {n_pairs} pairs, counts + DD/Dz/pi2 + r^2, NumPy oracle (single thread), "
                         f"{reps_cpu} passes, {cpu_s:.2f} s each"}
        log(f"[bench] cpu oracle {cpu['value']:.4f} Gpair*samples/s")
    line = base_line(args, value, ms_step,
                     f"plain per-pair haplotype LD: {n_var:,} SNPs x {n_hap} haplotypes, neutral SFS, no missing "
                     "data, positions uniform over 200 kb; r^2, DD, Dz, pi2 per pair within 20 kb, one population",
                     {"n_var": n_var, "n_hap": n_hap, "pairs": n_pairs, "pair_samples": pair_samples})
    line.update({"roofline": roof, "cpu_baseline": cpu})
    return line


def run_dense_r2(args, torch):
    from pg_gpu_amd import band, synthetic
    from pg_gpu_amd import _lib as L

    n_var = args.n_var or 40_000
    n_hap = args.n_samples or 10_000
    macs = 1  # haplotypes without missing data: h_i.h_j only
    log("[bench] generating workload")
    hap, pos = synthetic.dense_r2_panel(n_hap=n_hap, n_var=n_var)
    dev = L.device()
    hap_d = torch.as_tensor(hap, device=dev)
    plan = band.BandPlan(pos, np.inf, window=1e15, n_keys=1)
    pair_samples = plan.n_pairs * n_hap
    log(f"[bench] pairs={plan.n_pairs} tiles={plan.n_tiles} pair_samples={pair_samples:.4e}")
    fp4 = args.planes == "fp4" and args.method == "mfma"
    pk = band.pack_population(hap_d, None, "hap", fp4=fp4)
    assert not pk.has_missing
    cap = 1 << 22

    def step():
        p2 = band.pack_population(hap_d, None, "hap", validate=False, out=pk, fp4=fp4)
        return plan.r2(p2, n_hist=100, thresh=0.2, sparse_cap=cap, with_missing=False, method=args.method)

    ms_step, out = timed_steps(torch, step, args.steps, args.warmup)
    value = pair_samples / (ms_step * 1e-3) / 1e9
    log(f"[bench] step {ms_step:.3f} ms, {out['n_sparse']} pairs with r^2 > 0.2, "
        f"{int(out['nan_pairs'].sum())} NaN pairs")
    plan.set_timing(True)
    kms = []
    for _ in range(max(3, args.steps)):
        plan.r2(pk, n_hist=100, thresh=0.2, sparse_cap=cap, with_missing=False, method=args.method)
        kms.append(plan.kernel_ms())
    plan.set_timing(False)
    k_ms = float(np.mean(kms))
    roof = mfma_roofline(2.0 * macs * pair_samples, k_ms, "k_band_r2", macs, fp4=fp4)
    if args.method == "popcount":  # VALU variant: report against the int8 peak all the same, say so
        roof["ops"] = "pair*sample MACs x 2 as AND + popcount on the VALU (bit planes), against the int8 MFMA peak"
        roof["kernel"] = "k_band_r2<popcount>"
    log(f"[bench] k_band_r2 ({args.method}) {k_ms:.3f} ms -> {roof['achieved']:.1f} int8-equivalent TOP/s")
    cpu = None
    if not args.no_cpu:
        from oracle import ld_oracle as O
        ii, jj = band_pair_sample(pos, np.inf, args.cpu_pairs or 60_000)
        t2 = time.perf_counter()
        hist = np.zeros((1, 100), dtype=np.int64)
        for c0 in range(0, len(ii), 5000):
            a, b = ii[c0:c0 + 5000], jj[c0:c0 + 5000]
            c, _ = O.hap_counts(hap, a, b, None)
            r2 = O.r_squared_exact(c)
            hist += O.r2_histogram(r2, np.zeros(len(a), dtype=np.int64), 1, 100)[0]
        cpu_s = time.perf_counter() - t2
        cpu = {"value": len(ii) * n_hap / cpu_s / 1e9, "unit": "Gpair*samples/s", "cores": 1, "kind": "port",
               "sample": f"{len(ii)} random pairs of the {plan.n_pairs}, counts + r^2 + histogram, NumPy oracle "
                         f"(single thread), {cpu_s:.1f} s"}
        log(f"[bench] cpu oracle {cpu['value']:.4f} Gpair*samples/s over {cpu_s:.1f} s")
    line = base_line(args, value, ms_step,
                     f"dense all-pairs r^2: {n_var:,} SNPs x {n_hap:,} haplotypes, neutral SFS, no missing data; "
                     "r^2 of every pair from exact integer counts, 100-bin histogram on [0, 1] (NaN pairs counted "
                     "apart) + sparse (i, j, r^2) list for r^2 > 0.2",
                     {"n_var": n_var, "n_hap": n_hap, "pairs": plan.n_pairs, "pair_samples": pair_samples,
                      "n_sparse": out["n_sparse"], "method": args.method, "planes": "fp4" if fp4 else "int8"},
                     dtype="fp4" if fp4 else "int8")
    line.update({"roofline": roof, "cpu_baseline": cpu})
    return line


def run_banded_r2(args, torch):
    """r^2 of every pair within 100 kb on the windowed panel (2% missing data),
    histogram per 100 kb window of the left SNP: the banded missing-data r^2
    workload on which the MFMA and popcount tile products are compared."""
    from pg_gpu_amd import band, synthetic
    from pg_gpu_amd import _lib as L

    n_var = args.n_var or 100_000
    n_hap = args.n_samples or 2000
    macs = 4  # haplotypes with missing data: hh, hv, vh, vv
    log("[bench] generating workload")
    hap, pos, _ = synthetic.windowed_haplotype_panel(n_hap=n_hap, n_var=n_var)
    dev = L.device()
    hap_d = torch.as_tensor(hap, device=dev)
    plan = band.BandPlan(pos, 100_000.0, window=100_000.0, n_keys=100, origin=0.0)
    pair_samples = plan.n_pairs * n_hap
    log(f"[bench] pairs={plan.n_pairs} tiles={plan.n_tiles} pair_samples={pair_samples:.4e}")
    fp4 = args.planes == "fp4" and args.method == "mfma"
    pk = band.pack_population(hap_d, None, "hap", fp4=fp4)
    assert pk.has_missing
    cap = 1 << 22

    def step():
        p2 = band.pack_population(hap_d, None, "hap", validate=False, out=pk, fp4=fp4)
        return plan.r2(p2, n_hist=100, thresh=0.2, sparse_cap=cap, with_missing=True, method=args.method)

    ms_step, out = timed_steps(torch, step, args.steps, args.warmup)
    value = pair_samples / (ms_step * 1e-3) / 1e9
    log(f"[bench] step {ms_step:.3f} ms, {out['n_sparse']} pairs with r^2 > 0.2")
    plan.set_timing(True)
    kms = []
    for _ in range(max(3, args.steps)):
        plan.r2(pk, n_hist=100, thresh=0.2, sparse_cap=cap, with_missing=True, method=args.method)
        kms.append(plan.kernel_ms())
    plan.set_timing(False)
    k_ms = float(np.mean(kms))
    roof = mfma_roofline(2.0 * macs * pair_samples, k_ms, "k_band_r2", macs, fp4=fp4)
    if args.method == "popcount":
        roof["ops"] = "pair*sample MACs x 2 as AND + popcount on the VALU (bit planes), against the int8 MFMA peak"
        roof["kernel"] = "k_band_r2<popcount>"
    log(f"[bench] k_band_r2 ({args.method}, {'fp4' if fp4 else 'int8/bits'}) {k_ms:.3f} ms")
    hist = out["hist"].cpu().numpy()
    line = base_line(args, value, ms_step,
                     f"banded r^2 with missing data: {n_var:,} SNPs x {n_hap:,} haplotypes, 2% missing, pairs within "
                     "100 kb, 100-bin r^2 histogram per 100 kb window + sparse list for r^2 > 0.2",
                     {"n_var": n_var, "n_hap": n_hap, "pairs": plan.n_pairs, "pair_samples": pair_samples,
                      "method": args.method, "planes": "fp4" if fp4 else ("bits" if args.method == "popcount" else
                                                                          "int8"),
                      "hist_total": int(hist.sum())}, dtype="fp4" if fp4 else "int8")
    line.update({"roofline": roof, "cpu_baseline": None})
    return line


def run_ragged_regions(args, torch):
    from pg_gpu_amd import band, batch, synthetic
    from pg_gpu_amd import _lib as L

    n_regions = args.n_var or 512  # --n-var sets the region count here
    macs = 4
    edges = np.logspace(0, 6, 17)
    dev = L.device()
    log("[bench] generating workload on the device")
    datas, pos, rows = synthetic.ragged_regions(n_regions=n_regions, device=dev)
    rows_d = [[torch.as_tensor(np.asarray(r, dtype=np.int32), device=dev) for r in rr] for rr in rows]
    S = np.array([d.shape[0] for d in datas], dtype=np.int64)
    fp4 = args.planes == "fp4"
    packed = batch.pack_regions(datas, rows_d, fp4=fp4)  # validates once
    with_missing = any(p.has_missing for pl in packed for p in pl)
    t0 = time.perf_counter()
    b = batch.RegionBatch(pos, 1e6, edges).bind(packed, stat_set="self")
    plan_s = time.perf_counter() - t0
    per_region_pairs = np.array([len(p) * (len(p) - 1) // 2 for p in pos], dtype=np.int64)
    assert int(per_region_pairs.sum()) == b.n_pairs
    pair_samples = int((per_region_pairs * S).sum())
    log(f"[bench] regions={n_regions} pairs={b.n_pairs} tiles={b.n_tiles} pair_samples={pair_samples:.4e} "
        f"plan+bind {plan_s:.1f} s")
    choice = b.autotune(with_missing)
    log(f"[bench] autotuned schedule {choice}")
    out = torch.empty((n_regions, 16, 12), dtype=torch.float64, device=dev)
    kp = torch.empty((n_regions, 16), dtype=torch.float64, device=dev)

    packer = batch.PackBatch(datas, rows_d, packed)  # every population of every region in one launch

    def step_body():
        packer.run()
        b.sums(with_missing, out=out, key_pairs=kp)

    graph = batch.Graph().capture(step_body)

    def step():
        graph.launch()
        return out

    ms_step, res = timed_steps(torch, step, args.steps, args.warmup)
    res = res.clone()
    value = pair_samples / (ms_step * 1e-3) / 1e9
    log(f"[bench] graph step {ms_step:.3f} ms (one batched pack of 2,048 populations + one persistent pass)")
    b.set_timing(True)
    kms = []
    for _ in range(max(3, min(args.steps, 5))):
        b.sums(with_missing, out=out, key_pairs=kp)
        kms.append(b.kernel_ms())
    b.set_timing(False)
    k_ms = float(np.mean(kms))
    default_shape = n_regions == 512
    roof = mfma_roofline(2.0 * macs * pair_samples, k_ms, "k_batch_sums", macs,
                         RAGGED_TRAFFIC_BYTES.get(args.planes) if default_shape else None, fp4=fp4)
    valu = valu_issue_bound(RAGGED_SQ[args.planes], k_ms) if default_shape and args.planes in RAGGED_SQ else None
    log(f"[bench] k_batch_sums {k_ms:.3f} ms -> {roof['achieved']:.1f} TOP/s"
        + (f", VALU issue {valu['frac']:.2f} of the kernel time, MFMA busy {valu['mfma_busy_frac']:.2f}" if valu else ""))

    # the same work region by region: one band plan and one fused launch each
    plain_ms = check = None
    if not args.no_plain:
        plans = [band.BandPlan(p, 1e6, edges=edges) for p in pos]
        outs = [None] * n_regions

        def per_region():
            for r in range(n_regions):
                outs[r] = plans[r].sums(packed[r], stat_set="self", with_missing=with_missing)
        per_region()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        per_region()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t1) * 1e3
        ref = torch.stack(outs).cpu().numpy()
        check = float(np.max(np.abs(res.cpu().numpy() - ref)))
        log(f"[bench] region-by-region launches {plain_ms:.1f} ms, max |batch - per-region| {check:.3e}")
        del plans

    # error of every statistic against the FP64 oracle, in full, on the
    # regions of least work (all their pairs, every bin)
    oracle_err = None
    if not args.no_oracle:
        oracle_err = ragged_oracle_error(datas, pos, rows, S, edges, res.cpu().numpy(), n_check=args.oracle_regions)
        log(f"[bench] oracle check on regions {oracle_err['regions']}: max rel err {oracle_err['max_rel_err']:.2e}, "
            f"max scaled err {oracle_err['max_scaled_err']:.2e}")

    cpu = None
    if not args.no_cpu:
        from oracle import ld_oracle as O
        specs = [(t, [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(4)]
        t2 = time.perf_counter()
        n_ps = 0
        for r in range(4):
            h = datas[r].cpu().numpy()
            ii, jj = band_pair_sample(pos[r], np.inf, args.cpu_pairs or 300_000, seed=r)
            for c0 in range(0, len(ii), 20_000):
                a, bb = ii[c0:c0 + 20_000], jj[c0:c0 + 20_000]
                cl = [O.hap_counts(h, a, bb, rw) for rw in rows[r]]
                st = O.multi_pop_stats_hap([c for c, _ in cl], [nv for _, nv in cl], specs)
                O.keyed_sums(O.bin_index(pos[r][bb] - pos[r][a], edges), st, 16)
            n_ps += len(ii) * int(S[r])
        cpu_s = time.perf_counter() - t2
        cpu = {"value": n_ps / cpu_s / 1e9, "unit": "Gpair*samples/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_pairs or 300_000} random pairs from each of regions 0-3, 4-population counts + "
                         f"12 self statistics + bin sums, NumPy oracle (single thread), {cpu_s:.1f} s",
               "host": host_cpu()}
        log(f"[bench] cpu oracle {cpu['value']:.4f} Gpair*samples/s over {cpu_s:.1f} s")
    line = base_line(args, value, ms_step,
                     f"ragged persistent batch: {n_regions} independent regions, SNPs uniform 500-20,000, "
                     "haplotypes uniform 50-5,000 in 4 populations, 3% missing, positions uniform over 1 Mb, all "
                     "pairs; per-region 16 log-bin DD/Dz/pi2 sums per population (FP64) from one HIP graph "
                     "(one batched packing launch + one autotuned persistent launch)",
                     {"regions": n_regions, "pairs": b.n_pairs, "pair_samples": pair_samples, "tiles": b.n_tiles,
                      "schedule": choice, "planes": args.planes}, dtype="fp4" if fp4 else "int8")
    line.update({"roofline": roof, "plain_ms_per_step": plain_ms,
                 "speedup_vs_plain": (plain_ms / ms_step) if plain_ms else None,
                 "plain": "one band plan and one fused launch per region", "batch_vs_plain_max_abs": check,
                 "plan_bind_s": plan_s, "valu_roofline": valu, "oracle_error": oracle_err, "cpu_baseline": cpu})
    return line


RAGGED_STATS = [f"{t}_{p}" for t in ("DD", "Dz", "pi2") for p in range(4)]


def ragged_oracle_error(datas, pos, rows, S, edges, res, n_check=4):
    """Every statistic of the ragged batch against the FP64 NumPy oracle on
    the n_check regions of least work (all pairs, all 16 bins): per
    statistic the largest |gpu - oracle| / |oracle| over bins with
    |oracle| > 1e-6 Sum|terms|, and the largest |gpu - oracle| / Sum|terms|
    (the scale of a sum that cancels)."""
    from oracle import ld_oracle as O
    specs = [(t, [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(4)]
    work = np.array([len(p) * (len(p) - 1) // 2 * int(s) for p, s in zip(pos, S)])
    pick = [int(r) for r in np.argsort(work, kind="stable")[:n_check]]
    rel = np.zeros(len(specs))
    scaled = np.zeros(len(specs))
    n_pairs = 0
    for r in pick:
        h = datas[r].cpu().numpy()
        ii, jj = O.pairs_within_distance(pos[r], np.inf)
        n_pairs += len(ii)
        key = O.bin_index(pos[r][jj] - pos[r][ii], edges)
        want = np.zeros((len(edges) - 1, len(specs)))
        absw = np.zeros_like(want)
        for c0 in range(0, len(ii), 50_000):
            a, b = ii[c0:c0 + 50_000], jj[c0:c0 + 50_000]
            cl = [O.hap_counts(h, a, b, rw) for rw in rows[r]]
            st = O.multi_pop_stats_hap([c for c, _ in cl], [nv for _, nv in cl], specs)
            want += O.keyed_sums(key[c0:c0 + 50_000], st, len(edges) - 1)
            absw += O.keyed_sums(key[c0:c0 + 50_000], np.abs(st), len(edges) - 1)
        # the batch stores column s * P + p = statistic s of population p, as the specs above
        got = res[r]
        err = np.abs(got - want)
        ok = np.abs(want) > 1e-6 * absw
        rel = np.maximum(rel, np.max(np.where(ok, err / np.where(ok, np.abs(want), 1.0), 0.0), axis=0))
        scaled = np.maximum(scaled, np.max(err / np.maximum(absw, 1e-300), axis=0))
    return {"regions": pick, "pairs": int(n_pairs), "oracle": "NumPy FP64 (oracle/ld_oracle.py), all pairs",
            "max_rel_err": float(rel.max()), "max_scaled_err": float(scaled.max()),
            "max_rel_err_per_stat": dict(zip(RAGGED_STATS, map(float, rel))),
            "max_scaled_err_per_stat": dict(zip(RAGGED_STATS, map(float, scaled)))}


def host_cpu():
    """Model name and logical CPU count of the host the CPU baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here "
              "(the LD band would shard by SNP range, each GPU taking a contiguous block of left "
              "SNPs plus a max_dist halo); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calibration:
        n = 1 << 30
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {n} bytes 4 times")
        return
    run = {"windowed_hap": run_windowed_hap, "genotype_bins": run_genotype_bins, "plain_hap": run_plain_hap,
           "dense_r2": run_dense_r2, "ragged_regions": run_ragged_regions, "banded_r2": run_banded_r2}
    line = run[args.workload](args, torch)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
