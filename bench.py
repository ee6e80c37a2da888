"""Benchmark: batched windowed haplotype LD on one MI355X.

Workload: 100,000 SNPs x 2,000 haplotypes (two populations of 1,000), 2% MCAR
missing, heterogeneous positions over 10 Mb; all pairs within 100 kb; the 15
two-population moments statistics summed in FP64 per 100 kb window (keyed by
the left SNP).  One step = pack both populations' indicator planes from the
resident int8 matrix + one fused tiled int8-MFMA band pass + the per-window
reduction.

Prints one JSON line (see README of the harness contract): value is
Gpair*samples/s for the whole job; roofline is for the dominant kernel
(k_band_sums, int8 MFMA bound); cpu_baseline is the NumPy oracle on a bounded
random sample of the same pairs.
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
PEAK_HBM = 8.0e12        # bytes/s
MACS_PER_PAIR_SAMPLE = 4  # haplotypes with missing data: hh, hv, vh, vv
# HBM bytes per k_band_sums launch on the default workload, from rocprofv3
# FETCH_SIZE (2,187,585 KiB) and WRITE_SIZE (3,908 KiB) per dispatch, scaled
# by the same counters' reading of a 1 GiB device copy (--copy-calibration:
# FETCH x 2.000, WRITE x 1.000).  Summary and commands:
# profiles/k_band_sums_hap2pop_windows_100k_snps_2000_haps.md
TRAFFIC_BYTES_PER_LAUNCH = 4.484e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-var", type=int, default=100_000)
    ap.add_argument("--n-hap", type=int, default=2000)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-pairs", type=int, default=1_500_000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 device copies of 1 GiB (the HBM-counter reference of the traffic figure)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
    from pg_gpu_amd import band, synthetic, ld_pipeline as LP
    from pg_gpu_amd import _lib as L

    log("[bench] generating workload")
    hap, pos, pops = synthetic.config_c2(n_hap=args.n_hap, n_var=args.n_var)
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
    packed = [band.pack_population(hap_d, r, "hap") for r in rows_d]
    with_missing = any(p.has_missing for p in packed)

    def step():
        for k, r in enumerate(rows_d):
            band.pack_population(hap_d, r, "hap", validate=False, out=packed[k])
        return plan.sums(packed, with_missing=with_missing), packed

    for _ in range(args.warmup):
        out, packed = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        out, packed = step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    value = pair_samples / (ms_step * 1e-3) / 1e9
    log(f"[bench] fused step {ms_step:.3f} ms  ({wall:.3f} s wall for {args.steps})")

    # dominant kernel: k_band_sums alone, timed by the library with HIP events
    # recorded on the stream it is launched on, averaged over launches
    reps = max(3, args.steps)
    plan.set_timing(True)
    kms = []
    for _ in range(reps):
        plan.sums(packed, with_missing=with_missing)
        kms.append(plan.kernel_ms())
    plan.set_timing(False)
    band_ms = float(np.mean(kms))
    ops = 2.0 * MACS_PER_PAIR_SAMPLE * pair_samples
    achieved = ops / (band_ms * 1e-3)
    log(f"[bench] k_band_sums {band_ms:.3f} ms (min {min(kms):.3f}) -> {achieved / 1e12:.1f} int8 TOP/s")
    traffic = None
    if (args.n_var, args.n_hap) == (100_000, 2000):
        traffic = TRAFFIC_BYTES_PER_LAUNCH

    # plain per-pair path on identical inputs and outputs (speedup vs plain)
    plain_ms = None
    check = None
    if not args.no_plain:
        def plain_step():
            sums = torch.zeros((n_keys, 15), dtype=torch.float64, device=dev)
            pos_d = torch.as_tensor(pos.astype(np.float64), device=dev)
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
        ref = plain_step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.plain_steps):
            ref = plain_step()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t1) * 1e3 / args.plain_steps
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        check = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
        log(f"[bench] plain path {plain_ms:.1f} ms, max rel diff fused vs plain {check:.2e}")

    cpu = None
    if not args.no_cpu:
        from oracle import ld_oracle as O
        rng = np.random.default_rng(7)
        ii, jj = O.pairs_within_distance(pos, max_dist)
        sel = np.sort(rng.choice(len(ii), size=min(args.cpu_pairs, len(ii)), replace=False))
        ii, jj = ii[sel], jj[sel]
        chunk = 250_000
        t2 = time.perf_counter()
        acc = np.zeros((n_keys, 15))
        for c0 in range(0, len(ii), chunk):
            a, b = ii[c0:c0 + chunk], jj[c0:c0 + chunk]
            cl = [O.hap_counts(hap, a, b, r) for r in pops]
            st = O.two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
            acc += O.keyed_sums(O.window_key(pos[a], 0.0, window), st, n_keys)
        cpu_s = time.perf_counter() - t2
        cpu = {"value": len(ii) * sum(S) / cpu_s / 1e9, "unit": "Gpair*samples/s", "cores": 1, "kind": "port",
               "sample": f"{len(ii)} random pairs of the same band (of {plan.n_pairs}), both populations, "
                         f"counts + 15 statistics + window sums, NumPy oracle (single thread), {cpu_s:.1f} s"}
        log(f"[bench] cpu oracle {cpu['value']:.4f} Gpair*samples/s over {cpu_s:.1f} s")

    line = {
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
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "batched windowed haplotype LD: 100,000 SNPs x 2,000 haplotypes (2 pops x 1,000), "
                               "2% MCAR missing, gamma(0.5) heterogeneous positions over 10 Mb, pairs within "
                               "100 kb, 15 two-pop stats summed per 100 kb window (FP64)",
                   "n_var": args.n_var, "n_hap": args.n_hap, "pairs": plan.n_pairs, "pair_samples": pair_samples},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_INT8_OPS / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_INT8_OPS, "traffic": traffic,
                     "kernel": "k_band_sums", "kernel_ms": band_ms,
                     "ops": "int8 MFMA ops (1 MAC = 2 ops) of 4 MACs per pair*sample"},
        "plain_ms_per_step": plain_ms,
        "speedup_vs_plain": (plain_ms / ms_step) if plain_ms else None,
        "fused_vs_plain_max_rel": check,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
