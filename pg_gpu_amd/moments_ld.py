"""moments-compatible LD parsing -- the drop-in for
pg_gpu.moments_ld.compute_ld_statistics (moments_ld.py:40-388).

The per-bin sums come from one fused int8-MFMA band pass (band.BandPlan)
when a fused statistic set exists for the request (haplotypes and genotypes
with up to 4 populations); otherwise from the plain path (pair
chunks -> per-pair counts -> per-pair statistics -> keyed sums), which is
also what ``method="plain"`` selects.  Both run entirely in libpgld.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from . import band as B
from . import ld_pipeline as LP
from .genotype_matrix import GenotypeMatrix
from .haplotype_matrix import HaplotypeMatrix


def compute_ld_statistics(vcf_file=None, rec_map_file=None, pop_file=None, pops=None, r_bins=None,
                          bp_bins=None, use_genotypes=False, report=True, ac_filter=True,
                          haplotype_matrix=None, genotype_matrix=None, accessible_bed=None, method="auto"):
    """Same arguments and result layout as moments.LD.Parsing.
    compute_ld_statistics: {'bins': [(lo, hi)], 'sums': [per-bin arrays of
    ld_names(P)] + [H statistics], 'stats': (ld_names, het_names), 'pops'}."""
    if pops is None:
        pops = ["pop0", "pop1"]
    num_pops = len(pops)
    if num_pops < 1 or num_pops > 4:
        raise ValueError("1-4 populations supported")
    if r_bins is None and bp_bins is None:
        raise ValueError("Either r_bins or bp_bins must be provided")
    if accessible_bed is not None:
        raise NotImplementedError("accessible-site masks are outside this build's LD path")

    if use_genotypes:
        if genotype_matrix is not None:
            mat = genotype_matrix
        elif haplotype_matrix is not None:
            mat = GenotypeMatrix.from_haplotype_matrix(haplotype_matrix)
        else:
            if vcf_file is None:
                raise ValueError("vcf_file or genotype_matrix required")
            if pop_file is None:
                raise ValueError("pop_file is required when loading from VCF")
            if report:
                print(f"Loading {vcf_file} (genotypes) ...")
            mat = GenotypeMatrix.from_vcf(vcf_file)
            mat.load_pop_file(pop_file, pops=pops)
            if ac_filter:
                mat = mat.apply_biallelic_filter()
    else:
        if haplotype_matrix is not None:
            mat = haplotype_matrix
        else:
            if vcf_file is None:
                raise ValueError("vcf_file or haplotype_matrix is required")
            if pop_file is None:
                raise ValueError("pop_file is required when loading from VCF")
            if report:
                print(f"Loading {vcf_file} ...")
            mat = HaplotypeMatrix.from_vcf(vcf_file)
            mat.load_pop_file(pop_file, pops=pops)
            if ac_filter:
                mat = mat.apply_biallelic_filter()
    mat.transfer_to_gpu()

    pos = np.asarray(L.to_numpy(mat.positions), dtype=np.float64)
    if r_bins is not None:
        if rec_map_file is None:
            raise ValueError("rec_map_file required with r_bins")
        bins = np.asarray(r_bins, dtype=np.float64)
        gen = _interpolate_genetic_distances(pos, rec_map_file)
        max_bp = _max_bp_for_r_dist(pos, gen, float(bins[-1]))
    else:
        bins = np.asarray(bp_bins, dtype=np.float64)
        gen = None
        max_bp = float(bins[-1])
    if report:
        print(f"  Computing LD ({len(bins) - 1} bins, {num_pops} pops) ...")

    sums = _compute_ld_sums(mat, pops, bins, gen, max_bp, use_genotypes, method)
    het = _compute_heterozygosity(mat, pops, use_genotypes)
    if report:
        print("  Done.")
    names, hnames = LP.ld_names(num_pops), LP.het_names(num_pops)
    bin_tuples = [(float(bins[i]), float(bins[i + 1])) for i in range(len(bins) - 1)]
    sums_list = [sums[i] for i in range(len(bins) - 1)]
    sums_list.append(np.array([het[h] for h in hnames]))
    return {"bins": bin_tuples, "sums": sums_list, "stats": (names, hnames), "pops": pops}


def _interpolate_genetic_distances(positions, rec_map_file):
    """Map position -> Morgans by linear interpolation of a (pos, cM) map
    (moments_ld.py:211-223)."""
    mp, mv = [], []
    with open(rec_map_file) as f:
        for line in f:
            parts = line.strip().split()
            if len(parts) >= 2:
                try:
                    mp.append(float(parts[0]))
                    mv.append(float(parts[1]) / 100.0)
                except ValueError:
                    continue
    return np.interp(positions, np.array(mp), np.array(mv))


def _max_bp_for_r_dist(positions, gen_dists, max_r):
    """Conservative physical span for a recombination distance
    (moments_ld.py:226-245)."""
    chrom_len = float(positions[-1] - positions[0])
    if len(positions) < 2:
        return chrom_len
    bp = np.diff(positions).astype(np.float64)
    rd = np.diff(gen_dists)
    ok = bp > 0
    if not np.any(ok):
        return chrom_len
    rates = rd[ok] / bp[ok]
    min_rate = np.min(rates[rates > 0]) if np.any(rates > 0) else max_r / chrom_len
    return min(max_r / min_rate * 1.1, chrom_len)


def _union_biallelic_keep(mat, pops):
    """moments' union-of-populations biallelic filter for genotypes
    (moments_ld.py:256-285), from per-variant dosage sums."""
    seen = []
    for p in pops:
        for i in mat.sample_sets[p]:
            if i not in seen:
                seen.append(i)
    alt, valid = (a.cpu().numpy().astype(np.int64) for a in mat.allele_counts(seen))
    return np.nonzero((alt > 0) & (alt < 2 * valid) & (valid >= 2))[0]


def _compute_ld_sums(mat, pops, bins, gen, max_bp, use_genotypes, method="auto"):
    P = len(pops)
    if use_genotypes:
        keep = _union_biallelic_keep(mat, pops)
        mat = mat.get_subset(keep)
        if gen is not None:
            gen = gen[keep]
        data = mat.genotypes
    else:
        data = mat.haplotypes
    pos = np.asarray(L.to_numpy(mat.positions), dtype=np.float64)
    rows = [mat.sample_sets[p] for p in pops]
    fam = "geno" if use_genotypes else "hap"
    if method == "plain":
        return _plain_ld_sums(data, pos, rows, bins, gen, max_bp, use_genotypes)
    # fused band pass, or the per-pair path where no fused kernel covers the set
    sums, _ = B.ld_sums(data, fam, pos, rows, max_bp, edges=bins, key_coord=gen, abs_dist=gen is not None,
                        method=method)
    return sums.cpu().numpy()


def _plain_ld_sums(data, pos, rows, bins, gen, max_bp, use_genotypes):
    """The per-pair path: pair chunks -> counts -> statistics -> keyed sums
    (the structure of moments_ld.py:294-343)."""
    mode = L.KEY_BIN_ABS if gen is not None else L.KEY_BIN
    sums, _ = LP.plain_keyed_sums(data, "geno" if use_genotypes else "hap", pos, rows, max_bp, mode, len(bins) - 1,
                                  edges=np.asarray(bins, dtype=np.float64), key_coord=gen)
    return sums.cpu().numpy()


def _compute_heterozygosity(mat, pops, use_genotypes=False):
    """H_i_j from per-variant allele counts (moments_ld.py:346-388); for
    haplotypes the sample size is the constant population size, ignoring
    missing data, as the reference does."""
    alt, ref, size = [], [], []
    for p in pops:
        rows = mat.sample_sets[p]
        a, v = (x.cpu().numpy().astype(np.float64) for x in mat.allele_counts(rows))
        nh = 2.0 * v if use_genotypes else np.full_like(a, float(len(rows)))
        alt.append(a)
        ref.append(nh - a)
        size.append(nh)
    out = {}
    for a in range(len(pops)):
        for b in range(a, len(pops)):
            if a == b:
                v = np.sum(2.0 * ref[a] * alt[a] / (size[a] * (size[a] - 1)))
            else:
                v = np.sum((ref[a] * alt[b] + alt[a] * ref[b]) / (size[a] * size[b]))
            out[f"H_{a}_{b}"] = float(v)
    return out
