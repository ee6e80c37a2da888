"""HaplotypeMatrix -- the LD-path part of pg_gpu.haplotype_matrix
(haplotype_matrix.py:10-292, :475-789, :1320-1866).

Data: int8 (n_haplotypes, n_variants), values {-1 missing, 0, 1}; a NumPy
array on the host ('CPU') or a torch tensor on the GPU ('GPU').  VCF rows are
laid out as the reference does: sample s has haplotypes s and s + n_samples
(haplotype_matrix.py:326-329).  Accessible-site masks, zarr/tskit I/O and the
diversity methods of the reference are outside this build.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np

from . import _lib as L
from . import band as B
from . import ld_pipeline as LP
from . import vcf as V


def _xp_is_gpu(x):
    return L.is_tensor(x)


class HaplotypeMatrix:
    """Phased haplotypes with positions and population labels."""

    def __init__(self, genotypes, positions, chrom_start: int = None, chrom_end: int = None,
                 sample_sets: dict = None, n_total_sites: int = None, samples: list = None,
                 accessible_mask=None):
        if accessible_mask is not None:
            raise NotImplementedError("accessible-site masks are outside this build's LD path")
        if not (isinstance(genotypes, np.ndarray) or L.is_tensor(genotypes)):
            raise ValueError("genotypes must be a numpy array or a torch tensor")
        if (genotypes.numel() if L.is_tensor(genotypes) else genotypes.size) == 0:
            raise ValueError("genotypes cannot be empty")
        if (positions.numel() if L.is_tensor(positions) else np.size(positions)) == 0:
            raise ValueError("positions cannot be empty")
        if _xp_is_gpu(genotypes):
            self._device = "GPU"
            self._haplotypes = genotypes.to(L.torch().int8)
            self._positions = L.to_device(positions, np.int64)
        else:
            self._device = "CPU"
            self._haplotypes = np.asarray(genotypes).astype(np.int8, copy=False)
            self._positions = np.asarray(L.to_numpy(positions))
        self.chrom_start = chrom_start
        self.chrom_end = chrom_end
        self._sample_sets = None
        if sample_sets is not None:
            self.sample_sets = sample_sets
        self.n_total_sites = n_total_sites
        self.samples = samples

    # ------------------------------------------------------------------ data
    @property
    def haplotypes(self):
        return self._haplotypes

    @haplotypes.setter
    def haplotypes(self, value):
        self._haplotypes = value

    @property
    def positions(self):
        return self._positions

    @positions.setter
    def positions(self, value):
        self._positions = value

    @property
    def device(self):
        return self._device

    @property
    def sample_sets(self):
        """Population name -> list of haplotype rows; {'all': every row} by
        default (haplotype_matrix.py:137-150)."""
        if self._sample_sets is None:
            return {"all": list(range(self.num_haplotypes))}
        return self._sample_sets

    @sample_sets.setter
    def sample_sets(self, sample_sets: dict):
        if not isinstance(sample_sets, dict):
            raise ValueError("sample_sets must be a dictionary")
        for value in sample_sets.values():
            if not isinstance(value, list):
                raise ValueError("values in sample_sets must be lists")
        self._sample_sets = sample_sets

    def transfer_to_gpu(self):
        if self._device == "CPU":
            self._haplotypes = L.to_device(self._haplotypes, np.int8)
            self._positions = L.to_device(self._positions, np.int64)
            self._device = "GPU"

    def transfer_to_cpu(self):
        if self._device == "GPU":
            self._haplotypes = L.to_numpy(self._haplotypes)
            self._positions = L.to_numpy(self._positions)
            self._device = "CPU"

    def get_matrix(self):
        return self._haplotypes

    def get_positions(self):
        return self._positions

    @property
    def shape(self):
        return tuple(self._haplotypes.shape)

    @property
    def num_variants(self):
        return int(self._haplotypes.shape[1])

    @property
    def num_haplotypes(self):
        return int(self._haplotypes.shape[0])

    def __repr__(self):
        pos = L.to_numpy(self._positions)
        return (f"HaplotypeMatrix(shape={self.shape}, first_position={pos[0] if pos.size else None}, "
                f"last_position={pos[-1] if pos.size else None})")

    def _pos_np(self):
        return np.asarray(L.to_numpy(self._positions))

    # ------------------------------------------------------------------ I/O
    @classmethod
    def from_vcf(cls, path: str, region: str = None, samples: list = None, include_invariant: bool = False,
                 accessible_bed: str = None):
        """Load phased GT calls (haplotype_matrix.py:293-359); ``region`` is
        'chrom:start-end' (1-based inclusive) or 'chrom'."""
        if accessible_bed is not None:
            raise NotImplementedError("accessible-site masks are outside this build's LD path")
        d = V.read_vcf(path, samples=samples)
        gt, pos, chrom = d["gt"], d["positions"], d["chrom"]
        chrom_start, chrom_end = int(pos[0]), int(pos[-1])
        if region is not None:
            c, _, rng = region.partition(":")
            keep = np.ones(len(pos), dtype=bool)
            if c:
                keep &= chrom == c
            if rng:
                lo, _, hi = rng.partition("-")
                if lo:
                    chrom_start = int(lo.replace(",", ""))
                    keep &= pos >= chrom_start
                if hi:
                    chrom_end = int(hi.replace(",", ""))
                    keep &= pos <= chrom_end
            gt, pos = gt[keep], pos[keep]
            if len(pos) == 0:
                raise ValueError(f"No variants found in {path} for region {region}")
        n = gt.shape[1]
        hap = np.empty((2 * n, gt.shape[0]), dtype=np.int8)
        hap[:n, :] = gt[:, :, 0].T
        hap[n:, :] = gt[:, :, 1].T
        return cls(hap, pos, chrom_start, chrom_end, n_total_sites=gt.shape[0] if include_invariant else None,
                   samples=d["samples"])

    def load_pop_file(self, pop_file: str, pops: list = None):
        """Populations from a sample/pop file; sample s contributes rows s and
        s + n_samples (haplotype_matrix.py:475-512)."""
        if self.samples is None:
            raise ValueError("No sample names stored. Use from_vcf() to load data.")
        n = len(self.samples)
        pop_map = V.read_pop_file(pop_file)
        if pops is None:
            pops = sorted(set(pop_map.values()))
        sets = {p: [] for p in pops}
        for i, name in enumerate(self.samples):
            p = pop_map.get(name)
            if p in sets:
                sets[p].append(i)
                sets[p].append(i + n)
        self.sample_sets = sets

    # ------------------------------------------------------------------ subsets
    def _take(self, idx):
        if self._device == "GPU":
            it = L.to_device(idx, np.int64)
            return self._haplotypes[:, it], self._positions[it]
        idx = np.asarray(idx, dtype=np.int64)
        return self._haplotypes[:, idx], self._positions[idx]

    def get_subset(self, positions) -> "HaplotypeMatrix":
        idx = np.asarray(L.to_numpy(positions), dtype=np.int64)
        if idx.ndim != 1:
            raise ValueError("Positions must be a one-dimensional array.")
        if idx.size and (idx.min() < 0 or idx.max() >= self.num_variants):
            raise ValueError("Positions must be valid indices within the haplotype matrix.")
        h, p = self._take(idx)
        out = HaplotypeMatrix.__new__(HaplotypeMatrix)
        out._device, out._haplotypes, out._positions = self._device, h, p
        out.chrom_start, out.chrom_end = self.chrom_start, self.chrom_end
        out._sample_sets, out.n_total_sites, out.samples = self._sample_sets, self.n_total_sites, self.samples
        return out

    def get_subset_from_range(self, low: int, high: int) -> "HaplotypeMatrix":
        """Variants with low <= position < high.  As in the reference
        (haplotype_matrix.py:704), ``high`` may not exceed the number of
        variants -- a check on the count, not on positions."""
        pos = self._pos_np()
        if low < 0 or high > pos.size or low >= high:
            raise ValueError("Invalid range specified")
        idx = np.nonzero((pos >= low) & (pos < high))[0]
        sub = self.get_subset(idx)
        sub.chrom_start, sub.chrom_end = low, high
        return sub

    def allele_counts(self, rows=None):
        """(alt, valid) int32 per variant over the given rows (GPU)."""
        g = L.to_device(self._haplotypes, np.int8)
        rows_t, n_sel = LP._rows_tensor(rows, g.shape[0])
        alt = L.empty((g.shape[1],), np.int32)
        valid = L.empty((g.shape[1],), np.int32)
        L.call("pgld_allele_counts", L.ptr(g), g.shape[0], g.shape[1], L.ptr(rows_t), n_sel, L.ptr(alt),
               L.ptr(valid), L.stream_handle())
        return alt, valid

    def apply_biallelic_filter(self) -> "HaplotypeMatrix":
        """Keep sites with no value >= 2 and both a 0 and a 1 among all
        haplotypes (haplotype_matrix.py:724-789)."""
        g = L.to_device(self._haplotypes, np.int8)
        has_missing, bad = LP.scan_rows(g, None, g.shape[0], 1)
        x = L.to_numpy(self._haplotypes) if bad else None
        alt, valid = (a.cpu().numpy() for a in self.allele_counts())
        if bad:  # values >= 2 present: count per site on the host
            multi = (x >= 2).sum(axis=0)
            ones = (x == 1).sum(axis=0)
            zeros = (x == 0).sum(axis=0)
            keep = (multi == 0) & (zeros > 0) & (ones > 0)
        else:
            keep = (alt > 0) & (valid - alt > 0)
        idx = np.nonzero(keep)[0]
        sub = self.get_subset(idx)
        pos = sub._pos_np()
        if pos.size:
            sub.chrom_start, sub.chrom_end = int(pos[0]), int(pos[-1])
        return sub

    # ------------------------------------------------------------------ missing data
    def is_missing(self, axis=None):
        m = self._haplotypes < 0
        if axis is None:
            return m
        return m.sum(axis) if self._device == "CPU" else m.sum(dim=axis)

    def count_missing(self, axis=None):
        m = self._haplotypes < 0
        if axis is None:
            return int(m.sum())
        return m.sum(axis) if self._device == "CPU" else m.sum(dim=axis)

    # ------------------------------------------------------------------ LD
    def _all_pairs(self):
        m = self.num_variants
        iu = np.triu_indices(m, k=1)
        return iu[0].astype(np.int32), iu[1].astype(np.int32)

    def tally_gpu_haplotypes(self, pop=None):
        """Counts [n11, n10, n01, n00] for all m(m-1)/2 pairs (upper-triangle
        order) and n_valid or None (haplotype_matrix.py:1320-1434); exact for
        any sample count, with or without missing data."""
        self.transfer_to_gpu()
        rows = None
        if pop is not None:
            if self._sample_sets is None:
                raise ValueError("sample_sets must be defined to use pop parameter")
            if pop not in self._sample_sets:
                raise KeyError(f"Population key {pop} must exist in sample_sets")
            rows = self._sample_sets[pop]
        ii, jj = self._all_pairs()
        return LP.compute_counts_for_pairs(self._haplotypes, ii, jj, rows)

    def tally_gpu_haplotypes_two_pops(self, pop1: str, pop2: str):
        """(counts (N, 8), n_valid1, n_valid2) for all pairs
        (haplotype_matrix.py:1518-1589); n_valid are None without missing data."""
        self.transfer_to_gpu()
        if self._sample_sets is None:
            raise ValueError("sample_sets must be defined to use this function")
        if pop1 not in self._sample_sets or pop2 not in self._sample_sets:
            raise KeyError(f"Population keys {pop1} and {pop2} must exist in sample_sets")
        ii, jj = self._all_pairs()
        c1, n1 = LP.compute_counts_for_pairs(self._haplotypes, ii, jj, self._sample_sets[pop1])
        c2, n2 = LP.compute_counts_for_pairs(self._haplotypes, ii, jj, self._sample_sets[pop2])
        has_missing = LP.scan_rows(L.to_device(self._haplotypes, np.int8), None, self.num_haplotypes, 1)[0]
        if has_missing:
            T = L.torch()
            if n1 is None:
                n1 = c1.sum(dim=1, dtype=T.int32)
            if n2 is None:
                n2 = c2.sum(dim=1, dtype=T.int32)
            return L.torch().cat([c1, c2], dim=1), n1, n2
        return L.torch().cat([c1, c2], dim=1), None, None

    def tally_gpu_haplotypes_two_pops_with_missing(self, pop1: str, pop2: str):
        c, n1, n2 = self.tally_gpu_haplotypes_two_pops(pop1, pop2)
        T = L.torch()
        if n1 is None:
            n1 = c[:, :4].sum(dim=1, dtype=T.int32)
            n2 = c[:, 4:].sum(dim=1, dtype=T.int32)
        return c, n1, n2

    def _band_bin_sums(self, bp_bins, pops, stat_set="moments"):
        """Fused band pass keyed by distance bins: (sums (n_bins, S), pairs per bin)."""
        self.transfer_to_gpu()
        bins = np.asarray(bp_bins, dtype=np.float64)
        sums, kp = B.ld_sums(self._haplotypes, "hap", self._pos_np(), pops, float(bins[-1]), edges=bins,
                             stat_set=stat_set)
        return sums.cpu().numpy(), kp.cpu().numpy()

    def compute_ld_statistics_gpu_single_pop(self, bp_bins, raw=False, ac_filter=True, chunk_size="auto"):
        """{(lo, hi): (DD, Dz, pi2)} per distance bin [lo, hi), sums if raw
        else means over pairs (haplotype_matrix.py:1592-1716)."""
        if ac_filter:
            return self.apply_biallelic_filter().compute_ld_statistics_gpu_single_pop(
                bp_bins, raw=raw, ac_filter=False, chunk_size=chunk_size)
        bins = np.asarray(bp_bins, dtype=np.float64)
        sums, cnt = self._band_bin_sums(bins, [None])
        out = {}
        for k in range(len(bins) - 1):
            key = (float(bins[k]), float(bins[k + 1]))
            if raw:
                out[key] = tuple(float(x) for x in sums[k])
            else:
                out[key] = tuple(float(x) / cnt[k] for x in sums[k]) if cnt[k] > 0 else (0.0, 0.0, 0.0)
        return out

    def compute_ld_statistics_gpu_two_pops(self, bp_bins, pop1: str, pop2: str, raw=False, ac_filter=True,
                                           chunk_size="auto"):
        """{(lo, hi): OrderedDict of the 15 two-population statistics} per
        distance bin (haplotype_matrix.py:1719-1866)."""
        if ac_filter:
            return self.apply_biallelic_filter().compute_ld_statistics_gpu_two_pops(
                bp_bins, pop1, pop2, raw=raw, ac_filter=False, chunk_size=chunk_size)
        bins = np.asarray(bp_bins, dtype=np.float64)
        sums, cnt = self._band_bin_sums(bins, [self._sample_sets[pop1], self._sample_sets[pop2]])
        names = LP.ld_names(2)
        out = {}
        for k in range(len(bins) - 1):
            key = (float(bins[k]), float(bins[k + 1]))
            if raw:
                vals = sums[k]
            else:
                vals = sums[k] / cnt[k] if cnt[k] > 0 else np.zeros(15)
            out[key] = OrderedDict((n, float(v)) for n, v in zip(names, vals))
        return out

    def pairwise_LD_v(self):
        """Pairwise D = p_AB - p_i p_j of every variant pair as a dense (m, m)
        float64 GPU tensor, zero diagonal: per-site frequencies over all
        valid samples, joint frequency over the pair's valid samples (0 where
        no sample is valid), as _pairwise_ld_core (haplotype_matrix.py:
        1092-1136).  The joint counts come from the MFMA tiles."""
        self.transfer_to_gpu()
        T = L.torch()
        m = self.num_variants
        out = T.zeros((m, m), dtype=T.float64, device=L.device())
        if m < 2:
            return out
        plan = B.BandPlan(np.arange(m, dtype=np.float64), np.inf, window=1e300, n_keys=1)
        pk = B.pack_population(self._haplotypes, None, "hap", fp4=True)
        vals = plan.r2(pk, histogram=False, sparse_cap=0, per_pair=True, estimator="d_naive")["r2"]
        iu = T.triu_indices(m, m, offset=1, device=out.device)
        out[iu[0], iu[1]] = vals
        out[iu[1], iu[0]] = vals
        return out

    def pairwise_r2(self, estimator: str = "r2"):
        """Pairwise r^2 of every variant pair as a dense (m, m) float64 GPU
        tensor: per-site frequencies over all valid samples, joint frequency
        over the pair's valid samples, 0 where undefined, zero diagonal
        (haplotype_matrix.py:1139-1170).  Counts through the MFMA tiles."""
        if estimator not in ("r2", "rogers_huff"):
            raise ValueError(f"Unknown estimator: {estimator!r} (expected 'r2' or 'rogers_huff')")
        self.transfer_to_gpu()
        T = L.torch()
        m = self.num_variants
        if estimator == "rogers_huff":
            from .ld_statistics import rogers_huff_r
            r = rogers_huff_r(self)
            out = T.zeros((m, m), dtype=T.float64, device=L.device())
            if m >= 2:
                v = T.nan_to_num(r * r, nan=0.0)
                iu = T.triu_indices(m, m, offset=1, device=out.device)
                out[iu[0], iu[1]] = v
                out[iu[1], iu[0]] = v
            return out
        out = T.zeros((m, m), dtype=T.float64, device=L.device())
        if m < 2:
            return out
        plan = B.BandPlan(np.arange(m, dtype=np.float64), np.inf, window=1e300, n_keys=1)
        pk = B.pack_population(self._haplotypes, None, "hap", fp4=True)
        vals = plan.r2(pk, histogram=False, sparse_cap=0, per_pair=True, estimator="naive")["r2"]
        iu = T.triu_indices(m, m, offset=1, device=out.device)
        out[iu[0], iu[1]] = vals
        out[iu[1], iu[0]] = vals
        return out

    def locate_unlinked(self, size=100, step=20, threshold=0.1):
        """Variants in approximate linkage equilibrium: windows of `size`
        variants every `step`; within a window, walking the still-unlinked
        variants in order, each one drops every later variant whose
        frequency r^2 with it exceeds `threshold` (haplotype_matrix.py:
        1172-1231).  The window's r^2 matrix comes from the MFMA tiles; the
        greedy pruning runs on the host, as in the reference."""
        self.transfer_to_gpu()
        m = self.num_variants
        loc = np.ones(m, dtype=bool)
        for w_start in range(0, m, step):
            w_end = min(w_start + size, m)
            active_idx = np.nonzero(loc[w_start:w_end])[0] + w_start
            if len(active_idx) <= 1:
                continue
            r2 = self.get_subset(active_idx).pairwise_r2().cpu().numpy()
            n_active = len(active_idx)
            for i in range(n_active):
                if not loc[active_idx[i]]:
                    continue
                for j in range(i + 1, n_active):
                    if loc[active_idx[j]] and r2[i, j] > threshold:
                        loc[active_idx[j]] = False
        return loc

    def windowed_r_squared(self, bp_bins, percentile=50, pop=None, estimator: str = "r2"):
        """Percentiles of r^2 (from each pair's counts; NaN pairs excluded)
        per distance bin [b_k, b_{k+1}), and the pair count per bin
        (haplotype_matrix.py:1233-1317).  r^2 of the pairs within the last
        edge come from the MFMA tiles; the percentiles are taken on the
        host, as the reference does."""
        if estimator not in ("r2", "rogers_huff"):
            raise ValueError(f"Unknown estimator: {estimator!r} (expected 'r2' or 'rogers_huff')")
        if estimator == "rogers_huff" and pop is not None:
            raise NotImplementedError("windowed_r_squared(estimator='rogers_huff', pop=...) is not yet implemented; "
                                      "pass the full matrix or subset to the desired haplotypes first.")
        self.transfer_to_gpu()
        bins = np.asarray(bp_bins, dtype=np.float64)
        n_bins = len(bins) - 1
        rows = None
        if pop is not None:
            if self._sample_sets is None or pop not in self._sample_sets:
                raise KeyError(f"Population key {pop} must exist in sample_sets")
            rows = self._sample_sets[pop]
        pos = self._pos_np().astype(np.float64)
        plan = B.BandPlan(pos, float(bins[-1]), edges=bins)
        if estimator == "rogers_huff":
            from .ld_statistics import _dosage_matrix
            pk = B.pack_population(_dosage_matrix(self).genotypes, None, "geno", fp4=True)
            r = plan.r2(pk, histogram=False, sparse_cap=0, per_pair=True, estimator="rogers_huff_r")["r2"]
            vals = r * r
        else:
            pk = B.pack_population(self._haplotypes, rows, "hap", fp4=True)
            vals = plan.r2(pk, histogram=False, sparse_cap=0, per_pair=True, estimator="r2")["r2"]
        keys = L.empty((plan.n_pairs,), np.int32)
        pos_d = L.to_device(pos, np.float64)
        edges_d = L.to_device(bins, np.float64)
        off = 0
        for ii, jj in LP.iter_pairs_within_distance(pos, float(bins[-1]), 10**8):
            n = int(ii.shape[0])
            L.call("pgld_pair_keys", L.ptr(pos_d), L.ptr(ii), L.ptr(jj), n, L.KEY_BIN, L.ptr(edges_d), n_bins, 0.0,
                   1.0, L.ptr(keys[off:off + n]), L.stream_handle())
            off += n
        r2 = vals.cpu().numpy()
        k = keys.cpu().numpy()
        ok = (k >= 0) & ~np.isnan(r2)
        r2, k = r2[ok], k[ok]
        pct = np.atleast_1d(percentile)
        result = np.full((n_bins, len(pct)), np.nan)
        counts = np.bincount(k, minlength=n_bins)[:n_bins].astype(int)
        order = np.argsort(k, kind="stable")
        r2s, ks = r2[order], k[order]
        starts = np.searchsorted(ks, np.arange(n_bins + 1))
        for b in range(n_bins):
            if counts[b] > 0:
                result[b] = np.percentile(r2s[starts[b]:starts[b + 1]], pct)
        if result.shape[1] == 1:
            result = result[:, 0]
        return result, counts

    def windowed_ld_sums(self, window, max_dist=None, pops=None, origin=0.0, n_windows=None,
                         stat_set="moments"):
        """FP64 sums of the ld_names(P) statistics per window of the left SNP
        over every pair within max_dist (default: the window width), one fused
        band pass.  Returns (sums (n_windows, n_stats), pairs per window,
        statistic names)."""
        self.transfer_to_gpu()
        if pops is None:
            pops = list(self.sample_sets)
        rows = [self.sample_sets[p] if isinstance(p, str) else p for p in pops]
        md = float(window if max_dist is None else max_dist)
        sums, kp = B.ld_sums(self._haplotypes, "hap", self._pos_np(), rows, md, window=float(window),
                             n_keys=n_windows, origin=origin, stat_set=stat_set)
        P = len(rows)
        if stat_set == "moments" or P == 1:
            names = LP.ld_names(P)
        else:
            names = [f"DD_{p}_{p}" for p in range(P)] + [f"Dz_{p}_{p}_{p}" for p in range(P)] + \
                [f"pi2_{p}_{p}_{p}_{p}" for p in range(P)]
        return sums, kp, names
