"""GenotypeMatrix -- the LD-path part of pg_gpu.genotype_matrix
(genotype_matrix.py:15-390, :506-539).

Data: int8 (n_individuals, n_variants) dosages {-1 missing, 0, 1, 2} on the
host (NumPy) or the GPU (torch).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from . import band as B
from . import ld_pipeline as LP
from . import vcf as V


class GenotypeMatrix:
    """Diploid alt-allele dosages with positions and population labels."""

    def __init__(self, genotypes, positions, chrom_start=None, chrom_end=None, sample_sets=None,
                 n_total_sites=None, samples=None, accessible_mask=None):
        if accessible_mask is not None:
            raise NotImplementedError("accessible-site masks are outside this build's LD path")
        if not (isinstance(genotypes, np.ndarray) or L.is_tensor(genotypes)):
            raise ValueError("genotypes must be a numpy array or a torch tensor")
        if (genotypes.numel() if L.is_tensor(genotypes) else genotypes.size) == 0:
            raise ValueError("genotypes cannot be empty")
        if (positions.numel() if L.is_tensor(positions) else np.size(positions)) == 0:
            raise ValueError("positions cannot be empty")
        if L.is_tensor(genotypes):
            self._device = "GPU"
            self._genotypes = genotypes.to(L.torch().int8)
            self._positions = L.to_device(positions, np.int64)
        else:
            self._device = "CPU"
            self._genotypes = np.asarray(genotypes).astype(np.int8, copy=False)
            self._positions = np.asarray(L.to_numpy(positions))
        self.chrom_start, self.chrom_end = chrom_start, chrom_end
        self._sample_sets = sample_sets
        self.n_total_sites = n_total_sites
        self.samples = samples

    @property
    def genotypes(self):
        return self._genotypes

    @genotypes.setter
    def genotypes(self, value):
        self._genotypes = value

    @property
    def positions(self):
        return self._positions

    @property
    def device(self):
        return self._device

    @property
    def sample_sets(self):
        if self._sample_sets is None:
            return {"all": list(range(self.num_individuals))}
        return self._sample_sets

    @sample_sets.setter
    def sample_sets(self, value):
        if not isinstance(value, dict):
            raise ValueError("sample_sets must be a dictionary")
        self._sample_sets = value

    @property
    def num_individuals(self):
        return int(self._genotypes.shape[0])

    @property
    def num_variants(self):
        return int(self._genotypes.shape[1])

    @property
    def shape(self):
        return tuple(self._genotypes.shape)

    def transfer_to_gpu(self):
        if self._device == "CPU":
            self._genotypes = L.to_device(self._genotypes, np.int8)
            self._positions = L.to_device(self._positions, np.int64)
            self._device = "GPU"

    def transfer_to_cpu(self):
        if self._device == "GPU":
            self._genotypes = L.to_numpy(self._genotypes)
            self._positions = L.to_numpy(self._positions)
            self._device = "CPU"

    def _pos_np(self):
        return np.asarray(L.to_numpy(self._positions))

    @classmethod
    def from_haplotype_matrix(cls, hap_matrix):
        """Pair ADJACENT haplotype rows (0,1), (2,3), ... into dosages; a
        missing allele makes the genotype missing (genotype_matrix.py:248-305).
        Note the reference's VCF loader stores sample s at rows s and
        s + n_samples, so this pairing differs from the VCF sample layout;
        it is kept as the reference has it."""
        hap = hap_matrix.haplotypes
        n_hap = int(hap.shape[0])
        if n_hap % 2 != 0:
            raise ValueError(f"Need even number of haplotypes for diploid conversion, got {n_hap}")
        h = L.to_numpy(hap).astype(np.int16)
        h1, h2 = h[0::2], h[1::2]
        geno = (np.maximum(h1, 0) + np.maximum(h2, 0)).astype(np.int8)
        geno[(h1 < 0) | (h2 < 0)] = -1
        sets = None
        if hap_matrix._sample_sets is not None:
            sets = {k: sorted(set(i // 2 for i in v)) for k, v in hap_matrix._sample_sets.items()}
        gm = cls(geno, L.to_numpy(hap_matrix.positions), hap_matrix.chrom_start, hap_matrix.chrom_end,
                 sample_sets=sets, n_total_sites=hap_matrix.n_total_sites)
        if hap_matrix.device == "GPU":
            gm.transfer_to_gpu()
        return gm

    def to_haplotype_matrix(self):
        """Expand each individual to two adjacent haplotypes, hets as (1, 0)
        (genotype_matrix.py:307-341)."""
        from .haplotype_matrix import HaplotypeMatrix
        g = L.to_numpy(self._genotypes)
        hap = np.zeros((2 * g.shape[0], g.shape[1]), dtype=np.int8)
        miss = g < 0
        gg = np.maximum(g, 0)
        hap[0::2] = (gg >= 1).astype(np.int8)
        hap[1::2] = (gg >= 2).astype(np.int8)
        hap[0::2][miss] = -1
        hap[1::2][miss] = -1
        return HaplotypeMatrix(hap, self._pos_np(), self.chrom_start, self.chrom_end,
                               n_total_sites=self.n_total_sites)

    @classmethod
    def from_vcf(cls, path, include_invariant=False, accessible_bed=None):
        """Dosages of biallelic (0/1 only, both alleles seen) sites; a missing
        allele makes the genotype missing (genotype_matrix.py:343-390)."""
        if accessible_bed is not None:
            raise NotImplementedError("accessible-site masks are outside this build's LD path")
        d = V.read_vcf(path)
        gt, pos = d["gt"].astype(np.int16), d["positions"]
        called = gt >= 0
        n_alt = ((gt == 1) & called).sum(axis=(1, 2))
        n_ref = ((gt == 0) & called).sum(axis=(1, 2))
        n_other = ((gt > 1) & called).sum(axis=(1, 2))
        keep = (n_other == 0) & (n_alt > 0) & (n_ref > 0)
        gt, pos = gt[keep], pos[keep]
        geno = gt.sum(axis=2).astype(np.int8)
        geno[np.any(gt < 0, axis=2)] = -1
        geno = geno.T.copy()
        return cls(geno, pos, chrom_start=int(pos[0]), chrom_end=int(pos[-1]),
                   n_total_sites=geno.shape[1] if include_invariant else None, samples=d["samples"])

    def load_pop_file(self, pop_file, pops=None):
        if self.samples is None:
            raise ValueError("No sample names stored. Use from_vcf() to load data.")
        pop_map = V.read_pop_file(pop_file)
        if pops is None:
            pops = sorted(set(pop_map.values()))
        sets = {p: [] for p in pops}
        for i, name in enumerate(self.samples):
            p = pop_map.get(name)
            if p in sets:
                sets[p].append(i)
        self._sample_sets = sets

    def get_subset(self, idx):
        idx = np.asarray(L.to_numpy(idx), dtype=np.int64)
        if self._device == "GPU":
            it = L.to_device(idx, np.int64)
            g, p = self._genotypes[:, it], self._positions[it]
        else:
            g, p = self._genotypes[:, idx], self._positions[idx]
        out = GenotypeMatrix.__new__(GenotypeMatrix)
        out._device, out._genotypes, out._positions = self._device, g, p
        out.chrom_start, out.chrom_end = self.chrom_start, self.chrom_end
        out._sample_sets, out.n_total_sites, out.samples = self._sample_sets, self.n_total_sites, self.samples
        return out

    def allele_counts(self, rows=None):
        """(alt dosage sum, valid individuals) int32 per variant (GPU)."""
        g = L.to_device(self._genotypes, np.int8)
        rows_t, n_sel = LP._rows_tensor(rows, g.shape[0])
        alt = L.empty((g.shape[1],), np.int32)
        valid = L.empty((g.shape[1],), np.int32)
        L.call("pgld_allele_counts", L.ptr(g), g.shape[0], g.shape[1], L.ptr(rows_t), n_sel, L.ptr(alt),
               L.ptr(valid), L.stream_handle())
        return alt, valid

    def apply_biallelic_filter(self):
        """Keep sites with 0 < alt < 2 n_valid and n_valid >= 2
        (genotype_matrix.py:506-539)."""
        alt, valid = (a.cpu().numpy() for a in self.allele_counts())
        keep = (alt > 0) & (alt < 2 * valid) & (valid >= 2)
        return self.get_subset(np.nonzero(keep)[0])

    def ld_bin_sums(self, bp_bins, pops=None):
        """FP64 sums of the genotype ld_names(P) statistics per distance bin
        [b_k, b_{k+1}) over pairs within the last edge (one fused band pass);
        returns (sums, pairs per bin, names)."""
        self.transfer_to_gpu()
        if pops is None:
            pops = list(self.sample_sets)
        rows = [self.sample_sets[p] if isinstance(p, str) else p for p in pops]
        bins = np.asarray(bp_bins, dtype=np.float64)
        sums, kp = B.ld_sums(self._genotypes, "geno", self._pos_np(), rows, float(bins[-1]), edges=bins)
        return sums, kp, LP.ld_names(len(rows))
