"""Genotype LD statistics -- the public part of pg_gpu.ld_statistics_genotype
(sigma_d2_geno, ld_statistics_genotype.py:1518-1598) plus per-pair
single-population DD / Dz / pi2 from 9-way counts."""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .ld_pipeline import compute_genotype_counts_for_pairs, stats_from_counts


_ARITY = {"dd": 2, "dz": 3, "pi2": 4, "sigma_d2": 1}


def _one(counts, kind):
    spec = [("x", [(1.0, kind, (0,) * _ARITY[kind])])]
    return stats_from_counts("geno", [counts], [None], spec)[:, 0]


def dd_geno(counts):
    """Single-population genotype DD per pair (genotype_kernels.py:27-71)."""
    return _one(counts, "dd")


def dz_geno(counts):
    """Single-population genotype Dz per pair (genotype_kernels.py:235-272)."""
    return _one(counts, "dz")


def pi2_geno(counts):
    """Single-population genotype pi2 per pair (genotype_kernels.py:364-405)."""
    return _one(counts, "pi2")


def sigma_d2_geno(matrix, idx_i=None, idx_j=None):
    """Per-pair unbiased sigma_D^2 = DD / pi2 of diploid genotype counts; NaN
    where pi2 <= 0 or fewer than 4 individuals.  Raises ValueError on missing
    data and TypeError on other inputs (ld_statistics_genotype.py:1518-1598)."""
    from .genotype_matrix import GenotypeMatrix
    from .haplotype_matrix import HaplotypeMatrix
    if isinstance(matrix, GenotypeMatrix):
        gm = matrix
    elif isinstance(matrix, HaplotypeMatrix):
        gm = GenotypeMatrix.from_haplotype_matrix(matrix)
    else:
        raise TypeError(f"matrix must be a GenotypeMatrix or HaplotypeMatrix, got {type(matrix).__name__}")
    g = L.to_device(gm.genotypes, np.int8)
    m = g.shape[1]
    from .ld_pipeline import scan_rows
    has_missing, _ = scan_rows(g, None, g.shape[0], 2)
    if has_missing:
        raise ValueError("sigma_d2_geno does not support missing values (-1); filter or impute before calling")
    if idx_i is None and idx_j is None:
        iu = np.triu_indices(m, k=1)
        idx_i, idx_j = iu[0].astype(np.int32), iu[1].astype(np.int32)
    elif idx_i is None or idx_j is None:
        raise ValueError("idx_i and idx_j must be provided together")
    else:
        idx_i = L.to_numpy(idx_i)
        idx_j = L.to_numpy(idx_j)
        if np.shape(idx_i) != np.shape(idx_j):
            raise ValueError("idx_i and idx_j must have the same shape")
    counts, _ = compute_genotype_counts_for_pairs(g, idx_i, idx_j)
    return _one(counts, "sigma_d2")
