"""Batched multi-population genotype statistics -- the drop-in for
pg_gpu.genotype_kernels.compute_multi_pop_statistics_batch_geno
(genotype_kernels.py:881-936).  Formulas of genotype_kernels.py:27-492,
evaluated per pair from exact 9-way counts in one libpgld launch."""
from __future__ import annotations

from .ld_pipeline import generate_stat_specs, stats_from_counts


def compute_multi_pop_statistics_batch_geno(counts_per_pop, n_valid_per_pop, _unused=None, stat_specs=None):
    """(N, n_stats) float64 weighted spec sums (counts in n00..n22 order)."""
    if stat_specs is None:
        stat_specs = generate_stat_specs(len(counts_per_pop))
    return stats_from_counts("geno", list(counts_per_pop), list(n_valid_per_pop), stat_specs)
