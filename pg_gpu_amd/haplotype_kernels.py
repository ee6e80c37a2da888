"""Batched multi-population haplotype statistics -- the drop-in for
pg_gpu.haplotype_kernels.compute_multi_pop_statistics_batch_hap
(haplotype_kernels.py:584-640).  One libpgld launch evaluates every spec
term of every pair (fused-kernel forms of haplotype_kernels.py:69-304)."""
from __future__ import annotations

from .ld_pipeline import generate_stat_specs, stats_from_counts


def compute_multi_pop_statistics_batch_hap(counts_per_pop, n_valid_per_pop, ld_statistics_module=None,
                                           stat_specs=None):
    """(N, n_stats) float64 weighted spec sums."""
    if stat_specs is None:
        stat_specs = generate_stat_specs(len(counts_per_pop))
    return stats_from_counts("hap", list(counts_per_pop), list(n_valid_per_pop), stat_specs, pi2_form="pi2")
