"""pg_gpu_amd: MI355X-native (gfx950) LD pair statistics with the call
surface of kr-colab/pg_gpu's LD path.

Compute runs in libpgld.so (hand-written HIP kernels behind a C ABI, loaded
with ctypes); torch provides device memory and streams.  Importing the
package does not touch the GPU; the first compute call loads the library and
raises if it or the GPU is missing.
"""
from . import ld_pipeline, ld_statistics, moments_ld, windowed_analysis  # noqa: F401
from .genotype_matrix import GenotypeMatrix  # noqa: F401
from .haplotype_matrix import HaplotypeMatrix  # noqa: F401
from .ld_pipeline import (  # noqa: F401
    compute_counts_for_pairs, compute_genotype_counts_for_pairs, estimate_ld_chunk_size,
    generate_stat_specs, het_names, iter_pairs_within_distance, ld_names,
)

__version__ = "0.1.0"

__all__ = [
    "HaplotypeMatrix", "GenotypeMatrix", "ld_pipeline", "ld_statistics", "moments_ld", "windowed_analysis",
    "compute_counts_for_pairs", "compute_genotype_counts_for_pairs", "estimate_ld_chunk_size",
    "generate_stat_specs", "het_names", "iter_pairs_within_distance", "ld_names",
]
