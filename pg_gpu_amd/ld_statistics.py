"""LD statistics from haplotype counts -- the drop-in for pg_gpu.ld_statistics
(ld_statistics.py:13-246, :1030-1675).

Inputs are count arrays [n11, n10, n01, n00] per population (torch GPU
tensors or anything array-like); outputs are float64 torch GPU tensors.  All
arithmetic runs in libpgld (pgld_stats_pairs / pgld_r_family).
"""
from __future__ import annotations

import warnings
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

from . import _lib as L
from .ld_pipeline import stats_from_counts


def _split(counts, n_valid, pops_needed):
    """Per-population (N, 4) count tensors and n_valid tensors (or None),
    following _get_pop_data's n_valid conventions (ld_statistics.py:1087-1113)."""
    c = L.to_device(counts, np.int32)
    P = c.shape[1] // 4
    if c.shape[1] % 4 or P < 1:
        raise ValueError("counts must have 4 * n_pops columns")
    if max(pops_needed) >= P:
        raise IndexError("population index beyond the counts' columns")
    cl = [c[:, 4 * p:4 * p + 4].contiguous() for p in range(P)]
    if n_valid is None:
        nl = [None] * P
    elif isinstance(n_valid, tuple):
        nl = [n_valid[p] if p < len(n_valid) else None for p in range(P)]
    else:
        nv = L.to_device(n_valid, np.int32)
        if nv.dim() == 2:
            nl = [nv[:, p].contiguous() for p in range(P)]
        else:
            nl = [nv] * P
    return cl, nl


def _one(kind, counts, populations, n_valid, pi2_form="pi2_sym"):
    cl, nl = _split(counts, n_valid, list(populations))
    spec = [("x", [(1.0, kind, tuple(populations))])]
    out = stats_from_counts("hap", cl, nl, spec, pi2_form=pi2_form)
    return out[:, 0]


def dd(counts, populations: Optional[Union[Tuple[int, int], int]] = None, n_valid=None):
    """D^2: single population (None) or between populations (i, j)
    (ld_statistics.py:13-65)."""
    if populations is None:
        if n_valid is not None and L.is_tensor(n_valid) and n_valid.dim() == 2:
            n_valid = n_valid[:, 0]
        return _one("dd", counts, (0, 0), n_valid)
    a, b = populations
    return _one("dd", counts, (a, b), n_valid)


def dz(counts, populations: Optional[Tuple[int, int, int]] = None, n_valid=None):
    """Dz(i, j, k) (ld_statistics.py:68-97); None = single population."""
    if populations is None:
        populations = (0, 0, 0)
    return _one("dz", counts, tuple(populations), n_valid)


def pi2(counts, populations: Optional[Tuple[int, int, int, int]] = None, n_valid=None):
    """pi2(i, j, k, l) (ld_statistics.py:100-129); None = single population."""
    if populations is None:
        populations = (0, 0, 0, 0)
    return _one("pi2", counts, tuple(populations), n_valid)


def dd_within(counts, n_valid=None):
    return dd(counts, None, n_valid)


def dd_between(counts, pop1_idx: int, pop2_idx: int, n_valid=None):
    return dd(counts, (pop1_idx, pop2_idx), n_valid)


def _r_family(counts, n_valid, which):
    c = L.to_device(counts, np.int32)
    if c.dim() != 2 or c.shape[1] != 4:
        raise ValueError("counts must have shape (N, 4)")
    nv = None if n_valid is None else L.to_device(n_valid, np.int32)
    out = L.empty((c.shape[0],), np.float64)
    L.call("pgld_r_family", L.ptr(c), L.ptr(nv), c.shape[0], which, L.ptr(out), L.stream_handle())
    return out


def r(counts, n_valid=None):
    """Pearson r per pair; NaN where a locus is monomorphic (ld_statistics.py:167-191)."""
    return _r_family(counts, n_valid, 0)


def r_squared(counts, n_valid=None):
    """r^2 per pair; NaN where undefined (ld_statistics.py:194-217)."""
    return _r_family(counts, n_valid, 1)


def d_prime(counts, n_valid=None):
    """Lewontin's D' per pair; NaN where D_max is zero (ld_statistics.py:220-246)."""
    return _r_family(counts, n_valid, 2)


def compute_ld_statistics(counts, statistics: List[str] = ("dd", "dz", "pi2"),
                          populations: Optional[Dict[str, Union[Tuple, None]]] = None,
                          n_valid=None) -> Dict[str, object]:
    """Several statistics in one call (ld_statistics.py:1030-1076)."""
    populations = populations or {}
    out = {}
    for stat in statistics:
        if stat == "dd":
            out["dd"] = dd(counts, populations.get("dd"), n_valid)
        elif stat == "dz":
            out["dz"] = dz(counts, populations.get("dz"), n_valid)
        elif stat == "pi2":
            out["pi2"] = pi2(counts, populations.get("pi2"), n_valid)
        elif stat == "r":
            out["r"] = r(counts, n_valid)
        elif stat == "r_squared":
            out["r_squared"] = r_squared(counts, n_valid)
        else:
            raise ValueError(f"Unknown statistic: {stat}")
    return out


# Deprecated aliases (ld_statistics.py:1621-1675)
def DD(counts, n_valid=None):
    warnings.warn("DD() is deprecated. Use ld_statistics.dd() instead.", DeprecationWarning, stacklevel=2)
    return dd_within(counts, n_valid)


def DD_two_pops(counts, pop1_idx, pop2_idx, n_valid1=None, n_valid2=None):
    warnings.warn("DD_two_pops() is deprecated.", DeprecationWarning, stacklevel=2)
    nv = (n_valid1, n_valid2) if n_valid1 is not None and n_valid2 is not None else None
    return dd(counts, populations=(pop1_idx, pop2_idx), n_valid=nv)


def Dz_two_pops(counts, pop_indices, n_valid1=None, n_valid2=None):
    warnings.warn("Dz_two_pops() is deprecated.", DeprecationWarning, stacklevel=2)
    nv = (n_valid1, n_valid2) if n_valid1 is not None and n_valid2 is not None else None
    return dz(counts, populations=pop_indices, n_valid=nv)


def pi2_two_pops(counts, pop_indices, n_valid1=None, n_valid2=None):
    warnings.warn("pi2_two_pops() is deprecated.", DeprecationWarning, stacklevel=2)
    nv = (n_valid1, n_valid2) if n_valid1 is not None and n_valid2 is not None else None
    return pi2(counts, populations=pop_indices, n_valid=nv)


# ---------------------------------------------------------------------------
# ZnS and Omega (ld_statistics.py:534-862)
# ---------------------------------------------------------------------------

def _resolve_ld_estimator(estimator, is_hap_matrix):
    """'auto' -> 'sigma_d2' for a HaplotypeMatrix, 'rogers_huff' otherwise
    (ld_statistics.py:344-363)."""
    if estimator == "auto":
        return "sigma_d2" if is_hap_matrix else "rogers_huff"
    if estimator not in ("r2", "sigma_d2", "rogers_huff"):
        raise ValueError(f"Unknown estimator: {estimator!r} (expected one of 'auto', 'r2', 'sigma_d2', "
                         "'rogers_huff')")
    return estimator


def _segregating(mat, missing_data):
    """Sites kept for ZnS / Omega: optionally only sites without missing data,
    then sites with 0 < derived count < valid count (ld_statistics.py:249-278)."""
    mat.transfer_to_gpu()
    if missing_data == "exclude":
        alt, valid = (x.cpu().numpy() for x in mat.allele_counts())
        mat = mat.get_subset(np.nonzero(valid == mat.num_haplotypes)[0])
    if mat.num_variants == 0:
        return mat
    alt, valid = (x.cpu().numpy() for x in mat.allele_counts())
    seg = np.nonzero((alt > 0) & (alt < valid))[0]
    if len(seg) < mat.num_variants:
        mat = mat.get_subset(seg)
    return mat


def _genotype_sites(gm, missing_data):
    """GenotypeMatrix sites for ZnS / Omega: all sites, or with
    missing_data='exclude' those without missing genotypes
    (ld_statistics.py:940-951); no segregating filter, as the reference."""
    gm.transfer_to_gpu()
    if missing_data == "exclude":
        alt, valid = (x.cpu().numpy() for x in gm.allele_counts())
        gm = gm.get_subset(np.nonzero(valid == gm.num_individuals)[0])
    return gm


def _snp_sums(mat, estimator):
    """Per-SNP row / column sums of the estimator over all pairs of the
    matrix (one all-pairs band, MFMA tiles)."""
    from . import band as B
    pos = np.arange(mat.num_variants, dtype=np.float64)  # all pairs: positions play no part
    plan = B.BandPlan(pos, np.inf, window=1e300, n_keys=1)
    pk = B.pack_population(mat.haplotypes, None, "hap", fp4=True)
    out = plan.r2(pk, histogram=False, sparse_cap=0, estimator=estimator, snp_sums=True)
    return {k: out[k].cpu().numpy() for k in ("row_sums", "col_sums", "row_counts", "col_counts")}


def zns(r2_matrix_or_matrix, missing_data="include", estimator="auto"):
    """Kelly's ZnS, the mean pairwise LD over all SNP pairs
    (ld_statistics.py:679-731).  For a HaplotypeMatrix the pairs are never
    materialised: one pass of the MFMA tiles yields per-SNP sums.
    estimator 'sigma_d2' (the 'auto' choice) averages the unbiased
    D^2 / pi_2 over pairs with n >= 4 of all segregating sites (missing_data
    does not apply: the reference maps this estimator to its internal
    'project' mode, ld_statistics.py:721-726); 'r2' -- and 'rogers_huff',
    which the reference routes to the same naive path for haplotypes --
    averages the frequency r^2 over all pairs.  A precomputed square r^2
    array is averaged off the diagonal."""
    from .genotype_matrix import GenotypeMatrix
    from .haplotype_matrix import HaplotypeMatrix
    is_hm = isinstance(r2_matrix_or_matrix, HaplotypeMatrix)
    est = _resolve_ld_estimator(estimator, is_hm)
    if is_hm:
        mat = _segregating(r2_matrix_or_matrix, "include" if est == "sigma_d2" else missing_data)
        m = mat.num_variants
        if m < 2:
            return 0.0
        s = _snp_sums(mat, "sigma_d2" if est == "sigma_d2" else "naive")
        if est == "sigma_d2":
            n = int(s["row_counts"].sum())
            return float(s["row_sums"].sum() / n) if n else 0.0
        return float(2.0 * s["row_sums"].sum() / (m * (m - 1)))
    if isinstance(r2_matrix_or_matrix, GenotypeMatrix):
        if est == "sigma_d2":
            raise ValueError("estimator='sigma_d2' requires a HaplotypeMatrix, not a pre-computed r² array")
        gm = _genotype_sites(r2_matrix_or_matrix, missing_data)
        m = gm.num_variants
        if m < 2:
            return 0.0
        s = _all_pairs_values(gm, "dosage_r2", per_pair=False, snp_sums=True)
        return float(2.0 * s["row_sums"].sum().item() / (m * (m - 1)))
    if est == "sigma_d2":
        raise ValueError("estimator='sigma_d2' requires a HaplotypeMatrix, not a pre-computed r² array")
    M = np.asarray(L.to_numpy(r2_matrix_or_matrix), dtype=np.float64)
    m = M.shape[0]
    if m < 2:
        return 0.0
    return float((M.sum() - np.trace(M)) / (m * (m - 1)))


def _omega_from_sums(row_sums, col_sums):
    """Omega from per-SNP upper-triangle sums: for a split l, the left block
    sum is the prefix of column sums below l, the cross block the prefix of
    row sums below l minus the left block (ld_statistics.py:808-862)."""
    m = len(row_sums)
    if m < 5:
        return 0.0
    left_c = np.concatenate([[0.0], np.cumsum(col_sums)])
    rows_c = np.concatenate([[0.0], np.cumsum(row_sums)])
    total = rows_c[-1]
    l = np.arange(3, m - 1)
    left = left_c[l]
    cross = rows_c[l] - left
    right = total - left - cross
    n_left = l * (l - 1) // 2
    n_right = (m - l) * (m - l - 1) // 2
    n_cross = l * (m - l)
    n_w = n_left + n_right
    ok = (n_w > 0) & (n_cross > 0) & (cross > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        vals = np.where(ok, ((left + right) / np.where(n_w > 0, n_w, 1)) /
                        np.where(ok, cross / np.where(n_cross > 0, n_cross, 1), 1.0), 0.0)
    return float(vals.max()) if len(vals) else 0.0


def omega(r2_matrix_or_matrix, missing_data="include", estimator="auto"):
    """Kim and Nielsen's Omega: the largest ratio of mean LD within the two
    blocks of a split to mean LD across it, over the upper triangle
    (ld_statistics.py:761-862).  For a HaplotypeMatrix one pass of the MFMA
    tiles yields the per-SNP sums the split ratios need; 'rogers_huff' on a
    HaplotypeMatrix takes the frequency r^2 of pairwise_r2, as the reference
    (_resolve_r2_matrix, ld_statistics.py:926-970)."""
    from .genotype_matrix import GenotypeMatrix
    from .haplotype_matrix import HaplotypeMatrix
    is_hm = isinstance(r2_matrix_or_matrix, HaplotypeMatrix)
    est = _resolve_ld_estimator(estimator, is_hm)
    if is_hm:
        mat = _segregating(r2_matrix_or_matrix, missing_data)
        if mat.num_variants < 5:
            return 0.0
        s = _snp_sums(mat, "sigma_d2" if est == "sigma_d2" else "naive")
        return _omega_from_sums(s["row_sums"], s["col_sums"])
    if isinstance(r2_matrix_or_matrix, GenotypeMatrix):
        if est == "sigma_d2":
            raise ValueError("estimator='sigma_d2' requires a HaplotypeMatrix")
        gm = _genotype_sites(r2_matrix_or_matrix, missing_data)
        if gm.num_variants < 5:
            return 0.0
        s = _all_pairs_values(gm, "dosage_r2", per_pair=False, snp_sums=True)
        return _omega_from_sums(s["row_sums"].cpu().numpy(), s["col_sums"].cpu().numpy())
    if est == "sigma_d2":
        raise ValueError("estimator='sigma_d2' requires a HaplotypeMatrix")
    M = np.triu(np.asarray(L.to_numpy(r2_matrix_or_matrix), dtype=np.float64), k=1)
    return _omega_from_sums(M.sum(axis=1), M.sum(axis=0))


# ---------------------------------------------------------------------------
# Rogers-Huff r of diploid dosages (ld_statistics.py:366-531)
# ---------------------------------------------------------------------------

def _dosage_matrix(matrix):
    """GenotypeMatrix of 0/1/2 dosages: a HaplotypeMatrix is paired by
    adjacent rows (haplotypes 2s, 2s + 1); missing values and an odd
    haplotype count raise ValueError, other inputs TypeError
    (ld_statistics.py:366-407)."""
    from .genotype_matrix import GenotypeMatrix
    from .haplotype_matrix import HaplotypeMatrix
    if isinstance(matrix, HaplotypeMatrix):
        matrix.transfer_to_gpu()
        if matrix.count_missing() > 0:
            raise ValueError("rogers_huff_r: input HaplotypeMatrix contains missing values (-1). Rogers-Huff r "
                             "expects strict 0/1/2 dosage input; drop or impute missing sites first.")
        if matrix.num_haplotypes % 2 != 0:
            raise ValueError(f"rogers_huff_r: HaplotypeMatrix has an odd number of haplotypes "
                             f"({matrix.num_haplotypes}); cannot pair into diploids.")
        return GenotypeMatrix.from_haplotype_matrix(matrix)
    if isinstance(matrix, GenotypeMatrix):
        matrix.transfer_to_gpu()
        if bool((matrix.genotypes < 0).any()):
            raise ValueError("rogers_huff_r: input GenotypeMatrix contains missing values (-1). Rogers-Huff r "
                             "expects strict 0/1/2 dosage input; drop or impute missing sites first.")
        return matrix
    raise TypeError(f"rogers_huff_r: expected HaplotypeMatrix or GenotypeMatrix; got {type(matrix).__name__}")


def _all_pairs_values(gm, estimator, per_pair=True, snp_sums=False):
    from . import band as B
    m = gm.num_variants
    plan = B.BandPlan(np.arange(m, dtype=np.float64), np.inf, window=1e300, n_keys=1)
    pk = B.pack_population(gm.genotypes, None, "geno", fp4=True)
    return plan.r2(pk, histogram=False, sparse_cap=0, per_pair=per_pair, estimator=estimator, snp_sums=snp_sums)


def rogers_huff_r(matrix, tile_size: Optional[int] = None):
    """Signed Rogers-Huff (2008) r of every variant pair, condensed upper
    triangle (0,1), (0,2), ..., (n-2,n-1); NaN where either site is constant
    (ld_statistics.py:492-521).  The dosage products come from the MFMA tiles
    on genotype planes and r from exact integer sums; tile_size is accepted
    for signature compatibility."""
    gm = _dosage_matrix(matrix)
    if gm.num_variants < 2:
        return L.torch().empty((0,), dtype=L.torch().float64, device=L.device())
    return _all_pairs_values(gm, "rogers_huff_r")["r2"]


def rogers_huff_r_squared(matrix, tile_size: Optional[int] = None):
    """rogers_huff_r(matrix) ** 2 (ld_statistics.py:524-531)."""
    r = rogers_huff_r(matrix, tile_size)
    return r * r
