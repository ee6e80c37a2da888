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
