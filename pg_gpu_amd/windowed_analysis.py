"""Per-window LD columns of pg_gpu's windowed statistics.

``windowed_statistics_fused`` (windowed_analysis.py:1670-2058) computes many
per-window statistics; the LD ones are 'zns' (Kelly's ZnS with the unbiased
sigma_D^2 estimator, windowed_analysis.py:2040-2048 via
ld_statistics._zns_from_precomputed, ld_statistics.py:595-676) and 'omega'
(ld_statistics.omega of the window, :2050-2054).  The reference loops over
windows on the host and launches a tiled pass per window.  Here one MFMA pass
of pgld_band_ld_pairs covers every window at once: windows partition the
variants, so each variant is placed at window_rank * BIG + (pos - window
start) and a band of max_dist BIG / 2 holds exactly the pairs of a window;
the pass returns per-variant sigma_D^2 row / column sums and valid-pair
counts, from which each window's ZnS and Omega follow on the host.  Windows
that overlap (explicit window starts and stops) are dealt into layers of
disjoint windows, one pass per layer.

The other statistics of windowed_statistics_fused (diversity, divergence,
SFS, distance moments) are outside this build and raise NotImplementedError.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from . import band as B
from .ld_statistics import _omega_from_sums

LD_STATISTICS = ("zns", "omega")


def _window_results(chrom, starts_bp, stops_bp, n_variants):
    """The leading columns chrom, start, end, center, n_variants, window_id
    (windowed_analysis.py:66-82)."""
    start = np.asarray(starts_bp).astype(int)
    end = np.asarray(stops_bp).astype(int)
    n = len(start)
    return {
        "chrom": np.repeat(chrom if chrom is not None else 1, n),
        "start": start,
        "end": end,
        "center": (start + end) // 2,
        "n_variants": np.asarray(n_variants).astype(int),
        "window_id": np.arange(n, dtype=int),
    }


def _population_rows(matrix, population):
    """Haplotype rows of a population name or index list (_utils.py:9-40)."""
    if population is None:
        return None
    if isinstance(population, str):
        if matrix.sample_sets is None:
            raise ValueError("No sample_sets defined in haplotype matrix")
        if population not in matrix.sample_sets:
            raise ValueError(f"Population {population} not found in sample_sets")
        return list(matrix.sample_sets[population])
    return list(population)


def _layers(start, stop, windows):
    """Deal the windows (index ranges, sorted by start) into layers of
    pairwise disjoint ranges: each window goes to the first layer whose last
    window ends at or before its start."""
    order = sorted(windows, key=lambda w: (start[w], stop[w]))
    ends, layers = [], []
    for w in order:
        for k, end in enumerate(ends):
            if end <= start[w]:
                layers[k].append(w)
                ends[k] = stop[w]
                break
        else:
            layers.append([w])
            ends.append(stop[w])
    return layers


def _layer_sums(hap_rows, pos, keep, start, stop, layer):
    """sigma_D^2 row / column sums and valid-pair row counts of the kept
    variants of each window of one layer (disjoint windows), in one band
    pass.  Returns {window: (row_sums, col_sums, row_counts)}."""
    cols, coord = [], []
    span = 0.0
    for w in layer:
        idx = np.arange(start[w], stop[w])
        idx = idx[keep[idx]]
        cols.append(idx)
        if idx.size:
            span = max(span, float(pos[idx[-1]] - pos[idx[0]]))
    big = 4.0 * (span + 1.0)
    for rank, (w, idx) in enumerate(zip(layer, cols)):
        if idx.size:
            coord.append(rank * big + (pos[idx] - pos[idx[0]]))
    out = {w: (np.zeros(0), np.zeros(0), np.zeros(0, dtype=np.int64)) for w in layer}
    if not coord:
        return out
    sel = np.concatenate(cols)
    sub = hap_rows[:, L.to_device(sel, np.int64)]
    pk = B.pack_population(sub, None, "hap", fp4=True)
    # same-window pairs differ by at most span < max_dist; pairs of different
    # windows by at least big - span = 3 span + 4 > max_dist
    plan = B.BandPlan(np.concatenate(coord), 2.0 * (span + 1.0), window=1e300, n_keys=1)
    s = plan.r2(pk, histogram=False, sparse_cap=0, estimator="sigma_d2", snp_sums=True)
    rs, cs, rc = (s[k].cpu().numpy() for k in ("row_sums", "col_sums", "row_counts"))
    at = 0  # the selection lists the windows' variants window after window
    for w, idx in zip(layer, cols):
        out[w] = (rs[at:at + idx.size], cs[at:at + idx.size], rc[at:at + idx.size])
        at += idx.size
    return out


def windowed_statistics_fused(haplotype_matrix, bp_bins, statistics=("zns",), population=None, pop1=None,
                              pop2=None, per_base: bool = True, is_accessible=None, _win_starts=None,
                              _win_stops=None, missing_data="include", chrom=None):
    """Per-window ZnS and Omega with the window columns of
    windowed_statistics_fused (windowed_analysis.py:1670-2058).

    Windows are [bp_bins[k], bp_bins[k + 1]) (or the given starts and stops);
    a window's variants are those with start <= position < end.  Windows of
    fewer than 4 variants get NaN.  'zns' is the mean unbiased sigma_D^2 over
    the valid pairs of the window's segregating sites (0.0 below 2 sites or
    without a valid pair; missing_data does not apply, as in the reference);
    'omega' is Kim and Nielsen's Omega of the window's sigma_D^2 matrix over
    its segregating sites (sites with missing data dropped first when
    missing_data='exclude'; 0.0 below 5 sites).  per_base and is_accessible
    only normalise the diversity statistics and are accepted unused."""
    statistics = tuple(statistics)
    other = [s for s in statistics if s not in LD_STATISTICS]
    if other:
        raise NotImplementedError(f"windowed statistics {other} are outside this build's LD path "
                                  f"(supported: {LD_STATISTICS})")
    if missing_data not in ("include", "exclude"):
        raise ValueError("missing_data must be 'include' or 'exclude'")
    matrix = haplotype_matrix
    matrix.transfer_to_gpu()
    pos = np.asarray(L.to_numpy(matrix.positions), dtype=np.float64)
    if _win_starts is not None and _win_stops is not None:
        ws = np.asarray(_win_starts, dtype=np.float64)
        we = np.asarray(_win_stops, dtype=np.float64)
    else:
        bins = np.asarray(bp_bins, dtype=np.float64)
        ws, we = bins[:-1], bins[1:]
    start = np.searchsorted(pos, ws, side="left")
    stop = np.searchsorted(pos, we, side="left")
    results = _window_results(chrom, ws, we, stop - start)
    if not statistics:
        return results
    rows = _population_rows(matrix, population)
    hap = L.to_device(matrix.haplotypes, np.int8)
    if rows is not None:
        hap = hap[L.to_device(np.asarray(rows, dtype=np.int64), np.int64)]
    alt, valid = (x.cpu().numpy() for x in matrix.allele_counts(rows))
    n_rows = int(hap.shape[0])
    seg = (alt > 0) & (alt < valid)
    windows = [w for w in range(len(start)) if stop[w] - start[w] >= 4]
    layers = _layers(start, stop, windows)
    keep_om = seg & (valid == n_rows) if missing_data == "exclude" else seg
    zns = np.full(len(start), np.nan)
    om = np.full(len(start), np.nan)
    for layer in layers:
        sums = _layer_sums(hap, pos, seg, start, stop, layer)
        for w in layer:
            rs, cs, rc = sums[w]
            n = int(rc.sum())
            zns[w] = float(rs.sum() / n) if rs.size >= 2 and n else 0.0
            om[w] = _omega_from_sums(rs, cs) if rs.size >= 5 else 0.0
        if "omega" in statistics and not np.array_equal(keep_om, seg):
            sums = _layer_sums(hap, pos, keep_om, start, stop, layer)
            for w in layer:
                rs, cs, _ = sums[w]
                om[w] = _omega_from_sums(rs, cs) if rs.size >= 5 else 0.0
    if "zns" in statistics:
        results["zns"] = zns
    if "omega" in statistics:
        results["omega"] = om
    return results
