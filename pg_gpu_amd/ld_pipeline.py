"""LD pipeline: pair enumeration, per-pair counts, statistic names/specs and
batch statistics -- the drop-in for pg_gpu.ld_pipeline (ld_pipeline.py:1-482).

Arrays live on the GPU as torch tensors; every computation is a libpgld call.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


# ---------------------------------------------------------------------------
# Memory estimation (ld_pipeline.py:16-49)
# ---------------------------------------------------------------------------

def estimate_ld_chunk_size(n_haplotypes_per_pop, available_memory_bytes=None, num_pops=2):
    """Pairs per chunk = clamp(free * 0.5 / ((4 H P + 150 P) * 3), 1e5, 1e7)."""
    if available_memory_bytes is None:
        L.require_gpu()
        available_memory_bytes = int(L.torch().cuda.mem_get_info()[0] * 0.5)
    bytes_per_pair = (4 * n_haplotypes_per_pop * num_pops + 150 * num_pops) * 3
    chunk_size = available_memory_bytes // bytes_per_pair
    return max(100_000, min(chunk_size, 10_000_000))


# ---------------------------------------------------------------------------
# Pair generation (ld_pipeline.py:57-122)
# ---------------------------------------------------------------------------

def pairs_per_variant(positions, max_dist):
    """Device int64 tensor: #{j > v : pos[j] - pos[v] <= max_dist}."""
    pos = L.to_device(positions, np.float64)
    n = pos.shape[0]
    per = L.empty((n,), np.int64)
    L.call("pgld_pairs_per_variant", L.ptr(pos), n, float(max_dist), L.ptr(per), L.stream_handle())
    return per


def iter_pairs_within_distance(positions, max_dist, chunk_size):
    """Yield GPU chunks (idx_i, idx_j) of int32 pairs with i < j and
    pos[j] - pos[i] <= max_dist, ordered by (i, j).  Chunks are cut at anchor
    boundaries exactly as the reference does, so the chunk sequence (not only
    the pair set) matches."""
    n = len(positions)
    if n < 2:
        return
    per_d = pairs_per_variant(positions, max_dist)
    per = per_d.cpu().numpy()
    cumulative = np.cumsum(per)
    total = int(cumulative[-1])
    if total == 0:
        return
    chunk_size = max(1, int(chunk_size))
    n_chunks = (total + chunk_size - 1) // chunk_size
    targets = np.minimum(np.arange(1, n_chunks + 1, dtype=np.int64) * chunk_size, total)
    v_his = np.minimum(np.searchsorted(cumulative, targets, side="left") + 1, n)
    v_lo = 0
    for v_hi in v_his:
        v_hi = int(v_hi)
        if v_hi <= v_lo:
            continue
        counts_slice = per[v_lo:v_hi]
        n_pairs = int(counts_slice.sum())
        if n_pairs == 0:
            v_lo = v_hi
            continue
        excl = np.cumsum(counts_slice) - counts_slice
        excl_d = L.to_device(excl, np.int64)
        ii = L.empty((n_pairs,), np.int32)
        jj = L.empty((n_pairs,), np.int32)
        L.call("pgld_pairs_fill", L.ptr(excl_d), v_lo, v_hi - v_lo, n_pairs, L.ptr(ii), L.ptr(jj),
               L.stream_handle())
        yield ii, jj
        v_lo = v_hi


# ---------------------------------------------------------------------------
# Counts (ld_pipeline.py:130-263)
# ---------------------------------------------------------------------------

def _rows_tensor(pop_indices, n_rows):
    if pop_indices is None:
        return None, n_rows
    rows = np.asarray(L.to_numpy(pop_indices), dtype=np.int64)
    if rows.size and (rows.min() < 0 or rows.max() >= n_rows):
        raise IndexError("population row index out of range")
    return L.to_device(rows, np.int32), int(rows.size)


def scan_rows(data, rows_t, n_sel, max_value):
    """(has_missing, out_of_domain) for the selected rows (one small sync)."""
    flags = L.zeros((2,), np.int32)
    L.call("pgld_scan_rows", L.ptr(data), data.shape[0], data.shape[1], L.ptr(rows_t), n_sel, max_value,
           L.ptr(flags), L.stream_handle())
    f = flags.cpu().numpy()
    return bool(f[0]), bool(f[1])


def _counts(kind, data, idx_i, idx_j, pop_indices):
    g = L.to_device(data, np.int8)
    ii = L.to_device(idx_i, np.int32)
    jj = L.to_device(idx_j, np.int32)
    if ii.shape != jj.shape:
        raise ValueError("idx_i and idx_j must have the same shape")
    rows_t, n_sel = _rows_tensor(pop_indices, g.shape[0])
    max_value = 1 if kind == "hap" else 2
    has_missing, bad = scan_rows(g, rows_t, n_sel, max_value)
    if bad:
        raise ValueError("input values must lie in {-1, 0, 1}" if kind == "hap"
                         else "genotype values must lie in {-1, 0, 1, 2}")
    n = ii.shape[0]
    nc = 4 if kind == "hap" else 9
    counts = L.empty((n, nc), np.int32)
    n_valid = L.empty((n,), np.int32)
    name = "pgld_counts_pairs_hap" if kind == "hap" else "pgld_counts_pairs_geno"
    L.call(name, L.ptr(g), g.shape[0], g.shape[1], L.ptr(rows_t), n_sel, L.ptr(ii), L.ptr(jj), n,
           L.ptr(counts), L.ptr(n_valid), L.stream_handle())
    return counts, (n_valid if has_missing else None)


def compute_counts_for_pairs(haplotypes, idx_i, idx_j, pop_indices=None):
    """[n11, n10, n01, n00] per pair (int32, (N, 4)) and n_valid (int32, (N,))
    or None when the population has no missing data (ld_pipeline.py:130-175)."""
    return _counts("hap", haplotypes, idx_i, idx_j, pop_indices)


def compute_genotype_counts_for_pairs(genotypes, idx_i, idx_j, pop_indices=None):
    """9-way counts n00..n22 (int32, (N, 9)) and n_valid or None
    (ld_pipeline.py:220-263)."""
    return _counts("geno", genotypes, idx_i, idx_j, pop_indices)


# ---------------------------------------------------------------------------
# Names and moments averaging specs (ld_pipeline.py:357-457)
# ---------------------------------------------------------------------------

def ld_names(num_pops):
    """LD statistic names in moments.LD.Util.ld_names() order."""
    names = []
    for a in range(num_pops):
        for b in range(a, num_pops):
            names.append(f"DD_{a}_{b}")
    for a in range(num_pops):
        for b in range(num_pops):
            for c in range(b, num_pops):
                names.append(f"Dz_{a}_{b}_{c}")
    for a in range(num_pops):
        for b in range(a, num_pops):
            for c in range(a, num_pops):
                for d in range(c, num_pops):
                    if c == a == d and b != a:
                        continue
                    if a == c and d < b:
                        continue
                    names.append(f"pi2_{a}_{b}_{c}_{d}")
    return names


def het_names(num_pops):
    return [f"H_{a}_{b}" for a in range(num_pops) for b in range(a, num_pops)]


def generate_stat_specs(num_pops):
    """(name, [(weight, stat_type, pops), ...]) with moments' _call_sgc
    averaging weights 1, 1/2, 1/4, 1/8."""
    specs = []
    for name in ld_names(num_pops):
        kind, *rest = name.split("_")
        p = tuple(int(x) for x in rest)
        if kind == "DD":
            specs.append((name, [(1.0, "dd", p)]))
        elif kind == "Dz":
            a, b, c = p
            specs.append((name, [(1.0, "dz", p)] if b == c else
                          [(0.5, "dz", (a, b, c)), (0.5, "dz", (a, c, b))]))
        else:
            a, b, c, d = p
            if a == b and c == d:
                t = [(1.0, p)] if a == c else [(0.5, (a, b, c, d)), (0.5, (c, d, a, b))]
            elif a == b:
                t = [(0.25, q) for q in ((a, b, c, d), (a, b, d, c), (c, d, a, b), (d, c, a, b))]
            elif c == d:
                t = [(0.25, q) for q in ((a, b, c, d), (b, a, c, d), (c, d, a, b), (c, d, b, a))]
            else:
                t = [(0.125, q) for q in ((a, b, c, d), (a, b, d, c), (b, a, c, d), (b, a, d, c),
                                          (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a))]
            specs.append((name, [(w, "pi2", q) for w, q in t]))
    return specs


_TYPE = {"dd": L.STAT_DD, "dz": L.STAT_DZ, "pi2": L.STAT_PI2, "pi2_sym": L.STAT_PI2_SYM,
         "sigma_d2": L.STAT_SIGMA_D2}
_TERM_CACHE = {}


def term_table(specs, pi2_form="pi2"):
    """Device term table + offsets for a spec list (cached per spec list)."""
    key = (tuple((n, tuple((w, t, tuple(p)) for w, t, p in calls)) for n, calls in specs), pi2_form)
    hit = _TERM_CACHE.get(key)
    if hit is not None:
        return hit
    rows, off = [], [0]
    for _, calls in specs:
        for w, t, pops in calls:
            tt = pi2_form if t == "pi2" else t
            pp = list(pops) + [0] * (4 - len(pops))
            rows.append((_TYPE[tt], pp, 0, float(w)))
        off.append(len(rows))
    arr = np.zeros(len(rows), dtype=L.TERM_DTYPE)
    for k, (t, pp, _, w) in enumerate(rows):
        arr[k]["type"] = t
        arr[k]["pops"] = pp
        arr[k]["weight"] = w
    dev_terms = L.torch().as_tensor(arr.view(np.uint8), device=L.device())
    dev_off = L.to_device(np.asarray(off, dtype=np.int32), np.int32)
    _TERM_CACHE[key] = (dev_terms, dev_off, len(specs))
    return _TERM_CACHE[key]


def stats_from_counts(family, counts_list, n_valid_list, specs, pi2_form="pi2"):
    """(N, n_stats) float64 from per-population counts via pgld_stats_pairs."""
    counts_t = [L.to_device(c, np.int32) for c in counts_list]
    nv_t = [None if nv is None else L.to_device(nv, np.int32) for nv in (n_valid_list or [None] * len(counts_t))]
    n = counts_t[0].shape[0]
    terms, off, n_stats = term_table(specs, pi2_form)
    out = L.empty((n, n_stats), np.float64)
    cp, _keep1 = L.ptr_array(counts_t)
    nvp, _keep2 = L.ptr_array(nv_t)
    fam = L.FAMILY_HAP if family == "hap" else L.FAMILY_GENO
    L.call("pgld_stats_pairs", fam, len(counts_t), cp, nvp, n, L.ptr(terms), L.ptr(off), n_stats, L.ptr(out),
           L.stream_handle())
    return out


# ---------------------------------------------------------------------------
# Batch statistics (ld_pipeline.py:271-349)
# ---------------------------------------------------------------------------

_TWO_POP_SPECS = [(n, calls) for n, calls in generate_stat_specs(2)]


def compute_two_pop_statistics_batch(counts_pop1, counts_pop2, n_valid1, n_valid2, ld_statistics=None):
    """(N, 15) float64 in ld_names(2) order, with the fast-path forms of
    ld_pipeline.py:271-322 (pi2(i,i,k,k) symmetrised over the loci)."""
    return stats_from_counts("hap", [counts_pop1, counts_pop2], [n_valid1, n_valid2], _TWO_POP_SPECS,
                             pi2_form="pi2_sym")


def compute_single_pop_statistics_batch(counts, n_valid, ld_statistics=None):
    """(N, 3) float64: [DD, Dz, pi2] (ld_pipeline.py:330-349)."""
    return stats_from_counts("hap", [counts], [n_valid], generate_stat_specs(1))


class PopData:
    """Per-population float64 columns of a haplotype count array
    (ld_pipeline.py:465-482)."""
    __slots__ = ("c1", "c2", "c3", "c4", "n", "D", "pA", "qA", "pB", "qB")

    def __init__(self, counts, n_valid):
        c = L.to_device(counts, np.float64)
        self.c1, self.c2, self.c3, self.c4 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
        self.n = L.to_device(n_valid, np.float64) if n_valid is not None else c.sum(dim=1)
        self.D = self.c2 * self.c3 - self.c1 * self.c4
        self.pA = self.c1 + self.c2
        self.qA = self.c3 + self.c4
        self.pB = self.c1 + self.c3
        self.qB = self.c2 + self.c4


# ---------------------------------------------------------------------------
# Keyed sums through the per-pair path (the structure of moments_ld.py:294-343)
# ---------------------------------------------------------------------------

def self_stat_specs(num_pops):
    """The within-population set DD_p_p, Dz_p_p_p, pi2_p_p_p_p in the fused
    kernel's column order (statistic s of population p at s * P + p)."""
    if num_pops == 1:
        return generate_stat_specs(1)
    out = []
    for kind, t, k in (("DD", "dd", 2), ("Dz", "dz", 3), ("pi2", "pi2", 4)):
        for p in range(num_pops):
            out.append(("_".join([kind] + [str(p)] * k), [(1.0, t, (p,) * k)]))
    return out


def plain_keyed_sums(data, family, positions, rows, max_dist, key_mode, n_keys, edges=None, origin=0.0, width=1.0,
                     key_coord=None, stat_set="moments"):
    """FP64 sums per key of the moments (or within-population) statistics of
    every pair within max_dist, one pair chunk at a time: plain counts ->
    per-pair statistics -> pair keys -> keyed sums.  Keys as the fused band
    pass: a distance bin of key_coord (|difference| for KEY_BIN_ABS) or the
    window of the left SNP.  Returns (sums (n_keys, n_stats), pairs per key)
    as float64 device tensors."""
    T = L.torch()
    P = len(rows)
    if stat_set == "self":
        specs = self_stat_specs(P)
    else:
        specs = generate_stat_specs(P)
    pi2_form = "pi2_sym" if (family == "hap" and P == 2 and stat_set != "self") else "pi2"
    n_stats = len(specs)
    pos = np.asarray(L.to_numpy(positions), dtype=np.float64)
    coord = L.to_device(pos if key_coord is None else key_coord, np.float64)
    ed = L.to_device(np.zeros(1) if edges is None else edges, np.float64)
    acc = L.zeros((n_keys, n_stats + 1), np.float64)
    counts_fn = compute_counts_for_pairs if family == "hap" else compute_genotype_counts_for_pairs
    n_rows = data.shape[0]
    max_s = max(n_rows if r is None else len(r) for r in rows)
    chunk = estimate_ld_chunk_size(max_s, num_pops=P)
    for ii, jj in iter_pairs_within_distance(pos, max_dist, chunk):
        cl = [counts_fn(data, ii, jj, r) for r in rows]
        st = stats_from_counts(family, [c for c, _ in cl], [nv for _, nv in cl], specs, pi2_form=pi2_form)
        st = T.cat([st, T.ones((st.shape[0], 1), dtype=T.float64, device=st.device)], dim=1).contiguous()
        keys = L.empty((ii.shape[0],), np.int32)
        L.call("pgld_pair_keys", L.ptr(coord), L.ptr(ii), L.ptr(jj), ii.shape[0], key_mode, L.ptr(ed), n_keys,
               float(origin), float(width), L.ptr(keys), L.stream_handle())
        L.call("pgld_keyed_sums", L.ptr(st), L.ptr(keys), ii.shape[0], n_stats + 1, n_keys, L.ptr(acc),
               L.stream_handle())
    return acc[:, :n_stats].contiguous(), acc[:, n_stats].contiguous()
