"""CPU oracle for the LD pair-statistics path -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  It is a NumPy (int64 / float64) restatement of the algorithm of
kr-colab/pg_gpu's LD path; every function cites the reference file:line it
follows.  Nothing here touches a GPU.

Pinning (how each part is checked independently of this transcription):

* pair enumeration  -- brute-force O(m^2) set in tests (the reference's own
  oracle, tests/test_ld_pair_iteration.py:21-29) plus its edge cases.
* 4-way / 9-way counts -- definition-level brute force, the reference KATs
  (tests/test_haplotype_counting.py:43-130), marginal/symmetry invariants.
* r, r^2, D' -- closed-form KATs (tests/test_ld_r_squared.py:15-57,171-230).
* haplotype DD/Dz/pi2, every population case for 1-4 populations (the
  four-distinct-population pi2 included) -- brute-force U-statistics over
  ordered tuples of haplotypes (distinct within a population) in
  tests/ustat_bruteforce.py, and the frozen moments output on the IM example
  (tests/cache/moments_ld_test_im_4bins_deme0-deme1.pkl, decoded statically).
* genotype DD/Dz/pi2, 1-4 populations -- brute-force U-statistics over
  individuals with dosage kernels (distinctness within a population).
* ld_names / het_names / stat_specs -- a literal three-population name list
  written from moments' naming rule, the orbit form of that rule for 1-4
  populations, and each spec as the uniform average over its statistic's
  symmetry group.
* union_biallelic_keep_geno -- known answers (all-reference, all-alternative,
  one valid individual, duplicate rows across populations, rows in no
  population).
* heterozygosity, haplotype and genotype modes -- enumeration of allele pairs
  per site (missing haplotypes counted as reference over the constant |pop|).
* bin/window sums -- brute-force per-key sums over explicit pair lists.
* dense D / r^2 / sigma_D^2 / dosage r^2 matrices, ZnS, Omega -- equal to
  the pinned count D, r^2 / numerator ratio / Rogers-Huff r^2 where they
  coincide;
  Omega by brute force over partitions.
* locate_unlinked -- known answers and the reference's property tests
  (tests/test_ld_r_squared.py:124-146).
* per-window ZnS / Omega -- the pinned sigma_D^2 matrix, ZnS and Omega
  applied to each window's columns; the window ranges and the NaN / 0.0
  cases by known answers (bin edges, empty and short windows).
* tiled_window_sums_hap -- the per-pair window sums of this module.

Functions with no independent pin are marked "parity unpinned" in their
docstring; there are none at present except ``rogers_huff_r``'s parity with
scikit-allel (closed form only; scikit-allel is absent here).
"""
from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# Pair enumeration  (ld_pipeline.py:57-122)
# ---------------------------------------------------------------------------


def pairs_within_distance(positions, max_dist):
    """All (i, j), i < j, with pos[j] - pos[i] <= max_dist (inclusive),
    ordered by (i, j).  Follows ld_pipeline.py:85-121 (searchsorted
    side='right' upper bound, per-variant pair counts)."""
    pos = np.asarray(positions)
    n = pos.shape[0]
    if n < 2:
        e = np.empty(0, dtype=np.int64)
        return e, e.copy()
    j_max = np.searchsorted(pos, pos + max_dist, side="right")
    per = np.maximum(0, j_max - np.arange(n) - 1)
    total = int(per.sum())
    idx_i = np.repeat(np.arange(n, dtype=np.int64), per)
    start = np.cumsum(per) - per
    within = np.arange(total, dtype=np.int64) - np.repeat(start, per)
    idx_j = idx_i + 1 + within
    return idx_i, idx_j


def chunk_bounds(positions, max_dist, chunk_size):
    """Anchor-block chunk boundaries [(v_lo, v_hi)] of
    iter_pairs_within_distance (ld_pipeline.py:94-122): the exclusive upper
    anchor of chunk c is the smallest v with cumulative[v] >= (c+1)*chunk."""
    pos = np.asarray(positions)
    n = pos.shape[0]
    if n < 2:
        return []
    j_max = np.searchsorted(pos, pos + max_dist, side="right")
    per = np.maximum(0, j_max - np.arange(n) - 1)
    cum = np.cumsum(per)
    total = int(cum[-1])
    if total == 0:
        return []
    chunk_size = max(1, int(chunk_size))
    n_chunks = (total + chunk_size - 1) // chunk_size
    targets = np.minimum(np.arange(1, n_chunks + 1, dtype=np.int64) * chunk_size, total)
    v_his = np.minimum(np.searchsorted(cum, targets, side="left") + 1, n)
    out = []
    v_lo = 0
    for v_hi in v_his:
        v_hi = int(v_hi)
        if v_hi <= v_lo:
            continue
        if int(per[v_lo:v_hi].sum()) > 0:
            out.append((v_lo, v_hi))
        v_lo = v_hi
    return out


# ---------------------------------------------------------------------------
# Counts  (ld_pipeline.py:130-175 haplotype, :183-263 genotype)
# ---------------------------------------------------------------------------


def hap_counts(hap, idx_i, idx_j, rows=None):
    """[n11, n10, n01, n00] per pair over the selected rows, counting only
    samples valid (>= 0) at both sites; n10 needs x_j == 0 exactly
    (ld_pipeline.py:155-174).  n_valid is None when the selected submatrix
    has no -1 (ld_pipeline.py:158)."""
    x = np.asarray(hap)
    if rows is not None:
        x = x[np.asarray(rows, dtype=np.int64), :]
    xi = x[:, idx_i].astype(np.int16)
    xj = x[:, idx_j].astype(np.int16)
    has_missing = bool(np.any(x == -1))
    valid = (xi >= 0) & (xj >= 0)
    n11 = np.sum((xi == 1) & (xj == 1) & valid, axis=0)
    n10 = np.sum((xi == 1) & (xj == 0) & valid, axis=0)
    n01 = np.sum((xi == 0) & (xj == 1) & valid, axis=0)
    n00 = np.sum((xi == 0) & (xj == 0) & valid, axis=0)
    counts = np.stack([n11, n10, n01, n00], axis=1).astype(np.int64)
    n_valid = np.sum(valid, axis=0).astype(np.int64) if has_missing else None
    return counts, n_valid


def geno_counts(geno, idx_i, idx_j, rows=None):
    """9-way genotype counts per pair in order n00,n01,n02,n10,...,n22
    (combo = g_i*3 + g_j), valid at both sites (ld_pipeline.py:195-216).
    n_valid returned only if any value < 0 (ld_pipeline.py:253,263)."""
    g = np.asarray(geno)
    if rows is not None:
        g = g[np.asarray(rows, dtype=np.int64), :]
    gi = g[:, idx_i].astype(np.int16)
    gj = g[:, idx_j].astype(np.int16)
    valid = (gi >= 0) & (gj >= 0)
    out = np.zeros((len(idx_i), 9), dtype=np.int64)
    for a in range(3):
        for b in range(3):
            out[:, 3 * a + b] = np.sum((gi == a) & (gj == b) & valid, axis=0)
    has_missing = bool(np.any(g < 0))
    n_valid = np.sum(valid, axis=0).astype(np.int64) if has_missing else None
    return out, n_valid


# ---------------------------------------------------------------------------
# Haplotype r / r^2 / D'   (haplotype_kernels.py:20-62, ld_statistics.py:153-246)
# ---------------------------------------------------------------------------


def _rnd_inputs(counts, n_valid):
    c = np.asarray(counts, dtype=np.float64)
    a, b, cc, d = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    n = (a + b + cc + d) if n_valid is None else np.asarray(n_valid, dtype=np.float64)
    return a, b, cc, d, n


def r(counts, n_valid=None):
    """Pearson r (haplotype_kernels.py:20-30): NaN when denom <= 0."""
    a, b, c, d, n = _rnd_inputs(counts, n_valid)
    with np.errstate(divide="ignore", invalid="ignore"):
        pA, pB = (a + b) / n, (a + c) / n
        D = (a * d - b * c) / (n * n)
        den = pA * (1.0 - pA) * pB * (1.0 - pB)
        return np.where(den > 0.0, D / np.sqrt(np.where(den > 0, den, 1.0)), np.nan)


def r_squared(counts, n_valid=None):
    """r^2 (haplotype_kernels.py:32-42): NaN when denom <= 0."""
    a, b, c, d, n = _rnd_inputs(counts, n_valid)
    with np.errstate(divide="ignore", invalid="ignore"):
        pA, pB = (a + b) / n, (a + c) / n
        D = (a * d - b * c) / (n * n)
        den = pA * (1.0 - pA) * pB * (1.0 - pB)
        return np.where(den > 0.0, D * D / np.where(den > 0, den, 1.0), np.nan)


def r_squared_exact(counts):
    """r^2 of the dense/banded r^2 path: diff = n n11 - nA nB in int64 (n^2 D
    exactly), r^2 = (diff * diff) / ((nA (n - nA)) * (nB (n - nB))) with each
    factor converted to float64 and one division; NaN where the denominator is
    0.  The same value as r_squared (haplotype_kernels.py:32-42) up to
    rounding: that form rounds pA, pB and D before combining them."""
    c = np.asarray(counts, dtype=np.int64)
    n11 = c[:, 0]
    nA = c[:, 0] + c[:, 1]
    nB = c[:, 0] + c[:, 2]
    n = c.sum(axis=1)
    dd = (n * n11 - nA * nB).astype(np.float64)
    num = dd * dd
    den = (nA * (n - nA)).astype(np.float64) * (nB * (n - nB)).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(den > 0.0, num / np.where(den > 0.0, den, 1.0), np.nan)


def r2_histogram(r2, keys, n_keys, n_hist):
    """(hist (n_keys, n_hist), nan (n_keys,)) of r^2 values per key: bin
    min(floor(r2 * n_hist), n_hist - 1); keys outside [0, n_keys) dropped."""
    r2 = np.asarray(r2)
    keys = np.asarray(keys)
    ok = (keys >= 0) & (keys < n_keys)
    fin = ok & ~np.isnan(r2)
    b = np.minimum((r2[fin] * float(n_hist)).astype(np.int64), n_hist - 1)
    hist = np.zeros((n_keys, n_hist), dtype=np.int64)
    np.add.at(hist, (keys[fin], b), 1)
    nan = np.bincount(keys[ok & np.isnan(r2)], minlength=n_keys)[:n_keys]
    return hist, nan


def d_prime(counts, n_valid=None):
    """Lewontin D' (haplotype_kernels.py:44-62)."""
    a, b, c, d, n = _rnd_inputs(counts, n_valid)
    with np.errstate(divide="ignore", invalid="ignore"):
        pA = (a + b) / n
        qA = 1.0 - pA
        pB = (a + c) / n
        qB = 1.0 - pB
        D = (a * d - b * c) / (n * n)
        dmax = np.where(D >= 0.0, np.minimum(pA * qB, qA * pB), np.minimum(pA * pB, qA * qB))
        return np.where(dmax > 0.0, D / np.where(dmax > 0, dmax, 1.0), np.nan)


# ---------------------------------------------------------------------------
# Haplotype DD / Dz / pi2 -- per-call formulas
#
# Two transcriptions exist in the reference: the CuPy formulas of
# ld_statistics.py:1082-1617 (used by the 2-population fast path,
# ld_pipeline.py:271-322) and the fused kernels of haplotype_kernels.py:69-304
# (used for 1, 3 and 4 populations, moments_ld.py:331-333).  They agree per
# call except pi2(i,i,k,k), where the CuPy form is symmetrised over the two
# loci; after the moments averaging specs both give the same named statistic.
# ---------------------------------------------------------------------------


class HapPop:
    """Per-population float64 columns (ld_pipeline.py:465-482)."""

    def __init__(self, counts, n_valid=None):
        c = np.asarray(counts, dtype=np.float64)
        self.c1, self.c2, self.c3, self.c4 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
        self.n = (self.c1 + self.c2 + self.c3 + self.c4) if n_valid is None \
            else np.asarray(n_valid, dtype=np.float64)
        self.D = self.c2 * self.c3 - self.c1 * self.c4
        self.pA = self.c1 + self.c2
        self.qA = self.c3 + self.c4
        self.pB = self.c1 + self.c3
        self.qB = self.c2 + self.c4


def _div(num, den, ok):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(ok, num / np.where(ok, den, 1.0), 0.0)


def hap_dd(P, i, j):
    """DD(i,j): single-pop ld_statistics.py:1116-1128 / haplotype_kernels.py:69-79;
    between ld_statistics.py:1131-1169 / haplotype_kernels.py:81-89."""
    a = P[i]
    if i == j:
        c1, c2, c3, c4, n = a.c1, a.c2, a.c3, a.c4, a.n
        num = c1 * (c1 - 1) * c4 * (c4 - 1) + c2 * (c2 - 1) * c3 * (c3 - 1) - 2 * c1 * c2 * c3 * c4
        return _div(num, n * (n - 1) * (n - 2) * (n - 3), n >= 4)
    b = P[j]
    return _div(a.D * b.D, a.n * (a.n - 1) * b.n * (b.n - 1), (a.n >= 2) & (b.n >= 2))


def hap_dz(P, i, j, k):
    """Dz(i,j,k) in its five index-equality cases: ld_statistics.py:1172-1272,
    haplotype_kernels.py:95-160."""
    if i == j == k:
        p = P[i]
        c1, c2, c3, c4, n = p.c1, p.c2, p.c3, p.c4, p.n
        diff = c1 * c4 - c2 * c3
        num = (diff * ((c3 + c4) - (c1 + c2)) * ((c2 + c4) - (c1 + c3))
               + diff * ((c2 + c3) - (c1 + c4)) + 2 * (c2 * c3 + c1 * c4))
        return _div(num, n * (n - 1) * (n - 2) * (n - 3), n >= 4)
    if i == j:  # Dz(i,i,k)
        a, b = P[i], P[k]
        num = (a.qA - a.pA) * (-a.D) * (-b.c1 + b.c2 - b.c3 + b.c4)
        return _div(num, b.n * a.n * (a.n - 1) * (a.n - 2), (a.n >= 3) & (b.n >= 1))
    if i == k:  # Dz(i,j,i)
        a, b = P[i], P[j]
        num = (-a.c1 + a.c2 - a.c3 + a.c4) * (-a.D) * (b.qA - b.pA)
        return _div(num, b.n * a.n * (a.n - 1) * (a.n - 2), (a.n >= 3) & (b.n >= 1))
    if j == k:  # Dz(i,j,j)
        a, b = P[i], P[j]
        num = ((-a.D) * (-b.c1 + b.c2 + b.c3 - b.c4)
               + (-a.D) * (-b.c1 + b.c2 - b.c3 + b.c4) * (b.qA - b.pA))
        return _div(num, a.n * (a.n - 1) * b.n * (b.n - 1), (a.n >= 2) & (b.n >= 2))
    a, b, c = P[i], P[j], P[k]
    num = -(a.D * (b.pA - b.qA) * (c.c1 - c.c2 + c.c3 - c.c4))
    return _div(num, a.n * (a.n - 1) * b.n * c.n, (a.n >= 2) & (b.n >= 1) & (c.n >= 1))


def _pi2_case(i, j, k, l):
    """Case key of pi2(i,j,k,l), haplotype_kernels.py:460-486."""
    cnt = {}
    for p in (i, j, k, l):
        cnt[p] = cnt.get(p, 0) + 1
    nu, mc = len(cnt), max(cnt.values())
    if nu == 1:
        return "same", cnt
    if mc == 3:
        return "triple", cnt
    if i == j and k == l:
        return "iikk", cnt
    if i == j:
        return "iikl", cnt
    if k == l:
        return "ijkk", cnt
    if (i == k and j == l) or (i == l and j == k):
        return "ijij", cnt
    if nu == 3:
        return "shared", cnt
    return "alldiff", cnt


def hap_pi2_kernel(P, i, j, k, l):
    """pi2(i,j,k,l) as the fused kernels compute it (haplotype_kernels.py:167-304,
    case dispatch :460-573)."""
    case, cnt = _pi2_case(i, j, k, l)
    if case == "same":
        p = P[i]
        a, b, c, d, n = p.c1, p.c2, p.c3, p.c4, p.n
        num = (p.pA * p.pB * p.qB * p.qA - a * d * (-1 + a + 3 * b + 3 * c + d)
               - b * c * (-1 + 3 * a + b + c + 3 * d))
        return _div(num, n * (n - 1) * (n - 2) * (n - 3), n * (n - 1) * (n - 2) * (n - 3) > 0)
    if case == "triple":
        tp = [q for q, c in cnt.items() if c == 3][0]
        sp = [q for q, c in cnt.items() if c == 1][0]
        t, s = P[tp], P[sp]
        t1, t2, t3, t4 = t.c1, t.c2, t.c3, t.c4
        tpA, tqA = t1 + t2, t3 + t4
        num = (-(tpA * t4 * s.pB) - (t2 * tqA * s.pB) + (tpA * (t2 + t4) * tqA * s.pB)
               + (tpA * tqA * (-2.0 * s.c2 - 2.0 * s.c4)) + (tpA * t4 * s.qB)
               + (t2 * tqA * s.qB) + (tpA * (t1 + t3) * tqA * s.qB)) / 2.0
        den = s.n * t.n * (t.n - 1) * (t.n - 2)
        return _div(num, den, den > 0)
    if case == "iikk":
        a, b = P[i], P[k]
        den = a.n * (a.n - 1) * b.n * (b.n - 1)
        return _div(a.pA * a.qA * (b.pB * b.qB), den, den > 0)
    if case == "iikl":
        a, b, c = P[i], P[k], P[l]
        den = a.n * (a.n - 1) * b.n * c.n
        return _div(a.pA * a.qA * (b.pB * c.qB + b.qB * c.pB), 2.0 * den, den > 0)
    if case == "ijkk":
        a, b, c = P[i], P[j], P[k]
        den = c.n * (c.n - 1) * a.n * b.n
        return _div(c.pB * c.qB * (a.pA * b.qA + a.qA * b.pA), 2.0 * den, den > 0)
    if case == "ijij":
        a, b = P[i], P[j]
        return _pi2_ijij(a, b)
    if case == "shared":
        if i == k:
            s, x, y = i, j, l
        elif i == l:
            s, x, y = i, j, k
        elif j == k:
            s, x, y = j, i, l
        else:
            s, x, y = j, i, k
        return _pi2_shared(P[s], P[x], P[y])
    a, b, c, d = P[i], P[j], P[k], P[l]
    HA = a.pA * b.qA + a.qA * b.pA
    HB = c.pB * d.qB + c.qB * d.pB
    den = a.n * b.n * c.n * d.n
    return _div(HA * HB, 4.0 * den, den > 0)


def _pi2_ijij(A, B):
    """pi2(i,j,i,j): ld_statistics.py:1376-1434 == haplotype_kernels.py:243-266."""
    a1, a2, a3, a4 = A.c1, A.c2, A.c3, A.c4
    b1, b2, b3, b4 = B.c1, B.c2, B.c3, B.c4
    quart = ((a2 + a4) * (a3 + a4) * (b1 + b2) * (b1 + b3)
             + (a1 + a3) * (a3 + a4) * (b1 + b2) * (b2 + b4)
             + (a1 + a2) * (a2 + a4) * (b1 + b3) * (b3 + b4)
             + (a1 + a2) * (a1 + a3) * (b2 + b4) * (b3 + b4))
    cub = (-(a2 * a3 * b1) + a4 * b1 - a2 * a4 * b1 - a3 * a4 * b1 - a4 * a4 * b1 - a4 * b1 * b1
           + a3 * b2 - a1 * a3 * b2 - a3 * a3 * b2 - a1 * a4 * b2 - a3 * a4 * b2 - a3 * b1 * b2
           - a4 * b1 * b2 - a3 * b2 * b2
           + a2 * b3 - a1 * a2 * b3 - a2 * a2 * b3 - a1 * a4 * b3 - a2 * a4 * b3 - a2 * b1 * b3
           - a4 * b1 * b3 - a1 * b2 * b3 - a4 * b2 * b3 - a2 * b3 * b3
           + a1 * b4 - a1 * a1 * b4 - a1 * a2 * b4 - a1 * a3 * b4 - a2 * a3 * b4 - a2 * b1 * b4
           - a3 * b1 * b4 - a1 * b2 * b4 - a3 * b2 * b4 - a1 * b3 * b4 - a2 * b3 * b4 - a1 * b4 * b4)
    num = (quart + cub) / 4.0
    den = A.n * (A.n - 1) * B.n * (B.n - 1)
    return _div(num, den, (A.n >= 2) & (B.n >= 2))


def _pi2_shared(S, X, Y):
    """pi2 with one population shared between the two locus pairs:
    ld_statistics.py:1512-1584 == haplotype_kernels.py:218-241
    (cs = shared pop, ca = other pop of the first pair, cb = of the second)."""
    s1, s2, s3, s4 = S.c1, S.c2, S.c3, S.c4
    a1, a2, a3, a4 = X.c1, X.c2, X.c3, X.c4
    b1, b2, b3, b4 = Y.c1, Y.c2, Y.c3, Y.c4
    inner = (-1 + s3) * b1 + s3 * b2 - b3 + s3 * b3 + s3 * b4 + s1 * (b2 + b4)
    num = (s4 * s4 * (a1 + a2) * (b1 + b3)
           + s2 * s2 * (a3 + a4) * (b1 + b3)
           + (-1 + s1 + s3) * (s3 * (a1 + a2) + s1 * (a3 + a4)) * (b2 + b4)
           + s4 * (s1 * (a3 + a4) * (b1 + b3) + a1 * inner + a2 * inner)
           + s2 * (s4 * (a1 + a2 + a3 + a4) * (b1 + b3)
                   + s3 * (a1 * (b1 + b3) + a2 * (b1 + b3) + (a3 + a4) * (b2 + b4))
                   + (a3 + a4) * ((-1 + s1) * b1 - b3 + s1 * (b2 + b3 + b4)))) / 4.0
    den = S.n * (S.n - 1) * X.n * Y.n
    return _div(num, den, (S.n >= 2) & (X.n >= 1) & (Y.n >= 1))


def hap_pi2_python(P, i, j, k, l):
    """pi2 as the CuPy formulas compute it (ld_statistics.py:1299-1374): equal
    to hap_pi2_kernel except pi2(i,i,k,k), which is symmetrised over loci."""
    case, _ = _pi2_case(i, j, k, l)
    if case == "iikk":
        a, b = P[i], P[k]
        num1 = a.pA * a.qA * b.pB * b.qB
        num2 = b.pA * b.qA * a.pB * a.qB
        return _div(0.5 * (num1 + num2), a.n * (a.n - 1) * b.n * (b.n - 1),
                    (a.n >= 2) & (b.n >= 2))
    return hap_pi2_kernel(P, i, j, k, l)


# ---------------------------------------------------------------------------
# Statistic names and moments averaging specs  (ld_pipeline.py:357-457)
# ---------------------------------------------------------------------------


def ld_names(num_pops):
    """moments.LD.Util.ld_names order (ld_pipeline.py:357-376)."""
    out = []
    for a in range(num_pops):
        for b in range(a, num_pops):
            out.append(f"DD_{a}_{b}")
    for a in range(num_pops):
        for b in range(num_pops):
            for c in range(b, num_pops):
                out.append(f"Dz_{a}_{b}_{c}")
    for a in range(num_pops):
        for b in range(a, num_pops):
            for c in range(a, num_pops):
                for d in range(c, num_pops):
                    if c == a == d and b != a:
                        continue
                    if a == c and d < b:
                        continue
                    out.append(f"pi2_{a}_{b}_{c}_{d}")
    return out


def het_names(num_pops):
    return [f"H_{a}_{b}" for a in range(num_pops) for b in range(a, num_pops)]


def stat_specs(num_pops):
    """(name, [(weight, type, pops)]) following ld_pipeline.py:388-457."""
    specs = []
    for name in ld_names(num_pops):
        kind, *rest = name.split("_")
        p = tuple(int(x) for x in rest)
        if kind == "DD":
            specs.append((name, [(1.0, "dd", p)]))
        elif kind == "Dz":
            a, b, c = p
            if b == c:
                specs.append((name, [(1.0, "dz", p)]))
            else:
                specs.append((name, [(0.5, "dz", (a, b, c)), (0.5, "dz", (a, c, b))]))
        else:
            a, b, c, d = p
            if a == b and c == d:
                if a == c:
                    terms = [(1.0, (a, b, c, d))]
                else:
                    terms = [(0.5, (a, b, c, d)), (0.5, (c, d, a, b))]
            elif a == b:
                terms = [(0.25, (a, b, c, d)), (0.25, (a, b, d, c)),
                         (0.25, (c, d, a, b)), (0.25, (d, c, a, b))]
            elif c == d:
                terms = [(0.25, (a, b, c, d)), (0.25, (b, a, c, d)),
                         (0.25, (c, d, a, b)), (0.25, (c, d, b, a))]
            else:
                terms = [(0.125, t) for t in ((a, b, c, d), (a, b, d, c), (b, a, c, d),
                                              (b, a, d, c), (c, d, a, b), (d, c, a, b),
                                              (c, d, b, a), (d, c, b, a))]
            specs.append((name, [(w, "pi2", t) for w, t in terms]))
    return specs


def multi_pop_stats_hap(counts_list, n_valid_list, specs=None):
    """(N, n_stats) weighted sums per spec using the fused-kernel formulas
    (haplotype_kernels.py:584-640)."""
    P = [HapPop(c, nv) for c, nv in zip(counts_list, n_valid_list)]
    if specs is None:
        specs = stat_specs(len(P))
    cache = {}

    def call(t, pops):
        key = (t, pops)
        if key not in cache:
            if t == "dd":
                cache[key] = hap_dd(P, *pops)
            elif t == "dz":
                cache[key] = hap_dz(P, *pops)
            else:
                cache[key] = hap_pi2_kernel(P, *pops)
        return cache[key]

    N = P[0].n.shape[0]
    out = np.zeros((N, len(specs)))
    for s, (_, terms) in enumerate(specs):
        acc = np.zeros(N)
        for w, t, pops in terms:
            acc = acc + w * call(t, pops)
        out[:, s] = acc
    return out


def two_pop_stats(c0, c1, nv0=None, nv1=None):
    """The 15-column 2-population fast path (ld_pipeline.py:271-322), built on
    the CuPy formulas of ld_statistics.py."""
    P = [HapPop(c0, nv0), HapPop(c1, nv1)]
    dz = lambda *p: hap_dz(P, *p)
    pi = lambda *p: hap_pi2_python(P, *p)
    cols = [hap_dd(P, 0, 0), hap_dd(P, 0, 1), hap_dd(P, 1, 1),
            dz(0, 0, 0), 0.5 * (dz(0, 0, 1) + dz(0, 1, 0)), dz(0, 1, 1),
            dz(1, 0, 0), 0.5 * (dz(1, 0, 1) + dz(1, 1, 0)), dz(1, 1, 1),
            pi(0, 0, 0, 0), pi(0, 0, 0, 1), pi(0, 0, 1, 1), pi(0, 1, 0, 1), pi(0, 1, 1, 1),
            pi(1, 1, 1, 1)]
    return np.stack(cols, axis=1)


def tiled_window_sums_hap(hap, pos, pops, max_dist, window, n_keys, i_lo=0, i_hi=None, block=512):
    """BLAS-tiled exact CPU restatement of the windowed two-population band
    sums over the left SNPs [i_lo, i_hi): the reference's GEMM form of the
    counts (_tile_counts, ld_statistics.py:281-298) as float32 products of 0/1
    indicator planes (exact while counts < 2^24; multi-threaded BLAS), then the
    FP64 statistics of two_pop_stats and window-of-the-left-SNP sums.  The
    multi-core CPU baseline of the band pass; pinned against the per-pair
    oracle in the test suite."""
    x = np.asarray(hap)
    pos = np.asarray(pos)
    i_hi = x.shape[1] if i_hi is None else i_hi
    planes = []
    for r in pops:
        xr = x[np.asarray(r, dtype=np.int64), :]
        planes.append(((xr == 1).astype(np.float32), (xr >= 0).astype(np.float32), bool(np.any(xr == -1))))
    acc = np.zeros((n_keys, 15))
    for b0 in range(i_lo, i_hi, block):
        b1 = min(b0 + block, i_hi)
        j1 = int(np.searchsorted(pos, pos[b1 - 1] + max_dist, side="right"))
        ii, jj = np.meshgrid(np.arange(b0, b1), np.arange(b0, j1), indexing="ij")
        m = (jj > ii) & (pos[jj] - pos[ii] <= max_dist)
        if not m.any():
            continue
        cl = []
        for H, V, miss in planes:
            hh = H[:, b0:b1].T @ H[:, b0:j1]
            hv = H[:, b0:b1].T @ V[:, b0:j1]
            vh = V[:, b0:b1].T @ H[:, b0:j1]
            vv = V[:, b0:b1].T @ V[:, b0:j1]
            hh, hv, vh, vv = (a[m].astype(np.int64) for a in (hh, hv, vh, vv))
            c = np.stack([hh, hv - hh, vh - hh, vv - hv - vh + hh], axis=1)
            cl.append((c, vv if miss else None))
        st = two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
        acc += keyed_sums(window_key(pos[ii[m]], 0.0, window), st, n_keys)
    return acc


def single_pop_stats_hap(counts, n_valid=None):
    """[DD, Dz, pi2] (ld_pipeline.py:330-349)."""
    P = [HapPop(counts, n_valid)]
    return np.stack([hap_dd(P, 0, 0), hap_dz(P, 0, 0, 0), hap_pi2_kernel(P, 0, 0, 0, 0)], axis=1)


# ---------------------------------------------------------------------------
# Genotype statistics  (genotype_kernels.py:27-936)
# Moments order g1..g9 = n22, n21, n20, n12, n11, n10, n02, n01, n00
# (genotype_kernels.py:424-435).
# ---------------------------------------------------------------------------

# Monomial tables, exponents over (g1..g9).  Coefficients follow the
# polynomial numerators of genotype_kernels.py:38-65 (DD single, x16) and
# :131-157 (pi2 single correction, x16) and :211-232 (Dz single quartic, x4).
_DD_POLY = [
    (1, "24"), (-1, "224"), (4, "34"), (-4, "234"), (-4, "334"), (-1, "244"), (-4, "344"),
    (1, "15"), (-1, "115"), (1, "35"), (2, "135"), (-1, "335"), (-4, "345"), (-1, "155"),
    (-1, "355"),
    (4, "16"), (-4, "116"), (1, "26"), (-4, "126"), (-1, "226"), (2, "246"), (-4, "156"),
    (-4, "166"), (-1, "266"),
    (4, "27"), (-4, "227"), (16, "37"), (-16, "237"), (-16, "337"), (-4, "247"), (-16, "347"),
    (1, "57"), (2, "157"), (-4, "257"), (-18, "357"), (-1, "557"), (4, "67"), (8, "167"),
    (-16, "367"), (-4, "567"), (-4, "667"), (-4, "277"), (-16, "377"), (-1, "577"), (-4, "677"),
    (4, "18"), (-4, "118"), (4, "38"), (8, "138"), (-4, "338"), (1, "48"), (-4, "148"),
    (2, "248"), (-1, "448"), (-4, "158"), (-4, "358"), (1, "68"), (2, "268"), (-4, "368"),
    (2, "468"), (-1, "668"), (-16, "378"), (-4, "678"), (-4, "188"), (-4, "388"), (-1, "488"),
    (-1, "688"),
    (16, "19"), (-16, "119"), (4, "29"), (-16, "129"), (-4, "229"), (4, "49"), (-16, "149"),
    (8, "349"), (-4, "449"), (1, "59"), (-18, "159"), (-4, "259"), (2, "359"), (-4, "459"),
    (-1, "559"), (-16, "169"), (-4, "269"), (8, "279"), (2, "579"), (-16, "189"), (-4, "489"),
    (-16, "199"), (-4, "299"), (-4, "499"), (-1, "599"),
]

_PI2_CORR = [
    (13, "24"), (-16, "124"), (-11, "224"), (16, "34"), (-28, "134"), (-24, "234"),
    (-8, "334"), (-11, "244"), (-20, "344"), (-6, "5"), (12, "15"), (-4, "115"), (17, "25"),
    (-20, "125"), (-11, "225"), (12, "35"), (-28, "135"), (-20, "235"), (-4, "335"),
    (17, "45"), (-20, "145"), (-32, "245"), (-40, "345"), (-11, "445"), (11, "55"),
    (-16, "155"), (-17, "255"), (-16, "355"), (-17, "455"), (-6, "555"), (16, "16"), (-8, "116"),
    (13, "26"), (-24, "126"), (-11, "226"), (-28, "136"), (-16, "236"), (24, "46"),
    (-36, "146"), (-38, "246"), (-36, "346"), (-20, "446"), (17, "56"), (-40, "156"),
    (-32, "256"), (-20, "356"), (-42, "456"), (-17, "556"), (-20, "166"), (-11, "266"),
    (-20, "466"), (-11, "566"), (16, "27"), (-28, "127"), (-20, "227"), (16, "37"),
    (-48, "137"), (-44, "237"), (-16, "337"), (-24, "247"), (-44, "347"),
    (12, "57"), (-28, "157"), (-40, "257"), (-48, "357"), (-20, "457"), (-16, "557"),
    (16, "67"), (-48, "167"), (-48, "267"), (-44, "367"), (-36, "467"), (-40, "567"),
    (-20, "667"), (-8, "277"), (-16, "377"), (-4, "577"), (-8, "677"), (16, "18"), (-8, "118"),
    (24, "28"), (-36, "128"), (-20, "228"), (16, "38"), (-48, "138"), (-36, "238"), (-8, "338"),
    (13, "48"), (-24, "148"), (-38, "248"), (-48, "348"), (-11, "448"), (17, "58"), (-40, "158"),
    (-42, "258"), (-40, "358"), (-32, "458"), (-17, "558"), (13, "68"), (-48, "168"),
    (-38, "268"), (-24, "368"), (-38, "468"), (-32, "568"), (-11, "668"), (-28, "178"),
    (-36, "278"), (-44, "378"), (-16, "478"), (-20, "578"), (-24, "678"), (-20, "188"),
    (-20, "288"), (-20, "388"), (-11, "488"), (-11, "588"), (-11, "688"), (16, "19"), (-16, "119"),
    (16, "29"), (-44, "129"), (-20, "229"), (-48, "139"), (-28, "239"), (16, "49"),
    (-44, "149"), (-48, "249"), (-48, "349"), (-20, "449"), (12, "59"), (-48, "159"),
    (-40, "259"), (-28, "359"), (-40, "459"), (-16, "559"), (-44, "169"), (-24, "269"),
    (-36, "469"), (-20, "569"), (-48, "179"), (-48, "279"), (-48, "379"), (-28, "479"),
    (-28, "579"), (-28, "679"), (-44, "189"), (-36, "289"), (-28, "389"), (-24, "489"),
    (-20, "589"), (-16, "689"), (-16, "199"), (-8, "299"), (-8, "499"), (-4, "599"),
]

_DZ_QUARTIC = [
    (-1, "24"), (3, "124"), (1, "224"), (2, "34"), (4, "134"), (-1, "234"), (-4, "334"),
    (1, "244"), (2, "344"), (2, "15"), (-3, "115"), (-1, "125"), (2, "35"), (2, "135"),
    (-1, "235"), (-3, "335"), (-1, "145"), (1, "345"), (2, "16"), (-4, "116"), (-1, "26"),
    (-1, "126"), (1, "226"), (4, "136"), (3, "236"), (-2, "146"), (-2, "246"), (-2, "346"),
    (1, "156"), (-1, "356"), (2, "166"), (1, "266"), (2, "27"), (4, "127"), (2, "227"),
    (8, "37"), (4, "137"), (-4, "337"), (-1, "247"), (2, "57"), (2, "157"), (1, "257"),
    (2, "357"), (-1, "457"), (2, "67"), (-1, "267"), (-2, "467"), (1, "567"), (2, "667"),
    (-4, "277"), (-4, "377"), (-3, "577"), (-4, "677"), (2, "18"), (-4, "118"), (-2, "128"),
    (2, "38"), (-2, "238"), (-4, "338"), (-1, "48"), (-1, "148"), (-2, "248"), (-1, "348"),
    (1, "448"), (1, "158"), (1, "358"), (-1, "68"), (-1, "168"), (-2, "268"), (-1, "368"),
    (-2, "468"), (1, "668"), (4, "178"), (-2, "278"), (3, "478"), (-1, "578"), (-1, "678"),
    (2, "188"), (2, "388"), (1, "488"), (1, "688"), (8, "19"), (-4, "119"), (2, "29"),
    (2, "229"), (4, "139"), (4, "239"), (2, "49"), (-1, "249"), (2, "449"), (2, "59"),
    (2, "159"), (1, "259"), (2, "359"), (1, "459"), (-1, "269"), (-2, "469"), (-1, "569"),
    (4, "179"), (4, "379"), (4, "479"), (2, "579"), (4, "679"), (-2, "289"), (4, "389"),
    (-1, "489"), (-1, "589"), (3, "689"), (-4, "199"), (-4, "299"), (-4, "499"), (-3, "599"),
]


def eval_poly(table, g):
    """Evaluate a monomial table on float64 columns g[0..8] (= g1..g9)."""
    acc = np.zeros_like(g[0])
    for coef, mono in table:
        term = np.full_like(g[0], float(coef))
        for ch in mono:
            term = term * g[int(ch) - 1]
        acc = acc + term
    return acc


class GenoPop:
    """Per-population genotype features (genotype_kernels.py:412-492)."""

    def __init__(self, counts9, n_valid=None):
        c = np.asarray(counts9, dtype=np.float64)
        g = [c[:, 8 - t] for t in range(9)]  # g1=n22 ... g9=n00
        self.g = g
        g1, g2, g3, g4, g5, g6, g7, g8, g9 = g
        self.n = sum(g) if n_valid is None else np.asarray(n_valid, dtype=np.float64)
        self.D = (-(g2 / 2 + g3 + g5 / 4 + g6 / 2) * (g4 / 2 + g5 / 4 + g7 + g8 / 2)
                  + (g1 + g2 / 2 + g4 / 2 + g5 / 4) * (g5 / 4 + g6 / 2 + g8 / 2 + g9))
        self.pA = g1 + g2 + g3 + g4 / 2 + g5 / 2 + g6 / 2
        self.qA = g4 / 2 + g5 / 2 + g6 / 2 + g7 + g8 + g9
        self.pB = g1 + g2 / 2 + g4 + g5 / 2 + g7 + g8 / 2
        self.qB = g2 / 2 + g3 + g5 / 2 + g6 + g8 / 2 + g9
        n = self.n
        HL = g4 + g5 + g6
        HR = g2 + g5 + g8
        nn1 = np.maximum(4 * n * (n - 1), 1)
        self.BL = (4 * self.pA * self.qA - HL) / nn1
        self.BR = (4 * self.pB * self.qB - HR) / nn1
        xAB = g1 + g2 / 2 + g4 / 2 + g5 / 4
        xAb = g2 / 2 + g3 + g5 / 4 + g6 / 2
        xaB = g4 / 2 + g5 / 4 + g7 + g8 / 2
        xab = g5 / 4 + g6 / 2 + g8 / 2 + g9
        self.C00 = (self.pA * self.pB - xAB) / nn1
        self.C01 = (self.pA * self.qB - xAb) / nn1
        self.C10 = (self.qA * self.pB - xaB) / nn1
        self.C11 = (self.qA * self.qB - xab) / nn1
        self.X01, self.X10, self.X11 = xAb, xaB, xab
        self.Bsgn = xAB - xAb - xaB + xab
        P, R, Q, S = self.pA, self.qA, self.pB, self.qB
        L0 = g5 + 2 * g6 + 2 * g8 + 4 * g9
        L1 = g5 + 2 * g6
        M1 = g5 + 2 * g8
        negR3 = 3 * g4 + 3 * g5 + 3 * g6 + 4 * g7 + 4 * g8 + 4 * g9
        negR5 = 5 * g4 + 5 * g5 + 5 * g6 + 8 * g7 + 8 * g8 + 8 * g9
        beta5A = -2 * g2 - 4 * g3 + 4 * g4 + g5 - 2 * g6 + 8 * g7 + 4 * g8 + 4
        beta6A = -g5 - 2 * g6 - 4 * g7 - 4 * g8 - 4 * g9
        self.AQ3 = -(S * negR3) / 8 + P * S * R / 2 + (R * L0) / 8 + L1 / 8 - (P * L0) / 8
        self.AS3 = -(Q * negR5) / 8 + P * Q * R / 2 + (R * beta5A) / 8 + beta6A / 8 + (P * L0) / 8
        negL3 = 3 * g2 + 4 * g3 + 3 * g5 + 4 * g6 + 3 * g8 + 4 * g9
        negL5 = 5 * g2 + 8 * g3 + 5 * g5 + 8 * g6 + 5 * g8 + 8 * g9
        beta5B = 4 * g2 + 8 * g3 - 2 * g4 + g5 + 4 * g6 - 4 * g7 - 2 * g8 + 4
        beta6B = -4 * g3 - g5 - 4 * g6 - 2 * g8 - 4 * g9
        self.BP3 = -(R * negL3) / 8 + Q * S * R / 2 + (S * L0) / 8 + M1 / 8 - (Q * L0) / 8
        self.BR3 = -(P * negL5) / 8 + P * Q * S / 2 + (S * beta5B) / 8 + beta6B / 8 + (Q * L0) / 8
        # Dz precompute features (genotype_kernels.py:176-197)
        self.A = -g1 + g3 - g4 + g6 - g7 + g9
        self.B = -g1 - g2 - g3 + g7 + g8 + g9
        cc = -g1 + g3 + g7 - g9
        MA = (2 * g2 * g7 + 4 * g3 * g7 + 2 * g5 * g7 + 4 * g6 * g7 - 2 * g1 * g8 + 2 * g3 * g8
              - 2 * g4 * g8 + 2 * g6 * g8 - 4 * g1 * g9 - 2 * g2 * g9 - 4 * g4 * g9 - 2 * g5 * g9)
        MB = (2 * g3 * g4 + 2 * g3 * g5 - 2 * g1 * g6 - 2 * g2 * g6 + 4 * g3 * g7 + 2 * g6 * g7
              + 4 * g3 * g8 + 2 * g6 * g8 - 4 * g1 * g9 - 4 * g2 * g9 - 2 * g4 * g9 - 2 * g5 * g9)
        x00, x01 = xAB, xAb
        x10, x11 = xaB, xab
        self.Dx = x00 * x11 - x01 * x10
        self.Wjj = cc + self.A * self.B
        self.U = self.Dx * (1 + self.B) + 0.25 * MA
        self.V = self.Dx * (1 + self.A) + 0.25 * MB
        with np.errstate(divide="ignore", invalid="ignore"):
            self.p = self.pA / n
            self.r = self.qA / n
            self.q = self.pB / n
            self.s = self.qB / n


def _raw_div(num, den):
    with np.errstate(divide="ignore", invalid="ignore"):
        return num / den


def geno_dd(P, i, j):
    """DD genotype: single genotype_kernels.py:27-71 (0 if ns < 4), between
    :73-81."""
    if i == j:
        p = P[i]
        g = p.g
        ns = sum(g)
        numer = eval_poly(_DD_POLY, g) / 16.0 + p.D * p.D
        return np.where(ns < 4, 0.0, _raw_div(4.0 * numer, ns * (ns - 1) * (ns - 2) * (ns - 3)))
    a, b = P[i], P[j]
    den = a.n * (a.n - 1) * b.n * (b.n - 1)
    return _div(4.0 * a.D * b.D, den, den > 0)


def geno_dz(P, i, j, k):
    """Dz genotype cases genotype_kernels.py:199-272 (no guard except III)."""
    if i == j == k:
        p = P[i]
        g = p.g
        ns = sum(g)
        numer = 0.25 * eval_poly(_DZ_QUARTIC, g) + p.A * p.B * p.Dx
        return np.where(ns < 4, 0.0, _raw_div(2.0 * numer, ns * (ns - 1) * (ns - 2) * (ns - 3)))
    if i == j:
        a, b = P[i], P[k]
        return _raw_div(2.0 * a.U * b.A, b.n * a.n * (a.n - 1) * (a.n - 2))
    if i == k:
        a, b = P[i], P[j]
        return _raw_div(2.0 * a.V * b.B, b.n * a.n * (a.n - 1) * (a.n - 2))
    if j == k:
        a, b = P[i], P[j]
        return _raw_div(2.0 * a.Dx * b.Wjj, b.n * (b.n - 1) * a.n * (a.n - 1))
    a, b, c = P[i], P[j], P[k]
    return _raw_div(2.0 * a.Dx * b.B * c.A, a.n * (a.n - 1) * b.n * c.n)


def geno_pi2(P, i, j, k, l):
    """pi2 genotype cases genotype_kernels.py:279-405, dispatch :683-795."""
    cnt = {}
    for q in (i, j, k, l):
        cnt[q] = cnt.get(q, 0) + 1
    nu, mc = len(cnt), max(cnt.values())
    if nu == 1:
        p = P[i]
        g = p.g
        ns = sum(g)
        g1, g2, g3, g4, g5, g6, g7, g8, g9 = g
        a1 = g1 + g2 + g3 + 0.5 * g4 + 0.5 * g5 + 0.5 * g6
        a2 = g1 + 0.5 * g2 + g4 + 0.5 * g5 + g7 + 0.5 * g8
        a3 = 0.5 * g2 + g3 + 0.5 * g5 + g6 + 0.5 * g8 + g9
        a4 = 0.5 * g4 + 0.5 * g5 + 0.5 * g6 + g7 + g8 + g9
        corr = eval_poly(_PI2_CORR, g) / 16.0
        return np.where(ns < 4, 0.0,
                        _raw_div(a1 * a2 * a3 * a4 + corr, ns * (ns - 1) * (ns - 2) * (ns - 3)))
    if mc == 3:
        tp = [q for q, c in cnt.items() if c == 3][0]
        sp = [q for q, c in cnt.items() if c == 1][0]
        t, s = P[tp], P[sp]
        d = t.n * (t.n - 1) * (t.n - 2) * s.n
        if i == tp and j == tp:  # pi2(i,i;i,j) or (i,i;j,i)
            return _div(t.AQ3 * s.pB + t.AS3 * s.qB, d, d > 0)
        return _div(t.BP3 * s.pA + t.BR3 * s.qA, d, d > 0)
    if i == j and k == l:
        return P[i].BL * P[k].BR
    if i == j:
        return P[i].BL * 0.5 * (P[k].q * P[l].s + P[k].s * P[l].q)
    if k == l:
        return 0.5 * (P[i].p * P[j].r + P[i].r * P[j].p) * P[k].BR
    if (i == k and j == l) or (i == l and j == k):
        A, B = P[i], P[j]
        Pi, Qi, Ri, Si = A.pA, A.pB, A.qA, A.qB
        Pj, Qj, Rj, Sj = B.pA, B.pB, B.qA, B.qB
        X01i, X10i, X11i, Bi = A.X01, A.X10, A.X11, A.Bsgn
        X01j, X10j, X11j, Bj = B.X01, B.X10, B.X11, B.Bsgn
        PQi, PRi, PSi, QRi, RSi = Pi * Qi, Pi * Ri, Pi * Si, Qi * Ri, Ri * Si
        PQj, PSj, QRj, RSj = Pj * Qj, Pj * Sj, Qj * Rj, Rj * Sj
        Tc = Ri * (1.0 + X01i - X10i)
        Uc = Si * (1.0 + Ri)
        num = (-0.25 * RSi * Qj - 0.25 * X11i * PQj + 0.25 * RSi * PQj - 0.25 * Ri * PSj
               - 0.25 * PQi * X11j - 0.25 * RSi * X11j + 0.25 * PSi * X11j + 0.25 * QRi * X11j
               - 0.25 * PRi * Sj + 0.25 * Tc * Sj + 0.25 * X11i * PSj + 0.25 * QRi * PSj
               + 0.25 * Uc * Rj - 0.25 * X01i * QRj + 0.25 * PSi * QRj
               - 0.25 * PSi * Rj - 0.25 * Pi * RSj + 0.25 * X01i * RSj + 0.25 * PQi * RSj
               + 0.25 * X11i * Bj + 0.25 * Bi * X11j - 1.0 * X11i * X11j)
        d = A.n * (A.n - 1) * B.n * (B.n - 1)
        return _div(num, d, d > 0)
    if nu == 3:
        if i == k:
            s, a, b = i, j, l
        elif i == l:
            s, a, b = i, j, k
        elif j == k:
            s, a, b = j, i, l
        else:
            s, a, b = j, i, k
        S, A, B = P[s], P[a], P[b]
        return A.r * (S.C00 * B.s + S.C01 * B.q) + A.p * (S.C10 * B.s + S.C11 * B.q)
    a, b, c, d = P[i], P[j], P[k], P[l]
    return 0.25 * (a.p * b.r + a.r * b.p) * (c.q * d.s + c.s * d.q)


def multi_pop_stats_geno(counts_list, n_valid_list, specs=None):
    """(N, n_stats) genotype statistics (genotype_kernels.py:881-936)."""
    P = [GenoPop(c, nv) for c, nv in zip(counts_list, n_valid_list)]
    if specs is None:
        specs = stat_specs(len(P))
    cache = {}

    def call(t, pops):
        key = (t, pops)
        if key not in cache:
            fn = {"dd": geno_dd, "dz": geno_dz, "pi2": geno_pi2}[t]
            cache[key] = fn(P, *pops)
        return cache[key]

    N = P[0].n.shape[0]
    out = np.zeros((N, len(specs)))
    for s, (_, terms) in enumerate(specs):
        acc = np.zeros(N)
        for w, t, pops in terms:
            acc = acc + w * call(t, pops)
        out[:, s] = acc
    return out


def sigma_d2_geno_from_counts(counts9):
    """Per-pair genotype sigma_D^2 = DD_num / pi2_num, NaN where ns < 4 or
    pi2_num <= 0 (genotype_kernels.py:91-169)."""
    p = GenoPop(counts9)
    g = p.g
    ns = sum(g)
    dd_num = (eval_poly(_DD_POLY, g) / 16.0 + p.D * p.D) * 4.0
    g1, g2, g3, g4, g5, g6, g7, g8, g9 = g
    a1 = g1 + g2 + g3 + 0.5 * g4 + 0.5 * g5 + 0.5 * g6
    a2 = g1 + 0.5 * g2 + g4 + 0.5 * g5 + g7 + 0.5 * g8
    a3 = 0.5 * g2 + g3 + 0.5 * g5 + g6 + 0.5 * g8 + g9
    a4 = 0.5 * g4 + 0.5 * g5 + 0.5 * g6 + g7 + g8 + g9
    pi_num = a1 * a2 * a3 * a4 + eval_poly(_PI2_CORR, g) / 16.0
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(pi_num > 0.0, dd_num / np.where(pi_num > 0, pi_num, 1.0), np.nan)
    return np.where(ns < 4, np.nan, out)


def sigma_d2_hap_from_counts(counts):
    """Haplotype-tile sigma_D^2 = dd_num / pi2_num, 0 where n < 4 or
    pi2_num == 0 (ld_statistics.py:311-341)."""
    c = np.asarray(counts, dtype=np.float64)
    c1, c2, c3, c4 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    n = c1 + c2 + c3 + c4
    dd_num = c1 * (c1 - 1) * c4 * (c4 - 1) + c2 * (c2 - 1) * c3 * (c3 - 1) - 2 * c1 * c2 * c3 * c4
    pi_num = ((c1 + c2) * (c1 + c3) * (c2 + c4) * (c3 + c4)
              - c1 * c4 * (-1 + c1 + 3 * c2 + 3 * c3 + c4)
              - c2 * c3 * (-1 + 3 * c1 + c2 + c3 + 3 * c4))
    ok = (n >= 4) & (pi_num != 0)
    return _div(dd_num, pi_num, ok), n >= 4


# ---------------------------------------------------------------------------
# Binning and windows  (moments_ld.py:310-339; haplotype_matrix.py:1670-1689)
# ---------------------------------------------------------------------------


def bin_index(dist, bins):
    """digitize(dist, bins) - 1: bin k is [b_k, b_{k+1}); outside -> -1 or
    n_bins (moments_ld.py:314, :335)."""
    return np.digitize(np.asarray(dist, dtype=np.float64), np.asarray(bins, dtype=np.float64)) - 1


def keyed_sums(keys, values, n_keys):
    """FP64 sums of values (N, S) per key in [0, n_keys); other keys dropped."""
    keys = np.asarray(keys)
    ok = (keys >= 0) & (keys < n_keys)
    out = np.zeros((n_keys, values.shape[1]))
    np.add.at(out, keys[ok], values[ok])
    return out


def window_key(pos_i, origin, window):
    """Window of the left SNP of a pair: floor((pos_i - origin) / window)."""
    return np.floor((np.asarray(pos_i, dtype=np.float64) - origin) / window).astype(np.int64)


# ---------------------------------------------------------------------------
# moments-style heterozygosity (moments_ld.py:346-388)
# ---------------------------------------------------------------------------


def heterozygosity(mat_rows_list, use_genotypes=False):
    """H_i_j dict.  For haplotypes n_hap is the constant |pop| (missing data
    ignored, moments_ld.py:366-368); genotypes use 2 * n_valid."""
    alt, ref, size = [], [], []
    for X in mat_rows_list:
        X = np.asarray(X)
        if use_genotypes:
            v = X >= 0
            a = np.where(v, X, 0).astype(np.int64).sum(axis=0).astype(np.float64)
            nh = 2.0 * v.sum(axis=0).astype(np.float64)
        else:
            a = np.maximum(X, 0).astype(np.int64).sum(axis=0).astype(np.float64)
            nh = np.full_like(a, float(X.shape[0]))
        alt.append(a)
        ref.append(nh - a)
        size.append(nh)
    out = {}
    P = len(mat_rows_list)
    for a in range(P):
        for b in range(a, P):
            if a == b:
                v = np.sum(2.0 * ref[a] * alt[a] / (size[a] * (size[a] - 1)))
            else:
                v = np.sum((ref[a] * alt[b] + alt[a] * ref[b]) / (size[a] * size[b]))
            out[f"H_{a}_{b}"] = float(v)
    return out


# ---------------------------------------------------------------------------
# End-to-end moments parse (moments_ld.py:40-343), haplotype and genotype
# ---------------------------------------------------------------------------


def biallelic_filter_hap(hap):
    """Keep sites with no value >= 2 and at least one 0 and one 1
    (haplotype_matrix.py:751-759)."""
    hap = np.asarray(hap)
    keep = (np.sum(hap >= 2, axis=0) == 0) & (np.sum(hap == 0, axis=0) > 0) & \
        (np.sum(hap == 1, axis=0) > 0)
    return np.nonzero(keep)[0]


def union_biallelic_keep_geno(geno, pop_rows):
    """Union-of-pops biallelic filter of moments_ld.py:256-285."""
    geno = np.asarray(geno)
    seen = []
    for rows in pop_rows:
        for r_ in rows:
            if r_ not in seen:
                seen.append(r_)
    G = geno[np.asarray(seen, dtype=np.int64), :]
    v = G >= 0
    alt = np.where(v, G, 0).astype(np.int64).sum(axis=0)
    nv = v.sum(axis=0)
    keep = (alt > 0) & (alt < 2 * nv) & (nv >= 2)
    return np.nonzero(keep)[0]


def moments_ld_sums(data, positions, pop_rows, bins, use_genotypes=False, distances=None, max_dist=None,
                    absolute=False):
    """Per-bin FP64 sums of the ld_names(P) statistics (moments_ld.py:248-343).

    data: (n_rows, m) int8 matrix (already biallelic-filtered for haplotypes),
    pop_rows: list of row-index lists.  Haplotype mode with two populations
    uses the fast path; otherwise the spec-weighted fused-kernel formulas.
    ``distances`` (genetic map coordinates) key the bins by |d_j - d_i| over
    pairs within the physical span ``max_dist`` (moments_ld.py:152-160).
    ``absolute`` sums |statistic| instead: the scale against which a sum
    taken in another order is compared."""
    data = np.asarray(data)
    pos = np.asarray(positions)
    if use_genotypes:
        keep = union_biallelic_keep_geno(data, pop_rows)
        data = data[:, keep]
        pos = pos[keep]
        if distances is not None:
            distances = np.asarray(distances)[keep]
    bins = np.asarray(bins, dtype=np.float64)
    n_bins = len(bins) - 1
    P = len(pop_rows)
    if max_dist is None:
        max_dist = float(bins[-1]) if distances is None else np.inf
    ii, jj = pairs_within_distance(pos, max_dist)
    if distances is None:
        dist = pos[jj] - pos[ii]
    else:
        dist = np.abs(distances[jj] - distances[ii])
    key = bin_index(dist, bins)
    cl, nl = [], []
    for rows in pop_rows:
        if use_genotypes:
            c, nv = geno_counts(data, ii, jj, rows)
        else:
            c, nv = hap_counts(data, ii, jj, rows)
        cl.append(c)
        nl.append(nv)
    if use_genotypes:
        stats = multi_pop_stats_geno(cl, nl)
    elif P == 2:
        stats = two_pop_stats(cl[0], cl[1], nl[0], nl[1])
    else:
        stats = multi_pop_stats_hap(cl, nl)
    return keyed_sums(key, np.abs(stats) if absolute else stats, n_bins)


# ---------------------------------------------------------------------------
# Rogers-Huff r  (ld_statistics.py:366-521)
# ---------------------------------------------------------------------------


def rogers_huff_r(dosage):
    """Condensed upper-triangle Pearson r of (n_ind, m) dosages, NaN where a
    column is constant (ld_statistics.py:409-521).  Parity with
    scikit-allel.rogers_huff_r is unpinned here (closed form only)."""
    g = np.asarray(dosage, dtype=np.float64)
    n = g.shape[0]
    mu = g.mean(axis=0)
    ssd = ((g - mu) ** 2).sum(axis=0)
    cov = g.T @ g - n * np.outer(mu, mu)
    den = np.sqrt(np.outer(ssd, ssd))
    with np.errstate(divide="ignore", invalid="ignore"):
        rr = np.where(den > 0, cov / np.where(den > 0, den, 1.0), np.nan)
    iu = np.triu_indices(g.shape[1], k=1)
    return rr[iu]


# ---------------------------------------------------------------------------
# Dense r^2 / sigma_D^2 matrices, ZnS and Omega  (ld_statistics.py:249-862,
# haplotype_matrix.py:1092-1170)
# ---------------------------------------------------------------------------


def segregating_sites(hap):
    """Columns with 0 < derived count < valid count (ld_statistics.py:263-270)."""
    h = np.asarray(hap)
    dac = np.maximum(h, 0).astype(np.int64).sum(axis=0)
    nv = (h >= 0).sum(axis=0)
    return np.nonzero((dac > 0) & (dac < nv))[0]


def _naive_d_and_p(hap):
    """_pairwise_ld_core (haplotype_matrix.py:1092-1130): per-site p over
    every valid sample, p_AB over the pair's valid samples (0 where none),
    D = p_AB - p_i p_j."""
    h = np.asarray(hap)
    v = (h >= 0).astype(np.float64)
    hc = np.where(h >= 0, h, 0).astype(np.float64)
    nv = v.sum(axis=0)
    p = np.where(nv > 0, hc.sum(axis=0) / np.where(nv > 0, nv, 1.0), 0.0)
    jn = v.T @ v
    j11 = hc.T @ hc
    pab = np.where(jn > 0, j11 / np.where(jn > 0, jn, 1.0), 0.0)
    return pab - np.outer(p, p), p


def naive_d_matrix(hap):
    """pairwise_LD_v (haplotype_matrix.py:1132-1136): the D matrix of
    _pairwise_ld_core with a zero diagonal."""
    D, _ = _naive_d_and_p(hap)
    np.fill_diagonal(D, 0.0)
    return D


def naive_r2_matrix(hap):
    """pairwise_r2 (haplotype_matrix.py:1092-1170): per-site p over every
    valid sample, p_AB over the pair's valid samples, D = p_AB - p_i p_j,
    r^2 = D^2 / (p_i q_i p_j q_j), 0 where that is not > 0; diagonal 0."""
    D, p = _naive_d_and_p(hap)
    den = np.outer(p * (1 - p), p * (1 - p))
    with np.errstate(divide="ignore", invalid="ignore"):
        r2 = np.where(den > 0, (D ** 2) / np.where(den > 0, den, 1.0), 0.0)
    np.fill_diagonal(r2, 0.0)
    return r2


def sigma_d2_matrix(hap):
    """Full sigma_D^2 matrix and its valid mask (n >= 4), diagonal excluded
    (ld_statistics.py:734-758)."""
    h = np.asarray(hap)
    m = h.shape[1]
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    c, _ = hap_counts(h, ii.ravel(), jj.ravel(), None)
    sig, valid = sigma_d2_hap_from_counts(c)
    sig = sig.reshape(m, m)
    valid = valid.reshape(m, m)
    np.fill_diagonal(sig, 0.0)
    np.fill_diagonal(valid, False)
    return sig, valid


def zns_from_matrix(M, valid=None):
    """ZnS: mean over pairs i != j; with a valid mask, over valid pairs only
    (ld_statistics.py:534-731)."""
    M = np.asarray(M, dtype=np.float64)
    m = M.shape[0]
    if m < 2:
        return 0.0
    if valid is None:
        return float((M.sum() - np.trace(M)) / (m * (m - 1)))
    n = int(valid.sum())
    return float(M[valid].sum() / n) if n else 0.0


def omega_from_matrix(M):
    """Kim and Nielsen's Omega over the upper triangle (ld_statistics.py:
    761-862): the largest, over split points l = 3 .. m-2, of mean within /
    mean across where the split is valid (cross sum > 0) and 0.0 where it is
    not -- the maximum of those values, so a negative Omega (sigma_D^2 sums
    can be negative) when every split is valid and negative; by explicit
    block sums (no prefix sums)."""
    M = np.triu(np.asarray(M, dtype=np.float64), k=1)
    m = M.shape[0]
    if m < 5:
        return 0.0
    vals = []
    for l in range(3, m - 1):
        left = M[:l, :l].sum()
        right = M[l:, l:].sum()
        cross = M[:l, l:].sum()
        n_w = l * (l - 1) // 2 + (m - l) * (m - l - 1) // 2
        n_c = l * (m - l)
        ok = n_w > 0 and n_c > 0 and cross > 0
        vals.append(((left + right) / n_w) / (cross / n_c) if ok else 0.0)
    return float(max(vals))


def dosage_r2_matrix(geno):
    """_r2_matrix_diploid (ld_statistics.py:975-1021): dosages centred on
    each site's mean over its valid samples, missing entries zeroed after
    centring, r^2 = cov^2 / (var_i var_j), 0 where that denominator is 0;
    diagonal 0."""
    g = np.asarray(geno)
    v = (g >= 0).astype(np.float64)
    gc = np.where(g >= 0, g, 0).astype(np.float64)
    nv = v.sum(axis=0)
    mean = np.where(nv > 0, gc.sum(axis=0) / np.where(nv > 0, nv, 1.0), 0.0)
    gn = (gc - mean[None, :]) * v
    var = (gn ** 2).sum(axis=0)
    cov = gn.T @ gn
    den = np.sqrt(np.outer(var, var))
    with np.errstate(divide="ignore", invalid="ignore"):
        r2 = np.where(den > 0, (cov / np.where(den > 0, den, 1.0)) ** 2, 0.0)
    np.fill_diagonal(r2, 0.0)
    return r2


def locate_unlinked(hap, size=100, step=20, threshold=0.1):
    """Greedy sliding-window LD pruning of haplotype_matrix.py:1172-1231 on
    the naive r^2 of the window's still-unlinked variants."""
    h = np.asarray(hap)
    m = h.shape[1]
    loc = np.ones(m, dtype=bool)
    for w_start in range(0, m, step):
        w_end = min(w_start + size, m)
        active = np.nonzero(loc[w_start:w_end])[0] + w_start
        if len(active) <= 1:
            continue
        r2 = naive_r2_matrix(h[:, active])
        for i in range(len(active)):
            if not loc[active[i]]:
                continue
            for j in range(i + 1, len(active)):
                if loc[active[j]] and r2[i, j] > threshold:
                    loc[active[j]] = False
    return loc


# ---------------------------------------------------------------------------
# Per-window ZnS and Omega  (windowed_analysis.py:1658-1667, :2013-2058;
# ld_statistics.py:595-676)
# ---------------------------------------------------------------------------


def window_ranges(positions, starts_bp, stops_bp):
    """Variant index range [start, stop) of each window: searchsorted 'left'
    of the window's start and end in bp (windowed_analysis.py:1658-1667,
    :1740-1749), i.e. the variants with start <= pos < end."""
    pos = np.asarray(positions, dtype=np.float64)
    return (np.searchsorted(pos, np.asarray(starts_bp, dtype=np.float64), side="left"),
            np.searchsorted(pos, np.asarray(stops_bp, dtype=np.float64), side="left"))


def windowed_zns_omega(hap, positions, starts_bp, stops_bp, missing_data="include"):
    """The 'zns' and 'omega' columns of windowed_statistics_fused
    (windowed_analysis.py:2013-2058): NaN for windows of fewer than 4
    variants; otherwise ZnS = _zns_from_precomputed(..., use_projection=True)
    over the window's columns (segregating filter, mean sigma_D^2 over the
    valid pairs, 0.0 below 2 sites or without valid pairs), and Omega =
    ld_statistics.omega(window matrix, missing_data) with its default
    sigma_D^2 estimator (0.0 below 5 sites)."""
    h = np.asarray(hap)
    s, e = window_ranges(positions, starts_bp, stops_bp)
    zns = np.full(len(s), np.nan)
    om = np.full(len(s), np.nan)
    for w in range(len(s)):
        if e[w] - s[w] < 4:
            continue
        win = h[:, s[w]:e[w]]
        seg = segregating_sites(win)
        if len(seg) < 2:
            zns[w] = 0.0
        else:
            sig, valid = sigma_d2_matrix(win[:, seg])
            zns[w] = zns_from_matrix(sig, valid)
        wo = win
        if missing_data == "exclude":
            wo = win[:, np.nonzero((win >= 0).all(axis=0))[0]]
        seg = segregating_sites(wo)
        om[w] = omega_from_matrix(sigma_d2_matrix(wo[:, seg])[0]) if len(seg) >= 5 else 0.0
    return zns, om
