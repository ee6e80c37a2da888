"""CPU pins of the oracle: brute force, closed forms, the reference's known
answers and the frozen moments output.  No GPU."""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import ld_oracle as O
import ustat_bruteforce as U

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# pair enumeration (tests/test_ld_pair_iteration.py:21-115)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [2, 10, 50])
@pytest.mark.parametrize("factor", [0.1, 0.5, 1.0, 10.0])
def test_pairs_bruteforce(n, factor):
    rng = np.random.default_rng(n * 997 + int(factor * 1000))
    pos = np.unique(np.sort(rng.integers(0, 10 * n, size=n)))
    if len(pos) < 2:
        pytest.skip("collapsed")
    md = max(1, int((pos[-1] - pos[0]) * factor))
    ii, jj = O.pairs_within_distance(pos, md)
    exp = {(i, j) for i in range(len(pos)) for j in range(i + 1, len(pos)) if pos[j] - pos[i] <= md}
    assert set(zip(ii.tolist(), jj.tolist())) == exp
    assert len(ii) == len(exp)
    assert np.all(np.diff(ii) >= 0)


def test_pairs_edges_and_chunks():
    assert O.pairs_within_distance(np.array([]), 5)[0].size == 0
    assert O.pairs_within_distance(np.array([1]), 5)[0].size == 0
    assert O.pairs_within_distance(np.arange(10), 0)[0].size == 0
    ii, _ = O.pairs_within_distance(np.arange(25) * 10, 10**9)
    assert ii.size == 25 * 24 // 2
    rng = np.random.default_rng(12345)
    pos = np.unique(np.sort(rng.integers(0, 100_000, size=500)))
    total = O.pairs_within_distance(pos, 5000)[0].size
    for cs in (1, 100, 10_000, 1_000_000):
        b = O.chunk_bounds(pos, 5000, cs)
        per = np.searchsorted(pos, pos + 5000, side="right") - np.arange(len(pos)) - 1
        assert sum(int(np.maximum(per, 0)[lo:hi].sum()) for lo, hi in b) == total


# ---------------------------------------------------------------------------
# counts
# ---------------------------------------------------------------------------
def test_hap_counts_kat():
    # tests/test_haplotype_counting.py:56-112
    hap = np.array([[1, 0, 1, 0, 1, 1, 0, 0], [1, 1, 1, 0, 0, 0, 1, 0]], dtype=np.int8).T
    c, nv = O.hap_counts(hap, [0], [1], [0, 1, 4, 5])
    np.testing.assert_array_equal(c[0], [1, 2, 1, 0])
    assert nv is None
    c, _ = O.hap_counts(hap, [0], [1], [2, 3, 6, 7])
    np.testing.assert_array_equal(c[0], [1, 0, 1, 2])


def test_counts_definition_bruteforce():
    rng = np.random.default_rng(2)
    h = rng.integers(-1, 2, size=(23, 12)).astype(np.int8)
    g = rng.integers(-1, 3, size=(19, 12)).astype(np.int8)
    ii, jj = np.triu_indices(12, 1)
    ch, nvh = O.hap_counts(h, ii, jj)
    cg, nvg = O.geno_counts(g, ii, jj)
    for p, (i, j) in enumerate(zip(ii, jj)):
        exp = [0, 0, 0, 0]
        for s in range(23):
            a, b = h[s, i], h[s, j]
            if a >= 0 and b >= 0:
                exp[{(1, 1): 0, (1, 0): 1, (0, 1): 2, (0, 0): 3}[(a, b)]] += 1
        assert list(ch[p]) == exp and nvh[p] == sum(exp)
        e9 = [0] * 9
        for s in range(19):
            a, b = g[s, i], g[s, j]
            if a >= 0 and b >= 0:
                e9[3 * a + b] += 1
        assert list(cg[p]) == e9 and nvg[p] == sum(e9)
    # symmetry i <-> j swaps n10 and n01
    cs, _ = O.hap_counts(h, jj, ii)
    np.testing.assert_array_equal(cs[:, [0, 2, 1, 3]], ch)


# ---------------------------------------------------------------------------
# r family KATs (tests/test_ld_r_squared.py:15-57, :171-230)
# ---------------------------------------------------------------------------
def test_r_family_kats():
    c = np.array([[5, 0, 0, 5], [25, 25, 25, 25], [0, 5, 5, 0], [5, 5, 0, 0], [28, 52, 2, 18]])
    r2 = O.r_squared(c)
    np.testing.assert_allclose(r2[:3], [1.0, 0.0, 1.0], atol=1e-12)
    assert np.isnan(r2[3])
    np.testing.assert_allclose(O.r(c)[2], -1.0, rtol=1e-12)
    dp = O.d_prime(c)
    np.testing.assert_allclose(dp[[0, 1, 2, 4]], [1.0, 0.0, -1.0, 2.0 / 3.0], atol=1e-12)
    assert np.isnan(dp[3])


# ---------------------------------------------------------------------------
# DD / Dz / pi2 against brute-force U-statistics (every population case)
# ---------------------------------------------------------------------------
def _calls(P):
    out = set()
    for _, terms in O.stat_specs(P):
        for _, t, p in terms:
            out.add((t, p))
    return sorted(out)


def _hap_triple_normalised(p):
    """The reference evaluates every pi2 with one population three times in
    its (t, t | t, s) form (haplotype_kernels.py:268-289, ld_statistics.py:
    1323-1328): the U-statistic of that form is the pin."""
    cnt = {x: p.count(x) for x in set(p)}
    if len(cnt) == 2 and max(cnt.values()) == 3:
        t = [x for x, c in cnt.items() if c == 3][0]
        s = [x for x, c in cnt.items() if c == 1][0]
        return (t, t, t, s)
    return p


@pytest.mark.parametrize("P,seed", [(3, s) for s in range(5)] + [(4, s) for s in range(3)])
def test_hap_stats_bruteforce(P, seed):
    """Every (statistic, populations) call the moments specs make, up to four
    populations: P = 4 reaches the pi2 case with four distinct populations
    (haplotype_kernels.py:167-304, ld_statistics.py:1299-1617)."""
    rng = np.random.default_rng(seed + 1000 * (P - 3))
    pops = [rng.integers(0, 2, size=(rng.integers(4, 6), 2)) for _ in range(P)]
    H = [O.HapPop(U.hap_counts_of(p)[None, :]) for p in pops]
    for t, p in _calls(P):
        if t == "dd":
            got = O.hap_dd(H, *p)
        elif t == "dz":
            got = O.hap_dz(H, *p)
        else:
            got = O.hap_pi2_kernel(H, *p)
            p = _hap_triple_normalised(p)
        exp = U.hap_stat(pops, t, p)
        assert abs(float(got[0]) - exp) <= 1e-12 * max(1.0, abs(exp)), (t, p)


@pytest.mark.parametrize("P,seed", [(3, s) for s in range(4)] + [(4, s) for s in range(3)])
def test_geno_stats_bruteforce(P, seed):
    """As above for genotypes; P = 4 reaches the all-distinct pi2 case of
    genotype_kernels.py:279-405."""
    rng = np.random.default_rng(100 + seed + 1000 * (P - 3))
    pops = [rng.integers(0, 3, size=(rng.integers(4, 6), 2)) for _ in range(P)]
    G = [O.GenoPop(U.geno_counts_of(p)[None, :]) for p in pops]
    fn = {"dd": O.geno_dd, "dz": O.geno_dz, "pi2": O.geno_pi2}
    for t, p in _calls(P):
        got = float(fn[t](G, *p)[0])
        exp = U.geno_stat(pops, t, p)
        assert abs(got - exp) <= 1e-12 * max(1.0, abs(exp)), (t, p)


def test_hap_pi2_forms_agree_after_averaging():
    rng = np.random.default_rng(9)
    cl = [rng.integers(0, 30, size=(200, 4)) for _ in range(2)]
    a = O.two_pop_stats(cl[0], cl[1])
    b = O.multi_pop_stats_hap(cl, [None, None])
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def test_sigma_d2_geno_matches_ratio():
    rng = np.random.default_rng(5)
    c = rng.integers(0, 20, size=(300, 9))
    G = [O.GenoPop(c)]
    ratio = O.geno_dd(G, 0, 0) / O.geno_pi2(G, 0, 0, 0, 0)
    got = O.sigma_d2_geno_from_counts(c)
    ok = np.isfinite(got)
    np.testing.assert_allclose(got[ok], ratio[ok], rtol=1e-12)


def test_sigma_d2_hap_tile_matches_dd_over_pi2():
    rng = np.random.default_rng(6)
    c = rng.integers(0, 20, size=(300, 4))
    P = [O.HapPop(c)]
    s, valid = O.sigma_d2_hap_from_counts(c)
    ratio = O.hap_dd(P, 0, 0) / O.hap_pi2_kernel(P, 0, 0, 0, 0)
    ok = valid & np.isfinite(ratio) & (O.hap_pi2_kernel(P, 0, 0, 0, 0) != 0)
    np.testing.assert_allclose(s[ok], ratio[ok], rtol=1e-12)


# ---------------------------------------------------------------------------
# names / specs (ld_pipeline.py:357-457)
# ---------------------------------------------------------------------------
def test_names_counts_and_spec_weights():
    assert [len(O.ld_names(p)) for p in (1, 2, 3, 4)] == [3, 15, 45, 105]
    assert O.ld_names(2) == ["DD_0_0", "DD_0_1", "DD_1_1", "Dz_0_0_0", "Dz_0_0_1", "Dz_0_1_1", "Dz_1_0_0",
                             "Dz_1_0_1", "Dz_1_1_1", "pi2_0_0_0_0", "pi2_0_0_0_1", "pi2_0_0_1_1",
                             "pi2_0_1_0_1", "pi2_0_1_1_1", "pi2_1_1_1_1"]
    for P in (1, 2, 3, 4):
        for _, terms in O.stat_specs(P):
            assert abs(sum(w for w, _, _ in terms) - 1.0) < 1e-15


# ld_names(3) written out by hand from moments' naming rule: DD over unordered
# population pairs, Dz_i over unordered (j, k), pi2 over unordered pairs of
# unordered pairs (locus-A pair, locus-B pair), each family in lexicographic
# order of its smallest representative.
_LD_NAMES_3 = """
DD_0_0 DD_0_1 DD_0_2 DD_1_1 DD_1_2 DD_2_2
Dz_0_0_0 Dz_0_0_1 Dz_0_0_2 Dz_0_1_1 Dz_0_1_2 Dz_0_2_2
Dz_1_0_0 Dz_1_0_1 Dz_1_0_2 Dz_1_1_1 Dz_1_1_2 Dz_1_2_2
Dz_2_0_0 Dz_2_0_1 Dz_2_0_2 Dz_2_1_1 Dz_2_1_2 Dz_2_2_2
pi2_0_0_0_0 pi2_0_0_0_1 pi2_0_0_0_2 pi2_0_0_1_1 pi2_0_0_1_2 pi2_0_0_2_2
pi2_0_1_0_1 pi2_0_1_0_2 pi2_0_1_1_1 pi2_0_1_1_2 pi2_0_1_2_2
pi2_0_2_0_2 pi2_0_2_1_1 pi2_0_2_1_2 pi2_0_2_2_2
pi2_1_1_1_1 pi2_1_1_1_2 pi2_1_1_2_2 pi2_1_2_1_2 pi2_1_2_2_2 pi2_2_2_2_2
""".split()


def _orbit_rule_names(P):
    """The naming rule as orbits: a statistic is one name per orbit of its
    symmetry group, named by the orbit's lexicographically smallest tuple."""
    pairs = [(i, j) for i in range(P) for j in range(i, P)]
    dd = sorted(pairs)
    dz = sorted((i,) + jk for i in range(P) for jk in pairs)
    pi2 = sorted({min(a, b) + max(a, b) for a in pairs for b in pairs})
    return (["DD_%d_%d" % t for t in dd] + ["Dz_%d_%d_%d" % t for t in dz]
            + ["pi2_%d_%d_%d_%d" % t for t in pi2])


def _group_images(kind, t):
    """Images of a population tuple under the statistic's symmetry group:
    Dz_i_j_k is symmetric in (j, k); pi2_i_j_k_l in i <-> j, k <-> l and in
    exchanging the (i, j) and (k, l) pairs (moments averages over the group;
    DD_i_j is taken as one call because DD(i, j) = DD(j, i) itself)."""
    if kind == "DD":
        return [t]
    if kind == "Dz":
        i, j, k = t
        return [(i, j, k), (i, k, j)]
    i, j, k, l = t
    out = []
    for a in ((i, j), (j, i)):
        for b in ((k, l), (l, k)):
            out += [a + b, b + a]
    return out


def test_ld_names_literal_and_orbit_rule():
    assert O.ld_names(3) == _LD_NAMES_3
    for P in (1, 2, 3, 4):
        assert O.ld_names(P) == _orbit_rule_names(P)
        assert O.het_names(P) == ["H_%d_%d" % (i, j) for i in range(P) for j in range(i, P)]
    # 4 populations: 10 DD + 40 Dz + 55 pi2 (55 = unordered pairs of the 10 unordered pairs)
    n4 = O.ld_names(4)
    assert [sum(x.startswith(k) for x in n4) for k in ("DD", "Dz", "pi2")] == [10, 40, 55]
    assert len(set(n4)) == 105


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_stat_specs_are_group_averages(P):
    """Each spec is the uniform average of its statistic over the symmetry
    group images of the name's tuple (ld_pipeline.py:388-457 restates moments'
    averaging); weights are accumulated per distinct call so that repeated
    images (e.g. pi2_0_1_0_1) compare as their total weight."""
    kinds = {"DD": "dd", "Dz": "dz", "pi2": "pi2"}
    for name, terms in O.stat_specs(P):
        kind, *rest = name.split("_")
        t = tuple(int(x) for x in rest)
        imgs = _group_images(kind, t)
        want = {}
        for im in imgs:
            want[(kinds[kind], im)] = want.get((kinds[kind], im), 0.0) + 1.0 / len(imgs)
        got = {}
        for w, k, p in terms:
            got[(k, p)] = got.get((k, p), 0.0) + w
        assert got.keys() == want.keys(), name
        for key in got:
            assert got[key] == pytest.approx(want[key], abs=1e-15), (name, key)
    # DD between populations is symmetric, so one call stands for the orbit
    rng = np.random.default_rng(3)
    H = [O.HapPop(rng.integers(0, 9, size=(50, 4))) for _ in range(2)]
    np.testing.assert_array_equal(O.hap_dd(H, 0, 1), O.hap_dd(H, 1, 0))


def test_union_biallelic_keep_geno_known_answers():
    """moments_ld.py:256-285: over the union of the populations' individuals
    (each counted once), keep a site when 0 < alt < 2 n_valid and
    n_valid >= 2."""
    g = np.array([
        # s0 all ref, s1 all alt, s2 one valid, s3 poly, s4 poly only via pop B,
        # s5 alt only in a duplicated row, s6 all missing, s7 het everywhere
        [0, 2, -1, 0, 0, 0, -1, 1],   # row 0: pop A
        [0, 2, -1, 1, 0, 0, -1, 1],   # row 1: pop A and pop B (duplicate)
        [0, 2, 1, 2, 0, 0, -1, 1],    # row 2: pop A
        [0, 2, -1, 0, 1, 0, -1, 1],   # row 3: pop B
        [0, 2, -1, 0, 0, 2, -1, 1],   # row 4: in no population
    ], dtype=np.int8)
    keep = O.union_biallelic_keep_geno(g, [[0, 1, 2], [1, 3]])
    np.testing.assert_array_equal(keep, [3, 4, 7])
    # a duplicate row is counted once: alt 2 of n_valid 1 -> dropped
    g2 = np.array([[2, 0], [-1, 1], [-1, 0]], dtype=np.int8)
    np.testing.assert_array_equal(O.union_biallelic_keep_geno(g2, [[0, 1], [0, 2]]), [1])
    # two valid individuals, one het: 0 < 1 < 4 -> kept
    g3 = np.array([[1], [0], [-1]], dtype=np.int8)
    np.testing.assert_array_equal(O.union_biallelic_keep_geno(g3, [[0], [1, 2]]), [0])


def _allele_slots(col, use_genotypes):
    """The alleles a site contributes: each valid individual's dosage g as g
    alternative and 2 - g reference alleles; a haplotype row as one allele,
    with missing (-1) counted as reference over the constant |pop| of
    moments_ld.py:366-368."""
    out = []
    for x in col:
        if use_genotypes:
            if x >= 0:
                out += [1] * int(x) + [0] * (2 - int(x))
        else:
            out.append(1 if x == 1 else 0)
    return out


@pytest.mark.parametrize("use_genotypes", [False, True])
def test_heterozygosity_bruteforce_allele_pairs(use_genotypes):
    """H_i_i is the probability that two distinct alleles drawn from
    population i differ, H_i_j that an allele of i and one of j differ, both
    summed over sites (moments_ld.py:346-388); enumerated pair by pair."""
    rng = np.random.default_rng(44 + use_genotypes)
    hi = 3 if use_genotypes else 2
    mats = [rng.integers(-1, hi, size=(n, 6)).astype(np.int8) for n in (3, 4, 5)]
    for M in mats:
        M[0, :] = np.maximum(M[0, :], 0)
        M[1, :] = np.maximum(M[1, :], 0)     # at least two valid rows per site
    got = O.heterozygosity(mats, use_genotypes=use_genotypes)
    for a in range(3):
        for b in range(a, 3):
            tot = 0.0
            for s in range(6):
                A = _allele_slots(mats[a][:, s], use_genotypes)
                B = _allele_slots(mats[b][:, s], use_genotypes)
                if a == b:
                    diff = sum(A[x] != A[y] for x in range(len(A)) for y in range(len(A)) if x != y)
                    tot += diff / (len(A) * (len(A) - 1))
                else:
                    tot += sum(x != y for x in A for y in B) / (len(A) * len(B))
            assert got[f"H_{a}_{b}"] == pytest.approx(tot, rel=1e-13, abs=1e-15), (a, b)


# ---------------------------------------------------------------------------
# the frozen moments output on the IM example (Appendix-style golden vector)
# ---------------------------------------------------------------------------
def _im_example():
    from pg_gpu_amd import vcf
    d = vcf.read_vcf(os.path.join(GOLD, "im-parsing-example.vcf"))
    gt, pos, samples = d["gt"], d["positions"], d["samples"]
    ns = len(samples)
    hap = np.concatenate([gt[:, :, 0], gt[:, :, 1]], axis=1).T
    pm = vcf.read_pop_file(os.path.join(GOLD, "im_pop.txt"))
    rows = {"deme0": [], "deme1": []}
    for i, s in enumerate(samples):
        if pm.get(s) in rows:
            rows[pm[s]] += [i, i + ns]
    return hap, pos, [rows["deme0"], rows["deme1"]]


def test_moments_golden_im_example():
    G = json.load(open(os.path.join(GOLD, "moments_im_4bins.json")))
    hap, pos, pops = _im_example()
    keep = O.biallelic_filter_hap(hap)
    hap, pos = hap[:, keep], pos[keep]
    bins = np.array([G["bins"][0][0]] + [b[1] for b in G["bins"]])
    sums = O.moments_ld_sums(hap, pos, pops, bins)
    np.testing.assert_allclose(sums, np.array(G["sums"]), rtol=1e-10)
    H = O.heterozygosity([hap[r] for r in pops])
    np.testing.assert_allclose([H[k] for k in G["het_names"]], G["het"], rtol=1e-13)
    assert O.ld_names(2) == G["ld_names"]


def test_keyed_sums_bruteforce():
    rng = np.random.default_rng(1)
    keys = rng.integers(-2, 7, size=500)
    vals = rng.normal(size=(500, 3))
    got = O.keyed_sums(keys, vals, 5)
    for k in range(5):
        np.testing.assert_allclose(got[k], vals[keys == k].sum(axis=0), rtol=1e-13)
    # digitize semantics: [b_k, b_{k+1}); the last edge itself is outside
    e = np.array([1.0, 10.0, 100.0])
    np.testing.assert_array_equal(O.bin_index([0.5, 1.0, 9.99, 10.0, 99.9, 100.0], e), [-1, 0, 0, 1, 1, 2])


def test_r_squared_exact_matches_reference_form():
    """The exact-integer r^2 of the dense path equals the reference's float
    formula up to rounding, on random counts with and without monomorphic
    margins, and is 1 for identical and 0 for independent-by-construction
    columns."""
    rng = np.random.default_rng(11)
    c = rng.integers(0, 60, size=(4000, 4))
    c[:50, 1:3] = 0          # perfect association -> r^2 = 1
    c[50:80, [0, 1]] = 0     # site A monomorphic -> NaN
    got = O.r_squared_exact(c)
    want = O.r_squared(c)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-13, atol=1e-300)
    assert np.all(got[:50][~np.isnan(got[:50])] == 1.0)
    # n11 * n00 == n10 * n01 -> D = 0 -> r^2 = 0 exactly
    z = np.array([[6, 3, 4, 2], [10, 5, 2, 1]])
    np.testing.assert_array_equal(O.r_squared_exact(z), [0.0, 0.0])
    h, nan = O.r2_histogram(np.array([0.0, 0.0099, 0.01, 0.5, 1.0, np.nan, 0.3]),
                            np.array([0, 0, 0, 1, 1, 1, -1]), 2, 100)
    assert h[0, 0] == 2 and h[0, 1] == 1 and h[1, 50] == 1 and h[1, 99] == 1 and list(nan) == [0, 1]


def test_naive_r2_and_sigma_matrices_pins():
    """Dense matrices of the ZnS / Omega path: without missing data the naive
    r^2 equals the count-based r^2 (both are D^2 / (p q p q) on the same
    frequencies) and sigma_D^2 equals the tile ratio of the pinned DD and
    pi_2 numerators; Omega's explicit block sums are checked by brute force
    over pairs."""
    rng = np.random.default_rng(17)
    h = (rng.random((40, 12)) < rng.uniform(0.1, 0.9, size=12)).astype(np.int8)
    R = O.naive_r2_matrix(h)
    ii, jj = np.triu_indices(12, k=1)
    c, _ = O.hap_counts(h, ii, jj)
    w = O.r_squared(c)
    np.testing.assert_allclose(R[ii, jj], np.where(np.isnan(w), 0.0, w), rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(R, R.T)
    # the D matrix: without missing data the count D (n11 n00 - n10 n01) / n^2
    Dm = O.naive_d_matrix(h)
    np.testing.assert_allclose(Dm[ii, jj], (c[:, 0] * c[:, 3] - c[:, 1] * c[:, 2]) / 40.0 ** 2, rtol=1e-12,
                               atol=1e-15)
    np.testing.assert_array_equal(Dm, Dm.T)
    assert np.all(np.diag(Dm) == 0.0)
    S, valid = O.sigma_d2_matrix(h)
    s2, _ = O.sigma_d2_hap_from_counts(c)
    np.testing.assert_allclose(S[ii, jj], s2, rtol=1e-13, atol=1e-15)
    assert O.zns_from_matrix(R) == pytest.approx(2 * R[ii, jj].sum() / (12 * 11))
    # Omega by brute force over pairs
    m = 12
    best = 0.0
    for l in range(3, m - 1):
        w_s = w_n = c_s = c_n = 0
        for a, b in zip(ii, jj):
            if (a < l) == (b < l):
                w_s += R[a, b]
                w_n += 1
            else:
                c_s += R[a, b]
                c_n += 1
        if c_s > 0:
            best = max(best, (w_s / w_n) / (c_s / c_n))
    assert O.omega_from_matrix(R) == pytest.approx(best, rel=1e-12)
    assert list(O.segregating_sites(np.array([[0, 1, -1], [0, 0, 1], [0, 1, 1]]))) == [1]


def test_dosage_r2_equals_rogers_huff_without_missing():
    """Without missing data the dosage r^2 of the ZnS path is the square of
    the Rogers-Huff r (both are Pearson correlations of the same dosages),
    and Rogers-Huff r of a site with itself shifted is 1."""
    rng = np.random.default_rng(23)
    g = rng.integers(0, 3, size=(30, 9)).astype(np.int8)
    g[:, 4] = 1  # a constant site: r undefined, r^2 reported as 0
    r = O.rogers_huff_r(g)
    R2 = O.dosage_r2_matrix(g)
    ii, jj = np.triu_indices(9, k=1)
    np.testing.assert_allclose(R2[ii, jj], np.nan_to_num(r ** 2), rtol=1e-12, atol=1e-15)
    h = np.column_stack([g[:, 0], g[:, 0]])
    assert O.rogers_huff_r(h)[0] == pytest.approx(1.0)


def test_genotype_numerators_centred_moment_identity():
    """The HIP path (ld_math.h geno_dd_num / geno_dz_num / geno_pi2_num)
    evaluates the single-population genotype numerators in the moments of the
    centred dosages a' = a - 1, b' = b - 1 instead of the reference's
    monomial tables (genotype_kernels.py:27-71, :199-233, :279-300; restated
    in the oracle and pinned there by brute force).  The identities must hold
    exactly; integer counts, exact rational arithmetic."""
    from fractions import Fraction as F
    rng = np.random.default_rng(2024)
    for trial in range(400):
        hi = (4, 40, 3000)[trial % 3]
        g = [int(x) for x in rng.integers(0, hi, 9)]
        col = [np.array([x], dtype=np.int64) for x in g]
        n = sum(g)
        # g[3 ia + ib] has a' = 1 - ia, b' = 1 - ib
        R = [g[3 * a] + g[3 * a + 1] + g[3 * a + 2] for a in range(3)]
        B1 = [g[3 * a] - g[3 * a + 2] for a in range(3)]
        B2 = [g[3 * a] + g[3 * a + 2] for a in range(3)]
        Sb, Sb2 = sum(B1), sum(B2)
        Sa, Sab, Sab2 = R[0] - R[2], B1[0] - B1[2], B2[0] - B2[2]
        Sa2, Sa2b, Sa2b2 = R[0] + R[2], B1[0] + B1[2], B2[0] + B2[2]
        X, Y = Sb * Sb - Sb2, Sa * Sa - Sa2
        sab, lin = Sa * Sb * Sab, Sb * Sa2b + Sa * Sab2
        dd = F((n - 1) * ((n - 2) * Sab * Sab - n * Sa2b2 - 2 * sab + 2 * lin) + X * Y, 4)
        dz = F((n + 1) * (sab - lin) - (n - 1) * Sab * Sab + 2 * n * Sa2b2 - X * Y, 4)
        pi2 = F(n * (n - 1) * (n - 2) * (n - 3) - (n - 2) * (n - 3) * (X + Y) + X * Y - 4 * sab + 4 * lin
                + 2 * Sab * Sab - 6 * Sa2b2, 16)
        # the reference's forms
        xAB = F(g[0]) + F(g[1], 2) + F(g[3], 2) + F(g[4], 4)
        xAb = F(g[1], 2) + g[2] + F(g[4], 4) + F(g[5], 2)
        xaB = F(g[3], 2) + F(g[4], 4) + g[6] + F(g[7], 2)
        xab = F(g[4], 4) + F(g[5], 2) + F(g[7], 2) + g[8]
        D = xAB * xab - xAb * xaB
        assert D == F(n * Sab - Sa * Sb, 4)
        A = -g[0] + g[2] - g[3] + g[5] - g[6] + g[8]
        B = -g[0] - g[1] - g[2] + g[6] + g[7] + g[8]
        assert (A, B) == (-Sb, -Sa)
        pA = g[0] + g[1] + g[2] + F(g[3] + g[4] + g[5], 2)
        qA = F(g[3] + g[4] + g[5], 2) + g[6] + g[7] + g[8]
        pB = g[0] + F(g[1], 2) + g[3] + F(g[4], 2) + g[6] + F(g[7], 2)
        qB = F(g[1], 2) + g[2] + F(g[4], 2) + g[5] + F(g[7], 2) + g[8]
        assert dd == 4 * (F(int(O.eval_poly(O._DD_POLY, col)[0]), 16) + D * D), g
        assert dz == F(int(O.eval_poly(O._DZ_QUARTIC, col)[0]), 4) + A * B * D, g
        assert pi2 == pA * pB * qB * qA + F(int(O.eval_poly(O._PI2_CORR, col)[0]), 16), g


def test_haplotype_numerators_centred_moment_identity():
    """The HIP path (ld_math.h num_dd_single ... num_pi2_ijij) evaluates the one-
    and two-population haplotype numerators in the centred moments n, Sa, Sb,
    Sab (a' = 2a - 1, b' = 2b - 1) instead of the reference's count polynomials
    (restated in the oracle, pinned there by brute force).  For counts this
    small every numerator and denominator is exact in float64, so the
    statistics must agree exactly."""
    rng = np.random.default_rng(77)
    c = [rng.integers(0, 40, size=(500, 4)) for _ in range(2)]
    c[0][:7] = [[4, 0, 0, 0], [0, 0, 0, 5], [1, 1, 1, 1], [9, 0, 0, 3], [0, 6, 2, 0], [3, 3, 3, 3], [2, 0, 1, 1]]
    c = [x[x.sum(1) >= 4] for x in c]
    m = min(len(c[0]), len(c[1]))
    c = [x[:m] for x in c]
    P = [O.HapPop(x) for x in c]

    def mom(p):
        return p.n, p.pA - p.qA, p.pB - p.qB, (p.c1 + p.c4) - (p.c2 + p.c3), p.D

    def ff(n, k):
        out = np.ones_like(n)
        for t in range(k):
            out = out * (n - t)
        return out

    (n0, a0, b0, s0, D0), (n1, a1, b1, s1, D1) = mom(P[0]), mom(P[1])
    t0 = a0 * b0 * s0
    dd = 0.0625 * ((n0 - 1) * ((n0 - 2) * s0 * s0 - 2 * t0) + (n0 - 2) * (a0 * a0 + b0 * b0 - n0 * n0)
                   + a0 * a0 * b0 * b0)
    dz = 0.25 * ((n0 + 1) * t0 - (n0 - 1) * s0 * s0 + n0 * n0 - a0 * a0 - b0 * b0 - a0 * a0 * b0 * b0)
    pi2 = 0.0625 * ((n0 * n0 - 4 * n0 + 2) * (n0 * n0 - a0 * a0 - b0 * b0) + a0 * a0 * b0 * b0 - 4 * t0
                    + 2 * s0 * s0)
    f4 = ff(n0, 4)
    np.testing.assert_array_equal(dd / f4, O.hap_dd(P, 0, 0))
    np.testing.assert_array_equal(dz / f4, O.hap_dz(P, 0, 0, 0))
    np.testing.assert_array_equal(pi2 / f4, O.hap_pi2_kernel(P, 0, 0, 0, 0))
    np.testing.assert_array_equal(-(a0 * b1) * D0 / (n1 * ff(n0, 3)), O.hap_dz(P, 0, 0, 1))
    np.testing.assert_array_equal(-(a1 * b0) * D0 / (n1 * ff(n0, 3)), O.hap_dz(P, 0, 1, 0))
    np.testing.assert_array_equal(-D0 * (a1 * b1 - s1) / (ff(n0, 2) * ff(n1, 2)), O.hap_dz(P, 0, 1, 1))
    tri = 0.0625 * ((n0 - 2) * (n0 * (n0 * n1 - b0 * b1) - n1 * a0 * a0) + (a0 * b1) * (a0 * b0 - 2 * s0))
    np.testing.assert_array_equal(tri / (n1 * ff(n0, 3)), O.hap_pi2_kernel(P, 0, 0, 0, 1))
    sym = 0.03125 * (2 * n0 * n0 * n1 * n1 - n0 * n0 * (a1 * a1 + b1 * b1) - n1 * n1 * (a0 * a0 + b0 * b0)
                     + (a0 * b1) ** 2 + (a1 * b0) ** 2)
    np.testing.assert_allclose(sym / (ff(n0, 2) * ff(n1, 2)),
                               0.5 * (O.hap_pi2_kernel(P, 0, 0, 1, 1) + O.hap_pi2_kernel(P, 1, 1, 0, 0)),
                               rtol=1e-15, atol=0)
    Pq, Qq = a0 * a1, b0 * b1
    ijij = 0.0625 * ((n0 - 1) * (n1 - 1) * (n0 * n1 - Pq - Qq) + Pq * Qq - (a0 * b0) * s1 - (a1 * b1) * s0
                     + s0 * s1)
    np.testing.assert_array_equal(ijij / (ff(n0, 2) * ff(n1, 2)), O.hap_pi2_kernel(P, 0, 1, 0, 1))


def test_tiled_window_sums_equal_per_pair_oracle():
    """The BLAS-tiled CPU variant (float32 indicator GEMMs + FP64 statistics)
    gives the per-pair oracle's window sums."""
    rng = np.random.default_rng(31)
    n_hap, m = 120, 700
    hap = (rng.random((n_hap, m)) < rng.uniform(0.05, 0.6, m)).astype(np.int8)
    hap[rng.random((n_hap, m)) < 0.03] = -1
    pos = np.sort(rng.choice(np.arange(1, 70_000), m, replace=False)).astype(np.float64)
    pops = [list(range(0, 60)), list(range(60, 120))]
    md, width = 2500.0, 10_000.0
    n_keys = int(pos[-1] // width) + 1
    got = O.tiled_window_sums_hap(hap, pos, pops, md, width, n_keys, block=128)
    ii, jj = O.pairs_within_distance(pos, md)
    cl = [O.hap_counts(hap, ii, jj, r) for r in pops]
    st = O.two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
    want = O.keyed_sums(O.window_key(pos[ii], 0.0, width), st, n_keys)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    part = O.tiled_window_sums_hap(hap, pos, pops, md, width, n_keys, i_lo=100, i_hi=300, block=64)
    sel = (ii >= 100) & (ii < 300)
    np.testing.assert_allclose(part, O.keyed_sums(O.window_key(pos[ii[sel]], 0.0, width), st[sel], n_keys),
                               rtol=1e-12, atol=1e-12)


def test_locate_unlinked_known_answers():
    """Known answers of the greedy window pruning (haplotype_matrix.py:
    1172-1231): duplicates of an earlier kept variant in the same window are
    pruned, unlinked ones kept, r^2 equal to the threshold is kept (strict >),
    a variant pruned in one window is skipped later, and windows only compare
    variants within `size` of each other.  Also the reference's own property
    tests (tests/test_ld_r_squared.py:124-146)."""
    A = [0, 0, 0, 0, 1, 1, 1, 1]
    B = [0, 1, 0, 1, 0, 1, 0, 1]
    C = [0, 0, 1, 1, 0, 0, 1, 1]
    h = np.array([A, A, B, C, B, A], dtype=np.int8).T
    np.testing.assert_array_equal(O.locate_unlinked(h, size=6, step=6, threshold=0.1),
                                  [True, False, True, True, False, False])
    assert O.locate_unlinked(h, size=6, step=6, threshold=1.0).all()
    h3 = np.array([A, A, A], dtype=np.int8).T
    np.testing.assert_array_equal(O.locate_unlinked(h3, size=2, step=1, threshold=0.1), [True, False, True])
    np.testing.assert_array_equal(O.locate_unlinked(h3, size=3, step=1, threshold=0.1), [True, False, False])
    rng = np.random.RandomState(42)
    hap = rng.randint(0, 2, (20, 50)).astype(np.int8)
    loc = O.locate_unlinked(hap, size=10, step=5, threshold=0.1)
    assert loc.shape == (50,) and loc.dtype == bool and 0 < loc.sum() <= 50
    hap = rng.randint(0, 2, (20, 20)).astype(np.int8)
    assert O.locate_unlinked(hap, size=10, step=5, threshold=1.0).all()


def test_window_ranges_and_short_windows():
    """Known answers for the per-window LD columns (windowed_analysis.py:
    1658-1667, :2013-2058): a variant at a bin edge belongs to the window on
    its right, variants at or past the last edge to none; windows of fewer
    than 4 variants are NaN; windows of 4+ variants with fewer than 2
    segregating sites give ZnS 0.0, fewer than 5 give Omega 0.0."""
    pos = np.array([10, 20, 30, 40, 50, 60, 70, 80, 90, 100])
    s, e = O.window_ranges(pos, [10, 30, 100], [30, 100, 200])
    np.testing.assert_array_equal(s, [0, 2, 9])
    np.testing.assert_array_equal(e, [2, 9, 10])
    rng = np.random.default_rng(7)
    hap = rng.integers(0, 2, size=(12, 10)).astype(np.int8)
    hap[:, 2:6] = 0                      # window [30, 70): variants 2..5, of which
    hap[0, 2] = 1                        # only variant 2 segregates
    zns, om = O.windowed_zns_omega(hap, pos, [0, 30, 100, 0], [30, 70, 200, 1])
    assert np.isnan(zns[0]) and np.isnan(om[0])      # 2 variants
    assert np.isnan(zns[2]) and np.isnan(zns[3])     # 1 and 0 variants
    assert zns[1] == 0.0 and om[1] == 0.0           # 4 variants, 1 segregating
    zns, om = O.windowed_zns_omega(hap, pos, [30], [200])
    seg = O.segregating_sites(hap[:, 2:10])
    assert len(seg) == 5
    sig, valid = O.sigma_d2_matrix(hap[:, 2:10][:, seg])
    assert zns[0] == O.zns_from_matrix(sig, valid) and om[0] == O.omega_from_matrix(sig)


def test_omega_known_answers():
    """Omega's split rule (ld_statistics.py:848-862): the maximum over
    l = 3 .. m-2 of the within / across mean ratio where the cross sum is
    positive, 0.0 for other splits -- negative when every split has a
    positive cross sum and a negative within sum."""
    m = 5
    M = np.zeros((m, m))
    iu = np.triu_indices(m, k=1)
    M[iu] = -1.0
    M[:3, 3:] = 2.0          # cross block of the only split (l = 3): mean 2
    # within: 3 pairs left + 1 pair right, all -1 -> mean -1; omega = -1 / 2
    assert O.omega_from_matrix(M) == -0.5
    M[:3, 3:] = -2.0         # no valid split -> 0.0
    assert O.omega_from_matrix(M) == 0.0
    M = np.zeros((6, 6))
    M[np.triu_indices(6, k=1)] = 1.0
    assert O.omega_from_matrix(M) == 1.0  # uniform LD: every ratio 1
    assert O.omega_from_matrix(np.ones((4, 4))) == 0.0  # fewer than 5 sites

