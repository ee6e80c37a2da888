"""GPU parity: every libpgld path against the oracle on seeded inputs.

Counts must be bit-exact.  Per-pair statistics are evaluated from exact
integer numerators with one float64 division, so they must agree with the
oracle to a few ulp (rtol 1e-12).  Keyed sums are float64 reductions in a
different order than the oracle's: rtol 1e-12 against max(|sum|, sum of |v|).
"""
import numpy as np
import pytest

from oracle import ld_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from pg_gpu_amd import band, ld_pipeline as LP, ld_statistics as LS  # noqa: E402
from pg_gpu_amd import haplotype_kernels as HK, genotype_kernels as GK  # noqa: E402
from pg_gpu_amd import synthetic as SY  # noqa: E402
from pg_gpu_amd import _lib as L  # noqa: E402


def np_(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def hap_data(seed, n_hap, n_var, missing):
    rng = np.random.default_rng(seed)
    h = SY.neutral_haplotypes(n_hap, n_var, rng, missing=missing)
    pos = SY.uniform_positions(n_var, n_var * 100, rng)
    return h, pos


def geno_data(seed, n_ind, n_var, missing):
    rng = np.random.default_rng(seed)
    g = SY.hwe_genotypes(n_ind, n_var, rng, missing=missing)
    pos = SY.uniform_positions(n_var, n_var * 100, rng)
    return g, pos


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("chunk", [7, 1000, 10**7])
def test_pairs_match_oracle(chunk):
    rng = np.random.default_rng(5)
    pos = np.unique(np.sort(rng.integers(0, 20000, size=600)))
    md = 700
    got = [(np_(a), np_(b)) for a, b in LP.iter_pairs_within_distance(pos, md, chunk)]
    ii = np.concatenate([g[0] for g in got])
    jj = np.concatenate([g[1] for g in got])
    ei, ej = O.pairs_within_distance(pos, md)
    np.testing.assert_array_equal(ii, ei)
    np.testing.assert_array_equal(jj, ej)
    # chunk boundaries follow the reference's anchor blocks
    bounds = O.chunk_bounds(pos, md, chunk)
    assert len(bounds) == len(got)


def test_pairs_edge_cases():
    assert list(LP.iter_pairs_within_distance(np.array([], dtype=np.int64), 10, 5)) == []
    assert list(LP.iter_pairs_within_distance(np.array([3]), 10, 5)) == []
    assert list(LP.iter_pairs_within_distance(np.arange(10), 0, 4)) == []
    n = 25
    got = list(LP.iter_pairs_within_distance(np.arange(n) * 10, 10**9, 7))
    assert sum(int(a.shape[0]) for a, _ in got) == n * (n - 1) // 2


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("missing", [0.0, 0.05])
@pytest.mark.parametrize("n_hap", [37, 200])
def test_hap_counts_exact(missing, n_hap):
    h, pos = hap_data(11, n_hap, 3000, missing)
    ii, jj = O.pairs_within_distance(pos, 2500)
    # plus pairs spanning more than the LDS staging width (direct path)
    ii = np.concatenate([ii, [0, 5, 17]])
    jj = np.concatenate([jj, [2999, 2900, 2500]])
    rows = list(range(0, n_hap, 2))
    for sel in (None, rows):
        c, nv = LP.compute_counts_for_pairs(h, ii, jj, sel)
        ec, env = O.hap_counts(h, ii, jj, sel)
        np.testing.assert_array_equal(np_(c), ec)
        assert (nv is None) == (env is None)
        if env is not None:
            np.testing.assert_array_equal(np_(nv), env)


def test_hap_counts_kat():
    # tests/test_haplotype_counting.py:56-112 known answers; rows are ind0..ind3
    # first haplotypes then ind0..ind3 second haplotypes (the VCF (s, s+n) layout)
    hap = np.array([[1, 0, 1, 0, 1, 1, 0, 0], [1, 1, 1, 0, 0, 0, 1, 0]], dtype=np.int8).T
    c, _ = LP.compute_counts_for_pairs(hap, [0], [1], [0, 1, 4, 5])
    np.testing.assert_array_equal(np_(c)[0], [1, 2, 1, 0])
    c, _ = LP.compute_counts_for_pairs(hap, [0], [1], [2, 3, 6, 7])
    np.testing.assert_array_equal(np_(c)[0], [1, 0, 1, 2])


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_geno_counts_exact(missing):
    g, pos = geno_data(12, 53, 2600, missing)
    ii, jj = O.pairs_within_distance(pos, 3000)
    ii = np.concatenate([ii, [0]])
    jj = np.concatenate([jj, [2599]])
    for sel in (None, list(range(1, 53, 3))):
        c, nv = LP.compute_genotype_counts_for_pairs(g, ii, jj, sel)
        ec, env = O.geno_counts(g, ii, jj, sel)
        np.testing.assert_array_equal(np_(c), ec)
        assert (nv is None) == (env is None)


def test_counts_domain_error():
    h = np.zeros((4, 5), dtype=np.int8)
    h[1, 2] = 2
    with pytest.raises(ValueError):
        LP.compute_counts_for_pairs(h, [0], [2])


# ---------------------------------------------------------------------------
def _rand_counts(rng, N, P, n_max, kind="hap"):
    out = []
    for _ in range(P):
        n = rng.integers(0, n_max, size=N)
        k = 4 if kind == "hap" else 9
        w = rng.dirichlet(np.ones(k), size=N)
        c = np.floor(w * n[:, None]).astype(np.int64)
        out.append(c)
    return out


def assert_close(got, exp, rtol=1e-12, atol=1e-15):
    got, exp = np_(got), np.asarray(exp)
    fin = np.isfinite(exp)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    np.testing.assert_allclose(got[fin], exp[fin], rtol=rtol, atol=atol)


def test_two_pop_fast_path():
    rng = np.random.default_rng(3)
    c0, c1 = _rand_counts(rng, 4000, 2, 60)
    got = LP.compute_two_pop_statistics_batch(c0, c1, None, None)
    assert_close(got, O.two_pop_stats(c0, c1))


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_multi_pop_hap(P):
    rng = np.random.default_rng(10 + P)
    cl = _rand_counts(rng, 1500, P, 40)
    got = HK.compute_multi_pop_statistics_batch_hap(cl, [None] * P, None, LP.generate_stat_specs(P))
    assert_close(got, O.multi_pop_stats_hap(cl, [None] * P))


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_multi_pop_geno(P):
    rng = np.random.default_rng(20 + P)
    cl = _rand_counts(rng, 1200, P, 40, kind="geno")
    for c in cl:  # keep every population at n >= 4 (the reference divides unguarded)
        c[:, 0] += 4
    got = GK.compute_multi_pop_statistics_batch_geno(cl, [None] * P, None, LP.generate_stat_specs(P))
    assert_close(got, O.multi_pop_stats_geno(cl, [None] * P), rtol=1e-11, atol=1e-13)


def test_single_stat_api_and_n_valid():
    rng = np.random.default_rng(4)
    c0, c1 = _rand_counts(rng, 800, 2, 50)
    both = np.concatenate([c0, c1], axis=1)
    P = [O.HapPop(c0), O.HapPop(c1)]
    assert_close(LS.dd(c0), O.hap_dd([O.HapPop(c0)], 0, 0))
    assert_close(LS.dd(both, (0, 1)), O.hap_dd(P, 0, 1))
    assert_close(LS.dz(both, (0, 1, 1)), O.hap_dz(P, 0, 1, 1))
    assert_close(LS.pi2(both, (0, 0, 1, 1)), O.hap_pi2_python(P, 0, 0, 1, 1))
    assert_close(LS.pi2(both, (0, 1, 0, 1)), O.hap_pi2_python(P, 0, 1, 0, 1))
    nv = c0.sum(1)
    assert_close(LS.dd(c0, n_valid=nv), O.hap_dd([O.HapPop(c0, nv)], 0, 0))


def test_r_family():
    rng = np.random.default_rng(8)
    (c,) = _rand_counts(rng, 3000, 1, 80)
    c = np.concatenate([c, [[5, 0, 0, 5], [25, 25, 25, 25], [0, 5, 5, 0], [5, 5, 0, 0], [28, 52, 2, 18]]])
    assert_close(LS.r(c), O.r(c))
    assert_close(LS.r_squared(c), O.r_squared(c))
    assert_close(LS.d_prime(c), O.d_prime(c))
    np.testing.assert_allclose(np_(LS.d_prime(c))[-1], 2.0 / 3.0, rtol=1e-12)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("family,missing", [("hap", True), ("hap", False), ("geno", True), ("geno", False)])
def test_band_counts_exact(family, missing):
    if family == "hap":
        d, pos = hap_data(30, 173, 900, 0.03 if missing else 0.0)
    else:
        d, pos = geno_data(31, 141, 900, 0.03 if missing else 0.0)
    md = 4000
    plan = band.BandPlan(pos, md, edges=[0.0, 1e9])
    rows = list(range(0, d.shape[0], 1))
    pk = band.pack_population(d, rows, family=family)
    assert pk.has_missing == missing
    got = np_(plan.counts(pk))
    ii, jj = O.pairs_within_distance(pos, md)
    assert plan.n_pairs == len(ii)
    exp = (O.hap_counts if family == "hap" else O.geno_counts)(d, ii, jj, rows)[0]
    np.testing.assert_array_equal(got, exp)


def _sums_oracle(family, d, pos, pops, md, stat_set, key):
    ii, jj = O.pairs_within_distance(pos, md)
    cl = [(O.hap_counts if family == "hap" else O.geno_counts)(d, ii, jj, r)[0] for r in pops]
    P = len(pops)
    if family == "hap":
        if stat_set == "self" and P > 1:
            specs = [(f"{t}", [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(P)]
            stats = O.multi_pop_stats_hap(cl, [None] * P, specs)
        elif P == 2:
            stats = O.two_pop_stats(cl[0], cl[1])
        else:
            stats = O.multi_pop_stats_hap(cl, [None] * P)
    elif stat_set == "self" and P > 1:
        specs = [(f"{t}", [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(P)]
        stats = O.multi_pop_stats_geno(cl, [None] * P, specs)
    else:
        stats = O.multi_pop_stats_geno(cl, [None] * P)
    keys = key(ii, jj)
    return stats, keys


def _check_sums(got, stats, keys, n_keys):
    exp = O.keyed_sums(keys, stats, n_keys)
    scale = O.keyed_sums(keys, np.abs(stats), n_keys)
    got = np_(got)
    assert np.all(np.abs(got - exp) <= 1e-12 * np.maximum(scale, 1e-300) + 1e-300)


@pytest.mark.parametrize("P,missing,stat_set", [(1, True, "moments"), (2, True, "moments"),
                                                (2, False, "moments"), (3, True, "moments"),
                                                (4, True, "moments"), (4, False, "moments"),
                                                (4, True, "self"), (2, True, "self")])
def test_band_sums_windows_hap(P, missing, stat_set):
    d, pos = hap_data(40 + P, 40 * P + 17, 1300, 0.02 if missing else 0.0)
    n = d.shape[0]
    pops = [list(range(p, n, P)) for p in range(P)]
    md, width = 5000.0, 20000.0
    plan = band.BandPlan(pos, md, window=width)
    packed = [band.pack_population(d, r, "hap") for r in pops]
    sums, kp = plan.sums(packed, stat_set=stat_set, want_pairs=True)
    stats, keys = _sums_oracle("hap", d, pos, pops, md, stat_set,
                               lambda ii, jj: O.window_key(pos[ii], 0.0, width))
    _check_sums(sums, stats, keys, plan.n_keys)
    np.testing.assert_array_equal(np_(kp), np.bincount(keys, minlength=plan.n_keys))


@pytest.mark.parametrize("P,missing", [(1, True), (2, True), (2, False)])
def test_band_sums_bins(P, missing):
    d, pos = hap_data(60 + P, 50 * P + 9, 1100, 0.03 if missing else 0.0)
    n = d.shape[0]
    pops = [list(range(p, n, P)) for p in range(P)]
    edges = np.logspace(1, 4, 9)
    md = float(edges[-1])
    plan = band.BandPlan(pos, md, edges=edges)
    packed = [band.pack_population(d, r, "hap") for r in pops]
    sums = plan.sums(packed)
    stats, keys = _sums_oracle("hap", d, pos, pops, md, "moments",
                               lambda ii, jj: O.bin_index(pos[jj] - pos[ii], edges))
    _check_sums(sums, stats, keys, len(edges) - 1)


def _two_key_subtiles(pos, md, key_of_pair, n_var):
    """Count the 32 x 32 sub-tiles (rows and columns at multiples of 32) that
    lie strictly above the diagonal and inside the panel, whose farthest pair
    is within md and whose pairs span exactly two consecutive keys (the
    extreme pairs decide: keys are monotone in the left SNP and the
    distance)."""
    nb = n_var // 32
    n = 0
    for bi in range(nb):
        i0, i1 = 32 * bi, 32 * bi + 31
        for bj in range(bi + 1, nb):
            j0, j1 = 32 * bj, 32 * bj + 31
            if pos[j1] - pos[i0] > md:
                break
            k = sorted({int(key_of_pair(i0, j1)), int(key_of_pair(i1, j0))})
            n += len(k) == 2 and k[0] >= 0 and k[1] == k[0] + 1
    return n


@pytest.mark.parametrize("key", ["bins", "windows"])
@pytest.mark.parametrize("family,P,stat_set,missing,fp4", [
    ("hap", 1, "moments", True, True), ("hap", 1, "moments", False, True), ("hap", 3, "self", True, True),
    ("hap", 3, "self", False, False), ("hap", 2, "moments", True, True), ("hap", 2, "moments", False, False),
    ("geno", 1, "moments", True, True), ("geno", 1, "moments", False, False)])
def test_band_sums_two_key_subtiles(key, family, P, stat_set, missing, fp4):
    """Sub-tiles whose pairs fall in two consecutive bins or windows are
    summed into both keys at once: with SNPs 10 apart, bins 900 wide and
    windows 450 wide every sub-tile holds one or two keys, and 150 or more
    hold two.  Sums within 1e-12 of the oracle's per-pair statistics, pair counts
    exact."""
    n_var = 1216
    rng = np.random.default_rng(900 + 10 * P + missing)
    pos = 10.0 * np.arange(n_var) + rng.integers(0, 4, n_var)
    if family == "hap":
        d = SY.neutral_haplotypes(30 * P + 11, n_var, rng, missing=0.03 if missing else 0.0)
    else:
        d = SY.hwe_genotypes(41, n_var, rng, missing=0.03 if missing else 0.0)
    pops = [list(range(p, d.shape[0], P)) for p in range(P)]
    packed = [band.pack_population(d, r, family, fp4=fp4) for r in pops]
    if key == "bins":
        edges = 905.0 + 900.0 * np.arange(14)
        md = float(edges[-1])
        plan = band.BandPlan(pos, md, edges=edges)
        n_keys = len(edges) - 1
        kfun = lambda ii, jj: O.bin_index(pos[jj] - pos[ii], edges)  # noqa: E731
    else:
        md, width = 3000.0, 450.0
        plan = band.BandPlan(pos, md, window=width)
        n_keys = plan.n_keys
        kfun = lambda ii, jj: O.window_key(pos[ii], 0.0, width)  # noqa: E731
    assert _two_key_subtiles(pos, md, lambda i, j: kfun(np.array([i]), np.array([j]))[0], n_var) >= 150
    sums, kp = plan.sums(packed, stat_set=stat_set, want_pairs=True)
    stats, keys = _sums_oracle(family, d, pos, pops, md, stat_set, kfun)
    _check_sums(sums, stats, keys, n_keys)
    np.testing.assert_array_equal(np_(kp), np.bincount(keys[(keys >= 0) & (keys < n_keys)], minlength=n_keys))


@pytest.mark.parametrize("P,missing,stat_set,fp4", [
    (1, True, "moments", False), (1, False, "moments", False), (2, True, "moments", False),
    (3, True, "moments", True), (3, False, "moments", False), (4, True, "moments", True),
    (4, False, "moments", True), (3, True, "self", True), (4, True, "self", False)])
def test_band_sums_geno(P, missing, stat_set, fp4):
    """Genotype sets of 1-4 populations against the oracle; the three- and
    four-population moments (45 / 105 statistics, every case class of
    genotype_kernels.py:27-405) park their counts in the global workspace."""
    d, pos = geno_data(70 + P, 60 * P + 5, 800, 0.05 if missing else 0.0)
    n = d.shape[0]
    pops = [list(range(p, n, P)) for p in range(P)]
    edges = np.logspace(0, 5, 17)
    md = float(edges[-1])
    plan = band.BandPlan(pos, md, edges=edges)
    packed = [band.pack_population(d, r, "geno", fp4=fp4) for r in pops]
    sums, kp = plan.sums(packed, stat_set=stat_set, want_pairs=True)
    stats, keys = _sums_oracle("geno", d, pos, pops, md, stat_set,
                               lambda ii, jj: O.bin_index(pos[jj] - pos[ii], edges))
    _check_sums(sums, stats, keys, len(edges) - 1)
    nb = len(edges) - 1
    np.testing.assert_array_equal(np_(kp), np.bincount(keys[(keys >= 0) & (keys < nb)], minlength=nb))


@pytest.mark.parametrize("P", [3, 4])
def test_geno_moments_chunked_launches_equal_fp4_and_int8(P):
    """More tiles than one launch of the workspace-parked genotype moments
    covers (1,024): the launches continue the tile order, so FP4 and int8
    planes give the same sums bit for bit, and the sums match the per-pair
    path."""
    d, pos = geno_data(83, 72, 10000, 0.04)
    pops = [list(range(p, 72, P)) for p in range(P)]
    plan = band.BandPlan(pos, 2.5e5, window=3.0e5)
    assert plan.n_tiles > 4 * 1024
    s8, k8 = plan.sums([band.pack_population(d, r, "geno") for r in pops], want_pairs=True)
    s4, k4 = plan.sums([band.pack_population(d, r, "geno", fp4=True) for r in pops], want_pairs=True)
    np.testing.assert_array_equal(np_(s4), np_(s8))
    np.testing.assert_array_equal(np_(k4), np_(k8))
    ref, kref = band.ld_sums(d, "geno", pos, pops, 2.5e5, window=3.0e5, n_keys=plan.n_keys, method="plain")
    np.testing.assert_array_equal(np_(k8), np_(kref))
    ref = np_(ref)
    assert np.all(np.abs(np_(s8) - ref) <= 1e-10 * np.maximum(np.abs(ref), 1e-300) + 1e-9)


def test_band_sums_deterministic():
    d, pos = hap_data(80, 301, 1500, 0.02)
    pops = [list(range(0, 150)), list(range(150, 301))]
    plan = band.BandPlan(pos, 8000.0, window=10000.0)
    packed = [band.pack_population(d, r, "hap") for r in pops]
    a = np_(plan.sums(packed))
    b = np_(plan.sums(packed))
    np.testing.assert_array_equal(a, b)


# ---------------------------------------------------------------------------
# r^2 through the MFMA tiles: identical FP64 operations as the oracle, so the
# per-pair values, histograms and sparse lists must match exactly
# ---------------------------------------------------------------------------
def _r2_oracle(d, pos, md):
    ii, jj = O.pairs_within_distance(pos, md)
    c, _ = O.hap_counts(d, ii, jj)
    return ii, jj, O.r_squared_exact(c)


@pytest.mark.parametrize("missing", [False, True])
def test_band_r2_windows_exact(missing):
    d, pos = hap_data(91, 150, 700, 0.03 if missing else 0.0)
    d[:, 5] = 0  # a monomorphic site -> NaN pairs
    md, width = 6000.0, 15000.0
    plan = band.BandPlan(pos, md, window=width)
    pk = band.pack_population(d, None, "hap")
    assert pk.has_missing == missing
    out = plan.r2(pk, n_hist=50, thresh=0.1, per_pair=True)
    ii, jj, r2 = _r2_oracle(d, pos, md)
    got = np_(out["r2"])
    nan = np.isnan(r2)
    assert nan.any()
    np.testing.assert_array_equal(np.isnan(got), nan)
    np.testing.assert_array_equal(got[~nan], r2[~nan])
    h, hn = O.r2_histogram(r2, O.window_key(pos[ii], 0.0, width), plan.n_keys, 50)
    np.testing.assert_array_equal(np_(out["hist"]), h)
    np.testing.assert_array_equal(np_(out["nan_pairs"]), hn)
    sel = ~nan & (r2 > 0.1)
    assert out["n_sparse"] == int(sel.sum()) > 0
    np.testing.assert_array_equal(np_(out["sparse_i"]), ii[sel])
    np.testing.assert_array_equal(np_(out["sparse_j"]), jj[sel])
    np.testing.assert_array_equal(np_(out["sparse_r2"]), r2[sel].astype(np.float32))


def test_band_r2_all_pairs_bins_and_capacity():
    d, pos = hap_data(92, 260, 500, 0.0)
    edges = np.array([0.0, 1000.0, 5000.0, 20000.0, 1e9])
    plan = band.BandPlan(pos, np.inf, edges=edges)
    assert plan.n_pairs == 500 * 499 // 2
    pk = band.pack_population(d, None, "hap")
    full = plan.r2(pk, n_hist=100, thresh=0.05)
    ii, jj, r2 = _r2_oracle(d, pos, np.inf)
    h, hn = O.r2_histogram(r2, O.bin_index(pos[jj] - pos[ii], edges), 4, 100)
    np.testing.assert_array_equal(np_(full["hist"]), h)
    np.testing.assert_array_equal(np_(full["nan_pairs"]), hn)
    n_hit = int((r2 > 0.05).sum())
    assert full["n_sparse"] == n_hit > 20
    # a capacity below the hit count keeps a subset of true hits and reports the total
    cut = plan.r2(pk, n_hist=100, thresh=0.05, sparse_cap=7, histogram=False)
    assert cut["n_sparse"] == n_hit and len(cut["sparse_i"]) == 7
    hits = set(zip(ii[r2 > 0.05].tolist(), jj[r2 > 0.05].tolist()))
    assert set(zip(np_(cut["sparse_i"]).tolist(), np_(cut["sparse_j"]).tolist())) <= hits


# ---------------------------------------------------------------------------
# bit-packed popcount variant: the same counts and r^2 as the MFMA tiles
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("family,missing", [("hap", True), ("hap", False), ("geno", True), ("geno", False)])
def test_popcount_counts_equal_mfma_and_oracle(family, missing):
    if family == "hap":
        d, pos = hap_data(95, 1100, 300, 0.03 if missing else 0.0)  # > 1,024 samples: two bit-row chunks
    else:
        d, pos = geno_data(96, 700, 300, 0.04 if missing else 0.0)
    plan = band.BandPlan(pos, 4000.0, window=1e9, n_keys=1)
    pk = band.pack_population(d, None, family)
    a = np_(plan.counts(pk, method="mfma"))
    b = np_(plan.counts(pk, method="popcount"))
    np.testing.assert_array_equal(a, b)
    ii, jj = O.pairs_within_distance(pos, 4000.0)
    exp = (O.hap_counts if family == "hap" else O.geno_counts)(d, ii, jj, None)[0]
    np.testing.assert_array_equal(b, exp)


@pytest.mark.parametrize("missing", [False, True])
def test_popcount_r2_equals_mfma(missing):
    d, pos = hap_data(97, 300, 600, 0.02 if missing else 0.0)
    plan = band.BandPlan(pos, np.inf, edges=np.array([0.0, 2000.0, 10000.0, 1e9]))
    pk = band.pack_population(d, None, "hap")
    m = plan.r2(pk, n_hist=40, thresh=0.05, per_pair=True, method="mfma")
    p = plan.r2(pk, n_hist=40, thresh=0.05, per_pair=True, method="popcount")
    for k in ("hist", "nan_pairs", "sparse_i", "sparse_j", "sparse_r2", "r2"):
        np.testing.assert_array_equal(np_(m[k]), np_(p[k]), err_msg=k)
    assert m["n_sparse"] == p["n_sparse"]


# ---------------------------------------------------------------------------
# plane packing against its definition
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("family,n_var,n_rows,missing", [("hap", 301, 150, 0.05), ("hap", 256, 300, 0.0),
                                                          ("geno", 259, 131, 0.05), ("geno", 128, 257, 0.0)])
def test_pack_planes_definition(family, n_var, n_rows, missing):
    if family == "hap":
        d, _ = hap_data(101, n_rows, n_var, missing)
    else:
        d, _ = geno_data(102, n_rows, n_var, missing)
    rows = np.sort(np.random.default_rng(3).choice(n_rows, size=n_rows - 7, replace=False))
    pk = band.pack_population(d, rows, family)
    x = d[rows].astype(np.int16)  # (samples, variants)
    want = [(x == 1), (x >= 0)] if family == "hap" else [(x >= 0), (x == 1), (x == 2)]
    planes, marg = np_(pk.planes), np_(pk.marg)
    assert planes.shape == (len(want), n_var, band.k_pad_for(len(rows)))
    for p, w in enumerate(want):
        np.testing.assert_array_equal(planes[p, :, :len(rows)], w.T.astype(np.int8))
        assert not planes[p, :, len(rows):].any()
        np.testing.assert_array_equal(marg[p], w.sum(axis=0))
    assert pk.has_missing == bool((x < 0).any())


@pytest.mark.parametrize("family,bad", [("hap", 2), ("hap", -2), ("geno", 3), ("geno", -5)])
def test_pack_planes_domain_error(family, bad):
    d, _ = (hap_data if family == "hap" else geno_data)(103, 70, 90, 0.0)
    d = d.copy()
    d[13, 77] = bad
    with pytest.raises(ValueError):
        band.pack_population(d, None, family)
    # the same value in a row outside the selected population is not an error
    band.pack_population(d, [r for r in range(70) if r != 13], family)


def test_band_passes_reject_k_pad_beyond_32768():
    """Operand offsets within a 64-row block are 24-bit products that assume
    k_pad <= 32,768 (32,767 samples packed): larger paddings are an argument
    error with a message, before any launch."""
    d, pos = hap_data(105, 40, 200, 0.0)
    plan = band.BandPlan(pos, 1e4, window=1e4)
    pk = band.pack_population(d, None, "hap", k_pad=32768 + 128)
    assert pk.k_pad == 32896
    with pytest.raises(L.PgldError, match="k_pad must be at most 32768"):
        plan.sums([pk])
    with pytest.raises(L.PgldError, match="k_pad must be at most 32768"):
        plan.r2(pk)


# ---------------------------------------------------------------------------
# FP4 planes and the FP4-MFMA fused sums
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("family,n_var,n_rows,missing", [("hap", 301, 150, 0.05), ("hap", 256, 600, 0.0),
                                                          ("geno", 259, 131, 0.05), ("geno", 130, 300, 0.0)])
def test_pack_planes_fp4_definition(family, n_var, n_rows, missing):
    """Nibble n of an FP4 row byte b holds sample 2b + n, rows of k_pad / 2
    bytes with k_pad a multiple of 256; every plane as 0x0 / 0x2 (the e2m1
    codes of 0 and 1) except the haplotype allele plane, which is centred:
    0x2 (+1) where x == 1, 0xA (-1) where x == 0, 0x0 where missing and in the
    padding.  Marginals as the int8 planes' (0/1 allele counts)."""
    if family == "hap":
        d, _ = hap_data(103, n_rows, n_var, missing)
    else:
        d, _ = geno_data(104, n_rows, n_var, missing)
    rows = np.sort(np.random.default_rng(5).choice(n_rows, size=n_rows - 3, replace=False))
    p8 = band.pack_population(d, rows, family)
    p4 = band.pack_population(d, rows, family, fp4=True)
    kp = band.k_pad_for(len(rows), fp4=True)
    assert p4.fp4 and p4.k_pad == kp and kp % 256 == 0
    b = np_(p4.planes).view(np.uint8)
    assert b.shape == (np_(p8.planes).shape[0], n_var, kp // 2)
    nib = np.stack([b & 0xF, b >> 4], axis=-1).reshape(b.shape[0], n_var, kp)
    want = np.zeros_like(nib)
    want[:, :, :p8.k_pad] = np_(p8.planes) * 2
    if family == "hap":
        assert set(np.unique(nib)) <= {0x0, 0x2, 0xA}
        # allele plane from the data itself: x == 1 -> 0x2, x == 0 -> 0xA, missing -> 0
        x = np.asarray(d)[rows].T  # (n_var, samples)
        want[0] = 0
        want[0, :, :len(rows)] = np.where(x == 1, 0x2, np.where(x == 0, 0xA, 0x0))
    else:
        assert set(np.unique(nib)) <= {0, 2}
    np.testing.assert_array_equal(nib, want)
    np.testing.assert_array_equal(np_(p4.marg), np_(p8.marg))
    assert p4.has_missing == p8.has_missing


@pytest.mark.parametrize("family,P,missing,stat_set,key", [
    ("hap", 1, True, "moments", "window"), ("hap", 2, True, "moments", "window"),
    ("hap", 2, False, "moments", "bins"), ("hap", 3, True, "moments", "bins"),
    ("hap", 4, True, "self", "window"), ("hap", 2, True, "self", "bins"),
    ("hap", 4, True, "moments", "window"), ("hap", 4, False, "moments", "window"),
    ("geno", 1, True, "moments", "bins"), ("geno", 1, False, "moments", "bins"),
    ("geno", 2, True, "moments", "window"), ("geno", 2, True, "self", "bins")])
def test_band_sums_fp4_equal_int8(family, P, missing, stat_set, key):
    """The FP4-MFMA fused pass forms the same exact counts as the int8 one,
    so every sum agrees bit for bit (population sizes that are not multiples
    of 256 pad differently in the two layouts)."""
    if family == "hap":
        d, pos = hap_data(110 + P, 70 * P + 13, 1200, 0.02 if missing else 0.0)
    else:
        d, pos = geno_data(120 + P, 90 * P + 7, 900, 0.04 if missing else 0.0)
    n = d.shape[0]
    pops = [list(range(p, n, P)) for p in range(P)]
    if key == "window":
        plan = band.BandPlan(pos, 6000.0, window=15000.0)
    else:
        edges = np.logspace(1, 4.5, 11)
        plan = band.BandPlan(pos, float(edges[-1]), edges=edges)
    p8 = [band.pack_population(d, r, family) for r in pops]
    p4 = [band.pack_population(d, r, family, fp4=True) for r in pops]
    s8, k8 = plan.sums(p8, stat_set=stat_set, want_pairs=True)
    s4, k4 = plan.sums(p4, stat_set=stat_set, want_pairs=True)
    np.testing.assert_array_equal(np_(s4), np_(s8))
    np.testing.assert_array_equal(np_(k4), np_(k8))
    with pytest.raises(ValueError):
        plan.sums([p8[0]] + p4[1:] if P > 1 else [p4[0], p8[0]], stat_set=stat_set)
    with pytest.raises(ValueError):
        plan.counts(p4[0])


@pytest.mark.parametrize("estimator", ["r2", "naive", "sigma_d2"])
def test_band_r2_fp4_equal_int8(estimator):
    """Dense r^2 on the FP4 super-tile kernel: histograms, sparse lists and
    per-pair values bit-identical to the int8 kernel."""
    d, pos = hap_data(130, 300, 700, 0.0)
    plan = band.BandPlan(pos, np.inf, window=1e300, n_keys=1)
    p8 = band.pack_population(d, None, "hap")
    p4 = band.pack_population(d, None, "hap", fp4=True)
    a = plan.r2(p8, n_hist=50, thresh=0.2, per_pair=True, estimator=estimator)
    b = plan.r2(p4, n_hist=50, thresh=0.2, per_pair=True, estimator=estimator)
    for k in ("hist", "nan_pairs", "sparse_i", "sparse_j", "sparse_r2", "r2"):
        np.testing.assert_array_equal(np_(b[k]), np_(a[k]), err_msg=k)
    assert b["n_sparse"] == a["n_sparse"]


@pytest.mark.parametrize("family,estimator,missing,snp_sums", [
    ("hap", "r2", 0.02, False), ("hap", "naive", 0.02, True), ("hap", "sigma_d2", 0.0, True),
    ("hap", "sigma_d2", 0.03, True), ("geno", "rogers_huff_r", 0.0, False),
    ("geno", "rogers_huff_r", 0.0, True), ("geno", "dosage_r2", 0.03, False)])
def test_band_r2_fp4_64x64_equal_int8(family, estimator, missing, snp_sums):
    """FP4 planes on the 64 x 64 r^2 kernel (missing data, genotype
    estimators, per-SNP sums, banded plans): every output bit-identical to
    the int8 planes."""
    d, pos = (hap_data(132, 270, 900, missing) if family == "hap" else geno_data(133, 150, 900, missing))
    plan = band.BandPlan(pos, 15_000.0, window=1e300, n_keys=1)
    p8 = band.pack_population(d, None, family)
    p4 = band.pack_population(d, None, family, fp4=True)
    kw = dict(n_hist=40, thresh=0.1, per_pair=not snp_sums, estimator=estimator, snp_sums=snp_sums)
    a = plan.r2(p8, **kw)
    b = plan.r2(p4, **kw)
    assert set(a) == set(b)
    for k in a:
        if isinstance(a[k], (int, float)):
            assert a[k] == b[k], k
        else:
            np.testing.assert_array_equal(np_(b[k]), np_(a[k]), err_msg=k)


def _oracle_band_sums(d, family, pos, pops, edges, stat_set):
    """Per-bin oracle sums (and sums of |terms|) of the moments or
    within-population statistics over the pairs within the last edge."""
    ii, jj = O.pairs_within_distance(pos, float(edges[-1]))
    cnt = O.hap_counts if family == "hap" else O.geno_counts
    cl = [cnt(d, ii, jj, r) for r in pops]
    P = len(pops)
    if stat_set == "self" and P > 1:
        specs = [(t, [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(P)]
    else:
        specs = None
    if family == "hap":
        st = (O.two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1]) if (P == 2 and specs is None)
              else O.multi_pop_stats_hap([c for c, _ in cl], [nv for _, nv in cl], specs))
    else:
        st = O.multi_pop_stats_geno([c for c, _ in cl], [nv for _, nv in cl], specs)
    key = O.bin_index(pos[jj] - pos[ii], edges)
    n = len(edges) - 1
    return O.keyed_sums(key, st, n), O.keyed_sums(key, np.abs(st), n)


@pytest.mark.parametrize("family,sizes,missing,stat_set", [
    ("hap", (650, 650), True, "moments"), ("hap", (1100,), False, "moments"),
    ("hap", (610, 700, 530, 900), True, "self"), ("hap", (520, 777, 600), True, "moments"),
    ("geno", (700,), True, "moments"), ("geno", (1030,), False, "moments"), ("geno", (620, 545), True, "moments"),
    ("geno", (513, 800), True, "self")])
def test_band_sums_fp4_multichunk_vs_oracle_and_int8(family, sizes, missing, stat_set):
    """Populations of 513-1,100 samples: 3-5 FP4 K chunks of 256 samples
    (odd and even counts, both register-stage pipelines), sums against the
    oracle and bit-identical to the int8 planes."""
    n = sum(sizes)
    if family == "hap":
        d, pos = hap_data(140 + len(sizes), n, 320, 0.03 if missing else 0.0)
        tol = 1e-12
    else:
        d, pos = geno_data(150 + len(sizes), n, 320, 0.04 if missing else 0.0)
        tol = 1e-11
    cut = np.concatenate([[0], np.cumsum(sizes)])
    pops = [list(range(cut[p], cut[p + 1])) for p in range(len(sizes))]
    print(f"fp4 multichunk {family} sizes={sizes} missing={missing} {stat_set}", flush=True)
    edges = np.array([0.0, 300.0, 1000.0, 2500.0, 4000.0])
    plan = band.BandPlan(pos, float(edges[-1]), edges=edges)
    p4 = [band.pack_population(d, r, family, fp4=True) for r in pops]
    p8 = [band.pack_population(d, r, family) for r in pops]
    assert min(p.k_pad for p in p4) >= 768
    s4 = np_(plan.sums(p4, stat_set=stat_set))
    s8 = np_(plan.sums(p8, stat_set=stat_set))
    np.testing.assert_array_equal(s4, s8)
    want, scale = _oracle_band_sums(d, family, pos, pops, edges, stat_set)
    err = np.max(np.abs(s4 - want) / np.maximum(scale, 1e-300))
    assert err <= tol, err


@pytest.mark.parametrize("family,estimator,n,missing,snp_sums", [
    ("hap", "r2", 700, 0.02, False), ("hap", "sigma_d2", 1100, 0.03, True), ("hap", "naive", 1300, 0.0, True),
    ("geno", "dosage_r2", 650, 0.03, False), ("geno", "dosage_r2", 1030, 0.0, True),
    ("geno", "rogers_huff_r", 700, 0.0, False)])
def test_band_r2_fp4_multichunk_equal_int8(family, estimator, n, missing, snp_sums):
    """The FP4 r^2 kernels over 3-6 K chunks (missing data, genotype
    estimators, per-SNP sums): every output bit-identical to int8 planes."""
    d, pos = (hap_data(160, n, 500, missing) if family == "hap" else geno_data(161, n, 500, missing))
    plan = band.BandPlan(pos, 12_000.0, window=1e300, n_keys=1)
    p8 = band.pack_population(d, None, family)
    p4 = band.pack_population(d, None, family, fp4=True)
    kw = dict(n_hist=40, thresh=0.1, per_pair=not snp_sums, estimator=estimator, snp_sums=snp_sums)
    a = plan.r2(p8, **kw)
    b = plan.r2(p4, **kw)
    for k in a:
        if isinstance(a[k], (int, float)):
            assert a[k] == b[k], k
        else:
            np.testing.assert_array_equal(np_(b[k]), np_(a[k]), err_msg=k)
    if family == "hap" and estimator != "sigma_d2":
        full = band.BandPlan(pos, np.inf, window=1e300, n_keys=1)
        a = full.r2(p8, n_hist=50, thresh=0.2, per_pair=True, estimator=estimator)
        b = full.r2(p4, n_hist=50, thresh=0.2, per_pair=True, estimator=estimator)
        for k in ("hist", "nan_pairs", "sparse_i", "sparse_j", "sparse_r2", "r2"):
            np.testing.assert_array_equal(np_(b[k]), np_(a[k]), err_msg=k)
