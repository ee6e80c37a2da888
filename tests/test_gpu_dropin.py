"""GPU parity of the drop-in layer: HaplotypeMatrix, GenotypeMatrix and
moments_ld.compute_ld_statistics against the frozen moments output of the
reference's IM example and against the oracle.

Bin sums are float64 reductions in a different order than the oracle's, so
they are compared with rtol 1e-10 on the golden vector (the precision it was
frozen at) and otherwise against max(|sum|, sum of |terms|): rtol 1e-12 for
haplotype statistics and 1e-11 for genotype statistics, whose per-pair values
already agree only to ~1e-11 (their polynomials cancel within a pair).
"""
import json
import os

import numpy as np
import pytest

from oracle import ld_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from pg_gpu_amd import GenotypeMatrix, HaplotypeMatrix, moments_ld  # noqa: E402
from pg_gpu_amd import ld_statistics_genotype as LSG  # noqa: E402
from pg_gpu_amd import synthetic as SY  # noqa: E402
from pg_gpu_amd import _lib as L  # noqa: E402
from pg_gpu_amd import band  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")
VCF = os.path.join(GOLD, "im-parsing-example.vcf")
POPS = os.path.join(GOLD, "im_pop.txt")


def np_(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def close_sums(got, want, rtol=1e-12, scale=None):
    """|got - want| <= rtol * scale, scale = max(|want|, 1) or the sum of
    |terms| where the sum cancels."""
    got, want = np.asarray(got), np.asarray(want)
    s = np.maximum(np.abs(want), 1.0) if scale is None else np.maximum(np.asarray(scale), 1.0)
    err = np.max(np.abs(got - want) / s) if want.size else 0.0
    assert err <= rtol, f"max scaled error {err:.3e}"


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", ["auto", "plain"])
def test_moments_golden_im_example(method):
    G = json.load(open(os.path.join(GOLD, "moments_im_4bins.json")))
    bins = np.array([G["bins"][0][0]] + [b[1] for b in G["bins"]])
    res = moments_ld.compute_ld_statistics(VCF, pop_file=POPS, pops=G["pops"], bp_bins=bins, report=False,
                                           method=method)
    assert res["stats"][0] == G["ld_names"] and list(res["stats"][1]) == G["het_names"]
    assert [list(b) for b in res["bins"]] == G["bins"]
    np.testing.assert_allclose(np.array(res["sums"][:-1]), np.array(G["sums"]), rtol=1e-10)
    np.testing.assert_allclose(res["sums"][-1], G["het"], rtol=1e-13)


def _im_geno():
    gm = GenotypeMatrix.from_vcf(VCF)
    gm.load_pop_file(POPS, pops=["deme0", "deme1"])
    return gm.apply_biallelic_filter()


@pytest.mark.parametrize("method", ["auto", "plain"])
def test_moments_genotypes_im_example(method):
    bins = np.logspace(2, 5, 4)
    gm = _im_geno()
    g = np_(gm.genotypes)
    pos = np_(gm.positions)
    rows = [gm.sample_sets["deme0"], gm.sample_sets["deme1"]]
    want = O.moments_ld_sums(g, pos, rows, bins, use_genotypes=True)
    scale = O.moments_ld_sums(g, pos, rows, bins, use_genotypes=True, absolute=True)
    res = moments_ld.compute_ld_statistics(genotype_matrix=gm, pops=["deme0", "deme1"], bp_bins=bins,
                                           use_genotypes=True, report=False, method=method)
    close_sums(np.array(res["sums"][:-1]), want, rtol=1e-11, scale=scale)
    H = O.heterozygosity([g[r] for r in rows], use_genotypes=True)
    np.testing.assert_allclose(res["sums"][-1], [H[k] for k in res["stats"][1]], rtol=1e-13)


@pytest.mark.parametrize("P,method", [(1, "auto"), (3, "auto"), (3, "plain"), (4, "auto"), (4, "plain")])
def test_moments_hap_pop_counts(P, method):
    """1, 3 and 4 populations on a synthetic matrix with missing data, fused
    (4 populations: the 105 statistics in three column groups) and per
    pair."""
    rng = np.random.default_rng(40 + P)
    h = SY.neutral_haplotypes(24 * P, 300, rng, missing=0.03)
    pos = SY.uniform_positions(300, 60000, rng)
    names = [f"p{k}" for k in range(P)]
    sets = {n: list(range(24 * k, 24 * (k + 1))) for k, n in enumerate(names)}
    hm = HaplotypeMatrix(h, pos, sample_sets=sets)
    bins = np.array([0.0, 500.0, 2000.0, 8000.0])
    res = moments_ld.compute_ld_statistics(haplotype_matrix=hm, pops=names, bp_bins=bins, report=False,
                                           method=method)
    want = O.moments_ld_sums(h, pos, [sets[n] for n in names], bins)
    scale = O.moments_ld_sums(h, pos, [sets[n] for n in names], bins, absolute=True)
    close_sums(np.array(res["sums"][:-1]), want, scale=scale)
    H = O.heterozygosity([h[sets[n]] for n in names])
    np.testing.assert_allclose(res["sums"][-1], [H[k] for k in res["stats"][1]], rtol=1e-13)


@pytest.mark.parametrize("P,method", [(3, "auto"), (4, "auto"), (3, "plain"), (4, "plain")])
def test_moments_genotype_three_and_four_pops(P, method):
    """Genotype moments statistics for 3 and 4 populations (45 / 105
    statistics, every case class of genotype_kernels.py:27-405 including the
    all-distinct pi2 at P = 4) through moments_ld.compute_ld_statistics,
    fused and per pair, against the oracle; missing data and the union
    biallelic filter."""
    rng = np.random.default_rng(60 + P)
    g = SY.hwe_genotypes(16 * P, 220, rng, missing=0.04)
    g[:, 5] = 0   # monomorphic: dropped by the union filter
    pos = SY.uniform_positions(220, 44_000, rng)
    names = [f"g{k}" for k in range(P)]
    sets = {n: list(range(16 * k, 16 * (k + 1))) for k, n in enumerate(names)}
    gm = GenotypeMatrix(g, pos, sample_sets=sets)
    bins = np.array([0.0, 400.0, 1500.0, 6000.0])
    res = moments_ld.compute_ld_statistics(genotype_matrix=gm, pops=names, bp_bins=bins, use_genotypes=True,
                                           report=False, method=method)
    rows = [sets[n] for n in names]
    want = O.moments_ld_sums(g, pos, rows, bins, use_genotypes=True)
    scale = O.moments_ld_sums(g, pos, rows, bins, use_genotypes=True, absolute=True)
    assert len(res["stats"][0]) == {3: 45, 4: 105}[P]
    close_sums(np.array(res["sums"][:-1]), want, rtol=1e-11, scale=scale)
    H = O.heterozygosity([g[r] for r in rows], use_genotypes=True)
    np.testing.assert_allclose(res["sums"][-1], [H[k] for k in res["stats"][1]], rtol=1e-13)


def test_moments_many_bins_fall_back_to_the_per_pair_path():
    """Three populations and 2-bp bins: a 64-SNP tile touches more bins than
    the fused kernel's LDS holds (PGLD_EUNSUPPORTED); moments_ld and the
    HaplotypeMatrix band sums take the per-pair path instead and still
    match the oracle."""
    rng = np.random.default_rng(91)
    n_var = 260
    h = SY.neutral_haplotypes(60, n_var, rng, missing=0.03)
    pos = np.arange(n_var, dtype=np.int64) * 2
    names = ["a", "b", "c"]
    sets = {n: list(range(20 * k, 20 * (k + 1))) for k, n in enumerate(names)}
    bins = np.arange(0.0, 125.0, 2.0)   # 62 bins
    rows = [sets[n] for n in names]
    with pytest.raises(L.PgldError) as e:
        band.ld_sums(h, "hap", pos, rows, float(bins[-1]), edges=bins, method="tiled")
    assert e.value.status == L.EUNSUPPORTED
    hm = HaplotypeMatrix(h, pos, sample_sets=sets)
    res = moments_ld.compute_ld_statistics(haplotype_matrix=hm, pops=names, bp_bins=bins, report=False)
    want = O.moments_ld_sums(h, pos, rows, bins)
    scale = O.moments_ld_sums(h, pos, rows, bins, absolute=True)
    close_sums(np.array(res["sums"][:-1]), want, scale=scale)
    sums, kp = band.ld_sums(h, "hap", pos, rows, float(bins[-1]), edges=bins, stat_set="self")
    ii, jj = O.pairs_within_distance(pos, float(bins[-1]))
    key = O.bin_index(pos[jj] - pos[ii], bins)
    nb = len(bins) - 1
    np.testing.assert_array_equal(np_(kp), np.bincount(key[(key >= 0) & (key < nb)], minlength=nb))


def test_population_above_the_fused_sample_limit_takes_the_per_pair_path():
    """A population of more than 32,767 haplotypes is beyond the fused
    kernels (int16 count stash); the band sums run per pair."""
    rng = np.random.default_rng(92)
    n_var = 40
    h = SY.neutral_haplotypes(33000, n_var, rng, missing=0.01)
    pos = SY.uniform_positions(n_var, 4000, rng)
    bins = np.array([0.0, 300.0, 1000.0, 4000.0])
    assert not band.fused_supported("hap", 1, "moments", [33000])
    hm = HaplotypeMatrix(h, pos)
    got = hm.compute_ld_statistics_gpu_single_pop(bins, raw=True, ac_filter=False)
    want = O.moments_ld_sums(h, pos, [list(range(33000))], bins)
    scale = O.moments_ld_sums(h, pos, [list(range(33000))], bins, absolute=True)
    close_sums(np.array([got[(float(bins[k]), float(bins[k + 1]))] for k in range(3)]), want, scale=scale)


@pytest.mark.parametrize("use_genotypes", [False, True])
def test_moments_r_bins(tmp_path, use_genotypes):
    """r_bins: pairs within the physical span derived from the map, keyed by
    the genetic distance |g_j - g_i| in Morgans."""
    rng = np.random.default_rng(77)
    n_var = 400
    pos = SY.uniform_positions(n_var, 200000, rng)
    mp = np.linspace(0, 210000, 12)
    cm = np.cumsum(rng.uniform(0.05, 0.4, size=12))
    rec = tmp_path / "rec.map"
    rec.write_text("pos\tcM\n" + "".join(f"{a:.1f}\t{b:.6f}\n" for a, b in zip(mp, cm)))
    r_bins = np.array([0.0, 1e-4, 5e-4, 2e-3])
    if use_genotypes:
        g = SY.hwe_genotypes(60, n_var, rng, missing=0.02)
        sets = {"a": list(range(30)), "b": list(range(30, 60))}
        mat = GenotypeMatrix(g, pos, sample_sets=sets)
        kw = dict(genotype_matrix=mat)
    else:
        g = SY.neutral_haplotypes(80, n_var, rng, missing=0.02)
        sets = {"a": list(range(40)), "b": list(range(40, 80))}
        mat = HaplotypeMatrix(g, pos, sample_sets=sets)
        kw = dict(haplotype_matrix=mat)
    gen = np.interp(pos.astype(np.float64), mp, cm / 100.0)
    max_bp = moments_ld._max_bp_for_r_dist(pos.astype(np.float64), gen, float(r_bins[-1]))
    res = moments_ld.compute_ld_statistics(rec_map_file=str(rec), pops=["a", "b"], r_bins=r_bins,
                                           use_genotypes=use_genotypes, report=False, **kw)
    kw = dict(use_genotypes=use_genotypes, distances=gen, max_dist=max_bp)
    want = O.moments_ld_sums(g, pos, [sets["a"], sets["b"]], r_bins, **kw)
    scale = O.moments_ld_sums(g, pos, [sets["a"], sets["b"]], r_bins, absolute=True, **kw)
    close_sums(np.array(res["sums"][:-1]), want, rtol=1e-11 if use_genotypes else 1e-12, scale=scale)


# ---------------------------------------------------------------------------
def test_haplotype_matrix_vcf_layout_and_filters():
    hm = HaplotypeMatrix.from_vcf(VCF)
    hm.load_pop_file(POPS, pops=["deme0", "deme1"])
    n = len(hm.samples)
    for rows in hm.sample_sets.values():
        assert all((r + n in rows) for r in rows if r < n)
    h = np_(hm.haplotypes)
    hm.transfer_to_gpu()
    alt, valid = (np_(a) for a in hm.allele_counts(hm.sample_sets["deme0"]))
    X = h[hm.sample_sets["deme0"]]
    np.testing.assert_array_equal(alt, np.maximum(X, 0).astype(np.int64).sum(axis=0))
    np.testing.assert_array_equal(valid, (X >= 0).sum(axis=0))
    f = hm.apply_biallelic_filter()
    np.testing.assert_array_equal(np_(f.positions), np_(hm.positions)[O.biallelic_filter_hap(h)])


def test_haplotype_matrix_tally_and_bin_stats():
    rng = np.random.default_rng(3)
    h = SY.neutral_haplotypes(64, 120, rng, missing=0.0)
    pos = SY.uniform_positions(120, 30000, rng)
    hm = HaplotypeMatrix(h, pos, sample_sets={"x": list(range(32)), "y": list(range(32, 64))})
    c, nv = hm.tally_gpu_haplotypes()
    iu = np.triu_indices(120, k=1)
    want, _ = O.hap_counts(h, iu[0], iu[1], None)
    np.testing.assert_array_equal(np_(c), want)
    c2, n1, n2 = hm.tally_gpu_haplotypes_two_pops("x", "y")
    np.testing.assert_array_equal(np_(c2)[:, :4], O.hap_counts(h, iu[0], iu[1], list(range(32)))[0])
    np.testing.assert_array_equal(np_(c2)[:, 4:], O.hap_counts(h, iu[0], iu[1], list(range(32, 64)))[0])
    bins = [0.0, 1000.0, 5000.0, 20000.0]
    one = hm.compute_ld_statistics_gpu_single_pop(bins, raw=True)
    keep = O.biallelic_filter_hap(h)
    args1 = (h[:, keep], pos[keep], [list(range(64))], np.array(bins))
    close_sums(np.array([one[k] for k in sorted(one)]), O.moments_ld_sums(*args1),
               scale=O.moments_ld_sums(*args1, absolute=True))
    two = hm.compute_ld_statistics_gpu_two_pops(bins, "x", "y", raw=True)
    args2 = (h[:, keep], pos[keep], [list(range(32)), list(range(32, 64))], np.array(bins))
    close_sums(np.array([list(two[k].values()) for k in sorted(two)]), O.moments_ld_sums(*args2),
               scale=O.moments_ld_sums(*args2, absolute=True))


def test_haplotype_matrix_windowed_sums():
    rng = np.random.default_rng(9)
    h = SY.neutral_haplotypes(50, 500, rng, missing=0.02)
    pos = SY.uniform_positions(500, 100000, rng)
    hm = HaplotypeMatrix(h, pos, sample_sets={"a": list(range(25)), "b": list(range(25, 50))})
    sums, kp, names = hm.windowed_ld_sums(10000.0, pops=["a", "b"])
    ii, jj = O.pairs_within_distance(pos, 10000.0)
    key = O.window_key(pos[ii], 0.0, 10000.0)
    cl = [O.hap_counts(h, ii, jj, r) for r in (list(range(25)), list(range(25, 50)))]
    st = O.two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
    nk = sums.shape[0]
    close_sums(np_(sums), O.keyed_sums(key, st, nk), scale=O.keyed_sums(key, np.abs(st), nk))
    np.testing.assert_array_equal(np_(kp), np.bincount(key, minlength=nk)[:nk])
    assert names == O.ld_names(2)


def test_genotype_matrix_roundtrip_and_bins():
    rng = np.random.default_rng(12)
    h = SY.neutral_haplotypes(60, 200, rng, missing=0.0)
    pos = SY.uniform_positions(200, 50000, rng)
    hm = HaplotypeMatrix(h, pos, sample_sets={"a": list(range(30)), "b": list(range(30, 60))})
    gm = GenotypeMatrix.from_haplotype_matrix(hm)
    g = np_(gm.genotypes)
    np.testing.assert_array_equal(g, h[0::2] + h[1::2])
    assert gm.sample_sets["a"] == list(range(15))
    back = gm.to_haplotype_matrix()
    np.testing.assert_array_equal(np_(back.haplotypes)[0::2] + np_(back.haplotypes)[1::2], g)
    bins = [0.0, 2000.0, 10000.0]
    sums, kp, names = gm.ld_bin_sums(bins, pops=["a", "b"])
    ii, jj = O.pairs_within_distance(pos, 10000.0)
    key = O.bin_index(pos[jj] - pos[ii], np.array(bins))
    cl = [O.geno_counts(g, ii, jj, r) for r in (list(range(15)), list(range(15, 30)))]
    st = O.multi_pop_stats_geno([c for c, _ in cl], [nv for _, nv in cl])
    close_sums(np_(sums), O.keyed_sums(key, st, 2), rtol=1e-11, scale=O.keyed_sums(key, np.abs(st), 2))
    np.testing.assert_array_equal(np_(kp), np.bincount(key[key >= 0], minlength=2)[:2])
    assert names == O.ld_names(2)


def test_sigma_d2_geno():
    rng = np.random.default_rng(21)
    g = SY.hwe_genotypes(40, 60, rng, missing=0.0)
    gm = GenotypeMatrix(g, np.arange(60) * 10)
    got = np_(LSG.sigma_d2_geno(gm))
    iu = np.triu_indices(60, k=1)
    c, _ = O.geno_counts(g, iu[0], iu[1], None)
    want = O.sigma_d2_geno_from_counts(c)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-12)
    gmiss = g.copy()
    gmiss[0, 0] = -1
    with pytest.raises(ValueError):
        LSG.sigma_d2_geno(GenotypeMatrix(gmiss, np.arange(60) * 10))
    with pytest.raises(TypeError):
        LSG.sigma_d2_geno(g)


# ---------------------------------------------------------------------------
# r^2-family estimators, ZnS, Omega, pairwise_r2, windowed_r_squared
# ---------------------------------------------------------------------------
from pg_gpu_amd import band, ld_statistics as LS  # noqa: E402


def _ld_panel(seed, n_hap=90, n_var=150, missing=0.03):
    rng = np.random.default_rng(seed)
    h = SY.neutral_haplotypes(n_hap, n_var, rng, missing=missing)
    h[:, 7] = 0  # a monomorphic site
    return h


@pytest.mark.parametrize("missing", [0.0, 0.04])
def test_ld_pair_estimators_exact(missing):
    h = _ld_panel(61, missing=missing)
    m = h.shape[1]
    plan = band.BandPlan(np.arange(m, dtype=np.float64), np.inf, window=1e300, n_keys=1)
    pk = band.pack_population(h, None, "hap")
    ii, jj = np.triu_indices(m, k=1)
    c, _ = O.hap_counts(h, ii, jj)
    naive = O.naive_r2_matrix(h)
    sig, sval = O.sigma_d2_matrix(h)
    for est, want, counted in (("r2", O.r_squared_exact(c), None),
                               ("naive", naive[ii, jj], np.ones(len(ii), bool)),
                               ("sigma_d2", sig[ii, jj], sval[ii, jj])):
        out = plan.r2(pk, histogram=False, sparse_cap=0, per_pair=True, estimator=est, snp_sums=True)
        got = np_(out["r2"])
        np.testing.assert_array_equal(np.isnan(got), np.isnan(want), err_msg=est)
        ok = ~np.isnan(want)
        np.testing.assert_array_equal(got[ok], want[ok], err_msg=est)
        if counted is None:
            counted = ok
        vals = np.where(counted, np.nan_to_num(want), 0.0)
        Mv = np.zeros((m, m))
        Mc = np.zeros((m, m), dtype=np.int64)
        Mv[ii, jj] = vals
        Mc[ii, jj] = counted
        np.testing.assert_allclose(np_(out["row_sums"]), Mv.sum(axis=1), rtol=1e-12, atol=1e-12, err_msg=est)
        np.testing.assert_allclose(np_(out["col_sums"]), Mv.sum(axis=0), rtol=1e-12, atol=1e-12, err_msg=est)
        np.testing.assert_array_equal(np_(out["row_counts"]), Mc.sum(axis=1))
        np.testing.assert_array_equal(np_(out["col_counts"]), Mc.sum(axis=0))


@pytest.mark.parametrize("missing", [0.0, 0.04])
def test_zns_omega_pairwise_r2(missing):
    h = _ld_panel(62, n_hap=70, n_var=120, missing=missing)
    hm = HaplotypeMatrix(h, np.arange(h.shape[1]) * 100)
    seg = O.segregating_sites(h)
    hs = h[:, seg]
    sig, valid = O.sigma_d2_matrix(hs)
    naive = O.naive_r2_matrix(hs)
    assert LS.zns(hm) == pytest.approx(O.zns_from_matrix(sig, valid), rel=1e-10)
    assert LS.zns(hm, estimator="r2") == pytest.approx(O.zns_from_matrix(naive), rel=1e-10)
    assert LS.omega(hm) == pytest.approx(O.omega_from_matrix(sig), rel=1e-10)
    assert LS.omega(hm, estimator="r2") == pytest.approx(O.omega_from_matrix(naive), rel=1e-10)
    # precomputed arrays go through the same host formulas
    assert LS.zns(naive) == pytest.approx(O.zns_from_matrix(naive), rel=1e-12)
    assert LS.omega(naive) == pytest.approx(O.omega_from_matrix(naive), rel=1e-12)
    full = np_(HaplotypeMatrix(h, np.arange(h.shape[1]) * 100).pairwise_r2())
    np.testing.assert_array_equal(full, O.naive_r2_matrix(h))
    # the D matrix of pairwise_LD_v (haplotype_matrix.py:1132-1136), same frequencies
    D = np_(HaplotypeMatrix(h, np.arange(h.shape[1]) * 100).pairwise_LD_v())
    np.testing.assert_allclose(D, O.naive_d_matrix(h), rtol=0, atol=1e-15)
    # sigma_d2 ZnS keeps every segregating site whatever missing_data says (the
    # reference's internal 'project' mode, ld_statistics.py:721-726); the r2
    # ZnS and Omega honour 'exclude'; 'rogers_huff' on a HaplotypeMatrix is the
    # frequency r^2 path (ld_statistics.py:721-731, :800-808)
    ex = HaplotypeMatrix(h, np.arange(h.shape[1]) * 100)
    assert LS.zns(ex, missing_data="exclude") == pytest.approx(O.zns_from_matrix(sig, valid), rel=1e-10)
    assert LS.zns(hm, estimator="rogers_huff") == pytest.approx(O.zns_from_matrix(naive), rel=1e-10)
    assert LS.omega(hm, estimator="rogers_huff") == pytest.approx(O.omega_from_matrix(naive), rel=1e-10)
    hx = h[:, np.nonzero((h >= 0).all(axis=0))[0]]
    hx = hx[:, O.segregating_sites(hx)]
    sigx, _ = O.sigma_d2_matrix(hx)
    assert LS.zns(hm, estimator="r2", missing_data="exclude") == pytest.approx(
        O.zns_from_matrix(O.naive_r2_matrix(hx)), rel=1e-10, abs=1e-300)
    assert LS.omega(hm, missing_data="exclude") == pytest.approx(O.omega_from_matrix(sigx), rel=1e-10, abs=1e-300)



def _close_or_nan(got, want, rtol):
    got, want = np.asarray(got), np.asarray(want)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=rtol, atol=1e-12)


@pytest.mark.parametrize("missing", [0.0, 0.03])
def test_windowed_zns_omega(missing):
    """Per-window ZnS / Omega of windowed_statistics_fused against the
    oracle's window loop: bins (short and empty windows included), a named
    population, overlapping explicit windows, missing_data='exclude'."""
    from pg_gpu_amd import windowed_analysis as WA
    rng = np.random.default_rng(64)
    h = _ld_panel(64, n_hap=80, n_var=400, missing=missing)
    pos = np.sort(rng.choice(np.arange(1, 200_000), size=h.shape[1], replace=False))
    pos[50:53] = pos[50]  # duplicate positions stay in one window
    h[:, 60:64] = 0       # a window stretch without segregating sites
    bins = np.array([0, 3_000, 20_000, 20_500, 60_000, 61_000, 120_000, 150_000, 200_000, 250_000])
    hm = HaplotypeMatrix(h, pos, sample_sets={"a": list(range(0, 50)), "b": list(range(50, 80))})
    got = WA.windowed_statistics_fused(hm, bins, statistics=("zns", "omega"))
    want = O.windowed_zns_omega(h, pos, bins[:-1], bins[1:])
    assert list(got)[:6] == ["chrom", "start", "end", "center", "n_variants", "window_id"]
    np.testing.assert_array_equal(got["n_variants"], np.diff(np.searchsorted(pos, bins)))
    _close_or_nan(got["zns"], want[0], 1e-10)
    _close_or_nan(got["omega"], want[1], 1e-10)
    sub = h[hm.sample_sets["b"]]
    got = WA.windowed_statistics_fused(hm, bins, statistics=("zns",), population="b")
    _close_or_nan(got["zns"], O.windowed_zns_omega(sub, pos, bins[:-1], bins[1:])[0], 1e-10)
    assert "omega" not in got
    starts = np.arange(0, 180_000, 10_000)
    stops = starts + 25_000  # windows overlap by 15 kb: three layers
    got = WA.windowed_statistics_fused(hm, None, statistics=("omega", "zns"), _win_starts=starts,
                                       _win_stops=stops, missing_data="exclude")
    want = O.windowed_zns_omega(h, pos, starts, stops, missing_data="exclude")
    _close_or_nan(got["zns"], want[0], 1e-10)
    _close_or_nan(got["omega"], want[1], 1e-10)
    with pytest.raises(NotImplementedError):
        WA.windowed_statistics_fused(hm, bins, statistics=("pi", "zns"))
    with pytest.raises(ValueError):
        WA.windowed_statistics_fused(hm, bins, statistics=("zns",), population="c")

def test_windowed_r_squared():
    h = _ld_panel(63, n_hap=80, n_var=300, missing=0.02)
    rng = np.random.default_rng(4)
    pos = SY.uniform_positions(300, 60000, rng)
    hm = HaplotypeMatrix(h, pos, sample_sets={"a": list(range(40)), "b": list(range(40, 80))})
    bins = [0, 500, 2000, 8000, 30000]
    res, cnt = hm.windowed_r_squared(bins, percentile=[10, 50, 90], pop="b")
    ii, jj = np.triu_indices(300, k=1)
    c, nv = O.hap_counts(h, ii, jj, list(range(40, 80)))
    r2 = O.r_squared(c, nv)
    k = O.bin_index(pos[jj] - pos[ii], np.array(bins, dtype=float))
    ok = (k >= 0) & (k < 4) & ~np.isnan(r2)
    for b in range(4):
        sel = r2[ok & (k == b)]
        assert cnt[b] == len(sel)
        np.testing.assert_allclose(res[b], np.percentile(sel, [10, 50, 90]), rtol=1e-12, atol=1e-15)
    res1, _ = hm.windowed_r_squared(bins)
    assert res1.shape == (4,)


def test_rogers_huff_r():
    rng = np.random.default_rng(71)
    g = SY.hwe_genotypes(45, 80, rng)
    g[:, 3] = 1  # constant site -> NaN
    gm = GenotypeMatrix(g, np.arange(80) * 10)
    got = np_(LS.rogers_huff_r(gm))
    want = O.rogers_huff_r(g)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(np_(LS.rogers_huff_r_squared(gm))[ok], want[ok] ** 2, rtol=1e-12, atol=1e-14)
    h = SY.neutral_haplotypes(60, 50, rng)
    hm = HaplotypeMatrix(h, np.arange(50) * 10)
    w2 = O.rogers_huff_r(h[0::2] + h[1::2])
    g2 = np_(LS.rogers_huff_r(hm))
    np.testing.assert_array_equal(np.isnan(g2), np.isnan(w2))
    np.testing.assert_allclose(g2[~np.isnan(w2)], w2[~np.isnan(w2)], rtol=1e-12, atol=1e-14)
    full = np_(hm.pairwise_r2(estimator="rogers_huff"))
    ii, jj = np.triu_indices(50, k=1)
    np.testing.assert_allclose(full[ii, jj], np.nan_to_num(w2 ** 2), rtol=1e-12, atol=1e-14)
    assert np.all(np.diag(full) == 0)
    bad = h.copy()
    bad[0, 0] = -1
    with pytest.raises(ValueError):
        LS.rogers_huff_r(HaplotypeMatrix(bad, np.arange(50) * 10))
    with pytest.raises(ValueError):
        LS.rogers_huff_r(HaplotypeMatrix(h[:59], np.arange(50) * 10))
    with pytest.raises(TypeError):
        LS.rogers_huff_r(g)


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_genotype_zns_omega(missing):
    rng = np.random.default_rng(72)
    g = SY.hwe_genotypes(50, 90, rng, missing=missing)
    gm = GenotypeMatrix(g, np.arange(90) * 10)
    R = O.dosage_r2_matrix(g)
    assert LS.zns(gm) == pytest.approx(O.zns_from_matrix(R), rel=1e-10)
    assert LS.omega(gm) == pytest.approx(O.omega_from_matrix(R), rel=1e-10)


def test_windowed_r_squared_rogers_huff():
    rng = np.random.default_rng(73)
    h = SY.neutral_haplotypes(60, 200, rng)
    pos = SY.uniform_positions(200, 40000, rng)
    hm = HaplotypeMatrix(h, pos)
    bins = [0, 1000, 5000, 20000]
    res, cnt = hm.windowed_r_squared(bins, percentile=[25, 75], estimator="rogers_huff")
    ii, jj = np.triu_indices(200, k=1)
    r2 = O.rogers_huff_r(h[0::2] + h[1::2]) ** 2
    k = O.bin_index(pos[jj] - pos[ii], np.array(bins, dtype=float))
    for b in range(3):
        sel = r2[(k == b) & ~np.isnan(r2)]
        assert cnt[b] == len(sel)
        np.testing.assert_allclose(res[b], np.percentile(sel, [25, 75]), rtol=1e-11, atol=1e-14)


def test_locate_unlinked():
    rng = np.random.default_rng(74)
    base = SY.neutral_haplotypes(50, 40, rng, missing=0.02)
    h = np.repeat(base, 3, axis=1)  # runs of identical sites: strongly linked
    h[:, 1::3] = SY.neutral_haplotypes(50, 40, rng, missing=0.02)
    hm = HaplotypeMatrix(h, np.arange(120) * 50)
    for size, step, thr in ((30, 10, 0.1), (12, 5, 0.5)):
        got = hm.locate_unlinked(size=size, step=step, threshold=thr)
        np.testing.assert_array_equal(got, O.locate_unlinked(h, size, step, thr))
    assert not got.all() and got.any()
