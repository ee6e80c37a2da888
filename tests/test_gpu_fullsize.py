"""GPU checks at the benchmark workloads' full sizes.

The oracle cannot run these sizes in seconds, so the full-size checks are
properties that do not depend on size: pair counts per key equal the closed
form from the positions (integers, bit-exact), repeated runs are bit-identical,
the fused band pass agrees with the plain per-pair path (two different GPU
algorithms) on a contiguous sub-panel, and the per-pair plain workload
(2,000 SNPs x 200 haplotypes, all pairs within 20 kb) is checked against the
oracle in full.
"""
import numpy as np
import pytest

from oracle import ld_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from pg_gpu_amd import band, batch, ld_pipeline as LP, ld_statistics as LS  # noqa: E402
from pg_gpu_amd import synthetic as SY  # noqa: E402
from pg_gpu_amd import _lib as L  # noqa: E402


EDGES = np.logspace(0, 6, 17)


def np_(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def pairs_per_key_window(pos, max_dist, width, n_keys):
    per = np.searchsorted(pos, pos + max_dist, side="right") - np.arange(len(pos)) - 1
    key = np.floor(pos / width).astype(np.int64)
    ok = key < n_keys
    return np.bincount(key[ok], weights=per[ok], minlength=n_keys)


def pairs_per_key_bins(pos, edges):
    """Pairs with distance in [e_k, e_{k+1}) for every bin, by searchsorted."""
    out = np.zeros(len(edges) - 1)
    for k in range(len(edges) - 1):
        lo = np.searchsorted(pos, pos + edges[k], side="left")
        hi = np.searchsorted(pos, pos + edges[k + 1], side="left")
        lo = np.maximum(lo, np.arange(len(pos)) + 1)
        out[k] = np.maximum(hi - lo, 0).sum()
    return out


def plain_sums(data, family, pos, rows, max_dist, n_keys, mode, edges=None, width=1.0, two_pop=False):
    dev = L.device()
    sums = torch.zeros((n_keys, 15 if two_pop else 3), dtype=torch.float64, device=dev)
    pos_d = torch.as_tensor(np.asarray(pos, dtype=np.float64), device=dev)
    e_d = None if edges is None else torch.as_tensor(edges, device=dev)
    specs = LP.generate_stat_specs(1)
    for ii, jj in LP.iter_pairs_within_distance(pos, max_dist, 5_000_000):
        if family == "geno":
            c, nv = LP.compute_genotype_counts_for_pairs(data, ii, jj, rows[0])
            st = LP.stats_from_counts("geno", [c], [nv], specs)
        else:
            cl = [LP.compute_counts_for_pairs(data, ii, jj, r) for r in rows]
            st = LP.compute_two_pop_statistics_batch(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
        keys = torch.empty(ii.shape[0], dtype=torch.int32, device=dev)
        L.call("pgld_pair_keys", L.ptr(pos_d), L.ptr(ii), L.ptr(jj), ii.shape[0], mode, L.ptr(e_d), n_keys, 0.0,
               width, L.ptr(keys), L.stream_handle())
        L.call("pgld_keyed_sums", L.ptr(st), L.ptr(keys), ii.shape[0], st.shape[1], n_keys, L.ptr(sums),
               L.stream_handle())
    return np_(sums)


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def windowed_panel():
    print("[fullsize] windowed haplotype panel 100,000 SNPs x 2,000 haplotypes", flush=True)
    hap, pos, pops = SY.windowed_haplotype_panel()
    return hap, pos, pops


def test_windowed_panel_full_size(windowed_panel):
    hap, pos, pops = windowed_panel
    hap_d = torch.as_tensor(hap, device=L.device())
    plan = band.BandPlan(pos, 100_000.0, window=100_000.0, n_keys=100)
    packed = [band.pack_population(hap_d, r, "hap") for r in pops]
    assert all(p.has_missing for p in packed)
    s1, kp = plan.sums(packed, want_pairs=True)
    s2 = plan.sums(packed)
    np.testing.assert_array_equal(np_(s1), np_(s2))  # fixed-order reduction
    want = pairs_per_key_window(pos.astype(np.float64), 100_000.0, 100_000.0, 100)
    np.testing.assert_array_equal(np_(kp), want)
    assert int(np_(kp).sum()) == plan.n_pairs == 111_173_550
    # the pair counts of the MFMA path equal the plain kernel's on every pair
    # of the first 3,000 SNPs' band
    sub = slice(0, 3000)
    sp = band.BandPlan(pos[sub], 100_000.0, window=100_000.0, n_keys=100)
    pk = band.pack_population(hap_d[:, sub].contiguous(), pops[0], "hap")
    got = np_(sp.counts(pk))
    ii, jj = next(iter(LP.iter_pairs_within_distance(pos[sub], 100_000.0, 10**9)))
    c, _ = LP.compute_counts_for_pairs(hap_d[:, sub].contiguous(), ii, jj, pops[0])
    np.testing.assert_array_equal(got, np_(c))


def test_windowed_panel_fused_vs_plain_subpanel(windowed_panel):
    hap, pos, pops = windowed_panel
    sub = slice(0, 20_000)
    h = torch.as_tensor(np.ascontiguousarray(hap[:, sub]), device=L.device())
    p = pos[sub]
    plan = band.BandPlan(p, 100_000.0, window=100_000.0, n_keys=100)
    fused = np_(plan.sums([band.pack_population(h, r, "hap") for r in pops]))
    ref = plain_sums(h, "hap", p, pops, 100_000.0, 100, L.KEY_WINDOW, width=100_000.0, two_pop=True)
    nz = np.abs(ref) > 0
    assert rel_err(fused[nz], ref[nz]) < 1e-9
    np.testing.assert_array_equal(fused[~nz], 0.0)


# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def genotype_panel():
    print("[fullsize] genotype panel 50,000 SNPs x 1,000 diploids", flush=True)
    return SY.genotype_panel()


def test_genotype_panel_full_size(genotype_panel):
    g, pos = genotype_panel
    edges = np.logspace(0, 6, 17)
    g_d = torch.as_tensor(g, device=L.device())
    plan = band.BandPlan(pos, 1e6, edges=edges)
    assert plan.n_pairs == 50_000 * 49_999 // 2
    packed = [band.pack_population(g_d, None, "geno")]
    s1, kp = plan.sums(packed, want_pairs=True)
    s2 = plan.sums(packed)
    np.testing.assert_array_equal(np_(s1), np_(s2))
    np.testing.assert_array_equal(np_(kp), pairs_per_key_bins(pos.astype(np.float64), edges))
    assert np.all(np.isfinite(np_(s1)))


def test_genotype_panel_fused_vs_plain_subpanel(genotype_panel):
    g, pos = genotype_panel
    sub = slice(0, 6000)
    edges = np.logspace(0, 6, 17)
    gd = torch.as_tensor(np.ascontiguousarray(g[:, sub]), device=L.device())
    p = pos[sub]
    plan = band.BandPlan(p, 1e6, edges=edges)
    fused = np_(plan.sums([band.pack_population(gd, None, "geno")]))
    ref = plain_sums(gd, "geno", p, [None], 1e6, 16, L.KEY_BIN, edges=edges)
    nz = np.abs(ref) > 0
    assert rel_err(fused[nz], ref[nz]) < 1e-9
    np.testing.assert_array_equal(fused[~nz], 0.0)


# ---------------------------------------------------------------------------
def test_plain_haplotype_panel_full_size_vs_oracle():
    """2,000 SNPs x 200 haplotypes, all pairs within 20 kb: counts bit-exact,
    DD/Dz/pi2 and r^2 per pair against the oracle, through the plain path and
    through the MFMA band counts."""
    hap, pos = SY.plain_haplotype_panel()
    hap_d = torch.as_tensor(hap, device=L.device())
    ii, jj = next(iter(LP.iter_pairs_within_distance(pos, 20_000, 10**9)))
    ei, ej = O.pairs_within_distance(pos, 20_000)
    np.testing.assert_array_equal(np_(ii), ei)
    np.testing.assert_array_equal(np_(jj), ej)
    c, nv = LP.compute_counts_for_pairs(hap_d, ii, jj)
    want_c, want_nv = O.hap_counts(hap, ei, ej, None)
    assert nv is None and want_nv is None
    np.testing.assert_array_equal(np_(c), want_c)
    st = np_(LP.compute_single_pop_statistics_batch(c, nv))
    want = O.multi_pop_stats_hap([want_c], [None])
    np.testing.assert_allclose(st, want, rtol=1e-12, atol=1e-15)
    r2 = np_(LS.r_squared(c))
    w2 = O.r_squared(want_c)
    np.testing.assert_array_equal(np.isnan(r2), np.isnan(w2))
    ok = ~np.isnan(w2)
    np.testing.assert_allclose(r2[ok], w2[ok], rtol=1e-12, atol=1e-15)
    plan = band.BandPlan(pos, 20_000.0, window=1e9, n_keys=1)
    mc = np_(plan.counts(band.pack_population(hap_d, None, "hap")))
    np.testing.assert_array_equal(mc, want_c)


# ---------------------------------------------------------------------------
def test_windowed_panel_fp4_equals_int8_full_size(windowed_panel):
    """The FP4 planes the benchmark runs on give the int8 planes' window sums
    bit for bit on the whole 100,000-SNP x 2,000-haplotype panel (1,000
    samples per population: 4 FP4 K chunks)."""
    hap, pos, pops = windowed_panel
    hap_d = torch.as_tensor(hap, device=L.device())
    plan = band.BandPlan(pos, 100_000.0, window=100_000.0, n_keys=100)
    s8, k8 = plan.sums([band.pack_population(hap_d, r, "hap") for r in pops], want_pairs=True)
    s4, k4 = plan.sums([band.pack_population(hap_d, r, "hap", fp4=True) for r in pops], want_pairs=True)
    np.testing.assert_array_equal(np_(s4), np_(s8))
    np.testing.assert_array_equal(np_(k4), np_(k8))


def test_genotype_panel_fp4_equals_int8_full_size(genotype_panel):
    """The same on the whole 50,000-SNP x 1,000-diploid genotype panel (all
    1.25e9 pairs, 16 log bins; 4 FP4 K chunks with three planes)."""
    g, pos = genotype_panel
    g_d = torch.as_tensor(g, device=L.device())
    plan = band.BandPlan(pos, 1e6, edges=np.logspace(0, 6, 17))
    s8 = np_(plan.sums([band.pack_population(g_d, None, "geno")]))
    s4 = np_(plan.sums([band.pack_population(g_d, None, "geno", fp4=True)]))
    np.testing.assert_array_equal(s4, s8)


# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def dense_panel():
    print("[fullsize] dense r^2 panel 40,000 SNPs x 10,000 haplotypes", flush=True)
    return SY.dense_r2_panel()


def _hap_r2_exact(n11, na, nb, S):
    """r^2 from integer counts in float64, the kernel's own operation order:
    (S n11 - nA nB)^2 / ((nA (S - nA)) (nB (S - nB)))."""
    d = (S * n11 - na * nb).astype(np.float64)
    den = (na * (S - na)).astype(np.float64) * (nb * (S - nb)).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(den > 0, (d * d) / den, np.nan)


def test_dense_r2_full_size(dense_panel):
    """All 7.9998e8 pairs of the 40,000 x 10,000 dense panel: the histogram
    holds every pair once; FP4 planes give the int8 planes' histogram and
    sparse list bit for bit; every listed pair (r^2 > 0.2) has the r^2 of its
    exact counts recomputed independently (float64 products of the 0/1
    columns); and on 40 sampled rows the list is complete -- every pair
    (i, j > i) of those rows with r^2 > 0.2 is in it, no other."""
    hap, pos = dense_panel
    dev = L.device()
    hap_d = torch.as_tensor(hap, device=dev)
    S, n = hap.shape
    plan = band.BandPlan(pos, np.inf, window=1e15, n_keys=1)
    assert plan.n_pairs == n * (n - 1) // 2
    kw = dict(n_hist=100, thresh=0.2, sparse_cap=1 << 22, with_missing=False)
    out4 = plan.r2(band.pack_population(hap_d, None, "hap", fp4=True), **kw)
    out8 = plan.r2(band.pack_population(hap_d, None, "hap"), **kw)
    h4, h8 = np_(out4["hist"]), np_(out8["hist"])
    np.testing.assert_array_equal(h4, h8)
    np.testing.assert_array_equal(np_(out4["nan_pairs"]), np_(out8["nan_pairs"]))
    assert int(h4.sum()) + int(np_(out4["nan_pairs"]).sum()) == plan.n_pairs
    assert out4["n_sparse"] == out8["n_sparse"] == len(out4["sparse_i"]) > 0
    for k in ("sparse_i", "sparse_j", "sparse_r2"):
        np.testing.assert_array_equal(np_(out4[k]), np_(out8[k]))
    si, sj, sr = (np_(out4[k]) for k in ("sparse_i", "sparse_j", "sparse_r2"))
    # listed pairs against counts recomputed independently (float64 dot
    # products of the 0/1 columns, exact)
    na = np.sum(hap, axis=0, dtype=np.int64)
    hf = hap_d.to(torch.float64)
    ti, tj = (torch.as_tensor(x.astype(np.int64), device=dev) for x in (si, sj))
    n11 = np_((hf[:, ti] * hf[:, tj]).sum(0).round()).astype(np.int64)
    want = _hap_r2_exact(n11, na[si], na[sj], S)
    assert np.all(want > 0.2)
    np.testing.assert_array_equal(sr, want.astype(np.float32))
    # completeness on sampled rows: all pairs (i, j > i) through a float64 GEMM
    rows = np.sort(np.random.default_rng(11).choice(n - 1, size=40, replace=False))
    n11_rows = np_((hf[:, rows].T @ hf).round()).astype(np.int64)  # 40 x n exact counts
    listed = set(zip(si.tolist(), sj.tolist()))
    for q, i in enumerate(rows):
        j = np.arange(i + 1, n)
        r2 = _hap_r2_exact(n11_rows[q, j], np.full(j.shape, na[i]), na[j], S)
        hit = set((int(i), int(x)) for x in j[r2 > 0.2])
        mine = set(p for p in listed if p[0] == i)
        assert hit == mine, f"row {i}: {len(hit ^ mine)} pairs differ"


def test_ragged_batch_full_size():
    """The benchmark's 512-region batch (500-20,000 SNPs x 50-5,000 haplotypes
    in 4 populations, 3 % missing, 3.5e10 pairs, 16 log bins): pair counts
    per (region, bin) equal the closed form from the positions; FP4 planes
    equal int8 planes bit for bit; the batch equals one fused launch per region
    bit for bit; the two regions of least work against the FP64 oracle."""
    print("[fullsize] ragged batch 512 regions", flush=True)
    dev = L.device()
    datas, pos, rows = SY.ragged_regions(n_regions=512, device=dev)
    rows_d = [[torch.as_tensor(np.asarray(r, dtype=np.int32), device=dev) for r in rr] for rr in rows]
    p4 = batch.pack_regions(datas, rows_d, fp4=True)
    b = batch.RegionBatch(pos, 1e6, EDGES).bind(p4, stat_set="self")
    assert b.n_pairs == sum(len(p) * (len(p) - 1) // 2 for p in pos)
    b.schedule(0, "blocked", False, kernel_variant=3)  # the schedule and variant the autotuner picks here
    s4, k4 = b.sums(True, want_pairs=True)
    s4, k4 = np_(s4), np_(k4)
    for r in range(len(pos)):
        np.testing.assert_array_equal(k4[r], pairs_per_key_bins(pos[r], EDGES))
    b8 = batch.RegionBatch(pos, 1e6, EDGES).bind(batch.pack_regions(datas, rows_d), stat_set="self")
    np.testing.assert_array_equal(np_(b8.sums(True)), s4)
    del b8
    for r in range(len(pos)):
        single = band.BandPlan(pos[r], 1e6, edges=EDGES)
        np.testing.assert_array_equal(np_(single.sums(p4[r], stat_set="self", with_missing=True)), s4[r])
    S = np.array([d.shape[0] for d in datas])
    work = np.array([len(p) * (len(p) - 1) // 2 * int(s) for p, s in zip(pos, S)])
    specs = [(t, [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(4)]
    for r in np.argsort(work, kind="stable")[:2]:
        h = np_(datas[r])
        ii, jj = O.pairs_within_distance(pos[r], np.inf)
        cl = [O.hap_counts(h, ii, jj, rw) for rw in rows[r]]
        st = O.multi_pop_stats_hap([c for c, _ in cl], [nv for _, nv in cl], specs)
        keys = O.bin_index(pos[r][jj] - pos[r][ii], EDGES)
        want = O.keyed_sums(keys, st, 16)
        scale = O.keyed_sums(keys, np.abs(st), 16)
        assert np.all(np.abs(s4[r] - want) <= 1e-12 * np.maximum(scale, 1e-300) + 1e-300), f"region {r}"
