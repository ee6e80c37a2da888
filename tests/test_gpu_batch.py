"""GPU parity of ragged region batches: one persistent launch over many
independent regions against the oracle and against the single-band path
region by region (bit-identical: same tiles, same per-key slot order), under
every schedule the autotuner may pick, and through a captured HIP graph
replayed several times."""
import os

import numpy as np
import pytest

from oracle import ld_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

from pg_gpu_amd import band, batch  # noqa: E402
from pg_gpu_amd import synthetic as SY  # noqa: E402
from pg_gpu_amd import _lib as L  # noqa: E402

EDGES = np.logspace(0, 6, 17)


def np_(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def self_specs(P):
    return [(f"{t}", [(1.0, t, (p,) * k)]) for t, k in (("dd", 2), ("dz", 3), ("pi2", 4)) for p in range(P)]


@pytest.fixture(scope="module")
def small_batch():
    datas, pos, rows = SY.ragged_regions(seed=31, n_regions=9, m_lo=40, m_hi=400, s_lo=24, s_hi=90, span=50_000)
    dev = L.device()
    d_dev = [torch.as_tensor(d, device=dev) for d in datas]
    rows_dev = [[torch.as_tensor(np.asarray(r, dtype=np.int32), device=dev) for r in rr] for rr in rows]
    return datas, pos, rows, d_dev, rows_dev


def oracle_region(d, pos, rows, P):
    ii, jj = O.pairs_within_distance(pos, np.inf)
    cl = [O.hap_counts(d, ii, jj, r) for r in rows]
    st = O.multi_pop_stats_hap([c for c, _ in cl], [nv for _, nv in cl], self_specs(P))
    keys = O.bin_index(pos[jj] - pos[ii], EDGES)
    return O.keyed_sums(keys, st, 16), O.keyed_sums(keys, np.abs(st), 16), np.bincount(keys[keys >= 0],
                                                                                         minlength=16)[:16]


def test_batch_self_stats_vs_oracle_and_single_band(small_batch):
    datas, pos, rows, d_dev, rows_dev = small_batch
    print("[batch] 9 regions, 4 populations, self statistics", flush=True)
    b = batch.RegionBatch(pos, 1e6, EDGES)
    packed = batch.pack_regions(d_dev, rows_dev)
    b.bind(packed, stat_set="self")
    got, kp = b.sums(with_missing=True, want_pairs=True)
    got, kp = np_(got), np_(kp)
    assert got.shape == (9, 16, 12)
    for r in range(9):
        want, scale, pairs = oracle_region(datas[r], pos[r], rows[r], 4)
        assert np.all(np.abs(got[r] - want) <= 1e-12 * np.maximum(scale, 1e-300) + 1e-300), f"region {r}"
        np.testing.assert_array_equal(kp[r], pairs)
        single = band.BandPlan(pos[r], 1e6, edges=EDGES)
        ref = np_(single.sums(packed[r], stat_set="self", with_missing=True))
        np.testing.assert_array_equal(got[r], ref)


def test_batch_schedules_and_autotune_are_bit_identical(small_batch):
    _, pos, _, d_dev, rows_dev = small_batch
    b = batch.RegionBatch(pos, 1e6, EDGES).bind(batch.pack_regions(d_dev, rows_dev), stat_set="self")
    ref = np_(b.sums(True))
    for grid in (0, 8, 256, 1024):
        for order in ("cost", "natural", "blocked"):
            for xcd in (False, True):
                b.schedule(grid, order, xcd)
                np.testing.assert_array_equal(np_(b.sums(True)), ref)
    choice = b.autotune(True)
    assert choice["ms"] is not None and choice["ms"] > 0 and choice["grid"] >= 0
    np.testing.assert_array_equal(np_(b.sums(True)), ref)
    again = batch.RegionBatch(pos, 1e6, EDGES).bind(batch.pack_regions(d_dev, rows_dev), stat_set="self")
    assert again.autotune(True)["ms"] is None  # answered from the cache



def test_batch_fp4_equals_int8(small_batch):
    """A region batch bound to FP4 planes gives the int8 binding's sums bit
    for bit, under every schedule; layouts cannot be mixed."""
    _, pos, _, d_dev, rows_dev = small_batch
    ref = np_(batch.RegionBatch(pos, 1e6, EDGES).bind(batch.pack_regions(d_dev, rows_dev), stat_set="self")
              .sums(True))
    p4 = batch.pack_regions(d_dev, rows_dev, fp4=True)
    b = batch.RegionBatch(pos, 1e6, EDGES).bind(p4, stat_set="self")
    for grid, order in ((0, "blocked"), (256, "cost"), (1024, "natural")):
        for kv in (0, 1, 2, 3):  # kernel variants: streamed / per-population K loop, rolled rows, lean
            b.schedule(grid, order, False, kernel_variant=kv)
            np.testing.assert_array_equal(np_(b.sums(True)), ref)
    choice = b.autotune(True)
    assert choice["kernel_variant"] in (0, 1, 2, 3)
    np.testing.assert_array_equal(np_(b.sums(True)), ref)
    mixed = batch.pack_regions(d_dev, rows_dev)
    mixed[0][0] = p4[0][0]
    with pytest.raises(ValueError):
        batch.RegionBatch(pos, 1e6, EDGES).bind(mixed, stat_set="self")

def test_batch_moments_two_pops_and_genotypes():
    datas, pos, rows = SY.ragged_regions(seed=32, n_regions=5, m_lo=60, m_hi=300, s_lo=40, s_hi=120, n_pops=2,
                                         span=30_000)
    dev = L.device()
    d_dev = [torch.as_tensor(d, device=dev) for d in datas]
    b = batch.RegionBatch(pos, 1e6, EDGES).bind(batch.pack_regions(d_dev, rows), stat_set="moments")
    got = np_(b.sums(True))
    for r in range(5):
        ii, jj = O.pairs_within_distance(pos[r], np.inf)
        cl = [O.hap_counts(datas[r], ii, jj, rw) for rw in rows[r]]
        st = O.two_pop_stats(cl[0][0], cl[1][0], cl[0][1], cl[1][1])
        keys = O.bin_index(pos[r][jj] - pos[r][ii], EDGES)
        want, scale = O.keyed_sums(keys, st, 16), O.keyed_sums(keys, np.abs(st), 16)
        assert np.all(np.abs(got[r] - want) <= 1e-12 * np.maximum(scale, 1e-300) + 1e-300)
    rng = np.random.default_rng(5)
    gl = [SY.hwe_genotypes(int(rng.integers(30, 70)), len(p), rng, missing=0.04) for p in pos]
    gb = batch.RegionBatch(pos, 1e6, EDGES).bind(
        batch.pack_regions([torch.as_tensor(g, device=dev) for g in gl], [[None]] * 5, family="geno"),
        stat_set="moments")
    gg = np_(gb.sums(True))
    for r in range(5):
        ii, jj = O.pairs_within_distance(pos[r], np.inf)
        c, nv = O.geno_counts(gl[r], ii, jj, None)
        st = O.multi_pop_stats_geno([c], [nv])
        keys = O.bin_index(pos[r][jj] - pos[r][ii], EDGES)
        want, scale = O.keyed_sums(keys, st, 16), O.keyed_sums(keys, np.abs(st), 16)
        assert np.all(np.abs(gg[r] - want) <= 1e-11 * np.maximum(scale, 1e-300) + 1e-300)


def test_graph_replays_match_eager(small_batch):
    """Pack every region + the batch pass captured as one graph; each replay,
    with eager work and synchronizations in between, matches the eager run."""
    _, pos, _, d_dev, rows_dev = small_batch
    packed = batch.pack_regions(d_dev, rows_dev)
    b = batch.RegionBatch(pos, 1e6, EDGES).bind(packed, stat_set="self")
    eager = np_(b.sums(True))
    out = torch.full((9, 16, 12), float("nan"), dtype=torch.float64, device=L.device())
    kp = torch.full((9, 16), -1.0, dtype=torch.float64, device=L.device())

    def step():
        batch.pack_regions(d_dev, rows_dev, out=packed, validate=False)
        b.sums(True, out=out, key_pairs=kp)

    g = batch.Graph().capture(step)
    for rep in range(4):
        out.fill_(float("nan"))
        kp.fill_(-1.0)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device=L.device()).sum().item()  # eager work between replays
        g.launch()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(np_(out), eager, err_msg=f"replay {rep}")
        assert np.all(np_(kp) >= 0)
    g.close()


def test_batch_fp4_multichunk_vs_oracle_and_int8():
    """Regions of 2,100-2,600 haplotypes in 4 populations (525-650 samples
    each: 3 FP4 K chunks): the FP4 binding equals the int8 binding bit for
    bit and the oracle region by region."""
    datas, pos, rows = SY.ragged_regions(seed=33, n_regions=3, m_lo=60, m_hi=200, s_lo=2100, s_hi=2600,
                                         span=40_000)
    print("[batch] 3 regions, 2,100-2,600 haplotypes, FP4 multi-chunk", flush=True)
    dev = L.device()
    d_dev = [torch.as_tensor(d, device=dev) for d in datas]
    p4 = batch.pack_regions(d_dev, rows, fp4=True)
    assert min(p.k_pad for pl in p4 for p in pl) >= 768
    b4 = batch.RegionBatch(pos, 1e6, EDGES).bind(p4, stat_set="self")
    b8 = batch.RegionBatch(pos, 1e6, EDGES).bind(batch.pack_regions(d_dev, rows), stat_set="self")
    got = np_(b4.sums(True))
    np.testing.assert_array_equal(got, np_(b8.sums(True)))
    for r in range(3):
        want, scale, _ = oracle_region(datas[r], pos[r], rows[r], 4)
        assert np.all(np.abs(got[r] - want) <= 1e-12 * np.maximum(scale, 1e-300) + 1e-300), f"region {r}"


def test_autotune_cache_on_disk_serves_a_second_process(tmp_path):
    """The autotuner's choice is written to its cache file; a second process
    with the same shape class reads it and times nothing (ms None), and
    gets the same choice and the same sums."""
    import json
    import subprocess
    import sys
    cache = str(tmp_path / "tune.txt")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys, json, numpy as np, torch\n"
        f"sys.path.insert(0, {root!r})\n"
        "from pg_gpu_amd import batch, synthetic as SY, _lib as L\n"
        "datas, pos, rows = SY.ragged_regions(seed=34, n_regions=6, m_lo=40, m_hi=300, s_lo=24, s_hi=600, "
        "span=50_000)\n"
        "d = [torch.as_tensor(x, device=L.device()) for x in datas]\n"
        "b = batch.RegionBatch(pos, 1e6, np.logspace(0, 6, 17)).bind(batch.pack_regions(d, rows, fp4=True), "
        "stat_set='self')\n"
        "c = b.autotune(True)\n"
        "s = b.sums(True).cpu().numpy()\n"
        "print(json.dumps({'choice': c, 'sum': float(np.abs(s).sum()), 'first': float(s.flat[0])}))\n")
    env = dict(os.environ, PGLD_TUNE_CACHE=cache)
    out = []
    for k in range(2):
        print(f"[batch] autotune cache process {k}", flush=True)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-3000:]
        out.append(json.loads(r.stdout.strip().splitlines()[-1]))
    assert out[0]["choice"]["ms"] is not None and out[0]["choice"]["ms"] > 0
    assert out[1]["choice"]["ms"] is None
    for k in ("grid", "order", "xcd_split", "kernel_variant"):
        assert out[0]["choice"][k] == out[1]["choice"][k]
    assert out[0]["sum"] == out[1]["sum"] and out[0]["first"] == out[1]["first"]
    lines = open(cache).read().split("\n")
    assert any("/P4/" in ln and "/fp41/" in ln for ln in lines)


@pytest.mark.parametrize("family,fp4", [("hap", True), ("hap", False), ("geno", True), ("geno", False)])
def test_pack_batch_equals_pack_population(family, fp4):
    """One batched launch packs every population of every region into the
    same bytes, margins and flags as one pack_population call each (ragged
    shapes, variant counts not a multiple of 4, a row list of None, missing
    data and values outside the domain flagged per population); also when
    replayed from a captured graph."""
    rng = np.random.default_rng(35 + fp4)
    dev = L.device()
    datas, rows = [], []
    for r, (m, s) in enumerate(((130, 70), (257, 300), (61, 9), (400, 513))):
        if family == "hap":
            d = SY.neutral_haplotypes(s, m, rng, missing=0.02 if r != 1 else 0.0)
        else:
            d = SY.hwe_genotypes(s, m, rng, missing=0.03 if r != 1 else 0.0)
        if r == 2:
            d[0, 3] = 3  # outside the domain of both families
        datas.append(torch.as_tensor(d, device=dev))
        cut = sorted(rng.choice(np.arange(1, s), size=2, replace=False))
        rows.append([None] if r == 3 else [list(range(0, cut[0])), list(range(cut[0], cut[1])),
                                            list(range(cut[1], s))])
    ref = [[band.pack_population(d, rw, family, validate=False, fp4=fp4) for rw in rr] for d, rr in zip(datas, rows)]
    out = [[band.pack_population(d, rw, family, validate=False, fp4=fp4) for rw in rr] for d, rr in zip(datas, rows)]
    for pl in out:
        for p in pl:  # stale contents the batch must overwrite
            p.planes.fill_(7)
            p.marg.fill_(-5)
            p.flags.fill_(9)
    pb = batch.PackBatch(datas, rows, out)

    def check():
        for a_l, b_l in zip(ref, out):
            for a, b in zip(a_l, b_l):
                np.testing.assert_array_equal(np_(b.planes), np_(a.planes))
                np.testing.assert_array_equal(np_(b.marg), np_(a.marg))
                np.testing.assert_array_equal(np_(b.flags), np_(a.flags))

    pb.run()
    check()
    for pl in out:
        for p in pl:
            p.marg.fill_(-5)
            p.flags.fill_(9)
    g = batch.Graph().capture(pb.run)
    for _ in range(3):
        g.launch()
        torch.cuda.synchronize()
        check()
        out[0][0].marg.fill_(11)  # eager work between replays
    assert int(np_(ref[2][0].flags)[1]) == 1 and int(np_(ref[1][0].flags)[0]) == 0


def test_batch_four_population_moments_and_refused_sets():
    """The 105 four-population haplotype moments over a region batch equal
    each region's own band pass bit for bit and the oracle within 1e-12 of
    sum |terms|; the three-population genotype moments, which park their
    counts in a plan's own workspace, are refused for region batches."""
    edges = np.array([0.0, 1e3, 1e4, 1e5, 1e6])
    datas, pos, rows = SY.ragged_regions(seed=33, n_regions=4, m_lo=60, m_hi=250, s_lo=60, s_hi=140, n_pops=4,
                                         span=40_000)
    dev = L.device()
    d_dev = [torch.as_tensor(d, device=dev) for d in datas]
    packed = batch.pack_regions(d_dev, rows, fp4=True)
    b = batch.RegionBatch(pos, 1e6, edges).bind(packed, stat_set="moments")
    got = np_(b.sums(True))
    assert got.shape == (4, 4, 105)
    for r in range(4):
        single = band.BandPlan(pos[r], 1e6, edges=edges)
        np.testing.assert_array_equal(got[r], np_(single.sums(packed[r], stat_set="moments", with_missing=True)))
        ii, jj = O.pairs_within_distance(pos[r], np.inf)
        cl = [O.hap_counts(datas[r], ii, jj, rw) for rw in rows[r]]
        st = O.multi_pop_stats_hap([c for c, _ in cl], [nv for _, nv in cl])
        keys = O.bin_index(pos[r][jj] - pos[r][ii], edges)
        want, scale = O.keyed_sums(keys, st, 4), O.keyed_sums(keys, np.abs(st), 4)
        assert np.all(np.abs(got[r] - want) <= 1e-12 * np.maximum(scale, 1e-300) + 1e-300), f"region {r}"
    rng = np.random.default_rng(6)
    gl = [torch.as_tensor(SY.hwe_genotypes(60, len(p), rng, missing=0.04), device=dev) for p in pos]
    g3 = [[list(range(k, 60, 3)) for k in range(3)] for _ in pos]
    with pytest.raises(L.PgldError) as e:
        gb = batch.RegionBatch(pos, 1e6, edges).bind(batch.pack_regions(gl, g3, family="geno", fp4=True),
                                                     stat_set="moments")
        gb.sums(True)
    assert e.value.status == L.EUNSUPPORTED
