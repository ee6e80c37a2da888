"""CPU tests of the host side: the C-ABI library's exports against the header,
the VCF/pop-file readers, the matrix classes' host logic and the helpers of
moments_ld.  Nothing here launches a kernel."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import ld_oracle as O
from pg_gpu_amd import GenotypeMatrix, HaplotypeMatrix, moments_ld, vcf
from pg_gpu_amd import _lib as L
from pg_gpu_amd import ld_pipeline as LP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HEADER = os.path.join(ROOT, "include", "pgld.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(pgld_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_entry_points():
    decl = header_functions()
    assert len(decl) >= 18
    assert set(decl) == set(L.exported_symbols())


def test_library_loads_and_exports_every_header_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libpgld.so not built (python -m pg_gpu_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, missing
    # the loader binds every signature without touching a device
    code = ("import pg_gpu_amd._lib as L; h = L.lib(); "
            "print(h.pgld_version(), len(L.exported_symbols()))")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr
    ver, n = r.stdout.split()
    assert int(ver) >= 10000 and int(n) == len(header_functions())


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-fsyntax-only", "-x", "c", "-std=c99", "-Wall", "-Werror", HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


# ---------------------------------------------------------------------------
def test_read_vcf_im_example():
    d = vcf.read_vcf(os.path.join(GOLD, "im-parsing-example.vcf"))
    gt, pos = d["gt"], d["positions"]
    assert gt.ndim == 3 and gt.shape[2] == 2 and gt.shape[0] == pos.shape[0] > 0
    assert np.all(np.diff(pos) >= 0)
    assert set(np.unique(gt)).issubset({-1, 0, 1, 2})  # the example holds multi-allelic sites
    pm = vcf.read_pop_file(os.path.join(GOLD, "im_pop.txt"))
    assert set(pm.values()) == {"deme0", "deme1"}
    assert set(pm) == set(d["samples"])


def test_read_vcf_parsing_cases(tmp_path):
    p = tmp_path / "t.vcf"
    p.write_text(
        "##fileformat=VCFv4.2\n"
        "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\ta\tb\n"
        "1\t10\t.\tA\tC\t.\tPASS\t.\tGT\t0|1\t1|1\n"
        "1\t20\t.\tA\tC\t.\tPASS\t.\tGT:DP\t./.:3\t0/1:5\n"
        "2\t5\t.\tA\tC,G\t.\tPASS\t.\tGT\t2|0\t.|1\n")
    d = vcf.read_vcf(str(p))
    assert list(d["samples"]) == ["a", "b"]
    np.testing.assert_array_equal(d["positions"], [10, 20, 5])
    np.testing.assert_array_equal(d["gt"][0], [[0, 1], [1, 1]])
    np.testing.assert_array_equal(d["gt"][1], [[-1, -1], [0, 1]])
    np.testing.assert_array_equal(d["gt"][2], [[2, 0], [-1, 1]])
    sub = vcf.read_vcf(str(p), samples=["b"])
    np.testing.assert_array_equal(sub["gt"][:, 0, :], d["gt"][:, 1, :])


def test_haplotype_matrix_host_logic():
    hm = HaplotypeMatrix.from_vcf(os.path.join(GOLD, "im-parsing-example.vcf"))
    n = len(hm.samples)
    assert hm.device == "CPU" and hm.num_haplotypes == 2 * n
    d = vcf.read_vcf(os.path.join(GOLD, "im-parsing-example.vcf"))
    np.testing.assert_array_equal(hm.haplotypes[:n], d["gt"][:, :, 0].T)
    np.testing.assert_array_equal(hm.haplotypes[n:], d["gt"][:, :, 1].T)
    hm.load_pop_file(os.path.join(GOLD, "im_pop.txt"), pops=["deme0"])
    rows = hm.sample_sets["deme0"]
    assert sorted(rows) == sorted([i for i in range(n)] + [i + n for i in range(n)]) or \
        all((r + n) in rows for r in rows if r < n)
    sub = hm.get_subset(np.array([0, 2, 4]))
    np.testing.assert_array_equal(sub.haplotypes, hm.haplotypes[:, [0, 2, 4]])
    pos = hm.positions
    # the reference bounds ``high`` by the variant count, not a position
    # (haplotype_matrix.py:704); that check is kept as it is
    lo, hi = int(pos[0]), hm.num_variants
    r = hm.get_subset_from_range(lo, hi)
    np.testing.assert_array_equal(r.positions, pos[(pos >= lo) & (pos < hi)])
    with pytest.raises(ValueError):
        hm.get_subset(np.array([hm.num_variants]))
    with pytest.raises(ValueError):
        HaplotypeMatrix(np.zeros((0, 3), np.int8), np.arange(3))
    with pytest.raises(ValueError):
        hm.get_subset_from_range(10, 5)
    with pytest.raises(ValueError):
        HaplotypeMatrix(np.zeros((2, 2), np.int8), np.arange(2)).load_pop_file("x")
    assert hm.count_missing() == int((hm.haplotypes < 0).sum())


def test_genotype_matrix_host_logic():
    h = np.array([[0, 1, -1, 1], [1, 1, 0, 0], [0, 0, 1, 1], [0, 1, 1, -1]], dtype=np.int8)
    hm = HaplotypeMatrix(h, np.arange(4) * 10, sample_sets={"x": [0, 1, 2, 3], "y": [2, 3]})
    gm = GenotypeMatrix.from_haplotype_matrix(hm)
    np.testing.assert_array_equal(gm.genotypes, [[1, 2, -1, 1], [0, 1, 2, -1]])
    assert gm.sample_sets == {"x": [0, 1], "y": [1]}
    with pytest.raises(ValueError):
        GenotypeMatrix.from_haplotype_matrix(HaplotypeMatrix(h[:3], np.arange(4)))
    back = gm.to_haplotype_matrix()
    np.testing.assert_array_equal(back.haplotypes, [[1, 1, -1, 1], [0, 1, -1, 0], [0, 1, 1, -1],
                                                    [0, 0, 1, -1]])


def test_genotype_matrix_from_vcf_filters(tmp_path):
    p = tmp_path / "g.vcf"
    p.write_text(
        "##fileformat=VCFv4.2\n"
        "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\ta\tb\n"
        "1\t10\t.\tA\tC\t.\tPASS\t.\tGT\t0|1\t1|1\n"
        "1\t20\t.\tA\tC\t.\tPASS\t.\tGT\t0|0\t0|0\n"
        "1\t30\t.\tA\tC,G\t.\tPASS\t.\tGT\t2|0\t0|1\n"
        "1\t40\t.\tA\tC\t.\tPASS\t.\tGT\t.|1\t0|0\n")
    gm = GenotypeMatrix.from_vcf(str(p))
    np.testing.assert_array_equal(gm.positions, [10, 40])
    np.testing.assert_array_equal(gm.genotypes, [[1, -1], [2, 0]])


# ---------------------------------------------------------------------------
def test_stat_tables_match_oracle():
    for P in (1, 2, 3, 4):
        assert LP.ld_names(P) == O.ld_names(P)
        assert LP.het_names(P) == O.het_names(P)
        got = LP.generate_stat_specs(P)
        want = O.stat_specs(P)
        assert len(got) == len(want)
        for a, b in zip(got, want):
            assert list(a) == list(b)


def test_chunk_size_estimate():
    # (4 H P + 150 P) * 3 bytes per pair, clamped to [1e5, 1e7] pairs
    est = LP.estimate_ld_chunk_size
    assert est(1000, available_memory_bytes=(4 * 1000 * 2 + 300) * 3 * 2_000_000, num_pops=2) == 2_000_000
    assert est(1000, available_memory_bytes=10, num_pops=2) == 100_000
    assert est(10, available_memory_bytes=10**15, num_pops=1) == 10_000_000


def test_genetic_map_helpers(tmp_path):
    rec = tmp_path / "m.txt"
    rec.write_text("pos cM\n0 0\n1000 1.0\n3000 1.5\n")
    pos = np.array([0.0, 500.0, 1000.0, 2000.0, 3000.0])
    g = moments_ld._interpolate_genetic_distances(pos, str(rec))
    np.testing.assert_allclose(g, [0, 0.005, 0.01, 0.0125, 0.015])
    # slowest rate is 0.005 M / 2000 bp -> 2.5e-6 M/bp
    assert moments_ld._max_bp_for_r_dist(pos, g, 1e-3) == pytest.approx(1e-3 / 2.5e-6 * 1.1)
    assert moments_ld._max_bp_for_r_dist(pos, g, 1.0) == 3000.0
    assert moments_ld._max_bp_for_r_dist(pos, np.zeros(5), 1e-3) == 3000.0


def test_moments_argument_errors():
    with pytest.raises(ValueError):
        moments_ld.compute_ld_statistics(pops=["a"] * 5, bp_bins=[0, 1])
    with pytest.raises(ValueError):
        moments_ld.compute_ld_statistics(pops=["a"])
    with pytest.raises(ValueError):
        moments_ld.compute_ld_statistics(pops=["a"], bp_bins=[0, 1])


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2 and "one GPU only" in (r.stdout + r.stderr)


def test_window_layers_are_disjoint_and_cover():
    """Overlapping windows are dealt into layers of disjoint index ranges,
    each window exactly once, every layer in increasing order."""
    from pg_gpu_amd.windowed_analysis import _layers
    rng = np.random.default_rng(3)
    start = np.sort(rng.integers(0, 500, size=60))
    stop = start + rng.integers(1, 80, size=60)
    ws = list(range(60))
    layers = _layers(start, stop, ws)
    assert sorted(w for layer in layers for w in layer) == ws
    for layer in layers:
        for a, b in zip(layer, layer[1:]):
            assert stop[a] <= start[b]
    assert _layers(start, stop, []) == []



@pytest.mark.parametrize("fam", ["hap", "geno"])
def test_grouped_four_population_statistics_cover_the_set_once(fam):
    """The generated three-group form of the 105 four-population moments:
    group k writes out[0 .. n_k - 1] once each, the groups' first statistics
    are GOFF, together they cover ld_names(4) in order, and each group states
    the same expression for a statistic as the one-function form (up to the
    names of its shared terms)."""
    import importlib.util
    path = os.path.join(os.path.dirname(__file__), "..", "pg_gpu_amd", "csrc", "gen_epilogue.py")
    spec = importlib.util.spec_from_file_location("gen_epilogue_under_test", path)
    G = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(G)
    names = G.ld_names(4)
    assert len(names) == 105
    text = G.emit_groups(fam, 4, names, False, "f", 3)
    goff = [int(x) for x in re.search(r"F_GOFF\[4\] = \{([^}]*)\}", text).group(1).split(",")]
    assert goff == [0, 35, 70, 105]
    full, _ = G.emit_set(fam, 4, names, False, "f")

    def expanded(code):
        # statement per output index with shared-term names replaced by their definitions
        defs = dict(re.findall(r"const double (\w+) = (.*);", code))
        out = {}
        for k, e in re.findall(r"out\[(\d+)\] = (.*);", code):
            for _ in range(4):
                e = re.sub(r"\b([ptrk]\d+)\b", lambda m: "(" + defs.get(m.group(1), m.group(1)) + ")", e)
            out[int(k)] = e
        return out

    whole = expanded(full)
    assert sorted(whole) == list(range(105))
    for g in range(3):
        body = text.split("void f_g%d(" % g)[1].split("\n}\n")[0]
        part = expanded(body)
        assert sorted(part) == list(range(goff[g + 1] - goff[g]))
        for k, e in part.items():
            assert e == whole[goff[g] + k], (g, k)
