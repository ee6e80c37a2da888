"""The allocation-free fused sums entry point (pgld_band_sums_fp4_noalloc)
against pgld_band_sums_fp4, through the C ABI: two calls in a row, the
second on a smaller band, bit-identical sums and pair counts; a plan beyond
the fixed workspace is rejected with PGLD_ECAPACITY."""
import ctypes as C

import numpy as np
import pytest

from pg_gpu_amd import _lib as L
from pg_gpu_amd import band, synthetic

pytestmark = pytest.mark.gpu


def _call_sums(name, plan, pops, stat_set, with_missing):
    P = len(pops)
    n_stats = 3 * P if (stat_set == L.SET_SELF and P > 1) else {1: 3, 2: 15, 3: 45, 4: 105}[P]
    sums = L.empty((plan.n_keys, n_stats), np.float64)
    kp = L.empty((plan.n_keys,), np.float64)
    planes, _k1 = L.ptr_array([p.planes for p in pops])
    margs, _k2 = L.ptr_array([p.marg for p in pops])
    kpad = (C.c_int64 * P)(*[p.k_pad for p in pops])
    ns = (C.c_int32 * P)(*[p.n_samples for p in pops])
    rc = getattr(L.lib(), name)(plan._h, pops[0].family, P, planes, margs, kpad, ns, int(with_missing), stat_set,
                                L.ptr(sums), L.ptr(kp), L.stream_handle())
    return rc, sums, kp


def _case(seed, n_hap, n_var, family, P, stat_set, window):
    rng = np.random.default_rng(seed)
    if family == "hap":
        data = synthetic.neutral_haplotypes(n_hap, n_var, rng, missing=0.03)
    else:
        data = synthetic.hwe_genotypes(n_hap, n_var, rng, missing=0.05)
    pos = synthetic.uniform_positions(n_var, 60 * n_var, rng)
    cut = np.linspace(0, n_hap, P + 1).astype(int)
    pops = [band.pack_population(data, list(range(cut[p], cut[p + 1])), family, fp4=True) for p in range(P)]
    if window:
        plan = band.BandPlan(pos, 30.0 * 64, window=25.0 * 64)
    else:
        plan = band.BandPlan(pos, 60.0 * 200, edges=np.logspace(0, np.log10(60.0 * 200), 9))
    return plan, pops


@pytest.mark.parametrize("family,P,stat_set,window", [("hap", 2, L.SET_MOMENTS, True), ("hap", 4, L.SET_SELF, False),
                                                      ("geno", 1, L.SET_MOMENTS, False)])
def test_noalloc_equals_fp4_sums_two_calls(family, P, stat_set, window):
    import torch
    for seed, n_hap, n_var in ((11, 700, 1500), (12, 300, 400)):   # second call: a smaller band
        plan, pops = _case(seed, n_hap, n_var, family, P, stat_set, window)
        print(f"noalloc case family={family} P={P} set={stat_set} seed={seed} n_var={n_var}", flush=True)
        rc0, s0, k0 = _call_sums("pgld_band_sums_fp4", plan, pops, stat_set, True)
        rc1, s1, k1 = _call_sums("pgld_band_sums_fp4_noalloc", plan, pops, stat_set, True)
        torch.cuda.synchronize()
        assert rc0 == 0 and rc1 == 0, L.lib().pgld_last_error()
        a, b = s0.cpu().numpy(), s1.cpu().numpy()
        assert np.array_equal(a, b), np.max(np.abs(a - b))
        assert np.array_equal(k0.cpu().numpy(), k1.cpu().numpy())
        assert k0.cpu().numpy().sum() == plan.n_pairs


def test_noalloc_rejects_plans_beyond_the_fixed_workspace():
    # 15 statistics + count = 16 doubles per slot; > 2^22 / 16 = 262,144 slots
    rng = np.random.default_rng(5)
    n_var = 64 * 1200
    pos = np.arange(n_var, dtype=np.float64)
    hap = synthetic.neutral_haplotypes(40, n_var, rng, missing=0.0)
    pops = [band.pack_population(hap, list(range(0, 20)), "hap", fp4=True),
            band.pack_population(hap, list(range(20, 40)), "hap", fp4=True)]
    plan = band.BandPlan(pos, 64.0 * 14, window=1.0)  # one key per SNP: many (tile, key) slots
    assert (plan.n_slots + 1) * 16 > (1 << 22)
    rc, _, _ = _call_sums("pgld_band_sums_fp4_noalloc", plan, pops, L.SET_MOMENTS, False)
    assert rc == -3
    assert b"fixed" in L.lib().pgld_last_error()


def test_noalloc_refuses_the_genotype_sets_that_park_counts_globally():
    # the 3-population genotype moments need the plan's own parked-count
    # workspace: the allocation-free entry point returns PGLD_EUNSUPPORTED,
    # the allocating one computes them
    import torch
    plan, pops = _case(13, 90, 300, "geno", 3, L.SET_MOMENTS, True)
    print("noalloc refusal: geno P=3 moments", flush=True)
    rc, _, _ = _call_sums("pgld_band_sums_fp4_noalloc", plan, pops, L.SET_MOMENTS, True)
    assert rc == L.EUNSUPPORTED
    assert b"workspace" in L.lib().pgld_last_error()
    rc0, s0, k0 = _call_sums("pgld_band_sums_fp4", plan, pops, L.SET_MOMENTS, True)
    torch.cuda.synchronize()
    assert rc0 == 0 and k0.cpu().numpy().sum() == plan.n_pairs
