"""Seeded synthetic inputs for the LD workloads (host NumPy, int8 matrices).

Each generator takes size arguments so the same recipe yields small parity
cases and the full benchmark shapes.  Nothing here is timed.

* neutral_haplotypes -- 0/1 haplotypes with derived-allele counts drawn from
  the neutral SFS, P(i) proportional to 1/i, optional MCAR missing (-1).
* heterogeneous_positions -- sorted distinct positions from a Poisson process
  whose rate varies per segment with Gamma(shape) heterogeneity.
* hwe_genotypes -- diploid 0/1/2 dosages from Beta(a, a) frequencies under
  Hardy-Weinberg, optional MCAR missing.
"""
from __future__ import annotations

import numpy as np


def neutral_haplotypes(n_hap, n_var, rng, missing=0.0, chunk=4096):
    """(n_hap, n_var) int8.  For each SNP a derived count k in 1..n_hap-1 is
    drawn with P(k) ~ 1/k; carriers are k haplotypes chosen uniformly without
    replacement.  A fraction ``missing`` of entries becomes -1 (MCAR)."""
    k_vals = np.arange(1, n_hap)
    w = 1.0 / k_vals
    w /= w.sum()
    ks = rng.choice(k_vals, size=n_var, p=w)
    out = np.empty((n_hap, n_var), dtype=np.int8)
    for v0 in range(0, n_var, chunk):
        v1 = min(n_var, v0 + chunk)
        u = rng.random((n_hap, v1 - v0), dtype=np.float32)
        # the k haplotypes with the smallest uniforms carry the derived allele
        order = np.argsort(u, axis=0)
        carry = (np.arange(n_hap)[:, None] < ks[v0:v1][None, :]).astype(np.int8)
        blk = np.empty((n_hap, v1 - v0), dtype=np.int8)
        np.put_along_axis(blk, order, carry, axis=0)
        out[:, v0:v1] = blk
    if missing > 0:
        for v0 in range(0, n_var, chunk):
            v1 = min(n_var, v0 + chunk)
            m = rng.random((n_hap, v1 - v0), dtype=np.float32) < missing
            blk = out[:, v0:v1]
            blk[m] = -1
    return out


def uniform_positions(n_var, span, rng):
    """n_var distinct sorted integer positions uniform in [0, span)."""
    return np.sort(rng.choice(int(span), size=n_var, replace=False)).astype(np.int64)


def heterogeneous_positions(n_var, span, rng, segment=10_000, shape=0.5):
    """Sorted distinct integer positions in [0, span): per-segment rate
    lambda ~ Gamma(shape, mean 1), multinomial SNP counts per segment,
    uniform positions within a segment."""
    n_seg = int(np.ceil(span / segment))
    lam = rng.gamma(shape, 1.0 / shape, size=n_seg)
    counts = rng.multinomial(n_var, lam / lam.sum())
    # a segment holds at most `segment` distinct positions
    over = counts - segment
    while np.any(over > 0):
        extra = int(over[over > 0].sum())
        counts = np.minimum(counts, segment)
        free = np.nonzero(counts < segment)[0]
        add = rng.multinomial(extra, lam[free] / lam[free].sum())
        counts[free] += add
        over = counts - segment
    pos = []
    for s, c in enumerate(counts):
        if c:
            lo = s * segment
            hi = min(span, lo + segment)
            pos.append(lo + rng.choice(hi - lo, size=c, replace=False))
    return np.sort(np.concatenate(pos)).astype(np.int64)


def hwe_genotypes(n_ind, n_var, rng, beta=0.5, missing=0.0):
    """(n_ind, n_var) int8 dosages, p ~ Beta(beta, beta), g ~ Binomial(2, p)."""
    p = rng.beta(beta, beta, size=n_var)
    g = rng.binomial(2, p[None, :], size=(n_ind, n_var)).astype(np.int8)
    if missing > 0:
        g[rng.random((n_ind, n_var)) < missing] = -1
    return g


def plain_haplotype_panel(seed=1001, n_hap=200, n_var=2000, span=200_000):
    """Plain haplotype case: 2,000 SNPs x 200 haplotypes, neutral SFS, no
    missing data, positions uniform over 200 kb; pairs within 20 kb."""
    rng = np.random.default_rng(seed)
    hap = neutral_haplotypes(n_hap, n_var, rng)
    pos = uniform_positions(n_var, span, rng)
    return hap, pos


def windowed_haplotype_panel(seed=1002, n_hap=2000, n_var=100_000, span=10_000_000, missing=0.02, segment=10_000):
    """Batched windowed haplotype LD: 100,000 SNPs x 2,000 haplotypes (two
    populations of 1,000), 2% MCAR missing, heterogeneous positions over
    10 Mb; pairs within 100 kb, 100 kb windows keyed by the left SNP."""
    rng = np.random.default_rng(seed)
    hap = neutral_haplotypes(n_hap, n_var, rng, missing=missing)
    pos = heterogeneous_positions(n_var, span, rng, segment=segment)
    half = n_hap // 2
    pops = [list(range(0, half)), list(range(half, n_hap))]
    return hap, pos, pops


def genotype_panel(seed=1003, n_ind=1000, n_var=50_000, span=1_000_000, missing=0.05):
    """Genotype LD: 50,000 SNPs x 1,000 diploids, 5% missing, HWE dosages
    from Beta(0.5, 0.5) frequencies; all pairs in a 1 Mb span, 16 log bins."""
    rng = np.random.default_rng(seed)
    g = hwe_genotypes(n_ind, n_var, rng, missing=missing)
    pos = uniform_positions(n_var, span, rng)
    return g, pos


def dense_r2_panel(seed=1004, n_hap=10_000, n_var=40_000, span=4_000_000):
    """Dense all-pairs r^2: 40,000 SNPs x 10,000 haplotypes, neutral SFS as
    the plain panel, no missing data, positions uniform over 4 Mb."""
    rng = np.random.default_rng(seed)
    hap = neutral_haplotypes(n_hap, n_var, rng)
    pos = uniform_positions(n_var, span, rng)
    return hap, pos


def ragged_regions(seed=1005, n_regions=512, m_lo=500, m_hi=20_000, s_lo=50, s_hi=5000, n_pops=4, missing=0.03,
                   span=1_000_000, device=None):
    """Independent regions of uneven size: SNP counts uniform in [m_lo, m_hi],
    haplotype counts uniform in [s_lo, s_hi] split into n_pops contiguous
    groups (population p = rows [p S / n_pops, (p + 1) S / n_pops)), MCAR
    missing data, positions uniform over `span`.

    On the host (device=None) each SNP carries exactly k derived alleles with
    P(k) ~ 1/k; on a device the matrices are drawn there (torch) with
    per-SNP frequency k / S, the same spectrum in expectation, so that ~10 GB
    of input need not cross PCIe.  Returns (matrices, positions, rows)."""
    rng = np.random.default_rng(seed)
    m = rng.integers(m_lo, m_hi + 1, size=n_regions)
    S = rng.integers(s_lo, s_hi + 1, size=n_regions)
    positions = [uniform_positions(int(mr), span, rng) for mr in m]
    rows = [[list(range(p * int(s) // n_pops, (p + 1) * int(s) // n_pops)) for p in range(n_pops)] for s in S]
    datas = []
    if device is None:
        for mr, s in zip(m, S):
            datas.append(neutral_haplotypes(int(s), int(mr), rng, missing=missing))
    else:
        import torch
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        for mr, s in zip(m, S):
            s, mr = int(s), int(mr)
            kv = np.arange(1, s)
            w = 1.0 / kv
            k = rng.choice(kv, size=mr, p=w / w.sum())
            p = torch.as_tensor(k / s, dtype=torch.float32, device=device)
            h = (torch.rand((s, mr), generator=gen, device=device) < p[None, :]).to(torch.int8)
            if missing > 0:
                h[torch.rand((s, mr), generator=gen, device=device) < missing] = -1
            datas.append(h)
    return datas, positions, rows
