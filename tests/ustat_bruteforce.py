"""Brute-force U-statistics for LD statistics on tiny samples.

This is an independent pin for the oracle's closed-form DD / Dz / pi2
formulas: each statistic is the average of a product kernel over ordered
tuples of samples, where samples drawn from the same population must be
distinct.  Nothing here follows the reference's polynomials; the kernels are
the textbook estimator definitions (Ragsdale & Gravel 2019, moments.LD):

  haplotypes (A, B in {0,1} at the two loci):
    f(a,b)   = A_a B_a (1-A_b)(1-B_b) - A_a (1-B_a)(1-A_b) B_b
    DD(i,j)  = U[ f(a,b) f(c,d) ],            a,b in i; c,d in j
    Dz(i,j,k)= U[ f(a,b) (1-2A_c)(1-2B_d) ],  a,b in i; c in j; d in k
    pi2      = U[ hA(a,b) hB(c,d) ],  hX(a,b) = (X_a(1-X_b) + X_b(1-X_a))/2

  genotypes (dosages x, y in {0,1,2}):
    k(a,b)   = (x_a - x_b)(y_a - y_b) / 4
    DD       = U[ k(a,b) k(c,d) ]
    Dz       = U[ k(a,b) (1-x_c)(1-y_d) ]
    pi2      = U[ hA(a,b) hB(c,d) ],  hX(a,b) = (X_a/2 (1-X_b/2) + X_b/2 (1-X_a/2))/2
"""
import itertools

import numpy as np


def _tuples(pops, pop_of_slot):
    """Ordered tuples (one sample index per slot) with distinct samples
    among slots drawn from the same population."""
    ranges = [range(len(pops[p])) for p in pop_of_slot]
    for t in itertools.product(*ranges):
        ok = True
        for s1 in range(len(t)):
            for s2 in range(s1 + 1, len(t)):
                if pop_of_slot[s1] == pop_of_slot[s2] and t[s1] == t[s2]:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            yield t


def _ustat(pops, pop_of_slot, kern):
    tot, cnt = 0.0, 0
    for t in _tuples(pops, pop_of_slot):
        samples = [pops[p][idx] for p, idx in zip(pop_of_slot, t)]
        tot += kern(*samples)
        cnt += 1
    return tot / cnt if cnt else 0.0


def hap_stat(pops, kind, idx):
    """pops: list of (n_p, 2) arrays of (A, B) haplotype alleles."""
    f = lambda a, b: a[0] * a[1] * (1 - b[0]) * (1 - b[1]) - a[0] * (1 - a[1]) * (1 - b[0]) * b[1]
    hA = lambda a, b: 0.5 * (a[0] * (1 - b[0]) + b[0] * (1 - a[0]))
    hB = lambda c, d: 0.5 * (c[1] * (1 - d[1]) + d[1] * (1 - c[1]))
    if kind == "dd":
        i, j = idx
        return _ustat(pops, (i, i, j, j), lambda a, b, c, d: f(a, b) * f(c, d))
    if kind == "dz":
        i, j, k = idx
        return _ustat(pops, (i, i, j, k),
                      lambda a, b, c, d: f(a, b) * (1 - 2 * c[0]) * (1 - 2 * d[1]))
    i, j, k, l = idx
    return _ustat(pops, (i, j, k, l), lambda a, b, c, d: hA(a, b) * hB(c, d))


def geno_stat(pops, kind, idx):
    """pops: list of (n_p, 2) arrays of (x, y) dosages."""
    kk = lambda a, b: (a[0] - b[0]) * (a[1] - b[1]) / 4.0
    hA = lambda a, b: 0.5 * (a[0] / 2 * (1 - b[0] / 2) + b[0] / 2 * (1 - a[0] / 2))
    hB = lambda c, d: 0.5 * (c[1] / 2 * (1 - d[1] / 2) + d[1] / 2 * (1 - c[1] / 2))
    if kind == "dd":
        i, j = idx
        return _ustat(pops, (i, i, j, j), lambda a, b, c, d: kk(a, b) * kk(c, d))
    if kind == "dz":
        i, j, k = idx
        return _ustat(pops, (i, i, j, k),
                      lambda a, b, c, d: kk(a, b) * (1 - c[0]) * (1 - d[1]))
    i, j, k, l = idx
    return _ustat(pops, (i, j, k, l), lambda a, b, c, d: hA(a, b) * hB(c, d))


def hap_counts_of(pop):
    """[n11, n10, n01, n00] of an (n, 2) haplotype-pair array."""
    A, B = pop[:, 0], pop[:, 1]
    return np.array([np.sum((A == 1) & (B == 1)), np.sum((A == 1) & (B == 0)),
                     np.sum((A == 0) & (B == 1)), np.sum((A == 0) & (B == 0))])


def geno_counts_of(pop):
    """9-way counts n00..n22 (index 3*x + y) of an (n, 2) dosage array."""
    out = np.zeros(9, dtype=np.int64)
    for x, y in pop:
        out[3 * int(x) + int(y)] += 1
    return out
