"""Generate epilogue.inc: straight-line statistic sets for the fused band
kernel.

For every population count P (1..4) and family (haplotype / genotype) this
emits a device function that takes the per-population values of one SNP
pair and writes the ld_names(P) statistics, each the weighted average of its
moments terms (ld_pipeline.py:357-457).  Every distinct (type, populations)
term is evaluated once and shared between the statistics that use it.  With
two haplotype populations the terms follow the CuPy fast-path forms
(ld_pipeline.py:271-322), i.e. pi2(i,i,k,k) symmetrised over the loci.

"self" sets list only the within-population statistics DD_i_i, Dz_i_i_i,
pi2_i_i_i_i per population (3P values).  The 105 four-population moments are
also emitted as three functions of 35 consecutive statistics each
(<name>_g0 .. _g2, first statistics in <NAME>_GOFF), which the fused kernel
sums one group at a time.

Run by the build before hipcc.
"""
import os
import sys


def ld_names(P):
    out = []
    for a in range(P):
        for b in range(a, P):
            out.append(("DD", (a, b)))
    for a in range(P):
        for b in range(P):
            for c in range(b, P):
                out.append(("Dz", (a, b, c)))
    for a in range(P):
        for b in range(a, P):
            for c in range(a, P):
                for d in range(c, P):
                    if c == a == d and b != a:
                        continue
                    if a == c and d < b:
                        continue
                    out.append(("pi2", (a, b, c, d)))
    return out


def terms_of(kind, p):
    if kind == "DD":
        return [(1.0, "dd", p)]
    if kind == "Dz":
        a, b, c = p
        if b == c:
            return [(1.0, "dz", p)]
        return [(0.5, "dz", (a, b, c)), (0.5, "dz", (a, c, b))]
    a, b, c, d = p
    if a == b and c == d:
        if a == c:
            return [(1.0, "pi2", p)]
        return [(0.5, "pi2", (a, b, c, d)), (0.5, "pi2", (c, d, a, b))]
    if a == b:
        return [(0.25, "pi2", t) for t in ((a, b, c, d), (a, b, d, c), (c, d, a, b), (d, c, a, b))]
    if c == d:
        return [(0.25, "pi2", t) for t in ((a, b, c, d), (b, a, c, d), (c, d, a, b), (c, d, b, a))]
    return [(0.125, "pi2", t) for t in ((a, b, c, d), (a, b, d, c), (b, a, c, d), (b, a, d, c),
                                        (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a))]


def pi2_case(q):
    i, j, k, l = q
    cnt = {}
    for x in q:
        cnt[x] = cnt.get(x, 0) + 1
    nu, mc = len(cnt), max(cnt.values())
    if nu == 1:
        return ("same", (i,))
    if mc == 3:
        tp = [x for x, c in cnt.items() if c == 3][0]
        sp = [x for x, c in cnt.items() if c == 1][0]
        return ("triple_b" if (i == tp and j == tp) else "triple_a", (tp, sp))
    if i == j and k == l:
        return ("iikk", (i, k))
    if i == j:
        return ("iikl", (i, k, l))
    if k == l:
        return ("ijkk", (i, j, k))
    if (i == k and j == l) or (i == l and j == k):
        return ("ijij", (i, j))
    if nu == 3:
        if i == k:
            return ("shared", (i, j, l))
        if i == l:
            return ("shared", (i, j, k))
        if j == k:
            return ("shared", (j, i, l))
        return ("shared", (j, i, k))
    return ("alldiff", (i, j, k, l))


def hap_call(t, q, sym):
    """(numerator expression, its power-of-two factor, reciprocal-product
    expression) of one haplotype term.  Numerators with a factor other than 1
    are the unscaled poly_* forms; the factor is folded into the reciprocal
    product by emit_set."""
    R = lambda p, k: "R[%d].i%d" % (p, k)
    H = lambda *ps: ", ".join("H[%d]" % x for x in ps)
    if t == "dd":
        a, b = q
        if a == b:
            return "poly_dd_single(%s)" % H(a), 0.0625, R(a, 4)
        return "poly_dd_between(%s)" % H(a, b), 0.0625, "%s * %s" % (R(a, 2), R(b, 2))
    if t == "dz":
        a, b, c = q
        if a == b == c:
            return "poly_dz_iii(%s)" % H(a), 0.25, R(a, 4)
        if a == b:
            return "poly_dz_iij(%s)" % H(a, c), 0.25, "%s * %s" % (R(c, 1), R(a, 3))
        if a == c:
            return "poly_dz_iji(%s)" % H(a, b), 0.25, "%s * %s" % (R(b, 1), R(a, 3))
        if b == c:
            return "poly_dz_ijj(%s)" % H(a, b), 0.25, "%s * %s" % (R(a, 2), R(b, 2))
        return "num_dz_diff(%s)" % H(a, b, c), 1.0, "%s * %s * %s" % (R(a, 2), R(b, 1), R(c, 1))
    case, r = pi2_case(q)
    if case == "same":
        return "poly_pi2_iiii(%s)" % H(*r), 0.0625, R(r[0], 4)
    if case in ("triple_a", "triple_b"):
        tp, sp = r
        return "poly_pi2_triple(%s)" % H(tp, sp), 0.0625, "%s * %s" % (R(sp, 1), R(tp, 3))
    if case == "iikk":
        a, k = r
        if sym:  # symmetric in the two populations
            a, k = min(a, k), max(a, k)
            return "poly_pi2_iikk_sym(%s)" % H(a, k), 0.03125, "%s * %s" % (R(a, 2), R(k, 2))
        return "num_pi2_iikk(%s)" % H(a, k), 1.0, "%s * %s" % (R(a, 2), R(k, 2))
    if case == "iikl":
        a, k, l = r
        return "num_pi2_iikl(%s)" % H(a, k, l), 0.5, "%s * %s * %s" % (R(a, 2), R(k, 1), R(l, 1))
    if case == "ijkk":
        a, b, k = r
        return "num_pi2_ijkk(%s)" % H(a, b, k), 0.5, "%s * %s * %s" % (R(k, 2), R(a, 1), R(b, 1))
    if case == "ijij":
        a, b = min(r), max(r)  # the numerator is exactly symmetric in (i, j)
        return "poly_pi2_ijij(%s)" % H(a, b), 0.0625, "%s * %s" % (R(a, 2), R(b, 2))
    if case == "shared":
        s_, x, y = r
        return "num_pi2_shared(%s)" % H(s_, x, y), 1.0, "%s * %s * %s" % (R(s_, 2), R(x, 1), R(y, 1))
    a, b, c, d = r
    return "num_pi2_alldiff(%s)" % H(a, b, c, d), 0.25, "%s * %s * %s * %s" % (R(a, 1), R(b, 1), R(c, 1),
                                                                            R(d, 1))


def call_expr(fam, t, q, sym):
    """Genotype term expression.  The terms whose division the reference
    guards (0 for an invalid denominator: DD, Dz(i,i,i), pi2(i,i,i,i)) use the
    falling-factorial reciprocals GR[p] (hap_recips: one reciprocal per
    population, 0 where the factorial is 0); the unguarded ones keep their
    division, so a zero denominator still gives inf / NaN."""
    V = "H" if fam == "hap" else "G"
    pre = "hap" if fam == "hap" else "geno"
    if t == "dd":
        a, b = q
        if a == b:
            return "geno_dd_num(G[%d]) * GR[%d].i4" % (a, a)
        return "4.0 * geno_D(G[%d]) * geno_D(G[%d]) * (GR[%d].i2 * GR[%d].i2)" % (a, b, a, b)
    if t == "dz":
        a, b, c = q
        if a == b == c:
            return "2.0 * geno_dz_num(G[%d]) * GR[%d].i4" % (a, a)
        if a == b:
            return "%s_dz_iij(%s[%d], %s[%d])" % (pre, V, a, V, c)
        if a == c:
            return "%s_dz_iji(%s[%d], %s[%d])" % (pre, V, a, V, b)
        if b == c:
            return "%s_dz_ijj(%s[%d], %s[%d])" % (pre, V, a, V, b)
        return "%s_dz_diff(%s[%d], %s[%d], %s[%d])" % (pre, V, a, V, b, V, c)
    case, r = pi2_case(q)
    args = ", ".join("%s[%d]" % (V, x) for x in r)
    if case == "same":
        return "geno_pi2_num(G[%d]) * GR[%d].i4" % (r[0], r[0])
    if case in ("triple_a", "triple_b"):
        return "geno_pi2_%s(%s)" % ("iiij" if case == "triple_b" else "ijii", args)
    return "%s_pi2_%s(%s)" % (pre, case, args)


def emit_hap_set(P, names, sym, fname, first=0, count=None):
    """Haplotype statistics: each distinct numerator polynomial is evaluated
    once (p<k>), each distinct reciprocal product once (r<m>) and each distinct
    (reciprocal product x constant) once (k<m>).  A statistic whose terms share
    one reciprocal product and weight adds its (exact, integer-valued)
    numerators before the single scaling.  With first / count only statistics
    first .. first + count - 1 are written (to out[0 .. count - 1]), with only
    the terms they use; the scale function always covers the whole set."""
    polys, dens, coefs, lines, body, scales = {}, {}, {}, [], [], []
    if count is None:
        count = len(names) - first

    def poly(e):
        if e not in polys:
            polys[e] = "p%d" % len(polys)
            lines.append("  const double %s = %s;" % (polys[e], e))
        return polys[e]

    def den_name(den):
        key = " * ".join(sorted(den.split(" * ")))
        if key not in dens:
            if " * " not in key:
                dens[key] = key
            else:
                dens[key] = "r%d" % len(dens)
                lines.append("  const double %s = %s;" % (dens[key], key))
        return dens[key]

    def coef(den, c):
        r = den_name(den)
        if c == 1.0:
            return r
        key = (r, c)
        if key not in coefs:
            coefs[key] = "k%d" % len(coefs)
            lines.append("  const double %s = %r * %s;" % (coefs[key], c, r))
        return coefs[key]

    for s, (kind, p) in enumerate(names):
        weights = {}  # (numerator, factor, reciprocal product) -> summed weight
        for w, t, q in terms_of(kind, p):
            key = hap_call(t, q, sym)
            weights[key] = weights.get(key, 0.0) + w
        groups = {}  # (reciprocal product, weight x factor) -> numerator expressions
        for (num, fac, den), w in weights.items():
            groups.setdefault((" * ".join(sorted(den.split(" * "))), w * fac), []).append(num)
        # the smallest power-of-two factor of the statistic is left out of the
        # per-pair value and applied to the tile's partial sums (scales[s]);
        # scaling by a power of two commutes with FP64 rounding and addition,
        # so the sums are the same to the bit
        m = min(c for (_, c) in groups)
        scales.append(m)
        if not first <= s < first + count:
            continue
        parts = []
        for (den, c), nums in groups.items():
            names_ = [poly(x) for x in nums]
            total = names_[0] if len(names_) == 1 else "(%s)" % " + ".join(names_)
            parts.append("%s * %s" % (total, coef(den, c / m)))
        body.append("  out[%d] = %s;" % (s - first, " + ".join(parts)))
    head = "__device__ __forceinline__ void %s(const HapP* H, double* out) {" % fname
    whole = (first, count) == (0, len(names))
    used = [p for p in range(P) if whole or any("R[%d]" % p in ln for ln in lines + body)]
    pre = ["  HapR R[%d];" % P] + ["  R[%d] = hap_recips(H[%d].n);" % (p, p) for p in used]
    tail = ["}", scale_fn(fname, scales)] if whole else ["}"]
    return "\n".join([head] + pre + lines + body + tail), len(names)


def emit_groups(fam, P, names, sym, fname, n_groups):
    """The set split into n_groups functions fname_g<k> of consecutive
    statistics (the fused kernel sums one group at a time, so its running sums
    stay in registers), with their first statistics in fname_GOFF."""
    size = -(-len(names) // n_groups)
    offs = [min(len(names), k * size) for k in range(n_groups + 1)]
    parts = ["constexpr int %s_NG = %d;" % (fname.upper(), n_groups),
             "constexpr int %s_GOFF[%d] = {%s};" % (fname.upper(), n_groups + 1, ", ".join(map(str, offs)))]
    for k in range(n_groups):
        code, _ = emit_set(fam, P, names, sym, "%s_g%d" % (fname, k), offs[k], offs[k + 1] - offs[k])
        parts.append(code)
    return "\n\n".join(parts)


def scale_fn(fname, scales):
    """Device function: the factor applied to partial-sum column s of the set."""
    out = ["__device__ __forceinline__ double %s_scale(int s) {" % fname]
    if all(c == 1.0 for c in scales):
        out.append("  return 1.0;")
    else:
        out.append("  switch (s) {")
        for k, c in enumerate(scales):
            if c != 1.0:
                out.append("    case %d: return %r;" % (k, c))
        out.append("    default: return 1.0;")
        out.append("  }")
    out.append("}")
    return "\n".join(out)


def emit_set(fam, P, names, sym, fname, first=0, count=None):
    if fam == "hap":
        return emit_hap_set(P, names, sym, fname, first, count)
    if count is None:
        count = len(names) - first
    T = "GenoP"
    V = "G"
    calls, lines = {}, []
    body = []
    for s, (kind, p) in enumerate(names):
        if not first <= s < first + count:
            continue
        parts = []
        for w, t, q in terms_of(kind, p):
            e = call_expr(fam, t, q, sym)
            if e not in calls:
                calls[e] = "t%d" % len(calls)
                lines.append("  const double %s = %s;" % (calls[e], e))
            parts.append((w, calls[e]))
        merged = {}
        for w, v in parts:
            merged[v] = merged.get(v, 0.0) + w
        items = sorted(merged.items(), key=lambda kv: int(kv[0][1:]))
        if len(items) == 1 and items[0][1] == 1.0:
            body.append("  out[%d] = %s;" % (s - first, items[0][0]))
        else:
            expr = " + ".join("%r * %s" % (w, v) for v, w in items)
            body.append("  out[%d] = %s;" % (s - first, expr))
    head = "__device__ __forceinline__ void %s(const %s* %s, double* out) {" % (fname, T, V)
    used = [p for p in range(P) if any("GR[%d]" % p in ln for ln in lines)]
    pre = ["  HapR GR[%d];" % P] + ["  GR[%d] = hap_recips(G[%d].n);" % (p, p) for p in used] if used else []
    whole = (first, count) == (0, len(names))
    tail = ["}", scale_fn(fname, [1.0] * len(names))] if whole else ["}"]
    return "\n".join([head] + pre + lines + body + tail), len(names)


def self_names(P):
    return [("DD", (a, a)) for a in range(P)] + [("Dz", (a, a, a)) for a in range(P)] + \
        [("pi2", (a, a, a, a)) for a in range(P)]


def main(out_path):
    parts = ["// Generated by gen_epilogue.py; do not edit.", "#pragma once", ""]
    for fam in ("hap", "geno"):
        for P in (1, 2, 3, 4):
            code, n = emit_set(fam, P, ld_names(P), fam == "hap" and P == 2, "epi_%s_moments_%d" % (fam, P))
            parts.append("constexpr int EPI_%s_MOMENTS_%d_N = %d;" % (fam.upper(), P, n))
            parts.append(code)
            if P == 4:
                parts.append(emit_groups(fam, P, ld_names(P), False, "epi_%s_moments_4" % fam, 3))
            code, n = emit_set(fam, P, self_names(P), False, "epi_%s_self_%d" % (fam, P))
            parts.append("constexpr int EPI_%s_SELF_%d_N = %d;" % (fam.upper(), P, n))
            parts.append(code)
    text = "\n\n".join(parts) + "\n"
    old = open(out_path).read() if os.path.exists(out_path) else None
    if old != text:
        with open(out_path, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "epilogue.inc"))
