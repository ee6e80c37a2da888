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
pi2_i_i_i_i per population (3P values).

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


def call_expr(fam, t, q, sym):
    V = "H" if fam == "hap" else "G"
    pre = "hap" if fam == "hap" else "geno"
    if t == "dd":
        a, b = q
        if a == b:
            return "%s_dd_single(%s[%d])" % (pre, V, a)
        return "%s_dd_between(%s[%d], %s[%d])" % (pre, V, a, V, b)
    if t == "dz":
        a, b, c = q
        if a == b == c:
            return "%s_dz_iii(%s[%d])" % (pre, V, a)
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
        return "%s_pi2_iiii(%s)" % (pre, args)
    if case in ("triple_a", "triple_b"):
        if fam == "hap":
            return "hap_pi2_triple(%s)" % args
        return "geno_pi2_%s(%s)" % ("iiij" if case == "triple_b" else "ijii", args)
    if case == "iikk" and fam == "hap" and sym:
        return "hap_pi2_iikk_sym(%s)" % args
    return "%s_pi2_%s(%s)" % (pre, case, args)


def emit_set(fam, P, names, sym, fname):
    T = "HapP" if fam == "hap" else "GenoP"
    V = "H" if fam == "hap" else "G"
    calls, lines = {}, []
    body = []
    for s, (kind, p) in enumerate(names):
        parts = []
        for w, t, q in terms_of(kind, p):
            e = call_expr(fam, t, q, sym)
            if e not in calls:
                calls[e] = "t%d" % len(calls)
                lines.append("  const double %s = %s;" % (calls[e], e))
            parts.append((w, calls[e]))
        if len(parts) == 1 and parts[0][0] == 1.0:
            body.append("  out[%d] = %s;" % (s, parts[0][1]))
        else:
            expr = " + ".join("%r * %s" % (w, v) for w, v in parts)
            body.append("  out[%d] = %s;" % (s, expr))
    head = "__device__ __forceinline__ void %s(const %s* %s, double* out) {" % (fname, T, V)
    return "\n".join([head] + lines + body + ["}"]), len(names)


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
