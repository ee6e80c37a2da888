// Per-pair LD statistics from exact integer counts.
//
// Counts arrive as exact integers (int32 from the counting kernels) and are
// promoted to double; every numerator below is a polynomial with integer or
// dyadic (1/2, 1/4, 1/16) coefficients, so for populations of up to a few
// thousand samples it is evaluated exactly in FP64 and each statistic carries
// a single rounding: the final division.
//
// Haplotype formulas follow ld_statistics.py:1116-1617 and
// haplotype_kernels.py:69-304; genotype formulas follow
// genotype_kernels.py:27-492.  Invalid denominators give 0 (DD/Dz/pi2) or
// NaN (r, r^2, D', sigma_D^2) as the reference does.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>


namespace pgld {

#define PGLD_HD __host__ __device__ __forceinline__

// ---- statistic / case codes shared with the host -------------------------
enum StatType : int32_t {
  ST_DD = 0,
  ST_DZ = 1,
  ST_PI2 = 2,       // fused-kernel form (haplotype_kernels.py:167-304)
  ST_PI2_SYM = 3,   // CuPy form: pi2(i,i,k,k) symmetrised (ld_statistics.py:1337-1349)
  ST_SIGMA_D2 = 4,  // genotype sigma_D^2 (genotype_kernels.py:91-169)
};

PGLD_HD double safe_div(double num, double den) { return den > 0.0 ? num / den : 0.0; }

// ---------------------------------------------------------------------------
// Haplotype per-population values (ld_pipeline.py:465-482)
// ---------------------------------------------------------------------------
struct HapP {
  double c1, c2, c3, c4, n, D, pA, qA, pB, qB;
  // moments of the centred alleles a' = 2a - 1, b' = 2b - 1 (in {-1, 1})
  // over the pair-valid haplotypes: Sa = sum a', Sb = sum b', Sab = sum a' b';
  // Dp = n Sab - Sa Sb = -4 D
  double Sa, Sb, Sab, Dp;
};

PGLD_HD HapP hap_pop(double c1, double c2, double c3, double c4) {
  HapP p;
  p.c1 = c1; p.c2 = c2; p.c3 = c3; p.c4 = c4;
  p.n = c1 + c2 + c3 + c4;
  p.D = c2 * c3 - c1 * c4;
  p.pA = c1 + c2; p.qA = c3 + c4;
  p.pB = c1 + c3; p.qB = c2 + c4;
  p.Sa = p.pA - p.qA;
  p.Sb = p.pB - p.qB;
  p.Sab = (c1 + c4) - (c2 + c3);
  p.Dp = -4.0 * p.D;
  return p;
}

// The same values from integer counts: the sums and differences are formed in
// int32 (exact) and converted once; D from the moments (-(n Sab - Sa Sb) / 4,
// exact for counts below 2^26).
PGLD_HD HapP hap_pop_int(int c1, int c2, int c3, int c4) {
  HapP p;
  p.c1 = c1; p.c2 = c2; p.c3 = c3; p.c4 = c4;
  const int s12 = c1 + c2, s34 = c3 + c4, s13 = c1 + c3, s24 = c2 + c4;
  p.n = s12 + s34;
  p.pA = s12; p.qA = s34;
  p.pB = s13; p.qB = s24;
  p.Sa = s12 - s34;
  p.Sb = s13 - s24;
  p.Sab = (c1 + c4) - (c2 + c3);
  p.Dp = p.n * p.Sab - p.Sa * p.Sb;
  p.D = -0.25 * p.Dp;
  return p;
}

PGLD_HD double f4(double n) { return n * (n - 1.0) * (n - 2.0) * (n - 3.0); }
PGLD_HD double f3(double n) { return n * (n - 1.0) * (n - 2.0); }
PGLD_HD double f2(double n) { return n * (n - 1.0); }

// Numerators.  Each statistic below is num_x(...) / den_x(n...) with a
// denominator that is a product of per-population falling factorials.
//
// The reference writes them as polynomials in the four haplotype counts
// (haplotype_kernels.py:69-289, ld_statistics.py:1116-1434).  Those used by
// one and two populations are evaluated here in the centred moments (n, Sa,
// Sb, Sab) with D = c2 c3 - c1 c4 = -(n Sab - Sa Sb) / 4, where they collapse
// to a few products (the identities are exact and are checked against the
// reference's forms in the test suite).  Every intermediate value is an
// integer or a dyadic fraction of one, so for counts whose products stay
// below 2^53 they equal the reference's values exactly.
// The polynomials below are the numerators without their power-of-two factors
// (num = scale * poly, the scaling exact); the fused epilogue folds the
// factors into the per-pair reciprocal products.
PGLD_HD double poly_dd_single(const HapP& a) {  // x 1/16
  const double n = a.n, x2 = a.Sa * a.Sa, y2 = a.Sb * a.Sb, t = a.Sa * a.Sb * a.Sab;
  return (n - 1.0) * ((n - 2.0) * (a.Sab * a.Sab) - 2.0 * t) + (n - 2.0) * (x2 + y2 - n * n) + x2 * y2;
}
PGLD_HD double poly_dd_between(const HapP& a, const HapP& b) { return a.Dp * b.Dp; }  // x 1/16
PGLD_HD double poly_dz_iii(const HapP& a) {                                          // x 1/4
  const double n = a.n, x2 = a.Sa * a.Sa, y2 = a.Sb * a.Sb, t = a.Sa * a.Sb * a.Sab;
  return (n + 1.0) * t - (n - 1.0) * (a.Sab * a.Sab) + n * n - x2 - y2 - x2 * y2;
}
PGLD_HD double poly_dz_iij(const HapP& a, const HapP& b) { return (a.Sa * b.Sb) * a.Dp; }  // x 1/4
PGLD_HD double poly_dz_iji(const HapP& a, const HapP& b) { return (b.Sa * a.Sb) * a.Dp; }  // x 1/4
PGLD_HD double poly_dz_ijj(const HapP& a, const HapP& b) { return a.Dp * (b.Sa * b.Sb - b.Sab); }  // x 1/4
PGLD_HD double poly_pi2_iiii(const HapP& a) {  // x 1/16
  const double n = a.n, x2 = a.Sa * a.Sa, y2 = a.Sb * a.Sb, t = a.Sa * a.Sb * a.Sab;
  return (n * n - 4.0 * n + 2.0) * (n * n - x2 - y2) + x2 * y2 - 4.0 * t + 2.0 * (a.Sab * a.Sab);
}
PGLD_HD double poly_pi2_triple(const HapP& t, const HapP& s) {  // x 1/16
  return (t.n - 2.0) * (t.n * (t.n * s.n - t.Sb * s.Sb) - s.n * (t.Sa * t.Sa)) +
         (t.Sa * s.Sb) * (t.Sa * t.Sb - 2.0 * t.Sab);
}
PGLD_HD double poly_pi2_iikk_sym(const HapP& a, const HapP& b) {  // x 1/32
  const double na2 = a.n * a.n, nb2 = b.n * b.n;
  return 2.0 * na2 * nb2 - na2 * (b.Sa * b.Sa + b.Sb * b.Sb) - nb2 * (a.Sa * a.Sa + a.Sb * a.Sb) +
         (a.Sa * b.Sb) * (a.Sa * b.Sb) + (b.Sa * a.Sb) * (b.Sa * a.Sb);
}
PGLD_HD double poly_pi2_ijij(const HapP& A, const HapP& B) {  // x 1/16
  const double P = A.Sa * B.Sa, Q = A.Sb * B.Sb;
  return (A.n - 1.0) * (B.n - 1.0) * (A.n * B.n - P - Q) + P * Q - (A.Sa * A.Sb) * B.Sab - (B.Sa * B.Sb) * A.Sab +
         A.Sab * B.Sab;
}

PGLD_HD double num_dd_single(const HapP& a) { return 0.0625 * poly_dd_single(a); }
PGLD_HD double num_dd_between(const HapP& a, const HapP& b) { return 0.0625 * poly_dd_between(a, b); }
PGLD_HD double num_dz_iii(const HapP& a) { return 0.25 * poly_dz_iii(a); }
// Dz(i,i,k): a = pop i, b = pop k
PGLD_HD double num_dz_iij(const HapP& a, const HapP& b) { return 0.25 * poly_dz_iij(a, b); }
// Dz(i,j,i): a = pop i, b = pop j
PGLD_HD double num_dz_iji(const HapP& a, const HapP& b) { return 0.25 * poly_dz_iji(a, b); }
// Dz(i,j,j): a = pop i, b = pop j
PGLD_HD double num_dz_ijj(const HapP& a, const HapP& b) { return 0.25 * poly_dz_ijj(a, b); }
PGLD_HD double num_dz_diff(const HapP& a, const HapP& b, const HapP& c) {
  return -(a.D * (b.pA - b.qA) * (c.c1 - c.c2 + c.c3 - c.c4));
}
PGLD_HD double num_pi2_iiii(const HapP& a) { return 0.0625 * poly_pi2_iiii(a); }
// three samples of pop t, one of pop s (haplotype_kernels.py:268-289)
PGLD_HD double num_pi2_triple(const HapP& t, const HapP& s) { return 0.0625 * poly_pi2_triple(t, s); }
PGLD_HD double num_pi2_iikk(const HapP& a, const HapP& b) { return a.pA * a.qA * (b.pB * b.qB); }
PGLD_HD double num_pi2_iikk_sym(const HapP& a, const HapP& b) { return 0.03125 * poly_pi2_iikk_sym(a, b); }
// pi2(i,i,k,l), denominator 2 n_i(n_i-1) n_k n_l
PGLD_HD double num_pi2_iikl(const HapP& a, const HapP& b, const HapP& c) {
  return a.pA * a.qA * (b.pB * c.qB + b.qB * c.pB);
}
// pi2(i,j,k,k), denominator 2 n_k(n_k-1) n_i n_j
PGLD_HD double num_pi2_ijkk(const HapP& a, const HapP& b, const HapP& c) {
  return c.pB * c.qB * (a.pA * b.qA + a.qA * b.pA);
}
// pi2(i,j,i,j) (exactly symmetric in the two populations; the reference's 4
// quartic products and 36-monomial cubic correction, haplotype_kernels.py:
// 243-266, in centred moments)
PGLD_HD double num_pi2_ijij(const HapP& A, const HapP& B) { return 0.0625 * poly_pi2_ijij(A, B); }
// one population shared between the two locus pairs (haplotype_kernels.py:218-241)
PGLD_HD double num_pi2_shared(const HapP& S, const HapP& X, const HapP& Y) {
  const double s1 = S.c1, s2 = S.c2, s3 = S.c3, s4 = S.c4;
  const double xa = X.pA, xq = X.qA;
  const double b1 = Y.c1, b2 = Y.c2, b3 = Y.c3, b4 = Y.c4;
  double inner = (-1.0 + s3) * b1 + s3 * b2 - b3 + s3 * b3 + s3 * b4 + s1 * (b2 + b4);
  double num = s4 * s4 * xa * Y.pB + s2 * s2 * xq * Y.pB + (-1.0 + s1 + s3) * (s3 * xa + s1 * xq) * Y.qB +
               s4 * (s1 * xq * Y.pB + xa * inner) +
               s2 * (s4 * (xa + xq) * Y.pB + s3 * (xa * Y.pB + xq * Y.qB) +
                     xq * ((-1.0 + s1) * b1 - b3 + s1 * (b2 + b3 + b4)));
  return num * 0.25;
}
// denominator 4 n_a n_b n_c n_d
PGLD_HD double num_pi2_alldiff(const HapP& a, const HapP& b, const HapP& c, const HapP& d) {
  return (a.pA * b.qA + a.qA * b.pA) * (c.pB * d.qB + c.qB * d.pB);
}

// Division forms: one correctly rounded division per statistic, 0 where the
// denominator is 0 (ld_statistics.py:1124-1127).
PGLD_HD double hap_dd_single(const HapP& a) { return safe_div(num_dd_single(a), f4(a.n)); }
PGLD_HD double hap_dd_between(const HapP& a, const HapP& b) {
  return safe_div(num_dd_between(a, b), f2(a.n) * f2(b.n));
}
PGLD_HD double hap_dz_iii(const HapP& a) { return safe_div(num_dz_iii(a), f4(a.n)); }
PGLD_HD double hap_dz_iij(const HapP& a, const HapP& b) { return safe_div(num_dz_iij(a, b), b.n * f3(a.n)); }
PGLD_HD double hap_dz_iji(const HapP& a, const HapP& b) { return safe_div(num_dz_iji(a, b), b.n * f3(a.n)); }
PGLD_HD double hap_dz_ijj(const HapP& a, const HapP& b) { return safe_div(num_dz_ijj(a, b), f2(a.n) * f2(b.n)); }
PGLD_HD double hap_dz_diff(const HapP& a, const HapP& b, const HapP& c) {
  return safe_div(num_dz_diff(a, b, c), f2(a.n) * b.n * c.n);
}
PGLD_HD double hap_pi2_iiii(const HapP& a) { return safe_div(num_pi2_iiii(a), f4(a.n)); }
PGLD_HD double hap_pi2_triple(const HapP& t, const HapP& s) {
  return safe_div(num_pi2_triple(t, s), s.n * f3(t.n));
}
PGLD_HD double hap_pi2_iikk(const HapP& a, const HapP& b) { return safe_div(num_pi2_iikk(a, b), f2(a.n) * f2(b.n)); }
PGLD_HD double hap_pi2_iikk_sym(const HapP& a, const HapP& b) {
  return safe_div(num_pi2_iikk_sym(a, b), f2(a.n) * f2(b.n));
}
PGLD_HD double hap_pi2_iikl(const HapP& a, const HapP& b, const HapP& c) {
  return safe_div(num_pi2_iikl(a, b, c), 2.0 * f2(a.n) * b.n * c.n);
}
PGLD_HD double hap_pi2_ijkk(const HapP& a, const HapP& b, const HapP& c) {
  return safe_div(num_pi2_ijkk(a, b, c), 2.0 * f2(c.n) * a.n * b.n);
}
PGLD_HD double hap_pi2_ijij(const HapP& A, const HapP& B) {
  return safe_div(num_pi2_ijij(A, B), f2(A.n) * f2(B.n));
}
PGLD_HD double hap_pi2_shared(const HapP& S, const HapP& X, const HapP& Y) {
  return safe_div(num_pi2_shared(S, X, Y), f2(S.n) * X.n * Y.n);
}
PGLD_HD double hap_pi2_alldiff(const HapP& a, const HapP& b, const HapP& c, const HapP& d) {
  return safe_div(num_pi2_alldiff(a, b, c, d), 4.0 * a.n * b.n * c.n * d.n);
}

// The within-population haplotype set (DD_i_i, Dz_i_i_i, pi2_i_i_i_i) of one
// pair straight from its centred moments n, a = Sa, b = Sb, s = Sab, with
// m = n - 2, t = ab s, q = (ab)^2 and w = n^2 - a^2 - b^2 shared by the
// three numerators:
//   poly_dd_single  = (n-1)((n-2) s^2 - 2t) + (n-2)(a^2 + b^2 - n^2) + a^2 b^2
//                   = (m + 1)(m s^2 - 2t) - m w + q
//   poly_dz_iii     = (n+1) t - (n-1) s^2 + w - q
//   poly_pi2_iiii   = (m^2 - 2) w + q + 2 (s^2 - 2t)
// Every intermediate is an integer, so the numerators are exact (terms below
// 2^53) and equal to the poly_* forms above.  out[] receives the unscaled
// numerators times r = 1 / (n)_4 (hap_self_recip); the factors 1/16, 1/4,
// 1/16 are applied to the sums.
__device__ __forceinline__ void hap_self_stats(double n, double a, double b, double s, double r, double* out) {
  const double m = n - 2.0, m1 = n - 1.0;
  const double ab = a * b, s2 = s * s, t = ab * s, q = ab * ab;
  const double w = fma(-b, b, fma(-a, a, n * n));
  const double t2 = -2.0 * t;
  out[0] = fma(m1, fma(m, s2, t2), fma(-m, w, q)) * r;
  out[1] = fma(-m1, s2, fma(n + 1.0, t, w - q)) * r;
  out[2] = fma(fma(m, m, -2.0), w, fma(2.0, s2 + t2, q)) * r;
}

// 1 / (n)_4 = 1 / (h (h - 2)), h = (n - 1)(n - 2): v_rcp_f64 refined by one
// Newton step (within an ulp), 0 when n < 4 (the reference's "invalid
// denominator -> 0").  The fused kernels either call it per pair or read it
// from a table of n = 0 .. N filled by the same function, so both give the
// same bits.
__device__ __forceinline__ double hap_self_recip(double n) {
  const double m = n - 2.0, h = fma(m, m, m), g4 = fma(h, h, -2.0 * h);
  double r = 0.0;
  if (g4 > 0.0) {
    r = __builtin_amdgcn_rcp(g4);
    r = fma(fma(-g4, r, 1.0), r, r);
  }
  return r;
}

// Reciprocal form for the fused epilogue: the four falling-factorial
// reciprocals of one population from a single division (a few ulp from the
// division form); 0 where the factorial is 0, which reproduces the reference's
// "invalid denominator -> 0".
struct HapR {
  double i1, i2, i3, i4;  // 1/n, 1/(n)_2, 1/(n)_3, 1/(n)_4
};
__device__ __forceinline__ HapR hap_recips(double n) {
  HapR r;
  const double g2 = n * (n - 1.0), g3 = g2 * (n - 2.0), g4 = g3 * (n - 3.0);
  if (g4 > 0.0) {
    // v_rcp_f64 refined by one Newton step: within an ulp of 1 / g4
    double q = __builtin_amdgcn_rcp(g4);
    q = fma(fma(-g4, q, 1.0), q, q);
    r.i4 = q;
    r.i3 = r.i4 * (n - 3.0);
    r.i2 = r.i3 * (n - 2.0);
    r.i1 = r.i2 * (n - 1.0);
  } else {
    r.i4 = 0.0;
    r.i3 = g3 > 0.0 ? 1.0 / g3 : 0.0;
    r.i2 = g2 > 0.0 ? 1.0 / g2 : 0.0;
    r.i1 = n > 0.0 ? 1.0 / n : 0.0;
  }
  return r;
}

// The two-population haplotype moments set (15 statistics in the order and
// with the power-of-two factors of epi_hap_moments_2) straight from each
// population's centred moments n, a = Sa, b = Sb, s = Sab, sharing the
// products the polynomials have in common: with ab_k = a_k b_k, Dp_k = n_k s_k
// - ab_k, u = a0 b1 + a1 b0, X = n0 n1 - b0 b1 and w_k = n_k^2 - a_k^2 - b_k^2,
//   poly_dd_between    = Dp0 Dp1
//   poly_dz_iij + iji  = Dp0 u           (populations (0, 1); (1, 0): Dp1 u)
//   poly_dz_ijj(0, 1)  = Dp0 (ab1 - s1)
//   poly_pi2_triple(0, 1) = (n0 - 2)(n0 X - n1 a0^2) + (a0 b1)(ab0 - 2 s0)
//   poly_pi2_iikk_sym  = n0^2 w1 + n1^2 w0 + (a0 b1)^2 + (a1 b0)^2
//   poly_pi2_ijij      = (n0 - 1)(n1 - 1)(X - a0 a1) + a0 a1 b0 b1 - ab0 s1 - ab1 s0 + s0 s1
// and the within-population numerators of hap_self_stats.  Every
// intermediate is an integer, so the numerators equal the poly_* forms
// exactly (terms below 2^53) and the statistics are bit-identical to
// epi_hap_moments_2; the reciprocals are hap_recips'.
__device__ __forceinline__ void hap_two_pop_stats(const double* n, const double* a, const double* b, const double* s,
                                                  double* out) {
  double ab[2], w[2], nn[2], m[2], m1[2], Dp[2], self[2][3];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    m[k] = n[k] - 2.0;
    m1[k] = n[k] - 1.0;
    ab[k] = a[k] * b[k];
    nn[k] = n[k] * n[k];
    const double s2 = s[k] * s[k], t = ab[k] * s[k], q = ab[k] * ab[k];
    w[k] = fma(-b[k], b[k], fma(-a[k], a[k], nn[k]));
    const double t2 = -2.0 * t;
    self[k][0] = fma(m1[k], fma(m[k], s2, t2), fma(-m[k], w[k], q));
    self[k][1] = fma(-m1[k], s2, fma(n[k] + 1.0, t, w[k] - q));
    self[k][2] = fma(fma(m[k], m[k], -2.0), w[k], fma(2.0, s2 + t2, q));
    Dp[k] = fma(n[k], s[k], -ab[k]);
  }
  const HapR R0 = hap_recips(n[0]), R1 = hap_recips(n[1]);
  const double r1 = R0.i2 * R1.i2, r3 = R0.i3 * R1.i1, r4 = R0.i1 * R1.i3;
  const double a0b1 = a[0] * b[1], a1b0 = a[1] * b[0];
  const double u = a0b1 + a1b0;
  const double X = fma(-b[0], b[1], n[0] * n[1]);
  const double P = a[0] * a[1], Q = b[0] * b[1];
  const double tri01 = fma(m[0], fma(n[0], X, -n[1] * (a[0] * a[0])), a0b1 * fma(-2.0, s[0], ab[0]));
  const double tri10 = fma(m[1], fma(n[1], X, -n[0] * (a[1] * a[1])), a1b0 * fma(-2.0, s[1], ab[1]));
  const double iikk = fma(nn[0], w[1], fma(nn[1], w[0], fma(a0b1, a0b1, a1b0 * a1b0)));
  const double ijij = fma(m1[0] * m1[1], X - P, fma(P, Q, fma(-ab[0], s[1], fma(-ab[1], s[0], s[0] * s[1]))));
  out[0] = self[0][0] * R0.i4;
  out[1] = (Dp[0] * Dp[1]) * r1;
  out[2] = self[1][0] * R1.i4;
  out[3] = self[0][1] * R0.i4;
  out[4] = (Dp[0] * u) * r3;
  out[5] = (Dp[0] * (ab[1] - s[1])) * r1;
  out[6] = (Dp[1] * (ab[0] - s[0])) * r1;
  out[7] = (Dp[1] * u) * r4;
  out[8] = self[1][1] * R1.i4;
  out[9] = self[0][2] * R0.i4;
  out[10] = tri01 * r3;
  out[11] = iikk * r1;
  out[12] = ijij * r1;
  out[13] = tri10 * r4;
  out[14] = self[1][2] * R1.i4;
}

// ---------------------------------------------------------------------------
// Genotype per-population values.  Counts in moments order g[0..8] =
// g1..g9 = n22, n21, n20, n12, n11, n10, n02, n01, n00
// (genotype_kernels.py:424-435).
// ---------------------------------------------------------------------------
struct GenoP {
  double g[9];
  double n;
  double pA, qA, pB, qB;      // allele counts at the two loci
  double xAB, xAb, xaB, xab;  // same-individual joint allele counts
  // moments of the centred dosages a' = a - 1, b' = b - 1 (in {-1, 0, 1})
  // over the individuals: Sa = sum a', Sab = sum a' b', Sa2b = sum a'^2 b', ...
  double Sa, Sb, Sab, Sa2, Sb2, Sa2b, Sab2, Sa2b2;
};

PGLD_HD GenoP geno_pop(const double* cnt_n00_to_n22) {
  GenoP p;
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    p.g[t] = cnt_n00_to_n22[8 - t];
    s += p.g[t];
  }
  p.n = s;
  const double* g = p.g;
  p.pA = g[0] + g[1] + g[2] + 0.5 * (g[3] + g[4] + g[5]);
  p.qA = 0.5 * (g[3] + g[4] + g[5]) + g[6] + g[7] + g[8];
  p.pB = g[0] + 0.5 * g[1] + g[3] + 0.5 * g[4] + g[6] + 0.5 * g[7];
  p.qB = 0.5 * g[1] + g[2] + 0.5 * g[4] + g[5] + 0.5 * g[7] + g[8];
  p.xAB = g[0] + 0.5 * g[1] + 0.5 * g[3] + 0.25 * g[4];
  p.xAb = 0.5 * g[1] + g[2] + 0.25 * g[4] + 0.5 * g[5];
  p.xaB = 0.5 * g[3] + 0.25 * g[4] + g[6] + 0.5 * g[7];
  p.xab = 0.25 * g[4] + 0.5 * g[5] + 0.5 * g[7] + g[8];
  // g[3 ia + ib] has a' = 1 - ia, b' = 1 - ib
  double R[3], B1[3], B2[3];
#pragma unroll
  for (int ia = 0; ia < 3; ++ia) {
    R[ia] = g[3 * ia] + g[3 * ia + 1] + g[3 * ia + 2];
    B1[ia] = g[3 * ia] - g[3 * ia + 2];
    B2[ia] = g[3 * ia] + g[3 * ia + 2];
  }
  p.Sb = B1[0] + B1[1] + B1[2];
  p.Sb2 = B2[0] + B2[1] + B2[2];
  p.Sa = R[0] - R[2];
  p.Sab = B1[0] - B1[2];
  p.Sab2 = B2[0] - B2[2];
  p.Sa2 = R[0] + R[2];
  p.Sa2b = B1[0] + B1[2];
  p.Sa2b2 = B2[0] + B2[2];
  return p;
}

// The same values from integer counts: the sums are formed in int32 (exact,
// counts below 2^28) and converted once.
PGLD_HD GenoP geno_pop_int(const int* cnt_n00_to_n22) {
  GenoP p;
  int g[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    g[t] = cnt_n00_to_n22[8 - t];
    p.g[t] = g[t];
  }
  int Ri[3], B1[3], B2[3];
#pragma unroll
  for (int ia = 0; ia < 3; ++ia) {
    Ri[ia] = g[3 * ia] + g[3 * ia + 1] + g[3 * ia + 2];
    B1[ia] = g[3 * ia] - g[3 * ia + 2];
    B2[ia] = g[3 * ia] + g[3 * ia + 2];
  }
  const int n = Ri[0] + Ri[1] + Ri[2];
  p.n = n;
  // twice the allele counts, four times the joint ones (then scaled exactly)
  const int C1 = g[1] + g[4] + g[7];
  p.pA = 0.5 * (double)(2 * Ri[0] + Ri[1]);
  p.qA = 0.5 * (double)(Ri[1] + 2 * Ri[2]);
  p.pB = 0.5 * (double)(2 * (g[0] + g[3] + g[6]) + C1);
  p.qB = 0.5 * (double)(C1 + 2 * (g[2] + g[5] + g[8]));
  p.xAB = 0.25 * (double)(4 * g[0] + 2 * g[1] + 2 * g[3] + g[4]);
  p.xAb = 0.25 * (double)(2 * g[1] + 4 * g[2] + g[4] + 2 * g[5]);
  p.xaB = 0.25 * (double)(2 * g[3] + g[4] + 4 * g[6] + 2 * g[7]);
  p.xab = 0.25 * (double)(g[4] + 2 * g[5] + 2 * g[7] + 4 * g[8]);
  p.Sb = B1[0] + B1[1] + B1[2];
  p.Sb2 = B2[0] + B2[1] + B2[2];
  p.Sa = Ri[0] - Ri[2];
  p.Sab = B1[0] - B1[2];
  p.Sab2 = B2[0] - B2[2];
  p.Sa2 = Ri[0] + Ri[2];
  p.Sa2b = B1[0] + B1[2];
  p.Sa2b2 = B2[0] + B2[2];
  return p;
}

// D from genotype frequencies (_PopDataGeno.D_geno, genotype_kernels.py:441-446):
// xAB xab - xAb xaB = (n Sab - Sa Sb) / 4.
PGLD_HD double geno_D(const GenoP& p) { return (p.n * p.Sab - p.Sa * p.Sb) * 0.25; }

// The single-population numerators of genotype_kernels.py (DD :27-71, pi2
// :279-300 with its correction polynomial, Dz :199-233), polynomials of
// degree 4 in the nine genotype counts with a few hundred monomials, are
// written here in the centred moments, where they collapse to a handful of
// terms (X = Sb^2 - Sb2, Y = Sa^2 - Sa2; the identities are exact and are
// checked against the reference's tables in the test suite):
//   DD  = [(n-1)(n-2) Sab^2 - n(n-1) Sa2b2 - 2(n-1) Sa Sb Sab + 2(n-1)(Sb Sa2b + Sa Sab2) + X Y] / 4
//   Dz  = [(n+1)(Sa Sb Sab - Sb Sa2b - Sa Sab2) - (n-1) Sab^2 + 2n Sa2b2 - X Y] / 4
//   pi2 = [n(n-1)(n-2)(n-3) - (n-2)(n-3)(X + Y) + X Y - 4 Sa Sb Sab + 4 (Sb Sa2b + Sa Sab2)
//          + 2 Sab^2 - 6 Sa2b2] / 16
// Every intermediate value is an integer, so for counts whose products stay
// below 2^53 they equal the reference's values exactly.
PGLD_HD double geno_dd_num(const GenoP& p) {
  const double n = p.n, X = p.Sb * p.Sb - p.Sb2, Y = p.Sa * p.Sa - p.Sa2;
  const double sab = p.Sa * p.Sb * p.Sab, lin = p.Sb * p.Sa2b + p.Sa * p.Sab2;
  return 0.25 * ((n - 1.0) * ((n - 2.0) * (p.Sab * p.Sab) - n * p.Sa2b2 - 2.0 * sab + 2.0 * lin) + X * Y);
}
PGLD_HD double geno_dz_num(const GenoP& p) {
  const double n = p.n, X = p.Sb * p.Sb - p.Sb2, Y = p.Sa * p.Sa - p.Sa2;
  const double sab = p.Sa * p.Sb * p.Sab, lin = p.Sb * p.Sa2b + p.Sa * p.Sab2;
  return 0.25 * ((n + 1.0) * (sab - lin) - (n - 1.0) * (p.Sab * p.Sab) + 2.0 * n * p.Sa2b2 - X * Y);
}
PGLD_HD double geno_pi2_num(const GenoP& p) {
  const double n = p.n, X = p.Sb * p.Sb - p.Sb2, Y = p.Sa * p.Sa - p.Sa2;
  const double sab = p.Sa * p.Sb * p.Sab, lin = p.Sb * p.Sa2b + p.Sa * p.Sab2;
  const double n23 = (n - 2.0) * (n - 3.0);
  return 0.0625 * (n * (n - 1.0) * n23 - n23 * (X + Y) + X * Y - 4.0 * sab + 4.0 * lin + 2.0 * (p.Sab * p.Sab) -
                   6.0 * p.Sa2b2);
}
PGLD_HD double geno_dd_single(const GenoP& p) {
  return p.n < 4.0 ? 0.0 : geno_dd_num(p) / f4(p.n);
}
PGLD_HD double geno_dd_between(const GenoP& a, const GenoP& b) {
  return safe_div(4.0 * geno_D(a) * geno_D(b), f2(a.n) * f2(b.n));
}
// -g1 + g3 - g4 + g6 - g7 + g9 = -Sb and -g1 - g2 - g3 + g7 + g8 + g9 = -Sa
PGLD_HD double geno_Acoef(const GenoP& p) { return -p.Sb; }
PGLD_HD double geno_Bcoef(const GenoP& p) { return -p.Sa; }
PGLD_HD double geno_dz_iii(const GenoP& p) {
  if (p.n < 4.0) return 0.0;
  return 2.0 * geno_dz_num(p) / f4(p.n);
}
// Dz precompute features (genotype_kernels.py:176-197)
PGLD_HD double geno_Ucoef(const GenoP& p) {
  const double* g = p.g;
  double MA = 2. * g[1] * g[6] + 4. * g[2] * g[6] + 2. * g[4] * g[6] + 4. * g[5] * g[6] - 2. * g[0] * g[7] +
              2. * g[2] * g[7] - 2. * g[3] * g[7] + 2. * g[5] * g[7] - 4. * g[0] * g[8] - 2. * g[1] * g[8] -
              4. * g[3] * g[8] - 2. * g[4] * g[8];
  return geno_D(p) * (1. + geno_Bcoef(p)) + 0.25 * MA;
}
PGLD_HD double geno_Vcoef(const GenoP& p) {
  const double* g = p.g;
  double MB = 2. * g[2] * g[3] + 2. * g[2] * g[4] - 2. * g[0] * g[5] - 2. * g[1] * g[5] + 4. * g[2] * g[6] +
              2. * g[5] * g[6] + 4. * g[2] * g[7] + 2. * g[5] * g[7] - 4. * g[0] * g[8] - 4. * g[1] * g[8] -
              2. * g[3] * g[8] - 2. * g[4] * g[8];
  return geno_D(p) * (1. + geno_Acoef(p)) + 0.25 * MB;
}
PGLD_HD double geno_Wjj(const GenoP& p) {
  const double* g = p.g;
  return (-g[0] + g[2] + g[6] - g[8]) + geno_Acoef(p) * geno_Bcoef(p);
}
// The reference's multi-pop Dz kernels divide without a guard
// (genotype_kernels.py:199-233); a zero denominator gives inf/NaN there too.
PGLD_HD double geno_dz_iij(const GenoP& a, const GenoP& b) {
  return 2.0 * geno_Ucoef(a) * geno_Acoef(b) / (b.n * f3(a.n));
}
PGLD_HD double geno_dz_iji(const GenoP& a, const GenoP& b) {
  return 2.0 * geno_Vcoef(a) * geno_Bcoef(b) / (b.n * f3(a.n));
}
PGLD_HD double geno_dz_ijj(const GenoP& a, const GenoP& b) {
  return 2.0 * geno_D(a) * geno_Wjj(b) / (f2(b.n) * f2(a.n));
}
PGLD_HD double geno_dz_diff(const GenoP& a, const GenoP& b, const GenoP& c) {
  return 2.0 * geno_D(a) * geno_Bcoef(b) * geno_Acoef(c) / (f2(a.n) * b.n * c.n);
}
PGLD_HD double geno_pi2_iiii(const GenoP& p) {
  return p.n < 4.0 ? 0.0 : geno_pi2_num(p) / f4(p.n);
}
PGLD_HD double nn1(const GenoP& p) {
  double v = 4.0 * p.n * (p.n - 1.0);
  return v > 1.0 ? v : 1.0;
}
PGLD_HD double geno_BL(const GenoP& p) { return (4.0 * p.pA * p.qA - (p.g[3] + p.g[4] + p.g[5])) / nn1(p); }
PGLD_HD double geno_BR(const GenoP& p) { return (4.0 * p.pB * p.qB - (p.g[1] + p.g[4] + p.g[7])) / nn1(p); }
// triple coefficients (genotype_kernels.py:473-489)
PGLD_HD void geno_triple_A(const GenoP& p, double& AQ3, double& AS3) {
  const double* g = p.g;
  double P = p.pA, R = p.qA, Q = p.pB, S = p.qB;
  double L0 = g[4] + 2 * g[5] + 2 * g[7] + 4 * g[8];
  double L1 = g[4] + 2 * g[5];
  double negR3 = 3 * g[3] + 3 * g[4] + 3 * g[5] + 4 * g[6] + 4 * g[7] + 4 * g[8];
  double negR5 = 5 * g[3] + 5 * g[4] + 5 * g[5] + 8 * g[6] + 8 * g[7] + 8 * g[8];
  double beta5A = -2 * g[1] - 4 * g[2] + 4 * g[3] + g[4] - 2 * g[5] + 8 * g[6] + 4 * g[7] + 4;
  double beta6A = -g[4] - 2 * g[5] - 4 * g[6] - 4 * g[7] - 4 * g[8];
  AQ3 = -(S * negR3) / 8 + P * S * R / 2 + (R * L0) / 8 + L1 / 8 - (P * L0) / 8;
  AS3 = -(Q * negR5) / 8 + P * Q * R / 2 + (R * beta5A) / 8 + beta6A / 8 + (P * L0) / 8;
}
PGLD_HD void geno_triple_B(const GenoP& p, double& BP3, double& BR3) {
  const double* g = p.g;
  double P = p.pA, R = p.qA, Q = p.pB, S = p.qB;
  double L0 = g[4] + 2 * g[5] + 2 * g[7] + 4 * g[8];
  double M1 = g[4] + 2 * g[7];
  double negL3 = 3 * g[1] + 4 * g[2] + 3 * g[4] + 4 * g[5] + 3 * g[7] + 4 * g[8];
  double negL5 = 5 * g[1] + 8 * g[2] + 5 * g[4] + 8 * g[5] + 5 * g[7] + 8 * g[8];
  double beta5B = 4 * g[1] + 8 * g[2] - 2 * g[3] + g[4] + 4 * g[5] - 4 * g[6] - 2 * g[7] + 4;
  double beta6B = -4 * g[2] - g[4] - 4 * g[5] - 2 * g[7] - 4 * g[8];
  BP3 = -(R * negL3) / 8 + Q * S * R / 2 + (S * L0) / 8 + M1 / 8 - (Q * L0) / 8;
  BR3 = -(P * negL5) / 8 + P * Q * S / 2 + (S * beta5B) / 8 + beta6B / 8 + (Q * L0) / 8;
}
// pi2(t,t;t,s) and pi2(t,t;s,t): single pop in the B-locus pair
PGLD_HD double geno_pi2_iiij(const GenoP& t, const GenoP& s) {
  double AQ3, AS3;
  geno_triple_A(t, AQ3, AS3);
  return safe_div(AQ3 * s.pB + AS3 * s.qB, f3(t.n) * s.n);
}
// pi2(t,s;t,t) and pi2(s,t;t,t): single pop in the A-locus pair
PGLD_HD double geno_pi2_ijii(const GenoP& t, const GenoP& s) {
  double BP3, BR3;
  geno_triple_B(t, BP3, BR3);
  return safe_div(BP3 * s.pA + BR3 * s.qA, f3(t.n) * s.n);
}
PGLD_HD double geno_pi2_iikk(const GenoP& a, const GenoP& b) { return geno_BL(a) * geno_BR(b); }
PGLD_HD double geno_pi2_iikl(const GenoP& a, const GenoP& b, const GenoP& c) {
  double qb = b.pB / b.n, sb = b.qB / b.n, qc = c.pB / c.n, sc = c.qB / c.n;
  return geno_BL(a) * 0.5 * (qb * sc + sb * qc);
}
PGLD_HD double geno_pi2_ijkk(const GenoP& a, const GenoP& b, const GenoP& c) {
  double pa = a.pA / a.n, ra = a.qA / a.n, pb = b.pA / b.n, rb = b.qA / b.n;
  return 0.5 * (pa * rb + ra * pb) * geno_BR(c);
}
PGLD_HD double geno_pi2_ijij(const GenoP& A, const GenoP& B) {
  double Pi = A.pA, Qi = A.pB, Ri = A.qA, Si = A.qB;
  double Pj = B.pA, Qj = B.pB, Rj = B.qA, Sj = B.qB;
  double X01i = A.xAb, X10i = A.xaB, X11i = A.xab, Bi = A.xAB - A.xAb - A.xaB + A.xab;
  double X01j = B.xAb, X11j = B.xab, Bj = B.xAB - B.xAb - B.xaB + B.xab;
  double PQi = Pi * Qi, PRi = Pi * Ri, PSi = Pi * Si, QRi = Qi * Ri, RSi = Ri * Si;
  double PQj = Pj * Qj, PSj = Pj * Sj, QRj = Qj * Rj, RSj = Rj * Sj;
  double Tc = Ri * (1.0 + X01i - X10i), Uc = Si * (1.0 + Ri);
  (void)X01j;
  double num = 0.0;
  num += -0.25 * RSi * Qj - 0.25 * X11i * PQj + 0.25 * RSi * PQj - 0.25 * Ri * PSj;
  num += -0.25 * PQi * X11j - 0.25 * RSi * X11j + 0.25 * PSi * X11j + 0.25 * QRi * X11j;
  num += -0.25 * PRi * Sj + 0.25 * Tc * Sj + 0.25 * X11i * PSj + 0.25 * QRi * PSj;
  num += 0.25 * Uc * Rj - 0.25 * X01i * QRj + 0.25 * PSi * QRj;
  num += -0.25 * PSi * Rj - 0.25 * Pi * RSj + 0.25 * X01i * RSj + 0.25 * PQi * RSj;
  num += 0.25 * X11i * Bj + 0.25 * Bi * X11j - 1.0 * X11i * X11j;
  return safe_div(num, f2(A.n) * f2(B.n));
}
// one shared population s; a = other pop of the first pair, b = of the second
PGLD_HD double geno_pi2_shared(const GenoP& S, const GenoP& A, const GenoP& B) {
  double d = nn1(S);
  double C00 = (S.pA * S.pB - S.xAB) / d, C01 = (S.pA * S.qB - S.xAb) / d;
  double C10 = (S.qA * S.pB - S.xaB) / d, C11 = (S.qA * S.qB - S.xab) / d;
  double ra = A.qA / A.n, pa = A.pA / A.n, sb = B.qB / B.n, qb = B.pB / B.n;
  return ra * (C00 * sb + C01 * qb) + pa * (C10 * sb + C11 * qb);
}
PGLD_HD double geno_pi2_alldiff(const GenoP& a, const GenoP& b, const GenoP& c, const GenoP& d) {
  double pa = a.pA / a.n, ra = a.qA / a.n, pb = b.pA / b.n, rb = b.qA / b.n;
  double qc = c.pB / c.n, sc = c.qB / c.n, qd = d.pB / d.n, sd = d.qB / d.n;
  return 0.25 * (pa * rb + ra * pb) * (qc * sd + sc * qd);
}
PGLD_HD double geno_sigma_d2(const GenoP& p) {
  const double nan_v = __builtin_nan("");
  if (p.n < 4.0) return nan_v;
  double pn = geno_pi2_num(p);
  return pn > 0.0 ? geno_dd_num(p) / pn : nan_v;
}

// ---------------------------------------------------------------------------
// r, r^2, D' (haplotype_kernels.py:20-62)
// ---------------------------------------------------------------------------
PGLD_HD void hap_r_family(double a, double b, double c, double d, double n, double& r, double& r2,
                          double& dp) {
  const double nan_v = __builtin_nan("");
  double pA = (a + b) / n, pB = (a + c) / n;
  double qA = 1.0 - pA, qB = 1.0 - pB;
  double D = (a * d - b * c) / (n * n);
  double den = pA * qA * pB * qB;
  r = den > 0.0 ? D / sqrt(den) : nan_v;
  r2 = den > 0.0 ? D * D / den : nan_v;
  double dmax;
  if (D >= 0.0) {
    double v1 = pA * qB, v2 = qA * pB;
    dmax = v1 < v2 ? v1 : v2;
  } else {
    double v1 = pA * pB, v2 = qA * qB;
    dmax = v1 < v2 ? v1 : v2;
  }
  dp = dmax > 0.0 ? D / dmax : nan_v;
}

}  // namespace pgld
