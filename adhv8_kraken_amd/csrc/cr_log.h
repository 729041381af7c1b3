// Correctly rounded natural logarithm in binary64, usable on host and device.
//
// The SW2 friction coefficient evaluates log() on a depth that the
// finite-difference Jacobian perturbs by ~1.5e-8; a last-bit difference in
// log becomes a ~1e-8 relative difference in a Jacobian entry.  The host
// path links glibc's log, which is correctly rounded on all but a vanishing
// fraction of inputs; this routine returns the correctly rounded result on
// every input, so the two agree wherever glibc is correctly rounded.
//
// Method (Ziv's strategy, J.-M. Muller, "Elementary Functions", ch. 11):
//   x = 2^k * m, m in [sqrt(1/2), sqrt(2));  c = 1 + j/128 nearest to m;
//   log x = k*ln2 + log(c) + log1p((m - c)/c)
// evaluated in double-double (error < 2^-67 relative).  If the double-double
// result is not unambiguously rounded, a slow path recomputes log(m) by the
// atanh series 2*(s + s^3/3 + ...), s = (m-1)/(m+1), to ~2^-104.
// The tables log(c_j) and ln2 are generated at first use from the same
// atanh series, so no constant is copied from a library.
#pragma once

#include <stdint.h>
#include <string.h>
#include <math.h>

#ifdef __HIPCC__
#define CRL_FN __host__ __device__ inline
#else
#define CRL_FN static inline
#endif

namespace crlog {

struct dd { double hi, lo; };

CRL_FN dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
CRL_FN dd fast_two_sum(double a, double b) {
    double s = a + b;
    return {s, b - (s - a)};
}
CRL_FN dd two_prod(double a, double b) {
    double p = a * b;
    return {p, fma(a, b, -p)};
}
CRL_FN dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
CRL_FN dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
CRL_FN dd dd_mul_d(dd a, double b) {
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}
CRL_FN dd dd_div(dd a, dd b) {
    double q1 = a.hi / b.hi;
    dd r = dd_add(a, dd_mul_d(b, -q1));
    double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul_d(b, -q2));
    double q3 = r.hi / b.hi;
    dd q = fast_two_sum(q1, q2);
    return dd_add(q, dd{q3, 0.0});
}

CRL_FN uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
CRL_FN double from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

// log(m) for m in [0.7, 1.42] by 2*atanh(s), s = (m-1)/(m+1), in double-double.
CRL_FN dd log_series(dd m) {
    dd num = dd_add(m, dd{-1.0, 0.0});
    dd den = dd_add(m, dd{1.0, 0.0});
    dd s = dd_div(num, den);
    dd s2 = dd_mul(s, s);
    // sum_{n=0}^{27} s^(2n+1)/(2n+1); |s| <= 0.172 so s^56 < 2^-142
    dd acc = {0.0, 0.0};
    for (int n = 27; n >= 0; --n) {
        dd term = dd_div(dd{1.0, 0.0}, dd{(double)(2 * n + 1), 0.0});
        acc = dd_add(dd_mul(acc, s2), term);
    }
    acc = dd_mul(acc, s);
    return dd_add(acc, acc);
}

struct tables {
    dd ln2;
    dd logc[128];  // log(1 + (j - 40)/128), j = 0..95
    int ready;
};

CRL_FN void make_tables(tables *t) {
    // ln 2 = 2 atanh(1/3) = log_series(2) evaluated as 2*log_series(sqrt(2))? use m=2 directly:
    // s = 1/3, series converges (3^-2 per term): 60 terms.
    dd s = dd_div(dd{1.0, 0.0}, dd{3.0, 0.0});
    dd s2 = dd_mul(s, s);
    dd acc = {0.0, 0.0};
    for (int n = 60; n >= 0; --n) {
        dd term = dd_div(dd{1.0, 0.0}, dd{(double)(2 * n + 1), 0.0});
        acc = dd_add(dd_mul(acc, s2), term);
    }
    acc = dd_mul(acc, s);
    t->ln2 = dd_add(acc, acc);
    for (int j = 0; j < 96; ++j) {
        double c = 1.0 + (double)(j - 40) / 128.0;
        t->logc[j] = log_series(dd{c, 0.0});
    }
    t->ready = 1;
}

// Core: correctly rounded log of a positive normal or subnormal finite x.
CRL_FN double cr_log_t(double x, const tables *t) {
    uint64_t u = bits(x);
    if (!(x > 0.0)) {
        if (x == 0.0) return -INFINITY;
        return NAN;  // negative or NaN
    }
    if (u >= 0x7ff0000000000000ULL) return x;  // +inf
    int k = 0;
    if (u < 0x0010000000000000ULL) {  // subnormal
        x *= 18014398509481984.0;     // 2^54
        u = bits(x);
        k = -54;
    }
    int e = (int)((u >> 52) & 0x7ff) - 1023;
    double m = from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1,2)
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    k += e;
    if (k == 0 && m == 1.0) return 0.0;
    // table reduction
    int j = (int)floor((m - 1.0) * 128.0 + 0.5);
    double c = 1.0 + (double)j / 128.0;
    double d = m - c;                       // exact (Sterbenz)
    // r = d / c in double-double
    double r1 = d / c;
    double rem = fma(-r1, c, d);            // exact remainder
    double r2 = rem / c;
    dd r = fast_two_sum(r1, r2);
    // log1p(r) = r - r^2/2 + r^3 * P(r), |r| < 2^-7.5
    dd r2dd = dd_mul(r, r);
    dd half_r2 = {r2dd.hi * 0.5, r2dd.lo * 0.5};
    double rr = r.hi;
    double tail = 0.0;
    // r^3 * (1/3 - r/4 + r^2/5 - ... ) to r^13, Horner in double
    for (int n = 13; n >= 3; --n) tail = tail * rr + ((n & 1) ? 1.0 : -1.0) / (double)n;
    double r3 = r2dd.hi * rr;
    dd l1p = dd_add(r, dd{-half_r2.hi, -half_r2.lo});
    l1p = dd_add(l1p, dd{r3 * tail, 0.0});
    dd res = dd_add(t->logc[j + 40], l1p);
    dd kl = dd_mul_d(t->ln2, (double)k);
    res = dd_add(kl, res);
    // Ziv rounding test: error bound 2^-66 relative
    double err = fabs(res.hi) * 1.3552527156068805e-20;
    double a = res.hi + (res.lo + err);
    double b = res.hi + (res.lo - err);
    if (a == b) return a;
    // slow path
    dd lm = log_series(dd{m, 0.0});
    res = dd_add(kl, lm);
    return res.hi + res.lo;
}

}  // namespace crlog
