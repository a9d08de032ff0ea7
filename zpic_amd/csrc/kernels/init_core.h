/*
 * init_core.h - arithmetic of the device-side particle initialisation
 * (init.hip), written for both the device and the host compiler so that
 * the host can check it against the C library (tests/init_core_check.cpp):
 * the reference generator's multiply-with-carry halves as multiplicative
 * congruential generators with jump-ahead, and a correctly rounded double
 * log in double-double arithmetic.
 */
#ifndef ZK_INIT_CORE_H
#define ZK_INIT_CORE_H

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define ZK_HD __host__ __device__
#else
#define ZK_HD
#endif

namespace zk_init {

constexpr uint32_t MZ = 36969u * 65536u - 1u; /* 2422800383 */
constexpr uint32_t MW = 18000u * 65536u - 1u; /* 1179647999 */

ZK_HD inline uint32_t mulmod(uint32_t a, uint32_t b, uint32_t m)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) % m);
}

ZK_HD inline uint32_t powmod(uint32_t a, uint64_t e, uint32_t m)
{
    uint32_t r = 1u % m, b = a % m;
    while (e) {
        if (e & 1u) r = mulmod(r, b, m);
        b = mulmod(b, b, m);
        e >>= 1;
    }
    return r;
}

/* ---- double-double arithmetic (Dekker, Knuth) ---- */
struct dd {
    double hi, lo;
};

ZK_HD inline dd two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

ZK_HD inline dd quick_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

ZK_HD inline dd two_prod(double a, double b)
{
    const double p = a * b;
    return {p, __builtin_fma(a, b, -p)};
}

ZK_HD inline dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

ZK_HD inline dd dd_mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}

ZK_HD inline dd dd_div(dd a, dd b)
{
    const double q1 = a.hi / b.hi;
    dd r = dd_add(a, dd_mul({-q1, 0.0}, b));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul({-q2, 0.0}, b));
    const double q3 = r.hi / b.hi;
    dd q = quick_two_sum(q1, q2);
    return dd_add(q, {q3, 0.0});
}

/* 1 / (2k + 1) as double-double */
ZK_HD inline dd inv_odd(int k)
{
    const double d = (double)(2 * k + 1);
    const double h = 1.0 / d;
    return {h, __builtin_fma(-h, d, 1.0) / d};
}

/* log(x) for 0 < x < 1, x normal: correctly rounded double unless the
 * rounding test fails (then *amb is set and the nearest double to the
 * double-double value is returned). */
ZK_HD inline double log_cr(double x, int *amb)
{
    int e;
    double m = __builtin_frexp(x, &e); /* x = m 2^e, m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    /* f = (m - 1) / (m + 1); m - 1 is exact, m + 1 as double-double */
    const dd f = dd_div({m - 1.0, 0.0}, two_sum(m, 1.0));
    const dd f2 = dd_mul(f, f);
    /* series: sum_{k=0}^{K} f^2k / (2k + 1), |f| <= 0.1716, K = 22 */
    const int K = 22;
    dd s = inv_odd(K);
    for (int k = K - 1; k >= 0; k--) s = dd_add(dd_mul(s, f2), inv_odd(k));
    dd lm = dd_mul(f, s);
    lm.hi *= 2.0;
    lm.lo *= 2.0;
    /* e ln 2, ln 2 = L1 + L2 */
    const double L1 = 0x1.62e42fefa39efp-1, L2 = 0x1.abc9e3b39803fp-56;
    const dd el = dd_add(two_prod((double)e, L1), two_prod((double)e, L2));
    const dd r = dd_add(el, lm);
    /* rounding test: the double-double carries the value to about
     * |r| 2^-100; hi is correctly rounded unless lo is that close to half
     * an ulp of hi */
    int eh;
    (void)__builtin_frexp(r.hi, &eh); /* |hi| in [2^(eh-1), 2^eh): ulp 2^(eh-53) */
    const double half = __builtin_ldexp(1.0, eh - 54);
    const double tol = __builtin_fabs(r.hi) * 0x1p-98;
    const double alo = __builtin_fabs(r.lo);
    /* below a power of two the spacing toward zero is half as large */
    if (__builtin_fabs(alo - half) <= tol || __builtin_fabs(alo - 0.5 * half) <= tol) *amb = 1;
    return r.hi;
}

/* state after call n >= 1 of the generator, from the state after call 1 */
ZK_HD inline void mwc_at(uint32_t z1, uint32_t w1, uint64_t n, uint32_t &z, uint32_t &w)
{
    z = mulmod(powmod(36969u, n - 1, MZ), z1, MZ);
    w = mulmod(powmod(18000u, n - 1, MW), w1, MW);
}

ZK_HD inline uint32_t mwc_next(uint32_t &z, uint32_t &w)
{
    z = 36969u * (z & 0xffffu) + (z >> 16);
    w = 18000u * (w & 0xffffu) + (w >> 16);
    return (z << 16) + w;
}


} // namespace zk_init

#endif
