// Host-side check of zpic_amd/csrc/kernels/init_core.h (the arithmetic of
// the device-side particle initialisation), built with g++ by
// tests/test_init_core.py:
//  1. the jump-ahead of the multiply-with-carry halves equals stepping the
//     generator (em2d/random.c:48-53) call by call;
//  2. log_cr, the double-double log, is correctly rounded: on the polar
//     method's own s values and on values of every binade in (0, 1) it
//     equals the C library's log, and where the two differ log_cr is the
//     double nearer to the quadruple-precision logq (the C library's log is
//     not correctly rounded everywhere);
//  3. deviates drawn attempt-parallel (chunks of attempts counted, a prefix
//     sum of the counts, then the accepted attempts emitted as pairs, as
//     the device does) equal the sequential polar method of
//     em2d/random.c:67-101, for several chunk sizes and kept deviates, when
//     both use the C library's log; with log_cr the deviates that differ
//     are counted.
// Prints one JSON line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <quadmath.h>

#include "init_core.h"

using namespace zk_init;

static uint32_t seq_next(uint32_t &z, uint32_t &w)
{
    z = 36969u * (z & 0xffffu) + (z >> 16);
    w = 18000u * (w & 0xffffu) + (w >> 16);
    return (z << 16) + w;
}

// sequential polar method (em2d/random.c:67-101)
struct Seq {
    uint32_t z, w;
    int iset = 0;
    double gset = 0.0;
    double next()
    {
        if (iset) {
            iset = 0;
            return gset;
        }
        double v1, v2, rsq;
        do {
            v1 = (seq_next(z, w) + 0.5) / 2147483649.0 - 1.0;
            v2 = (seq_next(z, w) + 0.5) / 2147483649.0 - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (rsq == 0.0 || rsq >= 1.0);
        const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
        gset = v1 * fac;
        iset = 1;
        return v2 * fac;
    }
};

static long long bad_log = 0, glibc_diff = 0;

// a = log_cr(s): equal to the C library's log, or else nearer to logl(s)
static void check_log(double s, double a)
{
    const double g = std::log(s);
    if (a == g) return;
    glibc_diff++;
    const __float128 L = logq((__float128)s);
    if (fabsq((__float128)a - L) >= fabsq((__float128)g - L)) bad_log++;
}

int main()
{
    int fails_jump = 0;
    // 1. jump-ahead
    {
        uint32_t z0 = 67890u, w0 = 12345u;
        uint32_t z = z0, w = w0;
        seq_next(z, w);
        const uint32_t z1 = z, w1 = w;
        for (uint64_t n = 1; n <= 3000000; n++) {
            if (n > 1) seq_next(z, w);
            if (n % 99991 == 1 || n < 50) {
                uint32_t jz, jw;
                mwc_at(z1, w1, n, jz, jw);
                if (jz != z || jw != w) fails_jump++;
            }
        }
    }
    // 2. log
    long long nlog = 0, amb_log = 0;
    {
        uint32_t z = 67890u, w = 12345u;
        for (int t = 0; t < 3000000; t++) {
            const double u = (seq_next(z, w) + 0.5) / 2147483649.0 - 1.0;
            const double v = (seq_next(z, w) + 0.5) / 2147483649.0 - 1.0;
            const double s = u * u + v * v;
            if (s == 0.0 || s >= 1.0) continue;
            int amb = 0;
            const double a = log_cr(s, &amb);
            nlog++;
            check_log(s, a);
            amb_log += amb;
        }
        srand(7);
        for (int t = 0; t < 1000000; t++) {
            const double m = 0.5 + 0.5 * ((double)rand() / RAND_MAX) * (1.0 - 1e-16);
            const int e = -(rand() % 1000);
            const double s = std::ldexp(m, e);
            if (!(s > 0.0 && s < 1.0) || !std::isnormal(s)) continue;
            int amb = 0;
            const double a = log_cr(s, &amb);
            nlog++;
            check_log(s, a);
            amb_log += amb;
        }
    }
    // 3. attempt-parallel deviates
    long long bad_dev = 0, ndev = 0, cr_dev_diff = 0;
    for (int use_cr = 0; use_cr <= 1; use_cr++)
    for (int kept0 = 0; kept0 <= 1; kept0++)
        for (int per : {1, 3, 64}) {
            const uint32_t z0 = 1234567u, w0 = 7654321u;
            const long long D = 30001 + kept0; // deviates wanted
            Seq ref{z0, w0};
            std::vector<double> want(D);
            double kept_in = 0.0;
            if (kept0) {
                // a deviate kept from before: draw one pair's first deviate
                ref.next();
                kept_in = ref.gset;
            }
            const uint32_t zs = ref.z, ws = ref.w;
            for (long long q = 0; q < D; q++) want[q] = ref.next();
            // parallel: attempts counted per chunk of `per`, prefix, emit
            uint32_t z1 = zs, w1 = ws;
            seq_next(z1, w1);
            const long long pairs = (D - kept0 + 1) / 2;
            const long long nat = pairs * 2 + 1000;
            const long long nch = (nat + per - 1) / per;
            std::vector<long long> off(nch + 1, 0);
            for (long long c = 0; c < nch; c++) {
                uint32_t z, w;
                if (c == 0) {
                    z = zs;
                    w = ws;
                } else {
                    mwc_at(z1, w1, (uint64_t)(2 * c * per), z, w);
                }
                long long n = 0;
                for (long long t = c * per; t < std::min(nat, (c + 1) * per); t++) {
                    const double u = (mwc_next(z, w) + 0.5) / 2147483649.0 - 1.0;
                    const double v = (mwc_next(z, w) + 0.5) / 2147483649.0 - 1.0;
                    const double s = u * u + v * v;
                    n += !(s == 0.0 || s >= 1.0);
                }
                off[c + 1] = off[c] + n;
            }
            std::vector<double> got(D, NAN);
            if (kept0) got[0] = kept_in;
            for (long long c = 0; c < nch; c++) {
                uint32_t z, w;
                if (c == 0) {
                    z = zs;
                    w = ws;
                } else {
                    mwc_at(z1, w1, (uint64_t)(2 * c * per), z, w);
                }
                long long p = off[c];
                for (long long t = c * per; t < std::min(nat, (c + 1) * per) && p < pairs; t++) {
                    const double u = (mwc_next(z, w) + 0.5) / 2147483649.0 - 1.0;
                    const double v = (mwc_next(z, w) + 0.5) / 2147483649.0 - 1.0;
                    const double s = u * u + v * v;
                    if (s == 0.0 || s >= 1.0) continue;
                    int amb = 0;
                    const double L = use_cr ? log_cr(s, &amb) : std::log(s);
                    const double fac = std::sqrt(-2.0 * L / s);
                    const double dv[2] = {v * fac, u * fac};
                    for (int h = 0; h < 2; h++) {
                        const long long q = kept0 + 2 * p + h;
                        if (q < D) got[q] = dv[h];
                    }
                    p++;
                }
            }
            for (long long q = 0; q < D; q++) {
                if (use_cr) {
                    cr_dev_diff += !(got[q] == want[q]);
                } else {
                    ndev++;
                    bad_dev += !(got[q] == want[q]);
                }
            }
        }
    printf("{\"jump_fail\": %d, \"logs\": %lld, \"log_not_nearest\": %lld, \"log_differs_from_libc\": %lld, "
           "\"log_ambiguous\": %lld, \"deviates\": %lld, \"deviate_mismatch\": %lld, "
           "\"deviates_differing_with_log_cr\": %lld}\n",
           fails_jump, nlog, bad_log, glibc_diff, amb_log, ndev, bad_dev, cr_dev_diff);
    return 0;
}
