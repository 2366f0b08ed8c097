// CPU check of the exact bin-sum arithmetic (pylians3_amd/csrc/fixsum.h) against
// 128-bit integers:
//  * split: H*2^48 + L is the contribution x*2^E rounded to the nearest integer
//    (exactly x*2^E when that is an integer), with |L| <= 2^47, for both signs
//    across the whole range;
//  * sums of 10^5 contributions of a steep spectrum, in several orders and chunk
//    groupings, are identical and equal the 128-bit sum of the split values, and
//    the sum of |values| stays below 2^78;
//  * kFixMaxAdds additions of the largest low word do not overflow it;
//  * the conversion back to double is within two roundings of the exact sum.
#include "../../pylians3_amd/csrc/fixsum.h"

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <random>
#include <vector>

using namespace pyl;
using i128 = __int128;

// a split value: L is a signed low word
static i128 val(unsigned long long L, long long H) {
    return (i128)H * ((i128)1 << kFixSplit) + (i128)(long long)L;
}
// a normalised sum: slo in [0, 2^48)
static i128 sumval(unsigned long long slo, long long shi) {
    return (i128)shi * ((i128)1 << kFixSplit) + (i128)slo;
}

int main() {
    std::mt19937_64 rng(7);
    int bad = 0;
    std::uniform_real_distribution<double> ue(-40.0, 77.0);
    for (int i = 0; i < 2000000; ++i) {
        const int E = (int)(rng() % 120) - 20;
        double t = std::ldexp(1.0 + (double)(rng() >> 11) * 0x1p-53, (int)std::floor(ue(rng)));
        if (rng() & 1) t = -t;
        const double x = std::ldexp(t, -E);
        unsigned long long L;
        int H;
        fix_split(x, E, L, H);
        if (std::llabs((long long)L) > (1ll << 47)) ++bad;
        const i128 y = val(L, H);
        if (std::fabs(t) >= 0x1p52) {
            if (y != (i128)t) ++bad;                           // t is an integer
        } else {
            const long double d = (long double)y - (long double)t;
            if (!(std::fabs(d) <= 0.5L)) ++bad;
        }
    }
    for (double t : {0.0, -0.0, 0.5, -0.5, 1.5, -1.5, 0x1p48, -0x1p48, 0x1p48 - 1, -0x1p48 + 1,
                     0x1p52, -0x1p52, 0x1p77, -0x1p77, 1e-300, -1e-300}) {
        unsigned long long L;
        int H;
        fix_split(t, 0, L, H);
        const i128 y = val(L, H);
        const long double d = (long double)y - (long double)t;
        if (!(std::fabs(d) <= 0.5L) || std::llabs((long long)L) > (1ll << 47)) ++bad;
    }
    printf("split: %s\n", bad ? "FAIL" : "ok");

    int bad2 = 0;
    for (int trial = 0; trial < 20; ++trial) {
        const int n = 100000;
        std::vector<double> x(n);
        std::normal_distribution<double> g(0.0, 1.0);
        double bound = 0.0;
        for (int i = 0; i < n; ++i) {
            x[i] = std::exp(4.0 * g(rng)) * ((trial & 1) ? g(rng) : std::fabs(g(rng)));
            bound += std::fabs(x[i]);
        }
        const int E = fix_exponent(bound * 1.01);
        i128 ref = 0, absum = 0;
        std::vector<unsigned long long> L(n);
        std::vector<int> H(n);
        for (int i = 0; i < n; ++i) {
            fix_split(x[i], E, L[i], H[i]);
            const i128 y = val(L[i], H[i]);
            ref += y;
            absum += y < 0 ? -y : y;
        }
        if (absum >= ((i128)1 << kFixTop)) ++bad2;
        std::vector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        unsigned long long first_lo = 0;
        long long first_hi = 0;
        for (int order = 0; order < 4; ++order) {
            std::shuffle(perm.begin(), perm.end(), rng);
            std::vector<unsigned long long> clo;
            std::vector<long long> chi;
            const int chunk = 1 + (int)(rng() % 5000);
            for (int s0 = 0; s0 < n; s0 += chunk) {
                // a chunk as an LDS word pair would hold it: plain (wrapping) sums of
                // L and of H
                unsigned long long wl = 0;
                long long wh = 0;
                for (int i = s0; i < std::min(n, s0 + chunk); ++i) {
                    wl += L[perm[i]];
                    wh += H[perm[i]];
                }
                clo.push_back(wl);
                chi.push_back(wh);
            }
            unsigned long long slo = 0;
            long long shi = 0;
            for (int k = (int)clo.size() - 1; k >= 0; --k) fix_acc(slo, shi, clo[k], chi[k]);
            if (sumval(slo, shi) != ref) ++bad2;
            if (order == 0) {
                first_lo = slo;
                first_hi = shi;
            } else if (slo != first_lo || shi != first_hi) {
                ++bad2;
            }
            const double v = fix_value(slo, shi, E);
            const long double exact = (long double)ref * std::ldexp(1.0L, -E);
            if (std::fabs((long double)v - exact) > 4.5e-16L * std::fabs(exact) + 1e-300L) ++bad2;
        }
    }
    printf("order independence: %s\n", bad2 ? "FAIL" : "ok");

    // the low word's headroom: kFixMaxAdds additions of 2^47 stay below 2^63 in
    // magnitude
    int bad3 = 0;
    if ((i128)kFixMaxAdds * ((i128)1 << 47) >= ((i128)1 << 63)) ++bad3;
    bad3 += (fix_exponent(0.0) != 0) + (fix_exponent(1.0) != kFixTop - 1) +
            (fix_exponent(1.5) != kFixTop - 1) + (fix_exponent(2.0) != kFixTop - 2);
    printf("bounds: %s\n", bad3 ? "FAIL" : "ok");
    printf((bad || bad2 || bad3) ? "FAILURES\n" : "all ok\n");
    return (bad || bad2 || bad3) != 0;
}
