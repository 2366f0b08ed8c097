// Exact bin sums of the power-spectrum binning (csrc/pk.hip), kept apart so that the
// CPU tests can check the arithmetic against 128-bit integers.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>

namespace pyl {

// Every bin sum is an integer Y in units of 2^-E, held as two words that are only
// ever added to: Y = H * 2^48 + L, H a signed 32-bit word and L a signed 64-bit
// word (two's complement in an unsigned word).  A contribution x (a double, as the
// reference's loop forms it) is scaled, t = x 2^E, and split without conversions:
// adding M = 1.5 * 2^100 rounds t to a multiple of 2^48 whose quotient H sits in the
// low bits of the sum's bit pattern; the remainder t - H 2^48 (exact, |.| <= 2^47)
// plus 1.5 * 2^52 holds round(remainder) = L the same way.  Both words take plain,
// non-returning atomic adds.  Integer sums do not depend on the order of the
// additions, so a spectrum is the same bits on every run, however the waves'
// atomics interleave (double atomics were not: at 512^3 the last bits changed from
// call to call).  E is set per realisation from a bound on the sum of
// |contributions| -- 2 * (modes) * (largest |Re| or |Im| of the transform * largest
// MAS correction)^2, and the mode count for the phase sums -- so that every |sum|
// stays below 2^78 (|H| sums below 2^31) and a word L takes 2^15 contributions
// without overflow.  Each contribution is rounded to the nearest unit of 2^-E,
// below 2^-78 of that bound: a bin 10^14 times weaker than the bound keeps a
// relative error near 1e-10.
constexpr int kFixTop = 78;
constexpr int kFixSplit = 48;
constexpr int kFixMaxAdds = 1 << 15;   // additions into one low word before it could overflow

__host__ __device__ __forceinline__ unsigned long long fix_bits(double d) {
#ifdef __HIP_DEVICE_COMPILE__
    return (unsigned long long)__double_as_longlong(d);
#else
    unsigned long long u;
    std::memcpy(&u, &d, sizeof u);
    return u;
#endif
}

__host__ __device__ __forceinline__ void fix_split(double x, int E, unsigned long long &L, int &H) {
    constexpr double M1 = 0x1.8p100, M2 = 0x1.8p52;
    const double t = std::ldexp(x, E);
    const double s = t + M1;                           // M1 + H 2^48
    H = (int)(unsigned)(fix_bits(s) - fix_bits(M1));
    const double r = t - (s - M1);                     // exact, |r| <= 2^47
    L = fix_bits(r + M2) - fix_bits(M2);               // round(r), two's complement
}

// acc[i] += x (LDS: low words at lo_acc, high words at hi_acc)
__device__ __forceinline__ void fix_add(unsigned long long *lo_acc, int *hi_acc, int i, double x, int E) {
    if (x == 0.0) return;
    unsigned long long L;
    int H;
    fix_split(x, E, L, H);
    atomicAdd(&lo_acc[i], L);
    if (H) atomicAdd(&hi_acc[i], H);
}

// acc[i] += (L, H), a value already split
__device__ __forceinline__ void fix_add_parts(unsigned long long *lo_acc, int *hi_acc, int i,
                                              unsigned long long L, int H) {
    if (L) atomicAdd(&lo_acc[i], L);
    if (H) atomicAdd(&hi_acc[i], H);
}

// (slo, shi) += (L, H), L a signed low word: kept normalised (0 <= slo < 2^48 after
// each step), so that any number of terms can be summed
__host__ __device__ __forceinline__ void fix_acc(unsigned long long &slo, long long &shi,
                                                unsigned long long L, long long H) {
    constexpr unsigned long long mask = (1ull << kFixSplit) - 1;
    const long long l = (long long)L;
    const unsigned long long t = slo + ((unsigned long long)l & mask);
    shi += H + (l >> kFixSplit) + (long long)(t >> kFixSplit);
    slo = t & mask;
}

__host__ __device__ __forceinline__ double fix_value(unsigned long long lo, long long hi, int E) {
    return std::ldexp((double)hi, kFixSplit - E) + std::ldexp((double)lo, -E);
}

// E for sums of |contributions| up to `bound`: every such sum stays below 2^kFixTop
__host__ __device__ __forceinline__ int fix_exponent(double bound) {
    if (!(bound > 0.0) || !(bound < 1e300)) return 0;    // zero, or not finite
    int e;
    (void)std::frexp(bound, &e);                         // bound < 2^e
    return kFixTop - e;
}

}  // namespace pyl
