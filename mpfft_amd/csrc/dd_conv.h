// Conversions between P-bit limb values (mp_arith.h) and double-word
// numbers hi + lo, for the double-word form (dd_fft.hip): the exact split
// of a limb value into two binary64 words, and the correctly rounded P-bit
// value of the sum of two binary64 numbers.
//
// The header also compiles as plain C++ with MPFFT_CPU_BUILD defined, so
// the conversions can be checked against MPFR on a host without a GPU
// (tests/cpu_dd_check.cpp).
#pragma once
#include "mp_arith.h"

namespace mpfft {

struct DD {
    double hi, lo;
};

// P-bit limb value (53 <= P <= 106) times 2^-eoff (the transform's block
// exponent, an exact shift) -> hi + lo, exact.  The significand's top 53
// bits rounded to nearest (ties to even) give hi; the remainder, at most 53
// bits wide because P <= 106, is lo.  A scaled exponent e below -lim
// (binary64 could not hold lo; the value is below 2^-lim of the transform's
// largest input) is taken as zero; one above lim returns 1.
template <int L>
MPFFT_FN uint32_t dd_from_mp(DD &z, const Mp<L> &x, int lim, int32_t eoff)
{
    const int32_t e = x.e - eoff;
    if (x.zero || e < -lim) {
        z.hi = x.neg ? -0.0 : 0.0;
        z.lo = 0.0;
        return 0;
    }
    if (e > lim) {
        z.hi = 0.0;
        z.lo = 0.0;
        return 1;
    }
    // value = 0.m * 2^e; A = significand bits 1..64, B = bits 65..128
    const uint64_t A = (uint64_t)x.m[0] << 32 | (L > 1 ? x.m[1] : 0u);
    const uint64_t B = (uint64_t)(L > 2 ? x.m[2] : 0u) << 32 | (L > 3 ? x.m[3] : 0u);
    uint64_t h = A >> 11;                                        // bits 1..53
    int64_t r = (int64_t)(((A & 0x7ffull) << 42) | (B >> 22));  // bits 54..106, < 2^53
    const int64_t half = 1ll << 52;
    if (r > half || (r == half && (h & 1u))) {
        h += 1;
        r -= 2 * half;
    }
    const double hi = __builtin_ldexp((double)h, e - 53);
    const double lo = __builtin_ldexp((double)r, e - 106);
    z.hi = x.neg ? -hi : hi;
    z.lo = x.neg ? -lo : lo;
    return 0;
}

// a binary64 value as an exact limb value (53 bits <= 32 L)
template <int L>
MPFFT_FN void mp_from_double(Mp<L> &z, double v)
{
    if (v == 0.0) {
        set_zero(z, __builtin_signbit(v) ? 1u : 0u);
        return;
    }
    int ex = 0;
    const double f = __builtin_frexp(__builtin_fabs(v), &ex);  // [0.5, 1)
    const uint64_t m = (uint64_t)__builtin_ldexp(f, 53) << 11;  // 53 bits at the top
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = 0u;
    z.m[0] = (uint32_t)(m >> 32);
    z.m[1] = (uint32_t)m;
    z.e = ex;
    z.neg = v < 0.0 ? 1u : 0u;
    z.zero = 0u;
}

// z = a + b rounded to nearest (ties to even) at P bits, 1 <= P <= 32 L
// <= 128, for any two finite binary64 numbers; zeros as in IEEE 754 (x - x
// = +0, (-0) + (-0) = -0).  The larger magnitude's 53-bit significand sits
// at the top of a 128-bit window (bit 127 = its leading one) and the other
// is added or subtracted at its offset; what falls below the window is kept
// as a sticky bit (a subtracted one takes one unit more off the window and
// sets the sticky bit, so the window holds the floor of the exact value).
// The result is normalised (a carry shifts right, a cancellation left; with
// a sticky bit the offset is at least 76 bits, so at most one bit cancels)
// and rounded at bit 128 - P.
template <int L>
MPFFT_FN void dd_round_mp(Mp<L> &z, double a, double b, int P)
{
    if (__builtin_fabs(a) < __builtin_fabs(b)) {
        const double t = a;
        a = b;
        b = t;
    }
    if (a == 0.0) {
        set_zero(z, (__builtin_signbit(a) && __builtin_signbit(b)) ? 1u : 0u);
        return;
    }
    int ea = 0, eb = 0;
    const uint64_t A = (uint64_t)__builtin_ldexp(__builtin_frexp(__builtin_fabs(a), &ea), 53);
    uint64_t w1 = A << 11, w0 = 0;                 // a = 0.w 2^ea
    uint32_t sticky = 0;
    int32_t e = ea;
    if (b != 0.0) {
        const uint64_t B = (uint64_t)__builtin_ldexp(__builtin_frexp(__builtin_fabs(b), &eb), 53);
        const int sh = 75 - (ea - eb);             // b = B 2^sh in window units
        uint64_t b1 = 0, b0 = 0;
        uint32_t st = 0;
        if (sh >= 64) {
            b1 = B << (sh - 64);
        } else if (sh > 0) {
            b1 = B >> (64 - sh);
            b0 = B << sh;
        } else if (sh == 0) {
            b0 = B;
        } else if (sh > -64) {
            b0 = B >> -sh;
            st = (B & ((1ull << -sh) - 1)) != 0;
        } else {
            st = 1;
        }
        if (__builtin_signbit(a) == __builtin_signbit(b)) {
            w0 += b0;
            const uint64_t c = w0 < b0;
            const uint64_t o = w1;
            w1 += b1 + c;
            sticky = st;
            if (w1 < o || (c && w1 == o)) {       // carry out of bit 127
                sticky |= (uint32_t)(w0 & 1u);
                w0 = (w0 >> 1) | (w1 << 63);
                w1 = (w1 >> 1) | (1ull << 63);
                e += 1;
            }
        } else {
            const uint64_t d0 = b0 + st;           // the floor: one unit more when inexact
            const uint64_t c = (d0 < b0) ? 1u : 0u;
            const uint64_t br = w0 < d0;
            w0 -= d0;
            w1 -= b1 + c + br;
            sticky = st;
            if (w1 == 0 && w0 == 0) {              // exact cancellation
                set_zero(z, 0u);
                return;
            }
            const int s = w1 ? __builtin_clzll(w1) : 64 + __builtin_clzll(w0);
            if (s >= 64) {
                w1 = w0 << (s - 64);
                w0 = 0;
            } else if (s > 0) {
                w1 = (w1 << s) | (w0 >> (64 - s));
                w0 <<= s;
            }
            e -= s;
        }
    }
    // round at bit r = 128 - P: guard bit r - 1, the bits below it, the
    // last kept bit
    const int r = 128 - P;
    uint32_t g, lsb, below;
    if (r > 64) {
        g = (uint32_t)(w1 >> (r - 65)) & 1u;
        below = ((w1 & ((1ull << (r - 65)) - 1)) | w0) != 0;
        lsb = (uint32_t)(w1 >> (r - 64)) & 1u;
        w1 &= ~((1ull << (r - 64)) - 1);
        w0 = 0;
    } else if (r == 64) {
        g = (uint32_t)(w0 >> 63);
        below = (w0 << 1) != 0;
        lsb = (uint32_t)w1 & 1u;
        w0 = 0;
    } else {
        g = (uint32_t)(w0 >> (r - 1)) & 1u;
        below = (w0 & ((1ull << (r - 1)) - 1)) != 0;
        lsb = r < 64 ? (uint32_t)(w0 >> r) & 1u : 0u;
        w0 &= ~((1ull << r) - 1);
    }
    if (g && (below || sticky || lsb)) {
        if (r >= 64) {
            w1 += 1ull << (r - 64);
        } else {
            const uint64_t o = w0;
            w0 += 1ull << r;
            if (w0 < o) w1 += 1;
        }
        if (w1 == 0) {                            // rounded up to 2^128
            w1 = 1ull << 63;
            e += 1;
        }
    }
    const uint32_t m[4] = {(uint32_t)(w1 >> 32), (uint32_t)w1, (uint32_t)(w0 >> 32), (uint32_t)w0};
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = j < 4 ? m[j] : 0u;
    z.e = e;
    z.neg = __builtin_signbit(a) ? 1u : 0u;
    z.zero = 0u;
}

// (hi + lo) 2^eoff rounded once to P bits
template <int L>
MPFFT_FN void dd_to_mp(Mp<L> &z, const DD &x, int P, int32_t eoff)
{
    dd_round_mp(z, x.hi, x.lo, P);
    if (!z.zero) z.e += eoff;
}

}  // namespace mpfft
