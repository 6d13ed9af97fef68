// Fixed-point multiprecision arithmetic for the "fixed" form of the FFT.
//
// A real is a two's-complement integer X of W 32-bit words, word 0 least
// significant, with value X / 2^F, F = 32 (W - 1): one integer word (sign
// and 31 integer bits) above W - 1 fraction words.  Additions are exact
// (W-word carry chains); a product with a twiddle keeps the bits at and
// above 2^-F and drops the rest (error below 1.04 units of 2^-F).  No
// normalisation and no rounding step: that is where this form saves work
// against correctly rounded floating point, at the price of an absolute
// rather than relative error per operation (the FFT's own error is
// absolute-like, and F is chosen 32 bits beyond the floating precision).
//
// Like mp_arith.h, the header compiles as plain C++ with MPFFT_CPU_BUILD so
// the arithmetic can be checked on a host against exact integer arithmetic
// (tests/test_cpu_fx.py through tests/cpu_fx_check.cpp).
#pragma once
#include "mp_arith.h"

namespace mpfft {

template <int W>
struct Fx {
    uint32_t w[W];
};

// digits of a twiddle magnitude: |t| < 2, value T / 2^F, T < 2^(F+1).
// Digit width B: 29 bits while |X| has at most 63 digits (W <= 57), 28
// beyond (76 digits at W = 66); a product column then holds at most 63
// (resp. 76) terms below 2^(2B) plus a carry below 2^(64-B): below 2^64.
template <int W>
struct FxDig {
    static constexpr int F = 32 * (W - 1);
    static constexpr int B = (32 * W + kDigitBits - 1) / kDigitBits <= 63 ? kDigitBits : 28;
    static constexpr uint32_t MASK = (1u << B) - 1u;
    static constexpr int DX = (32 * W + B - 1) / B;        // |X| < 2^(32W-1)
    static constexpr int DT = (F + 1 + B - 1) / B;         // T < 2^(F+1)
    // columns below K0 are skipped.  Column c < K0 sums at most DX digit
    // products below 2^(2B) at weight 2^(B c), so the dropped part is below
    // DX 2^(2B) 2^(B (K0 - 1)) (1 + 2^(1-B)): at B = 29 (DX <= 63) below
    // 2^(29 K0 + 35.01), at B = 28 (DX <= 76) below 2^(28 K0 + 34.25); with
    // B K0 <= F - 40 that is under 2^(F - 4.99), 1/32 of the output's last
    // unit
    static constexpr int K0 = (F - 40) >= 0 ? (F - 40) / B : 0;
    static_assert(DX < (1 << (64 - 2 * B)) && DX <= 76, "column sums would overflow 64 bits");
};

template <int W>
MPFFT_FN void fx_add(Fx<W> &z, const Fx<W> &x, const Fx<W> &y)
{
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < W; j++) z.w[j] = add_cc(x.w[j], y.w[j], c);
}

template <int W>
MPFFT_FN void fx_sub(Fx<W> &z, const Fx<W> &x, const Fx<W> &y)
{
    uint32_t b = 0;
#pragma unroll
    for (int j = 0; j < W; j++) z.w[j] = sub_bb(x.w[j], y.w[j], b);
}

// z = x + (-1)^s * y, s per lane
template <int W>
MPFFT_FN void fx_addsub(Fx<W> &z, const Fx<W> &x, const Fx<W> &y, uint32_t s)
{
    // x - y = x + ~y + 1
    const uint32_t m = s ? 0xffffffffu : 0u;
    uint32_t c = s;
#pragma unroll
    for (int j = 0; j < W; j++) z.w[j] = add_cc(x.w[j], y.w[j] ^ m, c);
}

template <int W>
MPFFT_FN uint32_t fx_sign(const Fx<W> &x)
{
    return x.w[W - 1] >> 31;
}

// |x| (x != -2^(32W-1))
template <int W>
MPFFT_FN void fx_abs(Fx<W> &z, const Fx<W> &x, uint32_t neg)
{
    const uint32_t m = neg ? 0xffffffffu : 0u;
    uint32_t c = neg;
#pragma unroll
    for (int j = 0; j < W; j++) z.w[j] = add_cc(x.w[j] ^ m, 0, c);
}

// |X| (W words, LS first) -> DX digits of B bits, LS first
template <int W>
MPFFT_FN void fx_to_digits(const Fx<W> &x, uint32_t (&d)[FxDig<W>::DX])
{
    constexpr int B = FxDig<W>::B;
#pragma unroll
    for (int k = 0; k < FxDig<W>::DX; k++) {
        const int bit = B * k, wi = bit >> 5, s = bit & 31;
        const uint32_t lo = wi < W ? x.w[wi] : 0u;
        const uint32_t hi = wi + 1 < W ? x.w[wi + 1] : 0u;
        d[k] = (s == 0 ? lo : fshr(hi, lo, (uint32_t)s)) & FxDig<W>::MASK;
    }
}

// z = x * t / 2^F truncated toward zero in magnitude, t = (-1)^tneg T / 2^F
// given as DT digits of T.  x is clobbered (it holds |x| afterwards); z may
// be the same object as x.
template <int W>
MPFFT_FN void fx_mul(Fx<W> &z, Fx<W> &x, const uint32_t (&T)[FxDig<W>::DT], uint32_t tneg)
{
    constexpr int F = FxDig<W>::F, DX = FxDig<W>::DX, DT = FxDig<W>::DT, K0 = FxDig<W>::K0;
    constexpr int B = FxDig<W>::B;
    const uint32_t xneg = fx_sign(x);
    fx_abs(x, x, xneg);
    uint32_t A[DX];
    fx_to_digits(x, A);
#pragma unroll
    for (int j = 0; j < W; j++) z.w[j] = 0;
    uint64_t acc = 0;
#pragma unroll
    for (int c = K0; c < DX + DT; c++) {
        const int i0 = c - (DT - 1) > 0 ? c - (DT - 1) : 0;
        const int i1 = c < DX - 1 ? c : DX - 1;
#pragma unroll
        for (int i = i0; i <= i1; i++) acc = (uint64_t)A[i] * T[c - i] + acc;
        // digit c covers product bits [Bc, Bc + B) = result bits
        // [Bc - F, Bc - F + B); bits below result bit 0 are dropped
        const uint32_t v = (c == DX + DT - 1) ? (uint32_t)acc : ((uint32_t)acc & FxDig<W>::MASK);
        acc >>= B;
        const int b = B * c - F;
        if (b + B <= 0) continue;
        if (b < 0) {
            z.w[0] |= v >> (-b);
        } else {
            const int j = b >> 5, s = b & 31;
            if (j < W) z.w[j] |= v << s;
            if (s + B > 32 && j + 1 < W) z.w[j + 1] |= v >> (32 - s);
        }
    }
    fx_abs(z, z, xneg ^ tneg);           // negate when the signs differ
}

// ---- SoA rows: word j of element i at buf[j * M + i] -------------------------

template <int W>
MPFFT_FN void fx_load(Fx<W> &x, const uint32_t *__restrict__ buf, uint32_t M, uint32_t i)
{
#pragma unroll
    for (int j = 0; j < W; j++) x.w[j] = (buf + (size_t)j * M)[i];
}

template <int W>
MPFFT_FN void fx_store(const Fx<W> &x, uint32_t *__restrict__ buf, uint32_t M, uint32_t i)
{
#pragma unroll
    for (int j = 0; j < W; j++) (buf + (size_t)j * M)[i] = x.w[j];
}

// ---- conversions ---------------------------------------------------------------

// P-bit float (L limbs, MS first, value 0.m 2^e) scaled by 2^-eoff (the
// transform's block exponent: an exact shift) -> fixed point with F
// fraction bits, truncated toward zero.  Returns 1 when |x 2^-eoff| >= 2^lim
// (the caller's range limit; the result is then saturated to zero).
template <int L, int W>
MPFFT_FN uint32_t fx_from_mp(Fx<W> &z, const Mp<L> &x, int lim, int32_t eoff = 0)
{
    constexpr int F = FxDig<W>::F;
#pragma unroll
    for (int j = 0; j < W; j++) z.w[j] = 0;
    if (x.zero) return 0;
    const int32_t e = x.e - eoff;
    if (e > lim) return 1;
    // X = M * 2^sh, M = the L limbs as an integer, sh = e - 32L + F
    const int32_t sh = e - 32 * L + F;
    uint32_t t[W];
#pragma unroll
    for (int j = 0; j < W; j++) t[j] = j < L ? x.m[L - 1 - j] : 0u;   // LS first
    if (sh >= 0) {
        // left shift: sh <= lim - 32L + F + ... stays below 32 W - 1 bits
        int32_t s = sh;
#pragma unroll 1
        while (s >= 32) {
#pragma unroll
            for (int j = W - 1; j > 0; j--) t[j] = t[j - 1];
            t[0] = 0;
            s -= 32;
        }
        if (s) {
#pragma unroll
            for (int j = W - 1; j > 0; j--) t[j] = fshr(t[j], t[j - 1], 32u - (uint32_t)s);
            t[0] <<= s;
        }
    } else {
        int32_t s = -sh;
        if (s >= 32 * W) return 0;
#pragma unroll 1
        while (s >= 32) {
#pragma unroll
            for (int j = 0; j < W - 1; j++) t[j] = t[j + 1];
            t[W - 1] = 0;
            s -= 32;
        }
        if (s) {
#pragma unroll
            for (int j = 0; j < W - 1; j++) t[j] = fshr(t[j + 1], t[j], (uint32_t)s);
            t[W - 1] >>= s;
        }
    }
    Fx<W> m;
#pragma unroll
    for (int j = 0; j < W; j++) m.w[j] = t[j];
    fx_abs(z, m, x.neg);
    return 0;
}

// fixed point -> P-bit float, correctly rounded (to nearest, ties to even),
// times 2^eoff (the block exponent taken off by fx_from_mp); zero converts
// to +0
template <int W, int L, bool FULL>
MPFFT_FN void fx_to_mp(Mp<L> &z, const Fx<W> &x, int P, int32_t eoff = 0)
{
    constexpr int F = FxDig<W>::F;
    const uint32_t neg = fx_sign(x);
    Fx<W> a;
    fx_abs(a, x, neg);
    // MS-first copy with L + 1 window words after the leading word
    uint32_t t[W];
#pragma unroll
    for (int j = 0; j < W; j++) t[j] = a.w[W - 1 - j];
    uint32_t any = 0;
#pragma unroll
    for (int j = 0; j < W; j++) any |= t[j];
    if (!any) { set_zero(z, 0); return; }
    int32_t e = 32 * W - F;                      // exponent of the top word's bit 31, plus 1
#pragma unroll 1
    while (t[0] == 0) {
#pragma unroll
        for (int j = 0; j < W - 1; j++) t[j] = t[j + 1];
        t[W - 1] = 0;
        e -= 32;
    }
    const uint32_t lz = clz32(t[0]);
    if (lz) {
#pragma unroll
        for (int j = 0; j < W - 1; j++) t[j] = fshr(t[j], t[j + 1], 32u - lz);
        t[W - 1] <<= lz;
        e -= (int32_t)lz;
    }
    uint32_t win[L + 1], sticky = 0;
#pragma unroll
    for (int j = 0; j <= L; j++) win[j] = j < W ? t[j] : 0u;
#pragma unroll
    for (int j = L + 1; j < W; j++) sticky |= t[j];
    round_window<L, FULL>(win, sticky != 0, e, P);
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = win[j];
    z.e = e + eoff; z.neg = neg; z.zero = 0;
}

}  // namespace mpfft
