// Correctly rounded P-bit floating-point arithmetic on 32-bit limbs, for
// gfx950 (CDNA4).  Replaces the mpfr_add / mpfr_sub / mpfr_mul calls of the
// reference hot loop (lab/serial/mp/serial_mp_fft3.c:141-248) with
// register-resident limb code: every result equals MPFR's RNDN result at
// precision P (round to nearest, ties to even, unbounded exponent, signed
// zeros as in IEEE 754 / MPFR: x - x = +0, (-0) + (-0) = -0).
//
// Number layout in registers: m[0] is the most significant limb and has its
// top bit set for a nonzero value 0.m * 2^e; bits below position P are zero.
// L = number of limbs (compile time), 32(L-1) < P <= 32L.  FULL means
// P == 32L (no padding bits), which shortens the rounding step.
//
// Every routine is branch-light: the common path is straight-line code over
// the limb arrays with compile-time indices; data-dependent limb shifts
// (large exponent gaps, massive cancellation) run as bounded loops of
// one-limb moves, so no register array is ever indexed dynamically.
//
// The header also compiles as plain C++ with MPFFT_CPU_BUILD defined; that
// build exists only so the arithmetic can be checked against MPFR on a host
// without a GPU (tests/cpu_arith_check.cpp).
#pragma once
#include <stddef.h>
#include <stdint.h>
#ifdef MPFFT_CPU_BUILD
#define MPFFT_FN static inline
#else
#include <hip/hip_runtime.h>
#define MPFFT_FN __device__ __forceinline__
#endif

namespace mpfft {

constexpr int32_t kHdrZero = INT32_MIN;   // hdr of a signed zero (| sign bit)

#ifdef MPFFT_CPU_BUILD
// host check build only: how often a truncated product fell back to the
// exact one
static unsigned long mp_mul_fallbacks = 0;
#endif

template <int L>
struct Mp {
    uint32_t m[L];
    int32_t e;        // exponent (value = 0.m * 2^e) when !zero
    uint32_t neg;     // sign bit
    uint32_t zero;    // 1 for +-0
};

MPFFT_FN uint32_t fshr(uint32_t hi, uint32_t lo, uint32_t s)
{
    // low 32 bits of (hi:lo) >> s, s in [0, 31]
#ifdef MPFFT_CPU_BUILD
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#else
    return __builtin_amdgcn_alignbit(hi, lo, s);
#endif
}

MPFFT_FN uint32_t clz32(uint32_t v)   // v != 0
{
#ifdef MPFFT_CPU_BUILD
    return (uint32_t)__builtin_clz(v);
#else
    return (uint32_t)__clz((int)v);
#endif
}

// 32-bit add / subtract with carry (borrow) in and out: one
// v_addc_co_u32 / v_subb_co_u32 each in a chain
MPFFT_FN uint32_t add_cc(uint32_t a, uint32_t b, uint32_t &c)
{
#ifdef MPFFT_CPU_BUILD
    uint64_t s = (uint64_t)a + b + c;
    c = (uint32_t)(s >> 32);
    return (uint32_t)s;
#else
    unsigned int co;
    const uint32_t r = __builtin_addc(a, b, c, &co);
    c = co;
    return r;
#endif
}

MPFFT_FN uint32_t sub_bb(uint32_t a, uint32_t b, uint32_t &br)
{
#ifdef MPFFT_CPU_BUILD
    uint64_t s = (uint64_t)a - b - br;
    br = (uint32_t)(s >> 63);
    return (uint32_t)s;
#else
    unsigned int bo;
    const uint32_t r = __builtin_subc(a, b, br, &bo);
    br = bo;
    return r;
#endif
}

// Round the normalized window w[0..L] (w[0] has its top bit set) plus a
// sticky flag (nonzero remainder below w[L]) to P bits, ties to even.
// On mantissa overflow the result becomes 0.1000.. and e is incremented.
template <int L, bool FULL>
MPFFT_FN void round_window(uint32_t (&w)[L + 1], uint32_t sticky,
                           int32_t &e, int P)
{
    if (FULL) {
        uint32_t rbit = w[L] >> 31;
        uint32_t rest = (w[L] << 1) | sticky;
        uint32_t up = rbit & ((rest != 0) | (w[L - 1] & 1));
        uint32_t c = up;
        w[L - 1] = add_cc(w[L - 1], 0, c);
        if (c) {   // the increment carries past the last limb: rare
#pragma unroll
            for (int j = L - 2; j >= 0; j--) w[j] = add_cc(w[j], 0, c);
            if (c) { w[0] = 0x80000000u; e += 1; }
        }
    } else {
        // last kept bit: index P-1 from the top -> limb kl, bit kb (from LSB)
        const int kl = (P - 1) >> 5;
        const uint32_t kb = 31u - (uint32_t)((P - 1) & 31);
        // round bit: index P from the top -> limb rl, bit rb
        const int rl = P >> 5;
        const uint32_t rb = 31u - (uint32_t)(P & 31);
        uint32_t rbit = 0, rest = sticky, lsb = 0;
#pragma unroll
        for (int j = 0; j <= L; j++) {
            uint32_t v = w[j];
            uint32_t below = (rb == 0) ? 0u : (v & ((1u << rb) - 1u));
            if (j == rl) { rbit = (v >> rb) & 1u; rest |= below; }
            if (j > rl) rest |= v;
            if (j == kl) lsb = (v >> kb) & 1u;
        }
        uint32_t up = rbit & ((rest != 0) | lsb);
        uint32_t c = 0;
#pragma unroll
        for (int j = L; j >= 0; j--) {
            uint32_t keep = (j < kl) ? 0xffffffffu
                          : (j == kl) ? ~((1u << kb) - 1u) : 0u;
            uint32_t inc = (j == kl && up) ? (1u << kb) : 0u;
            w[j] = add_cc(w[j] & keep, inc, c);
        }
        if (c) { w[0] = 0x80000000u; e += 1; }
    }
}

template <int L>
MPFFT_FN void set_zero(Mp<L> &z, uint32_t neg)
{
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = 0;
    z.e = 0; z.neg = neg; z.zero = 1;
}

// z = (-1)^fx * x + (-1)^fy * y, correctly rounded to P bits.
// Zeros go through the same code as data (a zero operand has zero limbs and
// is never the larger-exponent operand unless both are zero); a zero result
// takes the sign xs & ys, which gives x - x = +0 and (-0) + (-0) = -0.
template <int L, bool FULL>
MPFFT_FN void mp_add(Mp<L> &z, const Mp<L> &x, uint32_t fx,
                     const Mp<L> &y, uint32_t fy, int P)
{
    const uint32_t xs = x.neg ^ fx, ys = y.neg ^ fy;
    // a = the operand with the larger exponent, b = the other one
    const bool sw = x.zero || (!y.zero && y.e > x.e);
    uint32_t aw[L + 1], bw[L + 1];
#pragma unroll
    for (int j = 0; j < L; j++) {
        aw[j] = sw ? y.m[j] : x.m[j];
        bw[j] = sw ? x.m[j] : y.m[j];
    }
    aw[L] = 0; bw[L] = 0;
    int32_t e = sw ? y.e : x.e;
    const uint32_t sa = sw ? ys : xs, sb = sw ? xs : ys;
    const bool bzero = sw ? x.zero : y.zero;
    const uint32_t d = bzero ? 0u : (uint32_t)(sw ? (y.e - x.e) : (x.e - y.e));

    // align b: shift right by d bits into the (L+1)-limb window, OR-ing
    // everything that falls off the window into sticky
    uint32_t sticky = 0;
    if (d >= 32u * (L + 1)) {
#pragma unroll
        for (int j = 0; j <= L; j++) bw[j] = 0;
        sticky = 1;
    } else {
        const uint32_t q = d >> 5, r = d & 31u;
#pragma unroll 1
        for (uint32_t t = 0; t < q; t++) {
            sticky |= bw[L];
#pragma unroll
            for (int j = L; j >= 1; j--) bw[j] = bw[j - 1];
            bw[0] = 0;
        }
        if (r) {
            sticky |= bw[L] << (32u - r);
#pragma unroll
            for (int j = L; j >= 1; j--) bw[j] = fshr(bw[j - 1], bw[j], r);
            bw[0] >>= r;
        }
        sticky = sticky != 0;
    }

    uint32_t sign = sa;
    if (sa == sb) {
        uint32_t c = 0;
#pragma unroll
        for (int j = L; j >= 0; j--) aw[j] = add_cc(aw[j], bw[j], c);
        if (c) {
            sticky |= aw[L] & 1u;
#pragma unroll
            for (int j = L; j >= 1; j--) aw[j] = fshr(aw[j - 1], aw[j], 1);
            aw[0] = (aw[0] >> 1) | 0x80000000u;
            e += 1;
        }
    } else {
        // exact value is aw - bw - delta with 0 < delta < 1 (guard-LSB units)
        // when sticky: subtract one more unit and keep sticky set
        uint32_t br = sticky;
#pragma unroll
        for (int j = L; j >= 0; j--) aw[j] = sub_bb(aw[j], bw[j], br);
        if (br) {
            // only possible for d == 0 (then sticky == 0): |b| > |a|
            uint32_t c = 1;
#pragma unroll
            for (int j = L; j >= 0; j--) aw[j] = add_cc(~aw[j], 0, c);
            sign ^= 1u;
        }
    }
    uint32_t any = sticky;
#pragma unroll
    for (int j = 0; j <= L; j++) any |= aw[j];
    if (any) {
        // normalize (more than one bit only after an exact subtraction)
#pragma unroll 1
        while (aw[0] == 0) {
#pragma unroll
            for (int j = 0; j < L; j++) aw[j] = aw[j + 1];
            aw[L] = 0;
            e -= 32;
        }
        const uint32_t lz = clz32(aw[0]);
        if (lz) {
#pragma unroll
            for (int j = 0; j < L; j++) aw[j] = fshr(aw[j], aw[j + 1], 32u - lz);
            aw[L] <<= lz;
            e -= (int32_t)lz;
        }
        round_window<L, FULL>(aw, sticky, e, P);
    }
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = aw[j];
    z.e = e;
    z.zero = any ? 0u : 1u;
    z.neg = any ? sign : (xs & ys);
}

// ---- multiplication ----------------------------------------------------
//
// The significands are re-cut into B-bit digits so that a whole product
// column  sum_i A_i * B_(c-i)  accumulates in one 64-bit register with
// v_mad_u64_u32 and no carry handling: a column has at most D terms below
// 2^(2B) each plus a carry-in below 2^(64-B), which stays below 2^64 while
// D < 2^(64-2B).  B = 29 up to 63 digits (L <= 57); B = 28 beyond (74
// digits at L = 64, where 256 would fit).  One mad per digit product; the
// carry moves on once per column.

constexpr int kDigitBits = 29;
constexpr uint32_t kDigitMask = (1u << kDigitBits) - 1u;

template <int L>
struct Digits {
    static constexpr int B = 32 * L <= 63 * kDigitBits ? kDigitBits : 28;
    static constexpr uint32_t MASK = (1u << B) - 1u;
    static constexpr int D = (32 * L + B - 1) / B;
    static_assert(D < (1 << (64 - 2 * B)), "column sums would overflow 64 bits");
};

// MS-first 32-bit limbs -> LS-first B-bit digits
template <int L>
MPFFT_FN void to_digits(const uint32_t (&m)[L], uint32_t (&d)[Digits<L>::D])
{
    constexpr int B = Digits<L>::B;
#pragma unroll
    for (int k = 0; k < Digits<L>::D; k++) {
        const int bit = B * k, w = bit >> 5, s = bit & 31;
        const uint32_t lo = w < L ? m[L - 1 - w] : 0u;
        const uint32_t hi = w + 1 < L ? m[L - 2 - w] : 0u;
        d[k] = (s == 0 ? lo : fshr(hi, lo, (uint32_t)s)) & Digits<L>::MASK;
    }
}

// Deposit digit c (value v, bits [Bc, Bc+B) of the 64L-bit product) into
// the window w[0..L] (product bits [32(L-1), 64L), MS-first) or into sticky.
template <int L>
MPFFT_FN void put_digit(uint32_t (&w)[L + 1], uint32_t &sticky, int c, uint32_t v)
{
    constexpr int B = Digits<L>::B;
    const int bit = B * c, j = bit >> 5, s = bit & 31;
    // low part -> product word j, high part -> word j + 1
    const uint32_t lo = v << s;
    if (j < L - 1) sticky |= lo;
    else if (j < 2 * L) w[2 * L - 1 - j] |= lo;
    if (s + B > 32) {
        const uint32_t hi = v >> (32 - s);
        if (j + 1 < L - 1) sticky |= hi;
        else if (j + 1 < 2 * L) w[2 * L - 2 - j] |= hi;
    }
}

// Exact product: every column, then one rounding step.
template <int L, bool FULL>
MPFFT_FN void mp_mul_exact(Mp<L> &z, const uint32_t (&A)[Digits<L>::D],
                           const uint32_t (&B)[Digits<L>::D], int32_t e, uint32_t sign, int P)
{
    constexpr int D = Digits<L>::D;
    uint32_t w[L + 1];
#pragma unroll
    for (int j = 0; j <= L; j++) w[j] = 0;
    uint32_t sticky = 0;
    uint64_t acc = 0;
#pragma unroll
    for (int c = 0; c < 2 * D - 1; c++) {
        const int i0 = c - (D - 1) > 0 ? c - (D - 1) : 0;
        const int i1 = c < D - 1 ? c : D - 1;
#pragma unroll
        for (int i = i0; i <= i1; i++) acc = (uint64_t)A[i] * B[c - i] + acc;
        put_digit<L>(w, sticky, c, (uint32_t)acc & Digits<L>::MASK);
        acc >>= Digits<L>::B;
    }
    put_digit<L>(w, sticky, 2 * D - 1, (uint32_t)acc);
    if (!(w[0] >> 31)) {
#pragma unroll
        for (int j = 0; j < L; j++) w[j] = fshr(w[j], w[j + 1], 31);
        w[L] <<= 1;
        e -= 1;
    }
    round_window<L, FULL>(w, sticky != 0, e, P);
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = w[j];
    z.e = e; z.neg = sign; z.zero = 0;
}

// Truncated product: columns below K0 are skipped.  Their sum is below
// 2^E, E = B*K0 + B + 7: column c < K0 holds at most K0 terms below 2^(2B),
// so the sum is below K0 2^(2B) 2^(B(K0-1)) 2^B / (2^B - 1) < 2^(BK0+B+7)
// for K0 <= 126 (at B = 29, E = 29*K0 + 36).  The exact product therefore
// lies in [T, T + 2^E) for the computed partial sum T.
// Rounding is monotone, hence RN(T) == RN(T + 2^E - 1) proves that RN(T) is
// the correctly rounded product; when the two could differ the caller
// computes the exact product instead.  K0 keeps E at least two bits below
// the window, whose bottom is bit Wb = 32(L-1).  At 32 limbs (36 digits)
// this skips columns 0..31, 528 of the 1296 digit products; at 64 limbs
// (74 digits of 28 bits) columns 0..69, 2485 of 5476.
template <int L>
struct Trunc {
    static constexpr int B = Digits<L>::B;
    static constexpr int K0 = (32 * L - 70) >= B ? (32 * L - 70) / B : 0;
    static constexpr int Wb = 32 * (L - 1);
    static constexpr int E = B * K0 + B + 7;
    static constexpr int J0 = (B * K0) >> 5;                // first word below the window
    static constexpr int NS = (L - 1) - J0;                 // words [J0, L-1) below the window
    static_assert(K0 <= 126, "skipped-column bound assumes K0 <= 126");
};

// (any bit set, all bits set) of the absolute bit range [lo, hi) of the
// words [J0, L-1) held in sub[]
template <int L>
MPFFT_FN void sub_range(const uint32_t (&sub)[Trunc<L>::NS > 0 ? Trunc<L>::NS : 1],
                        int lo, int hi, uint32_t &any, uint32_t &all)
{
    any = 0; all = 1;
#pragma unroll
    for (int k = 0; k < Trunc<L>::NS; k++) {
        const int b0 = 32 * (Trunc<L>::J0 + k);
        const int l = lo > b0 ? lo - b0 : 0, h = hi < b0 + 32 ? hi - b0 : 32;
        if (l >= h) continue;
        const uint32_t m = (h - l == 32) ? 0xffffffffu : (((1u << (h - l)) - 1u) << l);
        any |= (sub[k] & m) != 0;
        all &= (sub[k] & m) == m;
    }
}

template <int L, bool FULL>
MPFFT_FN bool mp_mul_trunc(Mp<L> &z, const uint32_t (&A)[Digits<L>::D],
                           const uint32_t (&B)[Digits<L>::D], int32_t e, uint32_t sign, int P)
{
    constexpr int D = Digits<L>::D, K0 = Trunc<L>::K0, Wb = Trunc<L>::Wb, E = Trunc<L>::E;
    constexpr int J0 = Trunc<L>::J0, NS = Trunc<L>::NS;
    static_assert(E <= Wb - 2, "bound must stay below the window");
    uint32_t w[L + 1], sub[NS > 0 ? NS : 1];
#pragma unroll
    for (int j = 0; j <= L; j++) w[j] = 0;
#pragma unroll
    for (int k = 0; k < (NS > 0 ? NS : 1); k++) sub[k] = 0;
    uint64_t acc = 0;
#pragma unroll
    for (int c = K0; c < 2 * D; c++) {
        uint32_t v;
        if (c < 2 * D - 1) {
            const int i0 = c - (D - 1) > 0 ? c - (D - 1) : 0;
            const int i1 = c < D - 1 ? c : D - 1;
#pragma unroll
            for (int i = i0; i <= i1; i++) acc = (uint64_t)A[i] * B[c - i] + acc;
            v = (uint32_t)acc & Digits<L>::MASK;
            acc >>= Digits<L>::B;
        } else {
            v = (uint32_t)acc;
        }
        // deposit digit c: word j of the product (window above L-2, sub below)
        const int bit = Digits<L>::B * c, j = bit >> 5, sh = bit & 31;
        const uint32_t lo = v << sh;
        if (j >= L - 1) { if (j < 2 * L) w[2 * L - 1 - j] |= lo; }
        else sub[j - J0] |= lo;
        if (sh + Digits<L>::B > 32) {
            const uint32_t hi = v >> (32 - sh);
            if (j + 1 >= L - 1) { if (j + 1 < 2 * L) w[2 * L - 2 - j] |= hi; }
            else sub[j + 1 - J0] |= hi;
        }
    }
    // predicates on T's bits below the window, for a normalization shift of
    // 0 (window bottom Wb) or 1 (window bottom Wb - 1)
    uint32_t nz_low, dummy, any0, all0, any1, all1;
    sub_range<L>(sub, Digits<L>::B * K0, E, nz_low, dummy);
    sub_range<L>(sub, E, Wb, dummy, all0);
    sub_range<L>(sub, E, Wb - 1, dummy, all1);
    sub_range<L>(sub, Digits<L>::B * K0, Wb, any0, dummy);
    sub_range<L>(sub, Digits<L>::B * K0, Wb - 1, any1, dummy);
    const uint32_t s = (w[0] >> 31) ^ 1u;
    if (s) {
#pragma unroll
        for (int j = 0; j < L; j++) w[j] = fshr(w[j], w[j + 1], 31);
        w[L] = (w[L] << 1) | (sub[NS - 1] >> 31);
        e -= 1;
    }
    const uint32_t sticky = s ? any1 : any0;
    // carry into the window when 2^E - 1 is added to T
    const uint32_t c = (s ? all1 : all0) & nz_low;
    // round bit, the window bits below it ("rest"), and the last kept bit
    uint32_t rbit, rest_zero, rest_ones, lsb;
    if (FULL) {
        rbit = w[L] >> 31;
        rest_zero = (w[L] << 1) == 0;
        rest_ones = (w[L] | 0x80000000u) == 0xffffffffu;
        lsb = w[L - 1] & 1u;
    } else {
        const int kl = (P - 1) >> 5, rl = P >> 5;
        const uint32_t kb = 31u - (uint32_t)((P - 1) & 31), rb = 31u - (uint32_t)(P & 31);
        const uint32_t bm = rb == 0 ? 0u : (1u << rb) - 1u;
        rbit = 0; rest_zero = 1; rest_ones = 1; lsb = 0;
#pragma unroll
        for (int j = 0; j <= L; j++) {
            const uint32_t v = w[j];
            if (j == rl) {
                rbit = (v >> rb) & 1u;
                rest_zero &= (v & bm) == 0;
                rest_ones &= (v & bm) == bm;
            }
            if (j > rl) { rest_zero &= v == 0; rest_ones &= v == 0xffffffffu; }
            if (j == kl) lsb = (v >> kb) & 1u;
        }
    }
    const uint32_t amb = c ? rest_ones : (rbit & rest_zero & (sticky ^ 1u) & (lsb ^ 1u));
    if (amb) return false;
    round_window<L, FULL>(w, sticky, e, P);
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = w[j];
    z.e = e; z.neg = sign; z.zero = 0;
    return true;
}

// z = x * y, correctly rounded to P bits.
template <int L, bool FULL>
MPFFT_FN void mp_mul(Mp<L> &z, const Mp<L> &x, const Mp<L> &y, int P)
{
    constexpr int D = Digits<L>::D;
    const uint32_t sign = x.neg ^ y.neg;
    if (x.zero | y.zero) { set_zero(z, sign); return; }
    uint32_t A[D], B[D];
    to_digits<L>(x.m, A);
    to_digits<L>(y.m, B);
    const int32_t e = x.e + y.e;
    if constexpr (Trunc<L>::K0 > 0) {
        if (mp_mul_trunc<L, FULL>(z, A, B, e, sign, P)) return;
#ifdef MPFFT_CPU_BUILD
        mp_mul_fallbacks++;
#endif
    }
    mp_mul_exact<L, FULL>(z, A, B, e, sign, P);
}

// Reference schoolbook product on 32-bit limbs (kept for cross-checks).
template <int L, bool FULL>
MPFFT_FN void mp_mul_limb32(Mp<L> &z, const Mp<L> &x, const Mp<L> &y, int P)
{
    const uint32_t sign = x.neg ^ y.neg;
    if (x.zero | y.zero) { set_zero(z, sign); return; }
    uint32_t w[L + 1];
    uint32_t sticky = 0;
    uint64_t acc = 0;
#pragma unroll
    for (int c = 0; c < 2 * L - 1; c++) {
        uint32_t ov = 0;
        const int i0 = c - (L - 1) > 0 ? c - (L - 1) : 0;
        const int i1 = c < L - 1 ? c : L - 1;
#pragma unroll
        for (int i = i0; i <= i1; i++) {
            const uint64_t p = (uint64_t)x.m[L - 1 - i] * (uint64_t)y.m[L - 1 - (c - i)];
            acc += p;
            ov += acc < p;
        }
        const uint32_t limb = (uint32_t)acc;
        if (c < L - 1) sticky |= limb;
        else w[2 * L - 1 - c] = limb;
        acc = (acc >> 32) | ((uint64_t)ov << 32);
    }
    w[0] = (uint32_t)acc;
    int32_t e = x.e + y.e;
    if (!(w[0] >> 31)) {
#pragma unroll
        for (int j = 0; j < L; j++) w[j] = fshr(w[j], w[j + 1], 31);
        w[L] <<= 1;
        e -= 1;
    }
    round_window<L, FULL>(w, sticky != 0, e, P);
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = w[j];
    z.e = e; z.neg = sign; z.zero = 0;
}

// o = c ? a : b, element-wise (keeps the arrays in registers)
template <int L>
MPFFT_FN void mp_select(Mp<L> &o, uint32_t c, const Mp<L> &a, const Mp<L> &b)
{
#pragma unroll
    for (int j = 0; j < L; j++) o.m[j] = c ? a.m[j] : b.m[j];
    o.e = c ? a.e : b.e;
    o.neg = c ? a.neg : b.neg;
    o.zero = c ? a.zero : b.zero;
}

// ---- SoA load/store: buffer = hdr int32[M] followed by limb uint32[L][M] ----

template <int L>
MPFFT_FN void mp_load(Mp<L> &x, const uint32_t *__restrict__ buf,
                      uint32_t M, uint32_t i)
{
    const int32_t h = (int32_t)buf[i];
    x.neg = (uint32_t)h & 1u;
    x.zero = (h & ~1) == kHdrZero;
    x.e = h >> 1;
    // limb j of every element starts at the uniform address buf + (j+1)*M;
    // the per-lane part is the 32-bit index i (saddr + voffset addressing)
#pragma unroll
    for (int j = 0; j < L; j++) x.m[j] = (buf + (size_t)(j + 1) * M)[i];
}

template <int L>
MPFFT_FN void mp_store(const Mp<L> &x, uint32_t *__restrict__ buf,
                       uint32_t M, uint32_t i)
{
    buf[i] = x.zero ? ((uint32_t)kHdrZero | x.neg) : (((uint32_t)x.e << 1) | x.neg);
#pragma unroll
    for (int j = 0; j < L; j++) (buf + (size_t)(j + 1) * M)[i] = x.zero ? 0u : x.m[j];
}

#ifndef MPFFT_CPU_BUILD
// Values parked in LDS between uses (keeps them out of the registers while a
// multiplication runs).  Slot-major layout: word j of slot k of thread t is
// at ((k * (L+1) + j) * nthreads + t), so a wave touches consecutive words.
template <int L>
MPFFT_FN void lds_put(uint32_t *lds, uint32_t nthr, uint32_t t, int k, const Mp<L> &x)
{
    uint32_t *p = lds + (uint32_t)k * (L + 1) * nthr + t;
    p[0] = x.zero ? ((uint32_t)kHdrZero | x.neg) : (((uint32_t)x.e << 1) | x.neg);
#pragma unroll
    for (int j = 0; j < L; j++) p[(1 + j) * nthr] = x.m[j];
}

template <int L>
MPFFT_FN void lds_get(Mp<L> &x, const uint32_t *lds, uint32_t nthr, uint32_t t, int k)
{
    const uint32_t *p = lds + (uint32_t)k * (L + 1) * nthr + t;
    const int32_t h = (int32_t)p[0];
    x.neg = (uint32_t)h & 1u;
    x.zero = (h & ~1) == kHdrZero;
    x.e = h >> 1;
#pragma unroll
    for (int j = 0; j < L; j++) x.m[j] = x.zero ? 0u : p[(1 + j) * nthr];
}

// element-major ("AoS") form: element i occupies W = round_up(L+1, 4) words
// starting at buf + i*W (hdr, limb 0, ..., limb L-1, padding), so a gather
// of one element reads whole 16-byte chunks
template <int L>
struct Aos {
    static constexpr int W = (L + 1 + 3) & ~3;
};

template <int L>
MPFFT_FN void mp_load_aos(Mp<L> &x, const uint32_t *__restrict__ buf, uint32_t i)
{
    constexpr int W = Aos<L>::W;
    uint32_t t[W];
#ifdef MPFFT_CPU_BUILD
    for (int k = 0; k < W; k++) t[k] = buf[(size_t)i * W + k];
#else
    const uint4 *p = reinterpret_cast<const uint4 *>(buf + (size_t)i * W);
#pragma unroll
    for (int k = 0; k < W / 4; k++) {
        const uint4 v = p[k];
        t[4 * k] = v.x; t[4 * k + 1] = v.y; t[4 * k + 2] = v.z; t[4 * k + 3] = v.w;
    }
#endif
    const int32_t h = (int32_t)t[0];
    x.neg = (uint32_t)h & 1u;
    x.zero = (h & ~1) == kHdrZero;
    x.e = h >> 1;
#pragma unroll
    for (int j = 0; j < L; j++) x.m[j] = t[1 + j];
}

// exchange a number with the partner lane (lane ^ 1) of the same wave
MPFFT_FN uint32_t swap_pair(uint32_t v)
{
    // DPP quad_perm [1,0,3,2]
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}

template <int L>
MPFFT_FN void mp_swap_pair(Mp<L> &o, const Mp<L> &x)
{
#pragma unroll
    for (int j = 0; j < L; j++) o.m[j] = swap_pair(x.m[j]);
    o.e = (int32_t)swap_pair((uint32_t)x.e);
    o.neg = swap_pair(x.neg);
    o.zero = swap_pair(x.zero);
}
#endif

}  // namespace mpfft
