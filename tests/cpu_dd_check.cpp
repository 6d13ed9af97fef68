// CPU check of the double-word conversions (mpfft_amd/csrc/dd_conv.h built
// with MPFFT_CPU_BUILD) against MPFR:
//  - dd_round_mp(a, b, P) equals mpfr_add of the two binary64 numbers at
//    P bits, MPFR_RNDN, on structured random pairs (normalised double words
//    |b| <= ulp(a)/2, ties at bit P, carries, cancellation down to exact
//    zero, gaps beyond the 128-bit window, subnormals, signed zeros), at
//    every limb count of the form (L = 2, 3, 4) and P from 53 to 106;
//  - dd_from_mp then dd_round_mp gives every random P-bit value back bit for
//    bit (the split is exact).
//
// Usage: cpu_dd_check <trials-per-case> <seed>; prints PASS or the first
// mismatches.
#define MPFFT_CPU_BUILD 1
#include "../mpfft_amd/csrc/dd_conv.h"
#include <gmp.h>
#include <mpfr.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

using namespace mpfft;

static uint64_t rng_state;
static uint64_t rnd64()
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static int rndint(int lo, int hi) { return lo + (int)(rnd64() % (uint64_t)(hi - lo + 1)); }

template <int L>
static void from_mpfr(Mp<L> &x, const mpfr_t r)
{
    x.neg = mpfr_signbit(r) ? 1 : 0;
    if (mpfr_zero_p(r)) {
        x.zero = 1;
        x.e = 0;
        for (int j = 0; j < L; j++) x.m[j] = 0;
        return;
    }
    x.zero = 0;
    x.e = (int32_t)mpfr_get_exp(r);
    mpz_t z;
    mpz_init(z);
    const mpfr_exp_t ex = mpfr_get_z_2exp(z, r);
    mpz_abs(z, z);
    const long sh = ex - ((long)x.e - 32 * L);
    if (sh >= 0) mpz_mul_2exp(z, z, sh);
    else mpz_tdiv_q_2exp(z, z, -sh);
    for (int j = L - 1; j >= 0; j--) {
        x.m[j] = (uint32_t)mpz_get_ui(z);
        mpz_tdiv_q_2exp(z, z, 32);
    }
    mpz_clear(z);
}

template <int L>
static bool same(const Mp<L> &x, const Mp<L> &y)
{
    if (x.zero != y.zero || x.neg != y.neg) return false;
    if (x.zero) return true;
    if (x.e != y.e) return false;
    for (int j = 0; j < L; j++)
        if (x.m[j] != y.m[j]) return false;
    return true;
}

static uint64_t mant53()
{
    switch (rnd64() % 6) {
    case 0: return 1ull << 52;                                  // power of two
    case 1: return (1ull << 53) - 1;                            // all ones
    case 2: return (1ull << 52) | 1u;
    default: return (1ull << 52) | (rnd64() >> 12);
    }
}

static double make(uint64_t m, int e, bool neg)
{
    const double v = std::ldexp((double)m, e - 53);            // 0.m 2^e
    return neg ? -v : v;
}

static long fails = 0, trials = 0;

// one (a, b) pair at precision P
template <int L>
static void check_pair(double a, double b, int P)
{
    trials++;
    Mp<L> z;
    dd_round_mp(z, a, b, P);
    mpfr_t x, y, r;
    mpfr_inits2(53, x, y, (mpfr_ptr)0);
    mpfr_init2(r, P);
    mpfr_set_d(x, a, MPFR_RNDN);
    mpfr_set_d(y, b, MPFR_RNDN);
    mpfr_add(r, x, y, MPFR_RNDN);
    Mp<L> ref;
    from_mpfr(ref, r);
    if (!same(z, ref) && fails++ < 10) {
        printf("MISMATCH L=%d P=%d a=%a b=%a: got zero=%u neg=%u e=%d m0=%08x m1=%08x; "
               "MPFR zero=%u neg=%u e=%d m0=%08x m1=%08x\n", L, P, a, b, z.zero, z.neg, z.e,
               z.m[0], z.m[1], ref.zero, ref.neg, ref.e, ref.m[0], ref.m[1]);
    }
    mpfr_clears(x, y, r, (mpfr_ptr)0);
}

template <int L>
static void pairs(int P, long n)
{
    for (long t = 0; t < n; t++) {
        const int ea = (t % 16 == 0) ? rndint(-1000, 1000) : rndint(-60, 60);
        const double a = make(mant53(), ea, rnd64() & 1);
        const bool nb = rnd64() & 1;
        double b = 0.0;
        switch (t % 8) {
        case 0: case 1:                       // normalised double word: |b| < ulp(a)/2
            b = make(mant53(), ea - 53 - rndint(1, 130), nb);
            break;
        case 2: {                             // |b| = ulp(a)/2 exactly
            b = std::ldexp(nb ? -1.0 : 1.0, ea - 54);
            break;
        }
        case 3: {                             // near a tie at bit P: (2k+1) 2^(ea-P-1) +- tiny
            const double h = std::ldexp((double)(2 * rndint(0, 7) + 1), ea - P - 1);
            const int tw = rndint(0, 2);
            const double tiny = tw == 0 ? 0.0 : std::ldexp(tw == 1 ? 1.0 : -1.0, ea - P - 40);
            b = (nb ? -h : h) + tiny;
            break;
        }
        case 4:                               // any gap, also beyond the window
            b = make(mant53(), ea - rndint(-3, 200), nb);
            break;
        case 5: {                             // cancellation: b close to -a
            const double ulp = std::ldexp(1.0, ea - 53);
            b = -a + (double)rndint(-3, 3) * ulp * (double)(1 + (rnd64() % 4));
            break;
        }
        case 6:                               // subnormal or zero b
            b = (rnd64() & 1) ? std::ldexp((double)(rnd64() >> 20), -1074) : (nb ? -0.0 : 0.0);
            break;
        default:                              // tiny a, or zeros
            if (rnd64() & 1) {
                check_pair<L>(nb ? -0.0 : 0.0, (rnd64() & 1) ? -0.0 : 0.0, P);
                b = make(mant53(), ea - 53 - rndint(1, 60), nb);
            } else {
                b = make(mant53(), ea - rndint(0, 60), nb);
            }
            break;
        }
        check_pair<L>(a, b, P);
        check_pair<L>(b, a, P);
    }
}

// random P-bit value -> dd_from_mp -> dd_round_mp: unchanged
template <int L>
static void split_round_trip(int P, long n)
{
    for (long t = 0; t < n; t++) {
        trials++;
        Mp<L> x;
        for (int j = 0; j < L; j++) x.m[j] = (uint32_t)rnd64();
        switch (t % 4) {
        case 0: for (int j = 0; j < L; j++) x.m[j] = 0xffffffffu; break;
        case 1: for (int j = 1; j < L; j++) x.m[j] = 0; break;
        default: break;
        }
        x.m[0] |= 0x80000000u;
        const int keep = P - 32 * (L - 1);                     // bits of the last limb
        if (keep < 32) x.m[L - 1] &= ~((1u << (32 - keep)) - 1);
        x.e = rndint(-200, 200);
        x.neg = rnd64() & 1;
        x.zero = 0;
        DD d;
        const int32_t eoff = rndint(-5, 5);
        x.e += eoff;
        if (dd_from_mp(d, x, 900, eoff)) {
            if (fails++ < 10) printf("range flag at e=%d\n", x.e);
            continue;
        }
        Mp<L> z;
        dd_to_mp(z, d, P, eoff);
        if (!same(z, x) && fails++ < 10)
            printf("ROUND TRIP L=%d P=%d e=%d m0=%08x: hi=%a lo=%a -> e=%d m0=%08x\n", L, P,
                   x.e, x.m[0], d.hi, d.lo, z.e, z.m[0]);
    }
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 20000;
    rng_state = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
    for (int P : {53, 54, 60, 63, 64}) {
        pairs<2>(P, n);
        split_round_trip<2>(P, n / 4);
    }
    for (int P : {65, 80, 95, 96}) {
        pairs<3>(P, n);
        split_round_trip<3>(P, n / 4);
    }
    for (int P : {97, 100, 105, 106}) {
        pairs<4>(P, n);
        split_round_trip<4>(P, n / 4);
    }
    printf("%ld checks, %ld mismatches\n", trials, fails);
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
