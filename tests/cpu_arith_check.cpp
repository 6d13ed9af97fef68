// CPU check of the device limb arithmetic (mpfft_amd/csrc/mp_arith.h built
// with MPFFT_CPU_BUILD) against MPFR, the library the reference calls
// (mpfr_add / mpfr_sub / mpfr_mul with MPFR_RNDN, serial_mp_fft3.c:141-248).
//
// Usage: cpu_arith_check <trials-per-case> <seed>
// Prints one line per (L, P, op) case and exits non-zero on any mismatch,
// printing the first offending operands in hex.
#define MPFFT_CPU_BUILD 1
#include "../mpfft_amd/csrc/mp_arith.h"
#include <gmp.h>
#include <mpfr.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>

using namespace mpfft;

static uint64_t rng_state;
static uint64_t rnd64()
{
    // splitmix64
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t rnd32() { return (uint32_t)(rnd64() >> 32); }

template <int L>
static void to_mpfr(mpfr_t r, const Mp<L> &x)
{
    if (x.zero) { mpfr_set_zero(r, x.neg ? -1 : 1); return; }
    mpz_t z; mpz_init(z);
    for (int j = 0; j < L; j++) { mpz_mul_2exp(z, z, 32); mpz_add_ui(z, z, x.m[j]); }
    int t = mpfr_set_z_2exp(r, z, (mpfr_exp_t)x.e - 32 * L, MPFR_RNDN);
    if (t) { fprintf(stderr, "to_mpfr inexact\n"); exit(3); }
    if (x.neg) mpfr_neg(r, r, MPFR_RNDN);
    mpz_clear(z);
}

template <int L>
static void from_mpfr(Mp<L> &x, const mpfr_t r)
{
    x.neg = mpfr_signbit(r) ? 1 : 0;
    if (mpfr_zero_p(r)) { x.zero = 1; x.e = 0; for (int j = 0; j < L; j++) x.m[j] = 0; return; }
    x.zero = 0;
    x.e = (int32_t)mpfr_get_exp(r);
    mpz_t z; mpz_init(z);
    mpfr_exp_t ex = mpfr_get_z_2exp(z, r);
    mpz_abs(z, z);
    long sh = ex - ((long)x.e - 32 * L);
    if (sh >= 0) mpz_mul_2exp(z, z, sh); else mpz_tdiv_q_2exp(z, z, -sh);
    for (int j = L - 1; j >= 0; j--) { x.m[j] = (uint32_t)mpz_get_ui(z); mpz_tdiv_q_2exp(z, z, 32); }
    mpz_clear(z);
}

// random P-bit significand with structured patterns
template <int L>
static void rand_mant(Mp<L> &x, int P)
{
    int kind = rnd32() % 8;
    for (int j = 0; j < L; j++) {
        switch (kind) {
        case 0: x.m[j] = 0; break;                 // 1000...0
        case 1: x.m[j] = 0xffffffffu; break;       // 1111...1
        case 2: x.m[j] = (rnd32() % 4 == 0) ? rnd32() : 0; break;   // sparse
        case 3: x.m[j] = (rnd32() % 4 == 0) ? rnd32() : 0xffffffffu; break;
        default: x.m[j] = rnd32(); break;
        }
    }
    x.m[0] |= 0x80000000u;
    // clear bits below position P (the padding may span several limbs)
    for (int j = 0; j < L; j++) {
        const int keep = P - 32 * j;                 // bits of limb j that are kept
        if (keep <= 0) x.m[j] = 0;
        else if (keep < 32) x.m[j] &= ~((1u << (32 - keep)) - 1u);
    }
    if (kind == 0 && (rnd32() & 1)) {                // 1000...01
        const int b = P - 1;
        x.m[b >> 5] |= 0x80000000u >> (b & 31);
    }
}

template <int L>
static void rand_pair(Mp<L> &x, Mp<L> &y, int P)
{
    rand_mant<L>(x, P);
    x.e = (int32_t)(rnd32() % 64) - 32;
    x.neg = rnd32() & 1; x.zero = 0;
    int mode = rnd32() % 10;
    y.neg = rnd32() & 1; y.zero = 0;
    if (mode == 0) {              // y close to x: cancellation
        y = x; y.neg = rnd32() & 1;
        int nb = 1 + rnd32() % 3;
        for (int t = 0; t < nb; t++) {
            int bit = rnd32() % P;     // flip a bit (not the leading one)
            if (bit == 0) continue;
            y.m[bit / 32] ^= 0x80000000u >> (bit % 32);
        }
    } else if (mode == 1) {       // exact tie candidates: y = 2^(e-P-1) * small
        for (int j = 0; j < L; j++) y.m[j] = 0;
        y.m[0] = 0x80000000u | ((rnd32() & 1) ? 0x40000000u : 0);
        if (P == 1) y.m[0] = 0x80000000u;
        y.e = x.e - P - (int32_t)(rnd32() % 3);
    } else {
        rand_mant<L>(y, P);
        static const int gaps[] = {0, 1, 2, 3, 31, 32, 33, 63, 64, 65};
        int g;
        int gm = rnd32() % 4;
        if (gm == 0) g = gaps[rnd32() % 10];
        else if (gm == 1) g = P - 2 + (int)(rnd32() % 5);
        else if (gm == 2) g = 32 * (L + 1) - 2 + (int)(rnd32() % 5);
        else g = rnd32() % (32 * L + 80);
        if (rnd32() & 1) g = -g;
        y.e = x.e - g;
    }
    int zm = rnd32() % 40;
    if (zm == 0) { x.zero = 1; for (int j = 0; j < L; j++) x.m[j] = 0; }
    if (zm == 1) { y.zero = 1; for (int j = 0; j < L; j++) y.m[j] = 0; }
    if (zm == 2) { x.zero = y.zero = 1; for (int j = 0; j < L; j++) x.m[j] = y.m[j] = 0; }
    if (rnd32() % 16 == 0) { std::swap(x, y); }
}

template <int L>
static bool same(const Mp<L> &a, const Mp<L> &b)
{
    if (a.zero != b.zero || a.neg != b.neg) return false;
    if (a.zero) return true;
    if (a.e != b.e) return false;
    for (int j = 0; j < L; j++) if (a.m[j] != b.m[j]) return false;
    return true;
}

template <int L>
static void dump(const char *tag, const Mp<L> &a)
{
    fprintf(stderr, "%s: zero=%u neg=%u e=%d m=", tag, a.zero, a.neg, a.e);
    for (int j = 0; j < L; j++) fprintf(stderr, "%08x", a.m[j]);
    fprintf(stderr, "\n");
}

template <int L, bool FULL>
static int run_case(int P, int op, long trials)
{
    mpfr_t a, b, c;
    mpfr_inits2(P, a, b, c, (mpfr_ptr)0);
    long bad = 0;
    for (long t = 0; t < trials; t++) {
        Mp<L> x, y, z, ref;
        rand_pair<L>(x, y, P);
        uint32_t fx = rnd32() & 1, fy = rnd32() & 1;
        to_mpfr<L>(a, x); to_mpfr<L>(b, y);
        if (op == 0) {
            if (fx) mpfr_neg(a, a, MPFR_RNDN);
            if (fy) mpfr_neg(b, b, MPFR_RNDN);
            mpfr_add(c, a, b, MPFR_RNDN);
            mp_add<L, FULL>(z, x, fx, y, fy, P);
        } else {
            mpfr_mul(c, a, b, MPFR_RNDN);
            mp_mul<L, FULL>(z, x, y, P);
        }
        from_mpfr<L>(ref, c);
        if (!same<L>(z, ref)) {
            if (bad == 0) {
                fprintf(stderr, "MISMATCH L=%d P=%d op=%d fx=%u fy=%u\n", L, P, op, fx, fy);
                dump<L>("x  ", x); dump<L>("y  ", y); dump<L>("got", z); dump<L>("ref", ref);
            }
            bad++;
        }
    }
    mpfr_clears(a, b, c, (mpfr_ptr)0);
    printf("L=%2d P=%4d %s trials=%ld mismatches=%ld fallbacks=%lu\n", L, P, op ? "mul" : "add",
           trials, bad, op ? mp_mul_fallbacks : 0ul);
    mp_mul_fallbacks = 0;
    return bad != 0;
}

template <int L>
static int run_L(long trials)
{
    int fail = 0;
    int Ps[5] = {32 * L, 32 * (L - 1) + 1, 32 * L - 11, 32 * L - 1, 32 * (L - 2) + 7};
    for (int pi = 0; pi < 5; pi++) {
        int P = Ps[pi];
        if (P < 2 || (pi > 0 && P == Ps[0])) continue;
        for (int op = 0; op < 2; op++) {
            if (P == 32 * L) fail |= run_case<L, true>(P, op, trials);
            else fail |= run_case<L, false>(P, op, trials);
        }
    }
    return fail;
}

int main(int argc, char **argv)
{
    long trials = argc > 1 ? atol(argv[1]) : 20000;
    rng_state = argc > 2 ? strtoull(argv[2], 0, 0) : 12345;
    int fail = 0;
    fail |= run_L<1>(trials);
    fail |= run_L<2>(trials);
    fail |= run_L<3>(trials);
    fail |= run_L<4>(trials);
    fail |= run_L<6>(trials);
    fail |= run_L<8>(trials);
    fail |= run_L<12>(trials / 2 + 1);
    fail |= run_L<16>(trials / 2 + 1);
    fail |= run_L<24>(trials / 4 + 1);
    fail |= run_L<32>(trials / 4 + 1);
    fail |= run_L<64>(trials / 8 + 1);
    printf(fail ? "FAIL\n" : "PASS\n");
    return fail;
}
