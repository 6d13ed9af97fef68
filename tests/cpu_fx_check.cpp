// CPU check of the fixed-point arithmetic of the fixed FFT form
// (mpfft_amd/csrc/fx_arith.h built with MPFFT_CPU_BUILD) against exact
// integer arithmetic (GMP) and MPFR rounding:
//   fx_add / fx_sub / fx_addsub  exact two's-complement sums
//   fx_mul                       |result - trunc(x t)| <= 1 unit of 2^-F,
//                                sign of the product
//   fx_from_mp                   trunc(x 2^F) exactly, range flag
//   fx_to_mp                     MPFR's RN rounding of X / 2^F at P bits
//
// Usage: cpu_fx_check <trials-per-case> <seed>; prints PASS on success.
#define MPFFT_CPU_BUILD 1
#include "../mpfft_amd/csrc/fx_arith.h"
#include <gmp.h>
#include <mpfr.h>
#include <cstdio>
#include <cstdlib>

using namespace mpfft;

static uint64_t rng_state;
static uint64_t rnd64()
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t rnd32() { return (uint32_t)(rnd64() >> 32); }

template <int W>
static void to_z(mpz_t z, const Fx<W> &x)     // two's complement -> integer
{
    mpz_set_ui(z, 0);
    for (int j = W - 1; j >= 0; j--) { mpz_mul_2exp(z, z, 32); mpz_add_ui(z, z, x.w[j]); }
    if (x.w[W - 1] >> 31) {
        mpz_t m; mpz_init(m); mpz_ui_pow_ui(m, 2, 32 * W); mpz_sub(z, z, m); mpz_clear(m);
    }
}

template <int W>
static void from_z(Fx<W> &x, const mpz_t v)   // integer -> two's complement (mod 2^32W)
{
    mpz_t z, m; mpz_init_set(z, v); mpz_init(m);
    mpz_ui_pow_ui(m, 2, 32 * W);
    mpz_fdiv_r(z, z, m);
    for (int j = 0; j < W; j++) { x.w[j] = (uint32_t)mpz_get_ui(z); mpz_tdiv_q_2exp(z, z, 32); }
    mpz_clear(z); mpz_clear(m);
}

// random value with |v| < 2^bits (two's complement), varied magnitudes
template <int W>
static void rand_fx(Fx<W> &x, int bits)
{
    mpz_t z; mpz_init(z);
    int nb = 1 + (int)(rnd64() % (uint64_t)bits);
    for (int k = 0; k < (nb + 31) / 32; k++) { mpz_mul_2exp(z, z, 32); mpz_add_ui(z, z, rnd32()); }
    mpz_tdiv_r_2exp(z, z, nb);
    if (rnd64() & 1) mpz_neg(z, z);
    from_z(x, z);
    mpz_clear(z);
}

static int fails = 0;

template <int W>
static void check_case(int trials)
{
    constexpr int F = FxDig<W>::F, DT = FxDig<W>::DT, B = FxDig<W>::B;
    mpz_t a, b, c, e, t;
    mpz_inits(a, b, c, e, t, (mpz_ptr)0);
    long worst = 0, nonexact = 0;
    for (int n = 0; n < trials; n++) {
        Fx<W> x, y, z;
        rand_fx(x, 32 * W - 3);
        rand_fx(y, 32 * W - 3);
        to_z(a, x); to_z(b, y);
        // sums
        fx_add(z, x, y); to_z(c, z); mpz_add(e, a, b);
        if (mpz_cmp(c, e)) { fails++; printf("fx_add mismatch W=%d\n", W); break; }
        fx_sub(z, x, y); to_z(c, z); mpz_sub(e, a, b);
        if (mpz_cmp(c, e)) { fails++; printf("fx_sub mismatch W=%d\n", W); break; }
        const uint32_t s = (uint32_t)(rnd64() & 1);
        fx_addsub(z, x, y, s); to_z(c, z);
        if (s) mpz_sub(e, a, b); else mpz_add(e, a, b);
        if (mpz_cmp(c, e)) { fails++; printf("fx_addsub mismatch W=%d\n", W); break; }
        // product with a twiddle magnitude T < 2^(F+1)
        uint32_t T[DT];
        mpz_set_ui(t, 0);
        const int tb = (n % 7 == 0) ? F + 1 : 1 + (int)(rnd64() % (uint64_t)(F + 1));
        for (int k = DT - 1; k >= 0; k--) {
            T[k] = rnd32() & FxDig<W>::MASK;
            if (B * k >= tb) T[k] = 0;
            else if (B * k + B > tb) T[k] &= (1u << (tb - B * k)) - 1u;
        }
        for (int k = DT - 1; k >= 0; k--) { mpz_mul_2exp(t, t, B); mpz_add_ui(t, t, T[k]); }
        const uint32_t tneg = (uint32_t)(rnd64() & 1);
        // |x| < 2^(32W-3) keeps |x t| / 2^F < 2^(32W-2): no overflow
        Fx<W> xx = x;
        fx_mul(z, xx, T, tneg);
        to_z(c, z);
        mpz_mul(e, a, t);
        if (tneg) mpz_neg(e, e);
        mpz_tdiv_q_2exp(e, e, F);                 // truncation toward zero
        mpz_sub(e, c, e);
        const long d = mpz_get_si(e);
        if (mpz_cmpabs_ui(e, 1) > 0) {
            fails++; printf("fx_mul error beyond one unit W=%d\n", W); break;
        }
        if (d) nonexact++;
        if (labs(d) > worst) worst = labs(d);
    }
    printf("W=%d (%d-bit digits) sums exact, products within %ld unit (%ld of %d not "
           "truncation-exact)\n", W, B, worst, nonexact, trials);
    mpz_clears(a, b, c, e, t, (mpz_ptr)0);
}

// conversions from and to P-bit floats, L limbs, W = L + 2; FULL: P == 32 L
template <int L, bool FULL>
static void check_conv(int trials, int P)
{
    constexpr int W = L + 2, F = FxDig<W>::F;
    mpfr_t r, q;
    mpfr_inits2(P, r, (mpfr_ptr)0);
    mpfr_init2(q, P);
    mpz_t z, e;
    mpz_inits(z, e, (mpz_ptr)0);
    for (int n = 0; n < trials; n++) {
        // float -> fixed: x = +-0.m 2^ex
        Mp<L> x;
        for (int j = 0; j < L; j++) x.m[j] = rnd32();
        x.m[0] |= 0x80000000u;
        if (P < 32 * L) x.m[L - 1] &= ~((1u << (32 * L - P)) - 1u);
        x.e = (int32_t)(rnd64() % 200) - 180 + (n % 3 == 0 ? (int32_t)(rnd64() % 40) - 20 : 0);
        x.neg = (uint32_t)(rnd64() & 1); x.zero = (n % 50 == 0);
        if (x.zero) for (int j = 0; j < L; j++) x.m[j] = 0;
        const int lim = 13;
        // a block exponent: the input carries eoff more, the conversion
        // takes it off again (exact), so the expected integer is unchanged
        const int32_t eoff = (n % 4 == 1) ? (int32_t)(rnd64() % (1u << 29)) - (1 << 28) : 0;
        Mp<L> xs = x;
        xs.e = x.e + eoff;
        Fx<W> f;
        const uint32_t over = fx_from_mp<L, W>(f, xs, lim, eoff);
        if (over != (uint32_t)(!x.zero && x.e > lim)) { fails++; printf("range flag L=%d\n", L); break; }
        if (!over) {
            // expected: trunc(+-M 2^(e - 32L + F))
            mpz_set_ui(e, 0);
            for (int j = 0; j < L; j++) { mpz_mul_2exp(e, e, 32); mpz_add_ui(e, e, x.m[j]); }
            const long sh = (long)x.e - 32L * L + F;
            if (sh >= 0) mpz_mul_2exp(e, e, sh); else mpz_tdiv_q_2exp(e, e, -sh);
            if (x.neg) mpz_neg(e, e);
            to_z(z, f);
            if (mpz_cmp(z, e)) { fails++; printf("fx_from_mp mismatch L=%d\n", L); break; }
        }
        // fixed -> float, against MPFR's rounding of X / 2^F
        Fx<W> g;
        rand_fx(g, 32 * W - 2);
        if (n % 11 == 0) {                       // exact ties: P+1 significant bits, last one set
            mpz_set_ui(e, 1); mpz_mul_2exp(e, e, P); mpz_add_ui(e, e, 1 + 2 * (rnd64() % 2));
            mpz_mul_2exp(e, e, (unsigned long)(rnd64() % 400));
            if (rnd64() & 1) mpz_neg(e, e);
            from_z(g, e);
        }
        Mp<L> y;
        const int32_t eout = (n % 4 == 2) ? (int32_t)(rnd64() % (1u << 29)) - (1 << 28) : 0;
        fx_to_mp<W, L, FULL>(y, g, P, eout);
        y.e -= eout;
        to_z(z, g);
        mpfr_set_z_2exp(q, z, -F, MPFR_RNDN);
        if (mpfr_zero_p(q)) {
            if (!y.zero) { fails++; printf("fx_to_mp zero L=%d\n", L); break; }
            continue;
        }
        mpz_set_ui(e, 0);
        for (int j = 0; j < L; j++) { mpz_mul_2exp(e, e, 32); mpz_add_ui(e, e, y.m[j]); }
        if (y.neg) mpz_neg(e, e);
        mpfr_set_z_2exp(r, e, (mpfr_exp_t)y.e - 32 * L, MPFR_RNDN);
        if (mpfr_cmp(r, q)) { fails++; printf("fx_to_mp mismatch L=%d P=%d\n", L, P); break; }
    }
    printf("L=%d P=%d conversions ok\n", L, P);
    mpfr_clears(r, q, (mpfr_ptr)0);
    mpz_clears(z, e, (mpz_ptr)0);
}

int main(int argc, char **argv)
{
    const int trials = argc > 1 ? atoi(argv[1]) : 2000;
    rng_state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
    check_case<3>(trials);
    check_case<4>(trials);
    check_case<5>(trials);
    check_case<6>(trials);
    check_case<8>(trials);
    check_case<10>(trials);
    check_case<14>(trials);
    check_case<18>(trials);
    check_case<26>(trials);
    check_case<34>(trials);
    check_case<66>(trials);                      // 28-bit digits
    check_conv<1, true>(trials, 32);
    check_conv<2, false>(trials, 53);
    check_conv<8, true>(trials, 256);
    check_conv<16, false>(trials, 500);
    check_conv<32, true>(trials, 1024);
    check_conv<64, true>(trials, 2048);          // the panel kernels' 66-word reals
    check_conv<64, false>(trials, 1500);
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
