/*
 * SoA-limb <-> mpfr_t conversion shared by the oracle and the reference
 * build drivers.  TEST INFRASTRUCTURE ONLY (see mpfft_oracle.c).
 *
 * An array of M reals at L = ceil(P/32) limbs is hdr int32[M] + limb
 * uint32[L*M]; limb j of element i is limb[j*M+i], j=0 most significant.
 * hdr = (exp<<1)|neg for a nonzero value 0.1b..b * 2^exp, INT32_MIN|neg for
 * a signed zero.
 */
#ifndef MPFFT_SOA_MPFR_H
#define MPFFT_SOA_MPFR_H
#include <stdint.h>
#include <gmp.h>
#include <mpfr.h>

#define ORC_HDR_ZERO ((int32_t)INT32_MIN)

/* ------------------------------------------------------------------ */

/* Load element i of an SoA array (L limbs) into x (any precision >= P of
 * the data: the value is exact).  Returns 0, or -1 for a malformed hdr. */
static inline int soa_get(mpfr_t x, const int32_t *hdr, const uint32_t *limb,
                   long L, long M, long i)
{
    int32_t h = hdr[i];
    int neg = h & 1;
    if ((h & ~1) == ORC_HDR_ZERO) {
        mpfr_set_zero(x, neg ? -1 : +1);
        return 0;
    }
    long e = (long)(h >> 1);
    /* build the integer significand sum_j limb[j] * 2^(32*(L-1-j)) */
    mpz_t z;
    mpz_init2(z, 32 * L);
    mpz_set_ui(z, 0);
    for (long j = 0; j < L; j++) {
        mpz_mul_2exp(z, z, 32);
        mpz_add_ui(z, z, limb[j * M + i]);
    }
    if (!(limb[i] & 0x80000000u)) { mpz_clear(z); return -1; }
    /* value = z * 2^(e - 32L) */
    int t = mpfr_set_z_2exp(x, z, e - 32 * L, MPFR_RNDN);
    mpz_clear(z);
    if (t != 0) return -1;            /* data wider than x: not allowed */
    if (neg) mpfr_neg(x, x, MPFR_RNDN);
    return 0;
}

/* Store x into element i of an SoA array with L limbs.  x must fit in
 * 32L bits (its precision <= 32L). */
static inline void soa_put(const mpfr_t x, int32_t *hdr, uint32_t *limb,
                    long L, long M, long i)
{
    int neg = mpfr_signbit(x) ? 1 : 0;
    if (mpfr_zero_p(x)) {
        hdr[i] = ORC_HDR_ZERO | neg;
        for (long j = 0; j < L; j++) limb[j * M + i] = 0;
        return;
    }
    long e = mpfr_get_exp(x);
    hdr[i] = (int32_t)(((uint32_t)e << 1) | (uint32_t)neg);
    mpz_t z;
    mpz_init(z);
    /* |x| * 2^(32L - e) is an integer < 2^(32L) */
    mpfr_exp_t ex = mpfr_get_z_2exp(z, x);  /* x = z * 2^ex */
    mpz_abs(z, z);
    long sh = ex - (e - 32 * L);
    if (sh >= 0) mpz_mul_2exp(z, z, sh);
    else mpz_tdiv_q_2exp(z, z, -sh);        /* exact: low bits are zero */
    for (long j = L - 1; j >= 0; j--) {
        limb[j * M + i] = (uint32_t)mpz_get_ui(z) & 0xffffffffu;
        mpz_tdiv_q_2exp(z, z, 32);
    }
    mpz_clear(z);
}

#endif
