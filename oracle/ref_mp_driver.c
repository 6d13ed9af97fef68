/*
 * Build driver for the reference's multiprecision kernel (fixture source).
 * TEST INFRASTRUCTURE ONLY.
 *
 * oracle/Makefile compiles this file with -DREF_KERNEL naming one of the
 * reference's own kernel files where it lies (lab/serial/mp/serial_mp_fft1.c
 * to serial_mp_fft4.c or lab/parallel/mp/parallel_mp_fft2.c; REF_4M2A marks
 * the 4m2a codes 1 and 2, whose lut() keeps only LUTr and LUTi) and -I pointing at
 * the reference root for mpFFT.h.  The reference is configured through
 * compile-time macros (mpFFT.c:12-41); here N and P are bound to variables so
 * one library serves every size.  Output goes to oracle/_ref/ only.
 *
 * mpfr_get_str_ndigits is an MPFR >= 4.1 function used by the reference's
 * code_info(); the runtime library here is libmpfr 4.1.0 but the only
 * installed mpfr.h is 4.0.2, so its prototype is declared below.
 */
#include "mpFFT.h"
#include <stdint.h>
#include <string.h>
#include "soa_mpfr.h"

size_t mpfr_get_str_ndigits(int b, mpfr_prec_t p);

static unsigned long g_N = 4;
static long g_P = 53;
static int g_in = 2;

#define N g_N
#define P g_P
#define M 1
#define IN g_in
#define R MPFR_RNDN
#define CC 1
#define CH 0
#define Q 64
#define PRA 0
#define PRB 0
#define PRC 0
#ifdef REF_V2
#define FFT mp_sr24_cp_fft_v2(N, 0, 1, 0, inr, ini, outr, outi);
#else
#define FFT mp_sr24_cp_fft(N, 1, 0, inr, ini, outr, outi);
#endif

#include REF_KERNEL

static void lut_clear(void)
{
    for (unsigned long k = 0; k < N / 4; k++) {
#ifdef REF_4M2A
        mpfr_clears(LUTr[k], LUTi[k], (mpfr_ptr)0);
#else
        mpfr_clears(LUTr[k], LUTi[k], LUTs[k], LUTd[k], (mpfr_ptr)0);
#endif
    }
}

/* Run the reference forward FFT once.
 * in_mode 0: input taken from (hdr_in, limb_in) SoA at P bits;
 * in_mode 2: the reference's own ic() with IN=2, written back to
 *            (hdr_in, limb_in).
 * Output SoA at P bits; cnt[0..3] = cmul, cadd, cfc, crb. */
int ref_mp_run(unsigned long n, long p, int in_mode,
               int32_t *hdr_in, uint32_t *limb_in,
               int32_t *hdr_out, uint32_t *limb_out, unsigned long *cnt)
{
    long L = (p + 31) / 32, MM = 2 * (long)n;
    g_N = n; g_P = p; g_in = in_mode;
    create_memory();
    lut();
    int rc = 0;
    if (in_mode == 2) {
        ic();
        for (unsigned long k = 0; k < n; k++) {
            soa_put(inr[k], hdr_in, limb_in, L, MM, (long)k);
            soa_put(ini[k], hdr_in, limb_in, L, MM, (long)(n + k));
        }
    } else {
        for (unsigned long k = 0; k < n && !rc; k++) {
            rc |= soa_get(inr[k], hdr_in, limb_in, L, MM, (long)k);
            rc |= soa_get(ini[k], hdr_in, limb_in, L, MM, (long)(n + k));
        }
    }
    cmul = cadd = cfc = crb = 0;
    if (!rc) {
#if defined(REF_PAR) && REF_PAR
        #pragma omp parallel num_threads(4)
        {
            #pragma omp single
            { FFT }
        }
#else
        FFT
#endif
        for (unsigned long k = 0; k < n; k++) {
            soa_put(outr[k], hdr_out, limb_out, L, MM, (long)k);
            soa_put(outi[k], hdr_out, limb_out, L, MM, (long)(n + k));
        }
    }
    cnt[0] = cmul; cnt[1] = cadd; cnt[2] = cfc; cnt[3] = crb;
    for (unsigned long k = 0; k < n; k++)
        mpfr_clears(inr[k], ini[k], outr[k], outi[k], (mpfr_ptr)0);
    lut_clear();
    clear_memory();
    mpfr_free_cache();
    return rc ? -2 : 0;
}

/* The reference LUT (lut(), 4 arrays of N/4 as SoA: r, i, s, d). */
int ref_mp_lut(unsigned long n, long p, int32_t *hdr, uint32_t *limb)
{
    long L = (p + 31) / 32, q = (long)n / 4;
    g_N = n; g_P = p;
    create_memory();
    lut();
    for (long k = 0; k < q; k++) {
        soa_put(LUTr[k], hdr, limb, L, (long)n, k);
        soa_put(LUTi[k], hdr, limb, L, (long)n, q + k);
#ifndef REF_4M2A
        soa_put(LUTs[k], hdr, limb, L, (long)n, 2 * q + k);
        soa_put(LUTd[k], hdr, limb, L, (long)n, 3 * q + k);
#endif
    }
    for (unsigned long k = 0; k < n; k++)
        mpfr_clears(inr[k], ini[k], outr[k], outi[k], (mpfr_ptr)0);
    lut_clear();
    clear_memory();
    mpfr_free_cache();
    return 0;
}

/* Mean seconds per reference FFT call over `reps` calls on the SoA input
 * (memory, LUT and conversion excluded, as the reference times only the
 * FFT loop, mpFFT.c:84-91). */
double ref_mp_time(unsigned long n, long p, const int32_t *hdr_in, const uint32_t *limb_in,
                   int reps)
{
    long L = (p + 31) / 32, MM = 2 * (long)n;
    g_N = n; g_P = p; g_in = 0;
    create_memory();
    lut();
    for (unsigned long k = 0; k < n; k++) {
        soa_get(inr[k], hdr_in, limb_in, L, MM, (long)k);
        soa_get(ini[k], hdr_in, limb_in, L, MM, (long)(n + k));
    }
    unsigned long long t0 = get_time_int();
    for (int r = 0; r < reps; r++) { FFT }
    double dt = 1e-9 * (double)(get_time_int() - t0) / reps;
    for (unsigned long k = 0; k < n; k++)
        mpfr_clears(inr[k], ini[k], outr[k], outi[k], (mpfr_ptr)0);
    lut_clear();
    clear_memory();
    mpfr_free_cache();
    return dt;
}
