/*
 * Build driver for the reference's double-precision conjugate-pair kernels
 * (fixture source).  TEST INFRASTRUCTURE ONLY.
 *
 * oracle/Makefile compiles this file with -DREF_KERNEL naming
 * lab/serial/double/serial_d_fft5.c, serial_d_fft6.c or serial_d_fft7.c
 * where they lie, and -DREF_CODE=5|6|7 selecting the matching entry point
 * (serial_d_main.c:56-64).  N is bound to a variable so one library serves
 * every size.  Output goes to oracle/_ref/ only.
 */
#include "mpFFT.h"
#include <stdint.h>

static unsigned long g_N = 4;
static int g_in = 2;

#define N g_N
#define M 1
#define IN g_in
#define CODE REF_CODE
#define CC 1
#define CH 0
#define PRA 0
#define PRB 0
#if REF_CODE == 7
#define FFT d_sr24_cp_fft_v2(N, 0, 1, 0, inr, ini, outr, outi);
#else
#define FFT d_sr24_cp_fft(N, 1, 0, inr, ini, outr, outi);
#endif

#include REF_KERNEL

/* Run the reference forward FFT once.  in_mode 0: input from xr/xi;
 * in_mode 2: the reference's ic() with IN=2, written back to xr/xi.
 * Also returns its LUT (N/4 cos, N/4 sin) and counters cmul, cadd, cfc, crb. */
int ref_d_run(unsigned long n, int in_mode, double *xr, double *xi,
              double *yr, double *yi, double *lr, double *li, unsigned long *cnt)
{
    g_N = n; g_in = in_mode;
    create_memory();
    lut();
    if (in_mode == 2) {
        ic();
        for (unsigned long k = 0; k < n; k++) { xr[k] = inr[k]; xi[k] = ini[k]; }
    } else {
        for (unsigned long k = 0; k < n; k++) { inr[k] = xr[k]; ini[k] = xi[k]; }
    }
    cmul = cadd = cfc = crb = 0;
    FFT
    for (unsigned long k = 0; k < n; k++) { yr[k] = outr[k]; yi[k] = outi[k]; }
    for (unsigned long k = 0; k < n / 4; k++) { lr[k] = LUTr[k]; li[k] = LUTi[k]; }
    cnt[0] = cmul; cnt[1] = cadd; cnt[2] = cfc; cnt[3] = crb;
    clear_memory();
    return 0;
}
