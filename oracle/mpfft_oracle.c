/*
 * mpfft oracle -- CPU restatement of the mpFFT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by the
 * test-suite, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * never by the product library (mpfft_amd/csrc), which must run on the GPU.
 *
 * What it restates (reference = mfrewer/mpFFT):
 *   - lut():            lab/serial/mp/serial_mp_fft3.c:71-82
 *   - mp_sr24_cp_fft:   lab/serial/mp/serial_mp_fft3.c:121-253 (CODE 3: 3m3a,
 *                       three butterflies, recursion break n<=4); the
 *                       alternative operation orders CODE 2
 *                       (lab/serial/mp/serial_mp_fft2.c:201-232: 4m2a, three
 *                       butterflies) and CODE 1 (serial_mp_fft1.c:181-200:
 *                       4m2a, one butterfly for every k)
 *   - test() round trip lab/serial/mp/serial_mp_fft3.c:346-427 (swap trick)
 *   - ic() IN=2:        lab/serial/mp/serial_mp_fft3.c:104-115
 *   - d_sr24_cp_fft:    lab/serial/double/serial_d_fft5.c:52-62, 93-229 (CODE 5)
 *                       lab/serial/double/serial_d_fft6.c:52-62, 93-429 (CODE 6;
 *                       CODE 7 = same arithmetic, shift-indexed twiddles)
 *
 * The reference recursion is depth-first; here the tree is flattened once
 * into leaves + per-size combine levels and executed breadth-first (the same
 * schedule the GPU uses).  Every output element sees exactly the same
 * sequence of correctly rounded MPFR operations as in the reference, so the
 * results are bit-identical (pinned against the compiled reference's outputs
 * in tests/golden).
 *
 * Exchange format ("SoA limbs"): an array of M reals at L = ceil(P/32) limbs
 * is hdr int32[M] + limb uint32[L*M]; limb j of element i lives at
 * limb[j*M+i], j=0 most significant.  hdr = (exp<<1)|neg for a nonzero value
 * 0.1b..b * 2^exp (MPFR convention), hdr = INT32_MIN|neg for a signed zero.
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <gmp.h>
#include <mpfr.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "soa_mpfr.h"

/* ------------------------------------------------------------------ */
/* Recursion tree flattened into leaves and combine levels              */
/* ------------------------------------------------------------------ */

typedef struct {
    long off;   /* first output index of the node */
    long n;     /* node size */
    long s;     /* input stride */
    long sn;    /* input offset (|sn| < s); element j = x[(sn + j*s) mod N] */
} orc_node;

typedef struct {
    long N, lgN, leaf_max;
    orc_node *leaf; long nleaf;
    long *lvl_off[64]; long lvl_cnt[64];  /* combine nodes of size 2^l */
} orc_tree;

static void tree_walk(orc_tree *t, long n, long s, long sn, long off, int count_only)
{
    /* reference recursion: serial_mp_fft3.c:125-182 (leaves n<=4), and
     * serial_d_fft6.c:97-390 (leaves n<=16) */
    if (n <= t->leaf_max) {
        if (!count_only) {
            orc_node *q = &t->leaf[t->nleaf];
            q->off = off; q->n = n; q->s = s; q->sn = sn;
        }
        t->nleaf++;
        return;
    }
    long l = 0; while ((1L << l) < n) l++;
    if (!count_only) t->lvl_off[l][t->lvl_cnt[l]] = off;
    t->lvl_cnt[l]++;
    tree_walk(t, n / 2, 2 * s, sn, off, count_only);
    tree_walk(t, n / 4, 4 * s, sn + s, off + n / 2, count_only);
    tree_walk(t, n / 4, 4 * s, sn - s, off + 3 * n / 4, count_only);
}

static int tree_build(orc_tree *t, long N, long leaf_max)
{
    memset(t, 0, sizeof(*t));
    t->N = N; t->leaf_max = leaf_max;
    while ((1L << t->lgN) < N) t->lgN++;
    tree_walk(t, N, 1, 0, 0, 1);
    t->leaf = (orc_node *)malloc(sizeof(orc_node) * (size_t)(t->nleaf ? t->nleaf : 1));
    for (int l = 0; l < 64; l++) {
        t->lvl_off[l] = (long *)malloc(sizeof(long) * (size_t)(t->lvl_cnt[l] ? t->lvl_cnt[l] : 1));
        t->lvl_cnt[l] = 0;
    }
    t->nleaf = 0;
    tree_walk(t, N, 1, 0, 0, 0);
    return 0;
}

static void tree_free(orc_tree *t)
{
    free(t->leaf);
    for (int l = 0; l < 64; l++) free(t->lvl_off[l]);
}

static inline long gidx(long N, long sn, long s, long j)
{
    long v = sn + j * s;          /* only j == 0 can be negative */
    return v < 0 ? v + N : v;
}

/* ------------------------------------------------------------------ */
/* LUT: exact replay of lut() (serial_mp_fft3.c:71-82)                  */
/* ------------------------------------------------------------------ */

typedef struct { long N; mpfr_t *r, *i, *s, *d; } orc_lut;

static void lut_build(orc_lut *T, long N, mpfr_prec_t P)
{
    long Q = N / 4;
    T->N = N;
    T->r = (mpfr_t *)malloc(sizeof(mpfr_t) * (size_t)(Q ? Q : 1));
    T->i = (mpfr_t *)malloc(sizeof(mpfr_t) * (size_t)(Q ? Q : 1));
    T->s = (mpfr_t *)malloc(sizeof(mpfr_t) * (size_t)(Q ? Q : 1));
    T->d = (mpfr_t *)malloc(sizeof(mpfr_t) * (size_t)(Q ? Q : 1));
    mpfr_t step;
    mpfr_init2(step, P);
    /* step = RN(RN(pi) / N): N is a power of two, the division is exact */
    mpfr_const_pi(step, MPFR_RNDN);
    mpfr_div_ui(step, step, (unsigned long)N, MPFR_RNDN);
    #pragma omp parallel
    {
        mpfr_t ang;
        mpfr_init2(ang, P);
        #pragma omp for schedule(dynamic, 64)
        for (long k = 0; k < Q; k++) {
            mpfr_inits2(P, T->r[k], T->i[k], T->s[k], T->d[k], (mpfr_ptr)0);
            /* the angle is RN(step * 2k): a rounded argument, as in the reference */
            mpfr_mul_ui(ang, step, (unsigned long)(2 * k), MPFR_RNDN);
            mpfr_cos(T->r[k], ang, MPFR_RNDN);
            mpfr_sin(T->i[k], ang, MPFR_RNDN);
            mpfr_add(T->s[k], T->r[k], T->i[k], MPFR_RNDN);
            mpfr_sub(T->d[k], T->r[k], T->i[k], MPFR_RNDN);
        }
        mpfr_clear(ang);
        mpfr_free_cache();
    }
    mpfr_clear(step);
}

static void lut_free(orc_lut *T)
{
    for (long k = 0; k < T->N / 4; k++)
        mpfr_clears(T->r[k], T->i[k], T->s[k], T->d[k], (mpfr_ptr)0);
    free(T->r); free(T->i); free(T->s); free(T->d);
}

/* ------------------------------------------------------------------ */
/* mp FFT, CODE 3 operation order, breadth-first                        */
/* ------------------------------------------------------------------ */

typedef struct { mpfr_t t, m1r, m1i, m2r, m2i, a1r, a1i, a2r, a2i; } orc_tmp;

static void tmp_init(orc_tmp *w, mpfr_prec_t P)
{
    mpfr_inits2(P, w->t, w->m1r, w->m1i, w->m2r, w->m2i,
                w->a1r, w->a1i, w->a2r, w->a2i, (mpfr_ptr)0);
}
static void tmp_clear(orc_tmp *w)
{
    mpfr_clears(w->t, w->m1r, w->m1i, w->m2r, w->m2i,
                w->a1r, w->a1i, w->a2r, w->a2i, (mpfr_ptr)0);
}

#define RN MPFR_RNDN

/* Leaves of size 1, 2, 4: serial_mp_fft3.c:125-177 */
static void mp_leaf(const orc_node *q, long N, mpfr_t *xr, mpfr_t *xi,
                    mpfr_t *yr, mpfr_t *yi, orc_tmp *w)
{
    long o = q->off;
    long i0 = gidx(N, q->sn, q->s, 0);
    if (q->n == 1) {
        mpfr_set(yr[o], xr[i0], RN);
        mpfr_set(yi[o], xi[i0], RN);
    } else if (q->n == 2) {
        long i1 = gidx(N, q->sn, q->s, 1);
        mpfr_add(yr[o], xr[i0], xr[i1], RN);
        mpfr_sub(yr[o + 1], xr[i0], xr[i1], RN);
        mpfr_add(yi[o], xi[i0], xi[i1], RN);
        mpfr_sub(yi[o + 1], xi[i0], xi[i1], RN);
    } else {
        long i1 = gidx(N, q->sn, q->s, 1), i2 = gidx(N, q->sn, q->s, 2),
             i3 = gidx(N, q->sn, q->s, 3);
        /* even pair -> y[0], y[1]; odd pair -> a1, a2; then the radix-2 merge */
        mpfr_add(yr[o], xr[i0], xr[i2], RN);
        mpfr_sub(yr[o + 1], xr[i0], xr[i2], RN);
        mpfr_add(yi[o], xi[i0], xi[i2], RN);
        mpfr_sub(yi[o + 1], xi[i0], xi[i2], RN);
        mpfr_add(w->a1r, xr[i1], xr[i3], RN);
        mpfr_sub(w->a2r, xr[i1], xr[i3], RN);
        mpfr_add(w->a1i, xi[i1], xi[i3], RN);
        mpfr_sub(w->a2i, xi[i1], xi[i3], RN);
        mpfr_sub(yr[o + 2], yr[o], w->a1r, RN);
        mpfr_add(yr[o], yr[o], w->a1r, RN);
        mpfr_sub(yr[o + 3], yr[o + 1], w->a2i, RN);
        mpfr_add(yr[o + 1], yr[o + 1], w->a2i, RN);
        mpfr_sub(yi[o + 2], yi[o], w->a1i, RN);
        mpfr_add(yi[o], yi[o], w->a1i, RN);
        mpfr_add(yi[o + 3], yi[o + 1], w->a2r, RN);
        mpfr_sub(yi[o + 1], yi[o + 1], w->a2r, RN);
    }
}

/* The 4-point tail shared by all butterflies: serial_mp_fft3.c:185-199 and
 * :234-248.  (m1, m2) are the twiddled odd quarters; results in place. */
static void mp_tail(mpfr_t *yr, mpfr_t *yi, long o0, long q, orc_tmp *w,
                    mpfr_t m1r, mpfr_t m1i, mpfr_t m2r, mpfr_t m2i)
{
    long o1 = o0 + q, o2 = o0 + 2 * q, o3 = o0 + 3 * q;
    mpfr_add(w->a1r, m1r, m2r, RN);
    mpfr_sub(w->a2r, m1r, m2r, RN);
    mpfr_add(w->a1i, m1i, m2i, RN);
    mpfr_sub(w->a2i, m1i, m2i, RN);
    mpfr_sub(yr[o2], yr[o0], w->a1r, RN);
    mpfr_add(yr[o0], yr[o0], w->a1r, RN);
    mpfr_sub(yr[o3], yr[o1], w->a2i, RN);
    mpfr_add(yr[o1], yr[o1], w->a2i, RN);
    mpfr_sub(yi[o2], yi[o0], w->a1i, RN);
    mpfr_add(yi[o0], yi[o0], w->a1i, RN);
    mpfr_add(yi[o3], yi[o1], w->a2r, RN);
    mpfr_sub(yi[o1], yi[o1], w->a2r, RN);
}

/* 4m2a twiddle products (serial_mp_fft2.c:219-230, the same sequence as
 * serial_mp_fft1.c:187-199): m1 = Z+ w, m2 = Z- conj(w), w = LUTr - i LUTi */
static void mp_4m2a(long p, long m, long kn, const orc_lut *T,
                    mpfr_t *yr, mpfr_t *yi, orc_tmp *w)
{
    mpfr_mul(w->m1r, T->r[kn], yr[p], RN);
    mpfr_mul(w->m2r, T->r[kn], yr[m], RN);
    mpfr_mul(w->m1i, T->r[kn], yi[p], RN);
    mpfr_mul(w->m2i, T->r[kn], yi[m], RN);
    mpfr_mul(w->t, T->i[kn], yi[p], RN); mpfr_add(w->m1r, w->m1r, w->t, RN);
    mpfr_mul(w->t, T->i[kn], yi[m], RN); mpfr_sub(w->m2r, w->m2r, w->t, RN);
    mpfr_mul(w->t, T->i[kn], yr[p], RN); mpfr_sub(w->m1i, w->m1i, w->t, RN);
    mpfr_mul(w->t, T->i[kn], yr[m], RN); mpfr_add(w->m2i, w->m2i, w->t, RN);
}

/* One butterfly k of a size-n combine at output offset off, in the
 * operation order of CODE 3 (serial_mp_fft3.c:184-251), CODE 2
 * (serial_mp_fft2.c:184-251) or CODE 1 (serial_mp_fft1.c:181-218). */
static void mp_butterfly(int code, long N, long n, long off, long k, const orc_lut *T,
                         mpfr_t *yr, mpfr_t *yi, orc_tmp *w)
{
    long q = n / 4;
    long p = off + k + 2 * q, m = off + k + 3 * q;   /* Z+ and Z- */
    if (code == 1) {
        /* CODE 1: the general 4m2a butterfly for every k, k = 0 included */
        mp_4m2a(p, m, k * (N / n), T, yr, yi, w);
    } else if (k == 0) {
        /* k = 0: no twiddle (:185-199); m1 = Z+, m2 = Z- */
        mpfr_set(w->m1r, yr[p], RN); mpfr_set(w->m1i, yi[p], RN);
        mpfr_set(w->m2r, yr[m], RN); mpfr_set(w->m2i, yi[m], RN);
    } else {
        long kn = k * (N / n);
        if (k == n / 8) {
            /* w^k = (1 - i) * c with c = LUTr[N/8] (:209-217) */
            mpfr_add(w->m1r, yr[p], yi[p], RN); mpfr_mul(w->m1r, w->m1r, T->r[kn], RN);
            mpfr_sub(w->m1i, yi[p], yr[p], RN); mpfr_mul(w->m1i, w->m1i, T->r[kn], RN);
            mpfr_sub(w->m2r, yr[m], yi[m], RN); mpfr_mul(w->m2r, w->m2r, T->r[kn], RN);
            mpfr_add(w->m2i, yr[m], yi[m], RN); mpfr_mul(w->m2i, w->m2i, T->r[kn], RN);
        } else if (code == 2) {
            mp_4m2a(p, m, kn, T, yr, yi, w);
        } else {
            /* 3m3a (:220-229): Z+ * conj-free twiddle, Z- * its conjugate */
            mpfr_mul(w->m1r, T->d[kn], yi[p], RN);
            mpfr_mul(w->m2r, T->d[kn], yi[m], RN);
            mpfr_mul(w->m1i, T->s[kn], yr[p], RN);
            mpfr_mul(w->m2i, T->s[kn], yr[m], RN);
            mpfr_add(w->t, yr[p], yi[p], RN); mpfr_mul(w->t, w->t, T->r[kn], RN);
            mpfr_sub(w->m1r, w->t, w->m1r, RN); mpfr_sub(w->m1i, w->t, w->m1i, RN);
            mpfr_sub(w->t, yr[m], yi[m], RN); mpfr_mul(w->t, w->t, T->r[kn], RN);
            mpfr_add(w->m2r, w->t, w->m2r, RN); mpfr_sub(w->m2i, w->m2i, w->t, RN);
        }
    }
    mp_tail(yr, yi, off + k, q, w, w->m1r, w->m1i, w->m2r, w->m2i);
}

static void mp_fft_run_code(int code, const orc_tree *t, const orc_lut *T, mpfr_prec_t P,
                            mpfr_t *xr, mpfr_t *xi, mpfr_t *yr, mpfr_t *yi, int nthr)
{
    long N = t->N;
#ifdef _OPENMP
    if (nthr <= 0) nthr = omp_get_max_threads();
#else
    nthr = 1;
#endif
    #pragma omp parallel num_threads(nthr)
    {
        orc_tmp w;
        tmp_init(&w, P);
        #pragma omp for schedule(static)
        for (long i = 0; i < t->nleaf; i++)
            mp_leaf(&t->leaf[i], N, xr, xi, yr, yi, &w);
        for (long l = 3; l <= t->lgN; l++) {
            long n = 1L << l, q = n / 4, cnt = t->lvl_cnt[l];
            long tot = cnt * q;
            #pragma omp for schedule(static)
            for (long b = 0; b < tot; b++) {
                long node = b / q, k = b % q;
                mp_butterfly(code, N, n, t->lvl_off[l][node], k, T, yr, yi, &w);
            }
        }
        tmp_clear(&w);
        mpfr_free_cache();
    }
}

static void mp_fft_run(const orc_tree *t, const orc_lut *T, mpfr_prec_t P,
                       mpfr_t *xr, mpfr_t *xi, mpfr_t *yr, mpfr_t *yi, int nthr)
{
    mp_fft_run_code(3, t, T, P, xr, xi, yr, yi, nthr);
}

static mpfr_t *mp_array(long M, mpfr_prec_t P)
{
    mpfr_t *a = (mpfr_t *)malloc(sizeof(mpfr_t) * (size_t)(M ? M : 1));
    for (long i = 0; i < M; i++) mpfr_init2(a[i], P);
    return a;
}
static void mp_array_free(mpfr_t *a, long M)
{
    for (long i = 0; i < M; i++) mpfr_clear(a[i]);
    free(a);
}

static int is_pow2(unsigned long N) { return N && !(N & (N - 1)); }

/* Public: forward FFT of SoA input at working precision Pw.
 * Input has Lin limbs (values exact at Pw bits, e.g. Pin <= Pw),
 * output has Lout = ceil(Pw/32) limbs.  nthr <= 0 means all threads.
 * Returns 0 or a negative error code. */
int orc_mp_fft_code(int code, unsigned long N, long Pw, long Lin,
                    const int32_t *hdr_in, const uint32_t *limb_in,
                    int32_t *hdr_out, uint32_t *limb_out, int nthr)
{
    if (!is_pow2(N) || Pw < 2 || code < 1 || code > 3) return -1;
    long Lout = (Pw + 31) / 32, M = 2 * (long)N;
    orc_tree t; tree_build(&t, (long)N, 4);
    orc_lut T; lut_build(&T, (long)N, Pw);
    mpfr_t *xr = mp_array((long)N, Pw), *xi = mp_array((long)N, Pw);
    mpfr_t *yr = mp_array((long)N, Pw), *yi = mp_array((long)N, Pw);
    int rc = 0;
    for (long i = 0; i < (long)N && !rc; i++) {
        rc |= soa_get(xr[i], hdr_in, limb_in, Lin, M, i);
        rc |= soa_get(xi[i], hdr_in, limb_in, Lin, M, (long)N + i);
    }
    if (!rc) {
        mp_fft_run_code(code, &t, &T, Pw, xr, xi, yr, yi, nthr);
        for (long i = 0; i < (long)N; i++) {
            soa_put(yr[i], hdr_out, limb_out, Lout, M, i);
            soa_put(yi[i], hdr_out, limb_out, Lout, M, (long)N + i);
        }
    }
    mp_array_free(xr, (long)N); mp_array_free(xi, (long)N);
    mp_array_free(yr, (long)N); mp_array_free(yi, (long)N);
    lut_free(&T); tree_free(&t);
    return rc ? -2 : 0;
}

int orc_mp_fft(unsigned long N, long Pw, long Lin,
               const int32_t *hdr_in, const uint32_t *limb_in,
               int32_t *hdr_out, uint32_t *limb_out, int nthr)
{
    return orc_mp_fft_code(3, N, Pw, Lin, hdr_in, limb_in, hdr_out, limb_out, nthr);
}

/* Timing entry for the CPU baseline: runs `reps` forward FFTs on data
 * already held as mpfr_t (conversion and LUT excluded), returns the mean
 * seconds per FFT. */
#include <time.h>
static double now_s(void)
{
    struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
double orc_mp_fft_time(unsigned long N, long P, long L,
                       const int32_t *hdr_in, const uint32_t *limb_in,
                       int reps, int nthr)
{
    if (!is_pow2(N) || reps < 1) return -1.0;
    long M = 2 * (long)N;
    orc_tree t; tree_build(&t, (long)N, 4);
    orc_lut T; lut_build(&T, (long)N, P);
    mpfr_t *xr = mp_array((long)N, P), *xi = mp_array((long)N, P);
    mpfr_t *yr = mp_array((long)N, P), *yi = mp_array((long)N, P);
    for (long i = 0; i < (long)N; i++) {
        soa_get(xr[i], hdr_in, limb_in, L, M, i);
        soa_get(xi[i], hdr_in, limb_in, L, M, (long)N + i);
    }
    double t0 = now_s();
    for (int r = 0; r < reps; r++) mp_fft_run(&t, &T, P, xr, xi, yr, yi, nthr);
    double dt = (now_s() - t0) / reps;
    mp_array_free(xr, (long)N); mp_array_free(xi, (long)N);
    mp_array_free(yr, (long)N); mp_array_free(yi, (long)N);
    lut_free(&T); tree_free(&t);
    return dt;
}

/* Public: the LUT as SoA (4 arrays of N/4: r, i, s, d, in that order,
 * M = N (4 * N/4) elements). */
int orc_mp_lut(unsigned long N, long P, int32_t *hdr, uint32_t *limb)
{
    if (!is_pow2(N) || N < 4) return -1;
    long L = (P + 31) / 32, Q = (long)N / 4, M = (long)N;
    orc_lut T; lut_build(&T, (long)N, P);
    for (long k = 0; k < Q; k++) {
        soa_put(T.r[k], hdr, limb, L, M, k);
        soa_put(T.i[k], hdr, limb, L, M, Q + k);
        soa_put(T.s[k], hdr, limb, L, M, 2 * Q + k);
        soa_put(T.d[k], hdr, limb, L, M, 3 * Q + k);
    }
    lut_free(&T);
    return 0;
}

/* Public: single correctly rounded ops on SoA operands (unit-test oracle
 * for the device limb arithmetic).  op: 0 add, 1 sub, 2 mul. */
int orc_mp_binop(int op, long count, long P,
                 const int32_t *ha, const uint32_t *la,
                 const int32_t *hb, const uint32_t *lb,
                 int32_t *hc, uint32_t *lc)
{
    long L = (P + 31) / 32;
    mpfr_t a, b, c;
    mpfr_inits2(P, a, b, c, (mpfr_ptr)0);
    int rc = 0;
    for (long i = 0; i < count; i++) {
        rc |= soa_get(a, ha, la, L, count, i);
        rc |= soa_get(b, hb, lb, L, count, i);
        if (op == 0) mpfr_add(c, a, b, RN);
        else if (op == 1) mpfr_sub(c, a, b, RN);
        else mpfr_mul(c, a, b, RN);
        soa_put(c, hc, lc, L, count, i);
    }
    mpfr_clears(a, b, c, (mpfr_ptr)0);
    return rc ? -2 : 0;
}

/* Public: IN=2 input of the reference, ic() serial_mp_fft3.c:104-115:
 * re[k] = 1/(1+k^2), im[k] = 1/(1+k^4), each step rounded at P. */
int orc_mp_input_in2(unsigned long N, long P, int32_t *hdr, uint32_t *limb)
{
    long L = (P + 31) / 32, M = 2 * (long)N;
    mpfr_t k2, v;
    mpfr_inits2(P, k2, v, (mpfr_ptr)0);
    for (unsigned long k = 0; k < N; k++) {
        mpfr_set_ui(k2, k * k, RN);
        mpfr_add_ui(v, k2, 1, RN);
        mpfr_ui_div(v, 1, v, RN);
        soa_put(v, hdr, limb, L, M, (long)k);
        mpfr_mul(k2, k2, k2, RN);
        mpfr_add_ui(v, k2, 1, RN);
        mpfr_ui_div(v, 1, v, RN);
        soa_put(v, hdr, limb, L, M, (long)N + (long)k);
    }
    mpfr_clears(k2, v, (mpfr_ptr)0);
    return 0;
}

/* Public: error of a P-bit result X against a higher-precision result X*
 * (same N, Lx and Ls limbs).  out[0] = max componentwise |X-X*| / ulp_P(X*)
 * over nonzero X*, out[1] = ||X-X*||_2 / ||X*||_2 / 2^-P,
 * out[2] = number of X* components that are zero while X is not. */
int orc_err_stats(unsigned long N, long P, long Lx, long Ls,
                  const int32_t *hx, const uint32_t *lx,
                  const int32_t *hs, const uint32_t *ls, double *out)
{
    long M = 2 * (long)N, Pbig = 32 * (Lx > Ls ? Lx : Ls) + 64;
    mpfr_t x, s, d, num, den, tmp;
    mpfr_inits2(Pbig, x, s, d, tmp, (mpfr_ptr)0);
    mpfr_inits2(256, num, den, (mpfr_ptr)0);
    mpfr_set_zero(num, 1); mpfr_set_zero(den, 1);
    double maxulp = 0.0, zbad = 0.0;
    int rc = 0;
    for (long i = 0; i < M; i++) {
        rc |= soa_get(x, hx, lx, Lx, M, i);
        rc |= soa_get(s, hs, ls, Ls, M, i);
        mpfr_sub(d, x, s, RN);  /* exact when the exponents are close, else
                                   rounded at 64 bits more than either holds */
        mpfr_sqr(tmp, d, RN); mpfr_add(num, num, tmp, RN);
        mpfr_sqr(tmp, s, RN); mpfr_add(den, den, tmp, RN);
        if (mpfr_zero_p(s)) { if (!mpfr_zero_p(d)) zbad += 1; continue; }
        /* ulp_P(s) = 2^(exp(s) - P) */
        mpfr_abs(d, d, RN);
        mpfr_mul_2si(d, d, P - mpfr_get_exp(s), RN);
        double u = mpfr_get_d(d, RN);
        if (u > maxulp) maxulp = u;
    }
    out[0] = maxulp;
    if (mpfr_zero_p(den)) out[1] = 0.0;
    else {
        mpfr_div(num, num, den, RN); mpfr_sqrt(num, num, RN);
        mpfr_mul_2si(num, num, P, RN);
        out[1] = mpfr_get_d(num, RN);
    }
    out[2] = zbad;
    mpfr_clears(x, s, d, tmp, num, den, (mpfr_ptr)0);
    return rc ? -2 : 0;
}

/* Public: componentwise error beyond correct rounding, for the forms whose
 * contract is "the exact transform within an absolute error A, then rounded
 * once to P bits": |X - X*| <= ulp_P(X)/2 + A for every component.
 * out[0] = max over components of max(0, |X - X*| - ulp_P(X)/2) / 2^ea
 * (ulp_P(0) = 0), i.e. the smallest A/2^ea the result satisfies;
 * out[1] = number of components with |X - X*| > ulp_P(X)/2.
 * X at P bits (Lx limbs), X* at higher precision (Ls limbs). */
int orc_err_excess(unsigned long N, long P, long Lx, long Ls,
                   const int32_t *hx, const uint32_t *lx,
                   const int32_t *hs, const uint32_t *ls, long ea, double *out)
{
    long M = 2 * (long)N, Pbig = 32 * (Lx > Ls ? Lx : Ls) + 64;
    mpfr_t x, s, d, h;
    mpfr_inits2(Pbig, x, s, d, h, (mpfr_ptr)0);
    double worst = 0.0, over = 0.0;
    int rc = 0;
    for (long i = 0; i < M; i++) {
        rc |= soa_get(x, hx, lx, Lx, M, i);
        rc |= soa_get(s, hs, ls, Ls, M, i);
        mpfr_sub(d, x, s, RN);  /* exact when the exponents are close, else
                                   rounded at 64 bits more than either holds */
        mpfr_abs(d, d, RN);
        if (!mpfr_zero_p(x)) {
            mpfr_set_ui_2exp(h, 1, mpfr_get_exp(x) - P - 1, RN);   /* ulp_P(X)/2 */
            mpfr_sub(d, d, h, RN);
        }
        if (mpfr_sgn(d) > 0) {
            over += 1;
            mpfr_mul_2si(d, d, -ea, RN);
            double u = mpfr_get_d(d, RN);
            if (u > worst) worst = u;
        }
    }
    out[0] = worst;
    out[1] = over;
    mpfr_clears(x, s, d, h, (mpfr_ptr)0);
    return rc ? -2 : 0;
}

/* Public: the reference's round-trip accuracy check (test(),
 * serial_mp_fft3.c:346-427).  Given the input x and the forward result X
 * (both SoA at P bits), run the forward FFT on swap(X) (re<->im), divide
 * by N, swap back and compare with x.  out[0..1] = rRMSE (re, im) as
 * doubles; out[2..3] = max abs error (re, im). */
int orc_mp_roundtrip(unsigned long N, long P,
                     const int32_t *hx, const uint32_t *lx,
                     const int32_t *hX, const uint32_t *lX, double *out)
{
    long L = (P + 31) / 32, M = 2 * (long)N;
    orc_tree t; tree_build(&t, (long)N, 4);
    orc_lut T; lut_build(&T, (long)N, P);
    mpfr_t *xr = mp_array((long)N, P), *xi = mp_array((long)N, P);
    mpfr_t *yr = mp_array((long)N, P), *yi = mp_array((long)N, P);
    int rc = 0;
    for (long i = 0; i < (long)N; i++) {   /* in := (out.im, out.re) */
        rc |= soa_get(xr[i], hX, lX, L, M, (long)N + i);
        rc |= soa_get(xi[i], hX, lX, L, M, i);
    }
    mp_fft_run(&t, &T, P, xr, xi, yr, yi, 0);
    mpfr_t v, e, sr, si, sdr, sdi, tmp, mr, mi;
    mpfr_inits2(P, v, e, sr, si, sdr, sdi, tmp, mr, mi, (mpfr_ptr)0);
    mpfr_set_zero(sr, 1); mpfr_set_zero(si, 1);
    mpfr_set_zero(sdr, 1); mpfr_set_zero(sdi, 1);
    mpfr_set_zero(mr, 1); mpfr_set_zero(mi, 1);
    for (long k = 0; k < (long)N; k++) {
        mpfr_div_ui(yr[k], yr[k], N, RN); mpfr_div_ui(yi[k], yi[k], N, RN);
        /* re part of x vs yi, im part vs yr */
        rc |= soa_get(v, hx, lx, L, M, k);
        mpfr_sub(e, v, yi[k], RN);
        mpfr_sqr(tmp, v, RN); mpfr_add(sr, sr, tmp, RN);
        mpfr_sqr(tmp, e, RN); mpfr_add(sdr, sdr, tmp, RN);
        mpfr_abs(e, e, RN); mpfr_max(mr, mr, e, RN);
        rc |= soa_get(v, hx, lx, L, M, (long)N + k);
        mpfr_sub(e, v, yr[k], RN);
        mpfr_sqr(tmp, v, RN); mpfr_add(si, si, tmp, RN);
        mpfr_sqr(tmp, e, RN); mpfr_add(sdi, sdi, tmp, RN);
        mpfr_abs(e, e, RN); mpfr_max(mi, mi, e, RN);
    }
    mpfr_div(tmp, sdr, sr, RN); mpfr_sqrt(tmp, tmp, RN); out[0] = mpfr_get_d(tmp, RN);
    mpfr_div(tmp, sdi, si, RN); mpfr_sqrt(tmp, tmp, RN); out[1] = mpfr_get_d(tmp, RN);
    out[2] = mpfr_get_d(mr, RN); out[3] = mpfr_get_d(mi, RN);
    mpfr_clears(v, e, sr, si, sdr, sdi, tmp, mr, mi, (mpfr_ptr)0);
    mp_array_free(xr, (long)N); mp_array_free(xi, (long)N);
    mp_array_free(yr, (long)N); mp_array_free(yi, (long)N);
    lut_free(&T); tree_free(&t);
    return rc ? -2 : 0;
}

/* Public: operation counts of the CODE 3 schedule (cmul, cadd, cfc, crb),
 * from the flattened tree (reference counters serial_mp_fft3.c:123-250). */
void orc_mp_counts(unsigned long N, unsigned long *c)
{
    orc_tree t; tree_build(&t, (long)N, 4);
    unsigned long mul = 0, add = 0, calls = 0;
    for (long i = 0; i < t.nleaf; i++) {
        long n = t.leaf[i].n;
        add += n == 4 ? 16 : n == 2 ? 4 : 0;
    }
    for (long l = 3; l <= t.lgN; l++) {
        long n = 1L << l, cnt = t.lvl_cnt[l];
        add += cnt * 12;                               /* k = 0 */
        add += cnt * (4 + 12); mul += cnt * 4;         /* k = n/8 */
        add += cnt * (n / 4 - 2) * (6 + 12);           /* general k */
        mul += cnt * (n / 4 - 2) * 6;
        calls += cnt;
    }
    c[0] = mul; c[1] = add; c[2] = calls + t.nleaf; c[3] = t.nleaf;
    tree_free(&t);
}

/* ------------------------------------------------------------------ */
/* double path: CODE 5 (serial_d_fft5.c) and CODE 6/7 (serial_d_fft6.c) */
/* ------------------------------------------------------------------ */

/* LUT and fixed constants: serial_d_fft5.c:52-62 / serial_d_fft6.c:52-62.
 * gcc -O3 (the reference's build) turns the cos(x), sin(x) pair of one
 * argument into one glibc sincos(x) call, and glibc's sincos sine differs
 * from its sin(x) by one ulp for some angles: call sincos explicitly so the
 * table matches that build whatever the compiler fuses. */
void orc_d_lut(unsigned long N, double *wr, double *wi, double *cst)
{
    for (unsigned long k = 0; k < N / 4; k++)
        sincos(2. * M_PI * k / N, &wi[k], &wr[k]);
    cst[0] = cos(2. * M_PI * 1 / 8);
    cst[1] = cos(2. * M_PI * 1 / 16);
    cst[2] = cos(2. * M_PI * 1 / 16) - cos(2. * M_PI * 3 / 16);
    cst[3] = cos(2. * M_PI * 3 / 16);
    cst[4] = cos(2. * M_PI * 1 / 16) + cos(2. * M_PI * 3 / 16);
}

/* leaves n = 1, 2, 4 shared by CODE 5 and CODE 6 (serial_d_fft5.c:97-147) */
static void d_leaf124(const orc_node *q, long N, const double *xr, const double *xi,
                      double *yr, double *yi)
{
    long o = q->off, i0 = gidx(N, q->sn, q->s, 0);
    if (q->n == 1) { yr[o] = xr[i0]; yi[o] = xi[i0]; return; }
    long i1 = gidx(N, q->sn, q->s, 1);
    if (q->n == 2) {
        yr[o] = xr[i0] + xr[i1]; yr[o + 1] = xr[i0] - xr[i1];
        yi[o] = xi[i0] + xi[i1]; yi[o + 1] = xi[i0] - xi[i1];
        return;
    }
    long i2 = gidx(N, q->sn, q->s, 2), i3 = gidx(N, q->sn, q->s, 3);
    double er0 = xr[i0] + xr[i2], er1 = xr[i0] - xr[i2];
    double or0 = xr[i1] + xr[i3], or1 = xr[i1] - xr[i3];
    double ei0 = xi[i0] + xi[i2], ei1 = xi[i0] - xi[i2];
    double oi0 = xi[i1] + xi[i3], oi1 = xi[i1] - xi[i3];
    yr[o] = er0 + or0; yr[o + 1] = er1 + oi1; yr[o + 2] = er0 - or0; yr[o + 3] = er1 - oi1;
    yi[o] = ei0 + oi0; yi[o + 1] = ei1 - or1; yi[o + 2] = ei0 - oi0; yi[o + 3] = ei1 + or1;
}

/* hard-coded 8-point leaf of CODE 6 (serial_d_fft6.c:148-212) */
static void d_leaf8(const orc_node *q, long N, const double *xr, const double *xi,
                    double *yr, double *yi, double w8)
{
    long o = q->off, I[8];
    for (int j = 0; j < 8; j++) I[j] = gidx(N, q->sn, q->s, j);
    double R[8], J[8];
    /* pairs (0,4) (2,6) (1,7) (3,5): sums/differences */
    static const int pa[4] = {0, 2, 1, 3}, pb[4] = {4, 6, 7, 5};
    double sr[4], dr[4], si[4], di[4];
    for (int u = 0; u < 4; u++) {
        sr[u] = xr[I[pa[u]]] + xr[I[pb[u]]]; dr[u] = xr[I[pa[u]]] - xr[I[pb[u]]];
        si[u] = xi[I[pa[u]]] + xi[I[pb[u]]]; di[u] = xi[I[pa[u]]] - xi[I[pb[u]]];
    }
    /* sr/dr[0]=a1/a2, [1]=a3/a4, [2]=a5/a6, [3]=a7/a8 */
    double m1r = w8 * (dr[2] + dr[3]), m1i = w8 * (di[2] + di[3]);
    double m2r = w8 * (sr[2] - sr[3]), m2i = w8 * (si[2] - si[3]);
    double b1r = sr[2] + sr[3], b1i = si[2] + si[3];
    double b2r = sr[0] + sr[1], b2i = si[0] + si[1];
    double b3r = sr[0] - sr[1], b3i = si[0] - si[1];
    double b4r = dr[2] - dr[3], b4i = di[2] - di[3];
    double b5r = dr[0] + m2r, b5i = di[0] + m2i;
    double b6r = dr[0] - m2r, b6i = di[0] - m2i;
    double b7r = dr[1] + m1r, b7i = di[1] + m1i;
    double b8r = dr[1] - m1r, b8i = di[1] - m1i;
    R[0] = b2r + b1r; R[1] = b5r + b7i; R[2] = b3r + b4i; R[3] = b6r - b8i;
    R[4] = b2r - b1r; R[5] = b6r + b8i; R[6] = b3r - b4i; R[7] = b5r - b7i;
    J[0] = b2i + b1i; J[1] = b5i - b7r; J[2] = b3i - b4r; J[3] = b6i + b8r;
    J[4] = b2i - b1i; J[5] = b6i - b8r; J[6] = b3i + b4r; J[7] = b5i + b7r;
    for (int j = 0; j < 8; j++) { yr[o + j] = R[j]; yi[o + j] = J[j]; }
}

/* hard-coded 16-point (Winograd) leaf of CODE 6 (serial_d_fft6.c:213-385) */
static void d_leaf16(const orc_node *q, long N, const double *xr, const double *xi,
                     double *yr, double *yi, const double *cst)
{
    const double w8 = cst[0], c1 = cst[1], c2 = cst[2], c3 = cst[3], c4 = cst[4];
    long o = q->off;
    double xR[16], xI[16];
    for (int j = 0; j < 16; j++) {
        long g = gidx(N, q->sn, q->s, j);
        xR[j] = xr[g]; xI[j] = xi[g];
    }
    /* stage a: u = x[j] + x[j+8], v = x[j] - x[j+8], j = 0..7
     * (imag of v[4] is formed as -x4 + x12, i.e. x12 - x4) */
    double uR[8], vR[8], uI[8], vI[8];
    for (int j = 0; j < 8; j++) {
        uR[j] = xR[j] + xR[j + 8]; vR[j] = xR[j] - xR[j + 8];
        uI[j] = xI[j] + xI[j + 8]; vI[j] = xI[j] - xI[j + 8];
    }
    vI[4] = -xI[4] + xI[12];
    /* names follow the a1..a16 numbering: a(2j+1) = u[j], a(2j+2) = v[j] */
#define A_R(k) (((k) & 1) ? uR[((k) - 1) / 2] : vR[((k) - 2) / 2])
#define A_I(k) (((k) & 1) ? uI[((k) - 1) / 2] : vI[((k) - 2) / 2])
    double b1r = A_R(1) - A_R(9), b2r = A_R(1) + A_R(9);
    double b3r = A_R(5) - A_R(13), b4r = A_R(5) + A_R(13);
    double b5r = b2r + b4r, b6r = b2r - b4r;
    double b7r = A_R(3) - A_R(11), b8r = A_R(3) + A_R(11);
    double b9r = A_R(7) - A_R(15), b10r = A_R(7) + A_R(15);
    double b11r = b8r + b10r, b12r = b8r - b10r;
    double b13r = A_R(4) + A_R(16), b14r = A_R(4) - A_R(16);
    double b15r = A_R(8) + A_R(12), b16r = A_R(12) - A_R(8);

    double b1i = A_I(1) - A_I(9), b2i = A_I(1) + A_I(9);
    double b3i = -A_I(5) + A_I(13), b4i = A_I(5) + A_I(13);
    double b5i = b2i + b4i, b6i = b2i - b4i;
    double b7i = A_I(3) - A_I(11), b8i = A_I(3) + A_I(11);
    double b9i = A_I(7) - A_I(15), b10i = A_I(7) + A_I(15);
    double b11i = b8i + b10i, b12i = -b8i + b10i;
    double b13i = A_I(4) + A_I(16), b14i = A_I(4) - A_I(16);
    double b15i = A_I(8) + A_I(12), b16i = A_I(12) - A_I(8);

    double m1r = w8 * (A_R(6) - A_R(14)), m2r = w8 * (A_R(6) + A_R(14));
    double m3r = w8 * (b7r - b9r), m4r = w8 * (b7r + b9r);
    double m5r = c1 * (b13r + b15r), m6r = -c2 * b13r, m7r = c4 * b15r;
    double m8r = c3 * (b14r + b16r), m9r = c4 * b14r, m10r = -c2 * b16r;

    double m1i = w8 * (A_I(6) - A_I(14)), m2i = -w8 * (A_I(6) + A_I(14));
    double m3i = w8 * (b7i - b9i), m4i = -w8 * (b7i + b9i);
    double m5i = -c1 * (b13i + b15i), m6i = c2 * b13i, m7i = -c4 * b15i;
    double m8i = c3 * (b14i + b16i), m9i = c4 * b14i, m10i = -c2 * b16i;
#undef A_R
#undef A_I
    double c1r = b1r + m3r, c2r = b1r - m3r, c3r = b3r + m4r, c4r = m4r - b3r;
    double c5r = vR[0] + m1r, c6r = vR[0] - m1r, c7r = m9r - m8r, c8r = m10r - m8r;
    double c9r = c5r + c7r, c10r = c5r - c7r, c11r = c6r + c8r, c12r = c6r - c8r;
    double c13r = vR[4] + m2r, c14r = vR[4] - m2r, c15r = m5r + m6r, c16r = m5r - m7r;
    double c17r = c13r + c15r, c18r = c13r - c15r, c19r = c14r + c16r, c20r = c14r - c16r;

    double c1i = b1i + m3i, c2i = b1i - m3i, c3i = b3i + m4i, c4i = m4i - b3i;
    double c5i = vI[0] + m1i, c6i = vI[0] - m1i, c7i = m9i - m8i, c8i = m10i - m8i;
    double c9i = c5i + c7i, c10i = c5i - c7i, c11i = c6i + c8i, c12i = c6i - c8i;
    double c13i = vI[4] + m2i, c14i = vI[4] - m2i, c15i = m5i + m6i, c16i = m5i - m7i;
    double c17i = c13i + c15i, c18i = c13i - c15i, c19i = c14i + c16i, c20i = c14i - c16i;

    double R[16] = {
        b5r + b11r, c9r - c17i, c1r - c3i, c12r + c20i,
        b6r - b12i, c11r - c19i, c2r - c4i, c10r + c18i,
        b5r - b11r, c10r - c18i, c2r + c4i, c11r + c19i,
        b6r + b12i, c12r - c20i, c1r + c3i, c9r + c17i };
    double J[16] = {
        b5i + b11i, c9i - c17r, c1i - c3r, c12i + c20r,
        b6i - b12r, c11i - c19r, c2i - c4r, c10i + c18r,
        b5i - b11i, c10i - c18r, c2i + c4r, c11i + c19r,
        b6i + b12r, c12i - c20r, c1i + c3r, c9i + c17r };
    for (int j = 0; j < 16; j++) { yr[o + j] = R[j]; yi[o + j] = J[j]; }
}

/* 4m2a butterfly, CODE 5 (3 butterflies, serial_d_fft5.c:154-227) and
 * CODE 6 (one butterfly for every k, serial_d_fft6.c:392-427) */
static void d_butterfly(int code, long N, long n, long off, long k,
                        const double *wr_t, const double *wi_t,
                        double *yr, double *yi)
{
    long q = n / 4;
    long i0 = off + k, i1 = i0 + q, i2 = i0 + 2 * q, i3 = i0 + 3 * q;
    double U0r = yr[i0], U1r = yr[i1], Zpr = yr[i2], Zmr = yr[i3];
    double U0i = yi[i0], U1i = yi[i1], Zpi = yi[i2], Zmi = yi[i3];
    double m1r, m1i, m2r, m2i;
    if (code == 5 && k == 0) {
        m1r = Zpr; m1i = Zpi; m2r = Zmr; m2i = Zmi;
    } else {
        long kn = k * (N / n);
        double wr = wr_t[kn], wi = wi_t[kn];
        if (code == 5 && k == n / 8) {
            m1r = wr * (Zpr + Zpi); m1i = wr * (Zpi - Zpr);
            m2r = wr * (Zmr - Zmi); m2i = wr * (Zmi + Zmr);
        } else {
            m1r = wr * Zpr + wi * Zpi; m1i = wr * Zpi - wi * Zpr;
            m2r = wr * Zmr - wi * Zmi; m2i = wr * Zmi + wi * Zmr;
        }
    }
    double a1r = m1r + m2r, a1i = m1i + m2i, a2r = m1r - m2r, a2i = m1i - m2i;
    yr[i0] = U0r + a1r; yr[i2] = U0r - a1r; yr[i1] = U1r + a2i; yr[i3] = U1r - a2i;
    yi[i0] = U0i + a1i; yi[i2] = U0i - a1i; yi[i1] = U1i - a2r; yi[i3] = U1i + a2r;
}

/* Public: double CP FFT, code 5 or 6 (7 gives the same values as 6). */
int orc_d_fft(int code, unsigned long N, const double *xr, const double *xi,
              double *yr, double *yi)
{
    if (!is_pow2(N) || (code != 5 && code != 6 && code != 7)) return -1;
    if (code == 7) code = 6;
    orc_tree t; tree_build(&t, (long)N, code == 5 ? 4 : 16);
    double *wr = (double *)malloc(sizeof(double) * (N / 4 + 1));
    double *wi = (double *)malloc(sizeof(double) * (N / 4 + 1));
    double cst[5];
    orc_d_lut(N, wr, wi, cst);
    for (long i = 0; i < t.nleaf; i++) {
        const orc_node *q = &t.leaf[i];
        if (q->n <= 4) d_leaf124(q, (long)N, xr, xi, yr, yi);
        else if (q->n == 8) d_leaf8(q, (long)N, xr, xi, yr, yi, cst[0]);
        else d_leaf16(q, (long)N, xr, xi, yr, yi, cst);
    }
    long lmin = code == 5 ? 3 : 5;
    for (long l = lmin; l <= t.lgN; l++) {
        long n = 1L << l;
        for (long c = 0; c < t.lvl_cnt[l]; c++)
            for (long k = 0; k < n / 4; k++)
                d_butterfly(code, (long)N, n, t.lvl_off[l][c], k, wr, wi, yr, yi);
    }
    free(wr); free(wi); tree_free(&t);
    return 0;
}
