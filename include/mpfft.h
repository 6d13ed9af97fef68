/*
 * mpfft -- multiprecision split-radix conjugate-pair FFT on AMD MI355X
 * (gfx950).  C ABI of libmpfft.so.
 *
 * Drop-in for the FFT kernels declared in mfrewer/mpFFT's mpFFT.h: the same
 * names and signatures (forward, out-of-place, split real/imaginary arrays,
 * N = 2^n).  The multiprecision results are bit-identical to the
 * reference's MPFR path (every operation correctly rounded at the arrays'
 * precision, in the reference's operation order); the double results are
 * bit-identical to the reference's double code.
 *
 * Two layers:
 *   1. the reference entry points (legacy names, void, synchronous); a
 *      failure is recorded per thread and read with mpfft_last_error();
 *   2. a plan API (int status) with device-resident execution for callers
 *      that keep data on the GPU.
 *
 * Device data format ("SoA limbs"), per transform of N points at L limbs:
 *   uint32 words [0, 2N)            hdr: (exp << 1) | sign for a nonzero
 *                                   value 0.1b..b * 2^exp, INT32_MIN | sign
 *                                   for a signed zero
 *   uint32 words [2N + j*2N, ...)   limb j (j = 0 most significant) of the
 *                                   2N reals; real parts first, then
 *                                   imaginary parts
 * i.e. 2N*(L+1) words per transform, transforms back to back.
 */
#ifndef MPFFT_H
#define MPFFT_H

#include <stddef.h>
#include <stdint.h>
#include <gmp.h>
#include <mpfr.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MPFFT_OK            0
#define MPFFT_ERR_SIZE     -1   /* N is not a power of two, or too large      */
#define MPFFT_ERR_PREC     -2   /* precision unsupported, or arrays disagree  */
#define MPFFT_ERR_VALUE    -3   /* NaN / Inf / exponent beyond +-2^29 in input */
#define MPFFT_ERR_SUBTREE  -4   /* legacy call with s != 1 or sn != 0         */
#define MPFFT_ERR_HIP      -5   /* HIP runtime error                          */
#define MPFFT_ERR_ARG      -6   /* null pointer or bad argument               */
#define MPFFT_ERR_NODEV    -7   /* no usable GPU                              */
#define MPFFT_ERR_FORM     -8   /* form not valid for this call               */
#define MPFFT_ERR_SCRATCH  -9   /* beyond the fixed workspace of a
                                   no-allocation entry point                  */

/* ---- reference entry points -------------------------------------------- */

/* mpFFT.h:38 / :66 (lab/serial/mp/serial_mp_fft3.c:121,
 * lab/parallel/mp/parallel_mp_fft2.c:121).  Supported call: the top-level
 * (n = N, s = 1, sn = 0), as every caller of the reference makes it
 * (mpFFT.c:47/50).  All four arrays must hold N elements of one precision
 * P (<= 2048 bits); outputs are written into outr/outi in place. */
void mp_sr24_cp_fft(unsigned long n, unsigned long s, long int sn,
                    mpfr_t *inr, mpfr_t *ini, mpfr_t *outr, mpfr_t *outi);

/* mpFFT.h:39 (lab/serial/mp/serial_mp_fft4.c:121): shift-indexed twiddles,
 * same results.  Supported call: (N, log2n = 0, s = 1, sn = 0). */
void mp_sr24_cp_fft_v2(unsigned long n, unsigned long log2n, unsigned long s,
                       long int sn, mpfr_t *inr, mpfr_t *ini, mpfr_t *outr,
                       mpfr_t *outi);

/* mpFFT.h:25 / :53 (lab/serial/double/serial_d_fft6.c:93, the CODE 6 op
 * order; CODE 5 is reachable through the plan API, MPFFT_FORM_D5). */
void d_sr24_cp_fft(unsigned long n, unsigned long s, long int sn,
                   double *inr, double *ini, double *outr, double *outi);

/* mpFFT.h:26 (lab/serial/double/serial_d_fft7.c:93-94): CODE 7, the same
 * arithmetic as CODE 6.  Supported call: (N, 0, 1, 0). */
void d_sr24_cp_fft_v2(unsigned long n, unsigned long log2n, unsigned long s,
                      long int sn, double *inr, double *ini, double *outr,
                      double *outi);

/* status of the last reference-entry-point call on this thread */
int mpfft_last_error(void);
const char *mpfft_strerror(int code);

/* ---- plan API ------------------------------------------------------------ */

typedef struct mpfft_plan_s *mpfft_plan;

#define MPFFT_FORM_MP  0   /* multiprecision, CODE 3 op order (3m3a), the
                              reference's default (serial_mp_fft3.c, and its
                              OpenMP twin parallel_mp_fft2.c bound by mpFFT.c) */
#define MPFFT_FORM_MP_C1 1 /* multiprecision, CODE 1 op order: 4m2a, one
                              butterfly for every k (serial_mp_fft1.c)      */
#define MPFFT_FORM_MP_C2 2 /* multiprecision, CODE 2 op order: 4m2a, special
                              butterflies k = 0 and n/8 (serial_mp_fft2.c)  */
#define MPFFT_FORM_FIXED 8 /* multiprecision "fixed" form: Stockham passes on
                              fixed-point reals with 32 (L+1) fraction bits
                              and a block exponent per transform; same
                              SoA-limb I/O as MPFFT_FORM_MP, results at the
                              plain form's accuracy or better instead of the
                              reference's bit pattern.  Up to 256 bits at
                              n >= 4096 (up to 64 bits at any n) as
                              MPFFT_FORM_FIXED_LANE; otherwise, at n >= 16
                              and a batch that fills its panels of 64
                              transforms to within a quarter (any multiple
                              of 64, or 48 of 64, ...; above 768 bits with
                              n <= 1024: any batch) as
                              MPFFT_FORM_FIXED_PANEL, else as
                              MPFFT_FORM_FIXED_LANE                        */
#define MPFFT_FORM_FIXED_LANE 9 /* the fixed form with per-lane twiddle
                              products (29-bit digit multiply-adds; 28-bit
                              at 2048 bits) at any batch: radix-16 passes
                              (four waves per butterfly) up to 256 bits,
                              radix 4 beyond; same format and error bound */
#define MPFFT_FORM_FIXED_PANEL 10 /* the fixed form with transforms
                              interleaved in panels of 64 and the twiddle
                              products on the int8 matrix cores; any batch
                              (a last, partial panel is padded), n >= 16
                              (MPFFT_ERR_SIZE below); same format and error
                              bound.  All three fixed forms and the plain op
                              orders take prec up to 2048 (64 limbs);
                              MPFFT_FORM_DD and _QD return MPFFT_ERR_PREC
                              beyond their own limits                     */
#define MPFFT_FORM_DD 11  /* double-word form, 53 <= prec <= 106: every
                              real is hi + lo (two binary64) and every
                              operation an error-free transformation on the
                              FP64 units (TwoSum, FMA residual products);
                              same SoA-limb I/O, the result rounded once to
                              prec bits; a block exponent per transform as
                              in the fixed form (a real more than 2^900
                              below the transform's largest counts as
                              zero); normwise error against the exact
                              transform below 2^-prec + (2 + 3 log2 n)
                              2^-104                                      */
#define MPFFT_FORM_QD 12  /* quad-word form, 53 <= prec <= 212: every real
                              c0 + c1 + c2 + c3 (four binary64), sums and
                              products by error-free transformations with
                              the order-3 terms rounded and renormalised;
                              same SoA-limb I/O, result rounded to prec
                              bits; a block exponent per transform (a real
                              more than 2^800 below the transform's
                              largest counts as zero); normwise error
                              below 2^-prec + (8 + 32 log2 n) 2^-212      */
#define MPFFT_FORM_AUTO 13 /* the chosen form: resolved at plan creation to
                              the fastest form whose error against a
                              2P-bit oracle measured no worse than the plain
                              form's, componentwise and normwise, on
                              MI355X: MPFFT_FORM_DD for 53 <= prec <= 106,
                              MPFFT_FORM_FIXED otherwise (which takes the
                              panel or the per-lane kernels); the resolved
                              form's input range applies.
                              mpfft_plan_form() names the kernels chosen  */
#define MPFFT_FORM_D5  5   /* double, CODE 5 op order (4m2a, 3 butterflies) */
#define MPFFT_FORM_D6  6   /* double, CODE 6/7 op order (1 butterfly)      */

/* Create a plan for `batch` transforms of size n at precision prec (bits;
 * ignored for the double forms).  Builds the twiddle table on the host
 * (MPFR, the reference's lut() sequence) and uploads it. */
int mpfft_plan_create(mpfft_plan *plan, unsigned long n, long prec, int form,
                      unsigned int batch);
int mpfft_plan_destroy(mpfft_plan plan);

/* limbs per real used by the device format, and words per transform */
int mpfft_plan_info(mpfft_plan plan, long *limbs, size_t *words_per_transform);

/* The form the plan runs (no counterpart in the reference): the one it was
 * created with, MPFFT_FORM_AUTO resolved, and MPFFT_FORM_FIXED as the
 * kernel family it took (MPFFT_FORM_FIXED_PANEL or _FIXED_LANE); a negative
 * status for a null plan. */
int mpfft_plan_form(mpfft_plan plan);

/* One transform through mpfr_t arrays (marshalling included).  The
 * significands are marshalled on a pool of up to 16 host threads (kept for
 * the process) into page-locked staging: the input in chunks, each uploaded
 * as soon as it is converted, the output in one copy.  Calls on different
 * plans share the pool one at a time. */
int mpfft_execute_mpfr(mpfft_plan plan, mpfr_t *inr, mpfr_t *ini,
                       mpfr_t *outr, mpfr_t *outi);

/* Wall times of the plan's last mpfft_execute_mpfr call, in ms:
 * ms[0] marshalling in (from the call's start to the end of the last
 * upload), ms[1] the transform on the GPU (with the device's change of the
 * staging layout), ms[2] download and conversion out, ms[3] total. */
int mpfft_plan_last_timing(mpfft_plan plan, double *ms);

/* `batch` transforms on device-resident SoA-limb buffers (d_in and d_out
 * hold batch * words_per_transform words each; they must not overlap).
 * Enqueued on `stream` (a hipStream_t, NULL = the plan's own stream);
 * returns without waiting. */
int mpfft_execute_limbs(mpfft_plan plan, const uint32_t *d_in, uint32_t *d_out,
                        void *stream);

/* Part of the transform for profiling: for the plain forms stage 0 is the
 * leaf pass, stage i (1 <= i < mpfft_plan_stages) the combine pass of node
 * size 2^(i+2); for the fixed form on lanes and the double- and quad-word
 * forms stage i is the i-th radix pass (the first also converts the input,
 * the last rounds the output); for the fixed form on panels stage 0
 * converts the input, the last stage the output, the ones between are the
 * passes (the four-step plan, mpfft_plan_panel_plan 2: stage i is its i-th
 * pass, the first also converting the input, the last rounding the output).
 * Running stages [0, mpfft_plan_stages) equals mpfft_execute_limbs. */
int mpfft_execute_limbs_range(mpfft_plan plan, const uint32_t *d_in, uint32_t *d_out,
                              void *stream, int stage_begin, int stage_end);
int mpfft_plan_stages(mpfft_plan plan);

/* The pass plan of the fixed form on panels (no counterpart in the
 * reference): 1 radix passes between an input and an output conversion
 * (the default), 2 the four-step plan (MPFFT_PANEL_PLAN=fs in the
 * environment at plan creation, for n = 16^m >= 4096 and at most 1024 bits:
 * n = n1 256, radix-16 passes over n1 and then over 256 points with the
 * twiddles w^(k1 n2) between, the input converted in the first pass and the
 * output in the last), 0 not a panel plan. */
int mpfft_plan_panel_plan(mpfft_plan plan);

/* Inverse transform, the way the reference's test() computes it
 * (serial_mp_fft3.c:367-385): out = swap(FFT(swap(in))) / N, where swap
 * exchanges real and imaginary parts; division by N is exact.  Same buffers
 * and stream contract as mpfft_execute_limbs. */
int mpfft_execute_inverse_limbs(mpfft_plan plan, const uint32_t *d_in, uint32_t *d_out,
                                void *stream);

/* Error of y against x per transform, the reference's accuracy measures
 * (serial_mp_fft3.c:387-404): aE = x - y at P bits, then
 * out[8b + 0/1] = rRMSE sqrt(sum aE^2 / sum x^2) of the real / imaginary
 * parts, out[8b + 2/3] = max |aE|, out[8b + 4..7] = log2 of those four (the
 * doubles underflow for P beyond ~1000 bits; the log2 values do not).  The
 * sums are accumulated in double, not at P bits.  Waits for `stream`. */
int mpfft_roundtrip_stats(mpfft_plan plan, const uint32_t *d_x, const uint32_t *d_y,
                          void *stream, double *out);

/* mpfft_roundtrip_stats without any allocation: the same statistics, bit
 * for bit, with the kernel's per-block partial sums in a fixed 4 MiB
 * workspace compiled into the library (one per device, held until the
 * process ends) instead of a buffer the plan allocates on first use, and
 * read back through a static host buffer.  No counterpart in the reference:
 * for callers that cannot allocate (serving loops and harnesses that
 * forbid hipMalloc, code that must not allocate after set-up).  Limit:
 * 2 ceil(2n / 256) * batch * 5 <= 2^19 doubles (n * batch <= 6.7 M points;
 * beyond: MPFFT_ERR_SCRATCH).  The workspace keeps the last call's partial
 * sums; calls serialize on a process-wide lock from launch to read back
 * (overlapping calls would otherwise race on it).  Waits for `stream`. */
int mpfft_roundtrip_stats_noalloc(mpfft_plan plan, const uint32_t *d_x, const uint32_t *d_y,
                                  void *stream, double *out);

/* The reference's test() on host data: forward transform of x, inverse of
 * the result, then mpfft_roundtrip_stats(x, round trip) into out[8 * batch]. */
int mpfft_roundtrip_host(mpfft_plan plan, const uint32_t *h_x, double *out);

/* The same from host SoA-limb buffers (copies in, runs, copies out). */
int mpfft_execute_limbs_host(mpfft_plan plan, const uint32_t *h_in, uint32_t *h_out);

/* Double forms: `batch` transforms of split arrays, re then im, 2N doubles
 * per transform; device-resident or host. */
int mpfft_execute_f64(mpfft_plan plan, const double *d_in, double *d_out, void *stream);
int mpfft_execute_f64_host(mpfft_plan plan, const double *h_in, double *h_out);

/* Wait for the plan's stream. */
int mpfft_plan_sync(mpfft_plan plan);

/* Wait for the plan's own stream and report input-range violations of the
 * fixed, double- and quad-word forms since the last check (MPFFT_ERR_VALUE);
 * MPFFT_OK otherwise.  After executes enqueued on a caller's stream, wait
 * for that stream first.  The host and mpfr_t entry points check by
 * themselves. */
int mpfft_plan_check(mpfft_plan plan);

/* ---- batched form --------------------------------------------------------- */

/* Many transforms of uneven size and precision in one graph launch.
 * Transforms of equal (n, prec) form a group that runs as one plan of that
 * group's batch; the groups run on up to four streams inside one HIP graph,
 * captured once per (d_in, d_out) pair and replayed by every execute.  With
 * MPFFT_FORM_FIXED a group runs on the panel or the per-lane kernels as an
 * MPFFT_FORM_FIXED plan of its size and count would, except that a group of
 * more than 64 whose last panel would be mostly padding runs its full
 * panels on the panel kernels and the rest on the per-lane kernels.  The
 * autotuner times every group alone, deals the groups over 1, 2 and 4
 * streams by those times and keeps the fastest graph.
 * (No counterpart in the reference, which runs one transform per call.) */
typedef struct mpfft_batch_s *mpfft_batch;

/* `count` transforms; transform i has size n[i] (a power of two) at prec[i]
 * bits.  form: MPFFT_FORM_MP, _MP_C1, _MP_C2 (the reference's op orders;
 * every transform bit-identical to a plan of that form), MPFFT_FORM_FIXED
 * (each group on the panel or the per-lane kernels by the rule of a fixed-
 * form plan of its size and count, full panels split off as above),
 * MPFFT_FORM_FIXED_LANE or _FIXED_PANEL
 * (one kernel family; groups with n < 16 run on the per-lane kernels),
 * MPFFT_FORM_DD or _QD, or MPFFT_FORM_AUTO (per group, the resolution of a
 * plan of that form).  Builds every group's plans (twiddle tables) before
 * it returns. */
int mpfft_batch_create(mpfft_batch *batch, unsigned int count, const unsigned long *n,
                       const long *prec, int form);
int mpfft_batch_destroy(mpfft_batch batch);

/* Packed device layout, the same for input and output: transform i is the
 * SoA-limb transform of n[i] points at mpfft_limbs_for_precision(prec[i])
 * limbs, at word `offset` of buffers of `words` words in all.  `form`: the
 * form that runs it (MPFFT_FORM_FIXED_LANE or MPFFT_FORM_FIXED_PANEL for
 * a group of an MPFFT_FORM_FIXED batch).  Null outputs are skipped. */
int mpfft_batch_words(mpfft_batch batch, size_t *words);
int mpfft_batch_info(mpfft_batch batch, unsigned int i, size_t *offset, long *limbs, int *form);

/* Time the schedules on these buffers and keep the fastest; d_out receives
 * the transforms of d_in.  Waits for the device. */
int mpfft_batch_tune(mpfft_batch batch, const uint32_t *d_in, uint32_t *d_out);

/* The schedule in use: streams of the graph and its measured time (ms) */
int mpfft_batch_schedule(mpfft_batch batch, int *streams, double *ms);

/* Every transform, d_in -> d_out (packed layout, no overlap), as one graph
 * launch on `stream` (a hipStream_t, NULL = the batch's own); tunes first on
 * the first call.  Returns without waiting.  A caller's stream must stay
 * alive until the next mpfft_batch_check, _tune or execute on other buffers
 * (each of which waits for it). */
int mpfft_batch_execute_limbs(mpfft_batch batch, const uint32_t *d_in, uint32_t *d_out,
                              void *stream);

/* Wait for the stream of the last execute and report input-range violations
 * of fixed-, double- and quad-word-form groups since the last check
 * (MPFFT_ERR_VALUE). */
int mpfft_batch_check(mpfft_batch batch);

/* ---- conversion ----------------------------------------------------------- */

/* mpfr_t arrays (re[n], im[n]) <-> one SoA-limb transform at `limbs` limbs. */
int mpfft_mpfr_to_limbs(unsigned long n, long limbs, mpfr_t *re, mpfr_t *im,
                        uint32_t *h_out);
int mpfft_limbs_to_mpfr(unsigned long n, long limbs, const uint32_t *h_in,
                        mpfr_t *re, mpfr_t *im);

/* Exact text form: `count` reals of a SoA-limb array (hdr[count], then
 * limbs[limbs][count]) as hexadecimal floats in the format MPFR prints for
 * "%Ra" (e.g. "0x1.8p-4", "-0x0p+0"), one per line.  Writes at most
 * `bufsize` bytes including the terminating NUL; *needed receives the size
 * required (MPFFT_ERR_ARG when buf is too small). */
int mpfft_limbs_to_hex(unsigned long count, long limbs, const uint32_t *h_in, char *buf,
                       size_t bufsize, size_t *needed);

/* Parse `count` whitespace-separated hexadecimal floats ("%Ra" format, as
 * accepted by mpfr_strtofr in base 16) into a SoA-limb array, rounding each
 * to `prec` bits, to nearest with ties to even.  MPFFT_ERR_VALUE on
 * malformed text or an exponent outside the accepted range. */
int mpfft_hex_to_limbs(unsigned long count, long limbs, long prec, const char *text,
                       uint32_t *h_out);

/* ---- device arithmetic check ---------------------------------------------- */

/* z[i] = x[i] op y[i] correctly rounded at prec (<= 2048) bits on the GPU; op 0 add,
 * 1 sub, 2 mul; x, y, z: host SoA-limb arrays of `count` reals at
 * mpfft_limbs_for_precision(prec) limbs (hdr[count] then limbs). */
int mpfft_mp_binop_host(int op, unsigned long count, long prec,
                        const uint32_t *x, const uint32_t *y, uint32_t *z);
/* Limbs per real of the device format for prec bits: the smallest of 1, 2,
 * 3, 4, 6, 8, 12, 16, 24, 32, 64 that holds prec bits; -1 outside [2, 2048].
 * (64 limbs: the plain op orders and the fixed form on panels.) */
long mpfft_limbs_for_precision(long prec);

#ifdef __cplusplus
}
#endif
#endif /* MPFFT_H */
