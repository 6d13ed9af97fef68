// Kernel argument blocks and launchers shared by the HIP sources
// (mp_fft.hip, d_fft.hip) and the host library (mpfft.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpfft {

// ---- multiprecision plain form (mp_fft.hip) ----

struct LevelArgs {
    uint32_t *data;             // out buffer of transform 0
    const uint32_t *lut;        // twiddle buffer: SoA of lut_m reals, rows of lut_e
                                // entries: r, s, d (code 3) or r, i (codes 1, 2);
                                // level n uses entries lut_off + k = LUT[k N/n]
    uint32_t lut_m, lut_e, lut_off;
    int code;                   // operation order: 3 (3m3a), 2 (4m2a), 1 (4m2a, no
                                // special butterflies)
    const uint32_t *node_off;   // output offsets of the size-n nodes
    uint64_t stride;            // words between consecutive transforms
    uint32_t N, n, cnt;         // size, node size, number of size-n nodes
    int P;
};

struct LeafArgs {
    const uint32_t *in;         // input buffer of transform 0 (SoA limbs)
    uint32_t *aos;              // element-major copy of the input: element i is
                                // aos_words consecutive words hdr, limb 0..L-1
    uint64_t aos_stride;        // words between transforms in aos
    uint32_t aos_words;         // words per element (L+1 rounded up to 4)
    uint32_t *data;             // out buffer of transform 0
    const int32_t *pos_sn;      // per output position p < N: sn of its leaf
    const uint32_t *pos_nr;     //   (leaf size n) | (p - leaf offset) << 8
    uint32_t npos;              // N, or 0 for no leaf launch
    uint64_t in_stride, stride;
    uint32_t N;
    uint32_t in_swap;           // 1: read element i from (i + N) mod 2N (re <-> im)
    int P;
};

// kernels exist for these limb counts; a precision P uses the smallest
// count >= ceil(P/32)
int limbs_for_precision(long P);

hipError_t mp_fft_launch(int L, int P, const LeafArgs &la, const LevelArgs *lv,
                         int nlev, uint32_t batch, hipStream_t st);
// in place: swap the real and imaginary halves and divide by 2^lgN (exact:
// an exponent shift), the closing step of the inverse by the swap trick
hipError_t mp_unswap_scale_launch(int L, uint32_t *data, uint64_t stride, uint32_t N,
                                  int32_t lgN, uint32_t batch, hipStream_t st);

// error statistics of y against x (aE = x - y at P bits): per transform and
// per block of 256 reals, kErrPartial doubles
constexpr int kErrPartial = 10;
constexpr uint32_t kErrBlock = 256;
hipError_t mp_err_launch(int L, int P, const uint32_t *x, const uint32_t *y, uint64_t stride,
                         uint32_t N, double *partial, uint32_t batch, hipStream_t st);

hipError_t mp_binop_launch(int L, int P, int op, uint32_t count, const uint32_t *x,
                           const uint32_t *y, uint32_t *z, hipStream_t st);

// ---- fixed form (fx_fft.hip) ----

struct FxArgs {
    const uint32_t *in;         // float SoA input of transform 0 (hdr + L limbs)
    uint64_t in_stride;
    uint32_t *out;              // float SoA output of transform 0
    uint64_t out_stride;
    uint32_t *buf0, *buf1;      // fixed-point ping-pong buffers (W rows of 2N words)
    uint64_t stride;            // words per transform in buf0 / buf1
    const uint32_t *tw;         // twiddles of the radix-4 passes, pass after pass,
                                // entry 3k + r - 1 = w^(r k N / (4 Ns)) = c + i d;
                                // reals c, d - c, c + d (3 ntw of them), each DT
                                // digit rows + a sign row
    uint32_t ntw;               // twiddle entries in all
    uint32_t *flag;             // set to nonzero when an input exceeds the range
    uint32_t N, batch;
    int lim;                    // inputs must have exponent <= lim (|x| < 2^lim)
    int P;
};

// fixed-point words per real for limb count L (F = 32 (L + 1) fraction bits)
constexpr int fx_words(int L) { return L + 2; }

// passes of the fixed form: a radix-2 pass when log2 N is odd, then radix 4
inline int fx_passes(uint32_t N)
{
    int lg = 0;
    while ((1u << lg) < N) lg++;
    return (lg & 1) + lg / 2;
}

// stages of the fixed form: one per pass (the first converts the input, the
// last the output); N = 1: input and output conversion
inline int fx_stages(uint32_t N) { return fx_passes(N) > 0 ? fx_passes(N) : 2; }

// stages [s0, s1) of the fixed form
hipError_t fx_fft_launch(int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st);

// ---- double path (d_fft.hip) ----

struct DTables {
    const int32_t *leaf_sn;     // per leaf: input offset
    const uint32_t *leaf_off;   //           output offset
    const uint32_t *leaf_n;     //           size (1, 2, 4, 8, 16)
    uint32_t nleaf;
    const uint32_t *lvl_off;    // node offsets of all levels, concatenated
    const uint32_t *lvl_begin;  // lvl_begin[l] .. lvl_begin[l+1]: level 2^l
    const double *wr, *wi;      // N/4 each
    double c[5];                // w8, w16_1, w16_2, w16_3, w16_4
    uint32_t N, lgN, lmin;      // combine levels lmin..lgN
    int code;                   // 5 or 6
};

// largest transform run from LDS by one workgroup (2N doubles)
constexpr uint32_t kDLdsMaxN = 4096;

hipError_t d_fft_launch(const DTables &t, const double *in, double *out, uint64_t stride,
                        uint32_t batch, hipStream_t st);

}  // namespace mpfft
