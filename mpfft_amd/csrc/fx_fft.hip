// The "fixed" form of the multiprecision FFT on its per-lane kernels: a
// Stockham (self-sorting, out-of-place) transform on fixed-point
// multiprecision reals (fx_arith.h), for callers that want the transform at
// the plain form's accuracy or better rather than the reference's bit
// pattern.  Passes (fxl_plan in kernels.h): radix 16 on four waves per
// butterfly up to kFxl16MaxW words per real (fx_r16_kernel), radix 4 with a
// DPP quad per butterfly beyond (fx_r4_kernel), a radix-2 pass first where
// log2 N is odd.
//
// Why it is cheaper than the plain form: every radix-4 pass covers two
// levels (a radix-16 pass four) with one read and one write of the data (the plain form reads and
// writes every element once per level), additions are exact W-word carry
// chains instead of aligned, normalised and rounded floating-point sums, and
// a product needs no normalisation or rounding test.  The cost is an
// absolute error of about 2^-F per operation, with F = 32 bits more than the
// floating precision.
//
// Data: two ping-pong buffers of W rows of 2N words per transform (real
// parts at [0, N), imaginary at [N, 2N)); word 0 least significant.  The
// first pass reads the plain form's SoA-limb layout and converts on the fly,
// the last pass rounds back to it at P bits, so the conversions cost no
// extra pass over the data.
//
// Twiddles: one table per pass, for radix 4 entry 3k + r - 1 = w^(r k) of
// that pass (k < Ns, r = 1..3), so the lanes of a wave read consecutive
// entries (radix 16: (r - 1) Ns + k, for the same reason).
//
// Radix-4 pass (radix R = 4 or 2, Ns = size of the sub-transforms already done):
//   for j < N/R, k = j mod Ns:
//     v_r = x[j + r N/R] * w^(r k), w = exp(-2 pi i / (R Ns))
//     y = DFT_R(v)
//     out[(j - k) R + k + r Ns] = y_r
// Lanes: R neighbouring lanes per j (a DPP quad); lane r loads v_r, forms
// its twiddle product, then the quad exchanges values twice (radix 4) or
// once (radix 2) through DPP.
#include "fx_arith.h"
#include "fx_lane.h"
#include "kernels.h"

#include <utility>

namespace mpfft {

namespace {

// LDS parking slot k of this lane: word j at lds[(k W + j) nt + t]
template <int W>
__device__ __forceinline__ void fx_put(uint32_t *lds, uint32_t nt, uint32_t t, int k, const Fx<W> &x)
{
#pragma unroll
    for (int j = 0; j < W; j++) lds[((uint32_t)k * W + j) * nt + t] = x.w[j];
}

template <int W>
__device__ __forceinline__ void fx_get(Fx<W> &x, const uint32_t *lds, uint32_t nt, uint32_t t, int k)
{
#pragma unroll
    for (int j = 0; j < W; j++) x.w[j] = lds[((uint32_t)k * W + j) * nt + t];
}

// twiddle real q of the table: DT digits then the sign word (SoA, MT reals)
template <int W>
__device__ __forceinline__ uint32_t tw_load(uint32_t (&T)[FxDig<W>::DT], const uint32_t *tw,
                                            uint32_t MT, uint32_t q)
{
#pragma unroll
    for (int j = 0; j < FxDig<W>::DT; j++) T[j] = (tw + (size_t)j * MT)[q];
    return (tw + (size_t)FxDig<W>::DT * MT)[q];
}

// Blocks per CU of the pass kernels: two up to 34 words per real; one at
// 66 words (2048 bits), where a lane's reals and the 76-digit product
// operands need more than 256 registers and the two LDS parking slots take
// 2 x 66 x 256 words (135 KB).
template <int W>
constexpr int kFxBlocks = W > 34 ? 1 : 2;

}  // namespace

// Pass I/O.  The first pass reads the float SoA input and converts on the
// fly (IO bit 1), the last pass rounds to P bits and writes the float SoA
// output (IO bit 2); the passes between read and write fixed point.
template <int L, int W, int IO>
__device__ __forceinline__ void fx_pass_load(Fx<W> &re, Fx<W> &im, const FxArgs &a,
                                             const uint32_t *src, uint32_t b, uint32_t i)
{
    const uint32_t N = a.N, M = 2 * N;
    if (IO & 1) {
        const uint32_t *in = a.in + a.in_stride * b;
        const int32_t eo = a.eoff[b];
        Mp<L> x;
        mp_load(x, in, M, i);
        uint32_t bad = fx_from_mp<L, W>(re, x, a.lim, eo);
        mp_load(x, in, M, N + i);
        bad |= fx_from_mp<L, W>(im, x, a.lim, eo);
        if (bad) atomicOr(a.flag, 1u);
    } else {
        fx_load(re, src, M, i);
        fx_load(im, src, M, N + i);
    }
}

template <int L, int W, bool FULL, int IO>
__device__ __forceinline__ void fx_pass_store(const Fx<W> &re, const Fx<W> &im, const FxArgs &a,
                                              uint32_t *dst, uint32_t b, uint32_t o)
{
    const uint32_t N = a.N, M = 2 * N;
    if (IO & 2) {
        uint32_t *out = a.out + a.out_stride * b;
        const int32_t eo = a.eoff[b];
        Mp<L> z;
        fx_to_mp<W, L, FULL>(z, re, a.P, eo);
        mp_store(z, out, M, o);
        fx_to_mp<W, L, FULL>(z, im, a.P, eo);
        mp_store(z, out, M, N + o);
    } else {
        fx_store(re, dst, M, o);
        fx_store(im, dst, M, N + o);
    }
}

// ---- block exponent -----------------------------------------------------------
//
// The header word of a nonzero real is (exp << 1) | sign, of a zero
// INT32_MIN | sign, so the arithmetic shift hdr >> 1 is the exponent of a
// nonzero real and -2^30 (below every exponent MPFR allows) for a zero.

constexpr int32_t kNoExp = -(1 << 30);

__device__ __forceinline__ int32_t wave_max(int32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// eoff[b] from the largest exponent m of transform b's inputs: 0 when all
// are zero; an m outside the limb format's input range is flagged
__device__ __forceinline__ void eoff_store(const FxArgs &a, uint32_t b, int32_t m)
{
    if (m == kNoExp) m = 0;                                  // all zero
    if (m >= kInputExpLimit || m <= -kInputExpLimit) {
        atomicOr(a.flag, 1u);
        m = 0;
    }
    a.eoff[b] = m;
}

// grid (eoff_nblk, batch): block c of transform b writes the largest
// exponent of its part of the 2N header words to eoff_part[b nblk + c];
// with one block per transform (2N <= 8192) it writes eoff[b] itself and
// the final kernel is not launched
__global__ void __launch_bounds__(256)
fx_eoff_part_kernel(FxArgs a)
{
    __shared__ int32_t red[4];
    const uint32_t M = 2 * a.N, nb = a.eoff_nblk, c = blockIdx.x, b = blockIdx.y;
    const uint32_t chunk = (M + nb - 1) / nb, i0 = c * chunk;
    const uint32_t i1 = i0 + chunk < M ? i0 + chunk : M;
    const int32_t *h = reinterpret_cast<const int32_t *>(a.in + a.in_stride * b);
    int32_t m0 = kNoExp, m1 = kNoExp;
    uint32_t i = i0 + threadIdx.x;
#pragma unroll 1
    for (; i + 256 < i1; i += 512) {
        const int32_t u = h[i] >> 1, v = h[i + 256] >> 1;
        m0 = u > m0 ? u : m0;
        m1 = v > m1 ? v : m1;
    }
    if (i < i1) {
        const int32_t u = h[i] >> 1;
        m0 = u > m0 ? u : m0;
    }
    int32_t m = wave_max(m0 > m1 ? m0 : m1);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = red[0];
#pragma unroll
        for (int w = 1; w < 4; w++) m = red[w] > m ? red[w] : m;
        if (nb == 1) eoff_store(a, b, m);
        else a.eoff_part[(size_t)b * nb + c] = m;
    }
}

// one thread per transform: eoff[b] from the partial maxima
__global__ void __launch_bounds__(256)
fx_eoff_final_kernel(FxArgs a)
{
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.batch) return;
    const int32_t *p = a.eoff_part + (size_t)b * a.eoff_nblk;
    int32_t m = kNoExp;
#pragma unroll 1
    for (uint32_t c = 0; c < a.eoff_nblk; c++) m = p[c] > m ? p[c] : m;
    eoff_store(a, b, m);
}

hipError_t fx_eoff_launch(const FxArgs &a, hipStream_t st)
{
    if (a.eoff_nblk != fx_eoff_blocks(a.N)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fx_eoff_part_kernel, dim3(a.eoff_nblk, a.batch), dim3(256), 0, st, a);
    if (a.eoff_nblk > 1)
        hipLaunchKernelGGL(fx_eoff_final_kernel, dim3((a.batch + 255) / 256), dim3(256), 0, st,
                           a);
    return hipGetLastError();
}

// float SoA (hdr + L limbs) -> fixed point; flags inputs beyond the range
template <int L, int W>
__global__ void __launch_bounds__(256)
fx_in_kernel(FxArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t M = 2 * a.N;
    if (i >= M) return;
    Mp<L> x;
    mp_load(x, a.in + a.in_stride * blockIdx.y, M, i);
    Fx<W> z;
    if (fx_from_mp<L, W>(z, x, a.lim, a.eoff[blockIdx.y])) atomicOr(a.flag, 1u);
    fx_store(z, a.buf0 + a.stride * blockIdx.y, M, i);
}

// fixed point -> float SoA at P bits, rounded to nearest
template <int W, int L, bool FULL>
__global__ void __launch_bounds__(256)
fx_out_kernel(FxArgs a, const uint32_t *src)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t M = 2 * a.N;
    if (i >= M) return;
    Fx<W> x;
    fx_load(x, src + a.stride * blockIdx.y, M, i);
    Mp<L> z;
    fx_to_mp<W, L, FULL>(z, x, a.P, a.eoff[blockIdx.y]);
    mp_store(z, a.out + a.out_stride * blockIdx.y, M, i);
}

// One radix-4 pass.  Blocks are ordered transform-fastest, so the blocks
// that read one twiddle for the different transforms of a batch run side
// by side and share it in L2.
template <int L, int W, bool FULL, int IO>
__global__ void __launch_bounds__(256, kFxBlocks<W>)
fx_r4_kernel(FxArgs a, const uint32_t *src, uint32_t *dst, uint32_t Ns, uint32_t toff)
{
    const uint32_t N = a.N, Q = N >> 2;
    const uint32_t b = blockIdx.x % a.batch, jb = blockIdx.x / a.batch;
    const uint32_t r = threadIdx.x & 3u;
    const uint32_t j = jb * (blockDim.x >> 2) + (threadIdx.x >> 2);
    if (j >= Q) return;                            // whole quads leave together
    src += a.stride * b;
    dst += a.stride * b;
    Fx<W> re, im;
    fx_pass_load<L, W, IO>(re, im, a, src, b, j + r * Q);
    const uint32_t k = j & (Ns - 1);
    if (r != 0 && k != 0) {
        // v_r *= w^t, t = r k N / (4 Ns), by three products (Gauss):
        //   k3 = im (c + d), k2 = re (d - c), k1 = (re + im) c,
        //   re' = k1 - k3, im' = k1 + k2   for w^t = c + i d.
        // re and im wait in LDS slots 0 and 1 and are replaced there by k2
        // and k3, so only s = re + im stays in registers across a product.
        // The pass's twiddles are stored in (k, r) order from toff, so the
        // lanes of a wave read consecutive entries.
        extern __shared__ uint32_t lds[];
        const uint32_t lt = threadIdx.x, nt = blockDim.x;
        const uint32_t q = toff + 3 * k + (r - 1), TT = a.ntw, MT = 3 * TT;
        Fx<W> s;
        fx_add(s, re, im);
        fx_put(lds, nt, lt, 0, re);
        fx_put(lds, nt, lt, 1, im);
        uint32_t T[FxDig<W>::DT];
#pragma unroll 1
        for (int it = 0; it < 3; it++) {
            // table reals: c at [0, TT), d - c at [TT, 2 TT), c + d at [2 TT, 3 TT)
            const uint32_t u = it == 0 ? 2u : (it == 1 ? 1u : 0u);
            const uint32_t tneg = tw_load<W>(T, a.tw, MT, u * TT + q);
            if (it < 2) fx_get(re, lds, nt, lt, 1 - it);    // im, then re
            else re = s;
            fx_mul(re, re, T, tneg);
            if (it < 2) fx_put(lds, nt, lt, 1 - it, re);     // k3 -> 1, k2 -> 0
        }
        // re = k1 here
        fx_get(im, lds, nt, lt, 1);                          // k3
        fx_sub(s, re, im);
        fx_get(im, lds, nt, lt, 0);                          // k2
        fx_add(im, re, im);
        re = s;
    }
    fx_dft4_quad(re, im, r);
    const uint32_t o = (j - k) * 4 + k + fx_dft4_rout(r) * Ns;
    fx_pass_store<L, W, FULL, IO>(re, im, a, dst, b, o);
}

// The radix-2 pass (first pass when log2 N is odd: Ns = 1, no twiddles).
template <int L, int W, bool FULL, int IO>
__global__ void __launch_bounds__(256, kFxBlocks<W>)
fx_r2_kernel(FxArgs a, const uint32_t *src, uint32_t *dst)
{
    const uint32_t N = a.N, H = N >> 1;
    const uint32_t b = blockIdx.x % a.batch, jb = blockIdx.x / a.batch;
    const uint32_t r = threadIdx.x & 1u;
    const uint32_t j = jb * (blockDim.x >> 1) + (threadIdx.x >> 1);
    if (j >= H) return;
    src += a.stride * b;
    dst += a.stride * b;
    Fx<W> re, im;
    fx_pass_load<L, W, IO>(re, im, a, src, b, j + r * H);
    fx_dft2_pair(re, im, r);
    fx_pass_store<L, W, FULL, IO>(re, im, a, dst, b, 2 * j + r);
}

// (re, im) *= c + i d, table entry q: three products (Gauss), k3 = im (c + d),
// k2 = re (d - c), k1 = (re + im) c; re' = k1 - k3, im' = k1 + k2
template <int W>
__device__ __forceinline__ void fx_cmul_tab(Fx<W> &re, Fx<W> &im, const uint32_t *tw,
                                            uint32_t TT, uint32_t q)
{
    const uint32_t MT = 3 * TT;
    uint32_t T[FxDig<W>::DT];
    Fx<W> s, k2, k3;
    fx_add(s, re, im);
    uint32_t tn = tw_load<W>(T, tw, MT, 2 * TT + q);
    fx_mul(k3, im, T, tn);
    tn = tw_load<W>(T, tw, MT, TT + q);
    fx_mul(k2, re, T, tn);
    tn = tw_load<W>(T, tw, MT, q);
    fx_mul(s, s, T, tn);
    fx_sub(re, s, k3);
    fx_add(im, s, k2);
}

// DFT4 in registers: v_q <- sum_r v_r (-i)^(r q), exact
template <int W>
__device__ __forceinline__ void fx_dft4_reg(Fx<W> (&re)[4], Fx<W> (&im)[4])
{
    Fx<W> t0r, t0i, t1r, t1i, t2r, t2i, t3r, t3i;
    fx_add(t0r, re[0], re[2]);
    fx_add(t0i, im[0], im[2]);
    fx_sub(t1r, re[0], re[2]);
    fx_sub(t1i, im[0], im[2]);
    fx_add(t2r, re[1], re[3]);
    fx_add(t2i, im[1], im[3]);
    fx_sub(t3r, re[1], re[3]);
    fx_sub(t3i, im[1], im[3]);
    fx_add(re[0], t0r, t2r);                     // t0 + t2
    fx_add(im[0], t0i, t2i);
    fx_sub(re[2], t0r, t2r);                     // t0 - t2
    fx_sub(im[2], t0i, t2i);
    fx_add(re[1], t1r, t3i);                     // t1 - i t3
    fx_sub(im[1], t1i, t3r);
    fx_sub(re[3], t1r, t3i);                     // t1 + i t3
    fx_add(im[3], t1i, t3r);
}

// Radix-16 pass of sub-transform size Ns, one butterfly on four waves (up
// to kFxl16MaxW words per real).  Butterfly g = 64 blockIdx + lane is j = g
// mod N/16 of transform g / (N/16) (small transforms share a block).  Wave w
// holds positions r = w + 4 r2 (r2 = 0 .. 3): it loads v_r = x[j + r N/16]
// (all four before converting any, IO bit 1), multiplies v_r by w^(r k N /
// (16 Ns)), k = j mod Ns (entry toff + (r - 1) Ns + k), takes the DFT4 over
// r2 and the internal twiddles w16^(w q1) (entries cent + w q1; w16^4 = -i
// exactly), and puts U_w(q1) into LDS slot 4 w + q1; after the barrier wave
// w takes the DFT4 over r1 of U_r1(w) and writes y_(w + 4 q2) to (j - k) 16 +
// k + (w + 4 q2) Ns (rounding to limbs, IO bit 2).  LDS: 16 slots of 2 W
// words x 64 lanes.
template <int L, int W, bool FULL, int IO>
__global__ void __launch_bounds__(256)
fx_r16_kernel(FxArgs a, const uint32_t *src, uint32_t *dst, uint32_t Ns, uint32_t toff,
              uint32_t cent, uint32_t lgS)
{
    extern __shared__ uint32_t xs[];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t g = (uint64_t)blockIdx.x * 64 + lane;
    const uint32_t N = a.N, M = 2 * N, S = N >> 4;
    const bool live = g < ((uint64_t)a.batch << lgS);
    const uint32_t b = live ? (uint32_t)(g >> lgS) : 0u, j = (uint32_t)g & (S - 1);
    const uint32_t k = j & (Ns - 1);
    auto slot = [&](uint32_t sl, uint32_t word) { return xs + ((sl * 2 * W + word) << 6) + lane; };
    Fx<W> re[4], im[4];
    if (live) {
        if constexpr ((IO & 1) != 0) {
            const uint32_t *in = a.in + a.in_stride * b;
            const int32_t eo = a.eoff[b];
            Mp<L> xr[4], xi[4];
#pragma unroll
            for (int r2 = 0; r2 < 4; r2++) {
                mp_load(xr[r2], in, M, j + (w + 4 * r2) * S);
                mp_load(xi[r2], in, M, N + j + (w + 4 * r2) * S);
            }
            uint32_t bad = 0;
#pragma unroll
            for (int r2 = 0; r2 < 4; r2++) {
                bad |= fx_from_mp<L, W>(re[r2], xr[r2], a.lim, eo);
                bad |= fx_from_mp<L, W>(im[r2], xi[r2], a.lim, eo);
            }
            if (bad) atomicOr(a.flag, 1u);
        } else {
            const uint32_t *sb = src + a.stride * b;
#pragma unroll
            for (int r2 = 0; r2 < 4; r2++) {
                fx_load(re[r2], sb, M, j + (w + 4 * r2) * S);
                fx_load(im[r2], sb, M, N + j + (w + 4 * r2) * S);
            }
        }
        if (k) {
#pragma unroll
            for (int r2 = 0; r2 < 4; r2++) {
                const uint32_t r = w + 4 * r2;
                if (r) fx_cmul_tab<W>(re[r2], im[r2], a.tw, a.ntw, toff + (r - 1) * Ns + k);
            }
        }
        fx_dft4_reg<W>(re, im);                    // [q1] = U_w(q1)
#pragma unroll
        for (int q1 = 1; q1 < 4; q1++) {
            const uint32_t m = w * q1;
            if (m == 4) {                          // -i (x + i y) = y - i x
                Fx<W> z;
#pragma unroll
                for (int t = 0; t < W; t++) z.w[t] = 0u;
                Fx<W> nre;
                fx_sub(nre, z, re[q1]);
                re[q1] = im[q1];
                im[q1] = nre;
            } else if (m) {
                fx_cmul_tab<W>(re[q1], im[q1], a.tw, a.ntw, cent + m);
            }
        }
#pragma unroll
        for (int q1 = 0; q1 < 4; q1++)
#pragma unroll
            for (int t = 0; t < W; t++) {
                *slot(4 * w + q1, t) = re[q1].w[t];
                *slot(4 * w + q1, W + t) = im[q1].w[t];
            }
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int r1 = 0; r1 < 4; r1++)
#pragma unroll
            for (int t = 0; t < W; t++) {
                re[r1].w[t] = *slot(4 * r1 + w, t);
                im[r1].w[t] = *slot(4 * r1 + w, W + t);
            }
        fx_dft4_reg<W>(re, im);                    // [q2] = y_(w + 4 q2)
    }
    if (Ns >= 64) {
        // a wave's outputs y_q lie on 64 consecutive positions of every row
        if (!live) return;
        const uint32_t o = (j - k) * 16 + k;
        uint32_t *db = dst + a.stride * b;
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++)
            fx_pass_store<L, W, FULL, IO>(re[q2], im[q2], a, db, b, o + (w + 4 * q2) * Ns);
        return;
    }
    // Ns < 64: the block's outputs fill 1024 consecutive positions of the
    // (transform, position) order, output q of lane l at e = 16 (l - k) + k
    // + q Ns, but a wave's stores would touch them 4 bytes in 4 Ns: they
    // go through LDS (row t of the words at t 1024, each 32-word group
    // rotated by its index so neither side conflicts on the banks) and
    // leave as 64 consecutive words per store
    auto swz = [](uint32_t e) { return (e & ~31u) | ((e + (e >> 5)) & 31u); };
    __syncthreads();                               // the slots have been read
    if (live) {
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++) {
            const uint32_t e = swz(16 * (lane - k) + k + (w + 4 * q2) * Ns);
#pragma unroll
            for (int t = 0; t < W; t++) {
                xs[t * 1024 + e] = re[q2].w[t];
                xs[(W + t) * 1024 + e] = im[q2].w[t];
            }
        }
    }
    __syncthreads();
    const uint32_t lgN = lgS + 4;
    const uint64_t total = (uint64_t)a.batch << lgN;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const uint32_t e = 256 * m + 64 * w + lane;
        const uint64_t G = (uint64_t)blockIdx.x * 1024 + e;
        if (G >= total) break;
        const uint32_t bb = (uint32_t)(G >> lgN), pos = (uint32_t)G & (N - 1), es = swz(e);
        Fx<W> yr, yi;
#pragma unroll
        for (int t = 0; t < W; t++) {
            yr.w[t] = xs[t * 1024 + es];
            yi.w[t] = xs[(W + t) * 1024 + es];
        }
        fx_pass_store<L, W, FULL, IO>(yr, yi, a, dst + a.stride * bb, bb, pos);
    }
}

// ---------------------------------------------------------------------------

// one pass of the plan (fxl_plan in kernels.h); IO: bit 1 converts the
// input, bit 2 the output
template <int L, int W, bool FULL, int IO>
static void launch_fx_pass(const FxArgs &a, const MwPass &P, uint32_t cent, const uint32_t *src,
                           uint32_t *dst, hipStream_t st)
{
    const uint32_t N = a.N;
    if (P.R == 2) {
        dim3 g(((N / 2) * 2 + 255) / 256 * a.batch);
        hipLaunchKernelGGL((fx_r2_kernel<L, W, FULL, IO>), g, dim3(256), 0, st, a, src, dst);
    } else if (P.R == 4) {
        dim3 g(((N / 4) * 4 + 255) / 256 * a.batch);
        const size_t sh = sizeof(uint32_t) * 2 * W * 256;     // two parking slots per lane
        hipLaunchKernelGGL((fx_r4_kernel<L, W, FULL, IO>), g, dim3(256), sh, st, a, src, dst,
                           P.Ns, P.toff);
    } else if constexpr (W <= kFxl16MaxW) {
        uint32_t lgS = 0;
        while ((2u << lgS) <= N >> 4) lgS++;
        const uint64_t bf = (uint64_t)a.batch << lgS;
        const size_t sh = sizeof(uint32_t) * 16 * 2 * W * 64;
        hipLaunchKernelGGL((fx_r16_kernel<L, W, FULL, IO>), dim3((uint32_t)((bf + 63) / 64)),
                           dim3(256), sh, st, a, src, dst, P.Ns, P.toff, cent, lgS);
    }
}

// Stages [s0, s1) of the transform: one per pass (buffer 0 -> 1 -> 0 ...),
// the first converting the input on the fly, the last the output.  N = 1
// has no pass: stage 0 converts in, stage 1 out.
template <int L, int W, bool FULL>
static hipError_t launch_fx(const FxArgs &a, int s0, int s1, hipStream_t st)
{
    static_assert(FxDig<W>::B == fx_lane_digit_bits(W), "host twiddle digits differ");
    const uint32_t N = a.N, M = 2 * N;
    MwPass ps[32];
    uint32_t nent = 0;
    const int npass = N > 1 ? fxl_plan(N, W, ps, &nent) : 0;
    auto in_range = [&](int s) { return s >= s0 && s < s1; };
    if (npass == 0) {
        dim3 gc((M + 255) / 256, a.batch);
        if (in_range(0)) hipLaunchKernelGGL((fx_in_kernel<L, W>), gc, dim3(256), 0, st, a);
        if (in_range(1))
            hipLaunchKernelGGL((fx_out_kernel<W, L, FULL>), gc, dim3(256), 0, st, a,
                               (const uint32_t *)a.buf0);
        return hipGetLastError();
    }
    uint32_t *src = a.buf1, *dst = a.buf0;
    for (int s = 0; s < npass; s++) {
        const MwPass &P = ps[s];
        // the radix-2 kernel has no twiddles; radix 16 only where the plan puts it
        if ((P.R == 2 && P.Ns != 1) || (P.R == 16 && W > kFxl16MaxW)) return hipErrorInvalidValue;
        if (in_range(s)) {
            const bool first = s == 0, last = s == npass - 1;
            if (first && last) launch_fx_pass<L, W, FULL, 3>(a, P, nent, src, dst, st);
            else if (first) launch_fx_pass<L, W, FULL, 1>(a, P, nent, src, dst, st);
            else if (last) launch_fx_pass<L, W, FULL, 2>(a, P, nent, src, dst, st);
            else launch_fx_pass<L, W, FULL, 0>(a, P, nent, src, dst, st);
        }
        std::swap(src, dst);
    }
    return hipGetLastError();
}

hipError_t fx_fft_launch(int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st)
{
    const bool full = (P == 32 * L);
#define CALLX(Lv) (full ? launch_fx<Lv, Lv + 2, true>(a, s0, s1, st) \
                        : launch_fx<Lv, Lv + 2, false>(a, s0, s1, st))
    switch (L) {
    case 1: return CALLX(1);
    case 2: return CALLX(2);
    case 3: return CALLX(3);
    case 4: return CALLX(4);
    case 6: return CALLX(6);
    case 8: return CALLX(8);
    case 12: return CALLX(12);
    case 16: return CALLX(16);
    case 24: return CALLX(24);
    case 32: return CALLX(32);
    case 64: return CALLX(64);
    default: return hipErrorInvalidValue;
    }
#undef CALLX
}

}  // namespace mpfft
