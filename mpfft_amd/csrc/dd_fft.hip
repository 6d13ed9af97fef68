// Double-word form (MPFFT_FORM_DD): the FFT on unevaluated sums of two
// binary64 numbers, hi + lo with |lo| <= ulp(hi) / 2, for precisions of 53 to
// 106 bits.
//
// The reference has no such form; it computes every operation with MPFR at
// P bits (lab/serial/mp/serial_mp_fft3.c:121-253).  Here each real is a
// double-word number and every operation is an error-free transformation
// on the FP64 units: TwoSum (Knuth) and Fast2Sum (Dekker) for sums, the FMA
// residual a*b - RN(a*b) for products, and their renormalisation.  FP64 is
// used only in those steps.  The algorithms and their error bounds are the
// double-word ones of Joldes, Muller and Popescu ("Tight and rigorous error
// bounds for basic building blocks of double-word arithmetic", ACM TOMS
// 44(2), 2017): DWPlusDW (their Algorithm 6, relative error below 3u^2) and
// DWTimesDW with two FMA corrections (Algorithm 11, below 5u^2), u = 2^-53.
//
// Structure: radix-4 Stockham passes, a radix-2 pass first when log2 N is
// odd, as the fixed form (fx_fft.hip).  The first pass reads the SoA-limb
// input and splits every real into hi + lo exactly (P <= 106); the last pass
// rounds hi + lo to P bits (correctly rounded: both parts are exact P-bit
// values, and their sum is rounded once by the limb adder of mp_arith.h)
// and writes SoA limbs.  In between, a real is one double2 {hi, lo}; element
// e of transform b (re: e < N, im: N + e) is at b * stride + e.
#include "kernels.h"
#include "mp_arith.h"

namespace mpfft {

namespace {

struct DD {
    double hi, lo;
};

struct CDD {
    DD re, im;
};

// Knuth's TwoSum: s + e == a + b exactly
__device__ __forceinline__ DD two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

// Dekker's Fast2Sum, |a| >= |b| (or a == 0): s + e == a + b exactly
__device__ __forceinline__ DD fast_two_sum(double a, double b)
{
    const double s = a + b;
    const double e = b - (s - a);
    return {s, e};
}

__device__ __forceinline__ DD dd_add(const DD &x, const DD &y)
{
    const DD s = two_sum(x.hi, y.hi);
    const DD t = two_sum(x.lo, y.lo);
    const DD v = fast_two_sum(s.hi, s.lo + t.hi);
    return fast_two_sum(v.hi, t.lo + v.lo);
}

__device__ __forceinline__ DD dd_neg(const DD &x) { return {-x.hi, -x.lo}; }

__device__ __forceinline__ DD dd_sub(const DD &x, const DD &y) { return dd_add(x, dd_neg(y)); }

__device__ __forceinline__ DD dd_mul(const DD &x, const DD &y)
{
    const double p = x.hi * y.hi;
    double e = __builtin_fma(x.hi, y.hi, -p);          // exact residual of the product
    e = __builtin_fma(x.hi, y.lo, e);
    e = __builtin_fma(x.lo, y.hi, e);
    return fast_two_sum(p, e);
}

__device__ __forceinline__ CDD c_add(const CDD &a, const CDD &b)
{
    return {dd_add(a.re, b.re), dd_add(a.im, b.im)};
}

__device__ __forceinline__ CDD c_sub(const CDD &a, const CDD &b)
{
    return {dd_sub(a.re, b.re), dd_sub(a.im, b.im)};
}

// a * w
__device__ __forceinline__ CDD c_mul(const CDD &a, const CDD &w)
{
    return {dd_sub(dd_mul(a.re, w.re), dd_mul(a.im, w.im)),
            dd_add(dd_mul(a.re, w.im), dd_mul(a.im, w.re))};
}

// a - i b and a + i b
__device__ __forceinline__ CDD c_sub_ib(const CDD &a, const CDD &b)
{
    return {dd_add(a.re, b.im), dd_sub(a.im, b.re)};
}

__device__ __forceinline__ CDD c_add_ib(const CDD &a, const CDD &b)
{
    return {dd_sub(a.re, b.im), dd_add(a.im, b.re)};
}

// P-bit limb value (53 <= P <= 106) -> hi + lo, exact.  The significand's
// top 53 bits rounded to nearest (ties to even) give hi; the remainder, at
// most 53 bits wide because P <= 106, is lo.  Returns 1 when the exponent
// lies outside [-lim, lim] (binary64 could not hold lo).
template <int L>
__device__ __forceinline__ uint32_t dd_from_mp(DD &z, const Mp<L> &x, int lim)
{
    if (x.zero) {
        z.hi = x.neg ? -0.0 : 0.0;
        z.lo = 0.0;
        return 0;
    }
    if (x.e > lim || x.e < -lim) {
        z.hi = 0.0;
        z.lo = 0.0;
        return 1;
    }
    // value = 0.m * 2^e; A = significand bits 1..64, B = bits 65..128
    const uint64_t A = (uint64_t)x.m[0] << 32 | (L > 1 ? x.m[1] : 0u);
    const uint64_t B = (uint64_t)(L > 2 ? x.m[2] : 0u) << 32 | (L > 3 ? x.m[3] : 0u);
    uint64_t h = A >> 11;                                        // bits 1..53
    int64_t r = (int64_t)(((A & 0x7ffull) << 42) | (B >> 22));  // bits 54..106, < 2^53
    const int64_t half = 1ll << 52;
    if (r > half || (r == half && (h & 1u))) {
        h += 1;
        r -= 2 * half;
    }
    const double hi = __builtin_ldexp((double)h, x.e - 53);
    const double lo = __builtin_ldexp((double)r, x.e - 106);
    z.hi = x.neg ? -hi : hi;
    z.lo = x.neg ? -lo : lo;
    return 0;
}

// a binary64 value as an exact limb value (53 bits <= 32 L)
template <int L>
__device__ __forceinline__ void mp_from_double(Mp<L> &z, double v)
{
    if (v == 0.0) {
        set_zero(z, __builtin_signbit(v) ? 1u : 0u);
        return;
    }
    int ex = 0;
    const double f = __builtin_frexp(__builtin_fabs(v), &ex);  // [0.5, 1)
    const uint64_t m = (uint64_t)__builtin_ldexp(f, 53) << 11;  // 53 bits at the top
#pragma unroll
    for (int j = 0; j < L; j++) z.m[j] = 0u;
    z.m[0] = (uint32_t)(m >> 32);
    z.m[1] = (uint32_t)m;
    z.e = ex;
    z.neg = v < 0.0 ? 1u : 0u;
    z.zero = 0u;
}

// hi + lo rounded once to P bits
template <int L, bool FULL>
__device__ __forceinline__ void dd_to_mp(Mp<L> &z, const DD &x, int P)
{
    Mp<L> a, b;
    mp_from_double(a, x.hi);
    mp_from_double(b, x.lo);
    mp_add<L, FULL>(z, a, 0u, b, 0u, P);
}

// element e of transform b: from the limb input (IN) or a double2 buffer
template <int L, bool IN>
__device__ __forceinline__ CDD dd_load(const FxArgs &a, const double2 *src, uint32_t b,
                                       uint32_t e)
{
    CDD v;
    if constexpr (IN) {
        const uint32_t *in = a.in + a.in_stride * b;
        Mp<L> x;
        uint32_t bad = 0;
        mp_load(x, in, 2 * a.N, e);
        bad |= dd_from_mp(v.re, x, a.lim);
        mp_load(x, in, 2 * a.N, a.N + e);
        bad |= dd_from_mp(v.im, x, a.lim);
        if (bad) atomicOr(a.flag, 1u);
    } else {
        const double2 *p = src + a.stride * b;
        const double2 r = p[e], i = p[a.N + e];
        v.re = {r.x, r.y};
        v.im = {i.x, i.y};
    }
    return v;
}

template <int L, bool FULL, bool OUT>
__device__ __forceinline__ void dd_store(const FxArgs &a, double2 *dst, uint32_t b, uint32_t e,
                                         const CDD &v)
{
    if constexpr (OUT) {
        uint32_t *out = a.out + a.out_stride * b;
        Mp<L> z;
        dd_to_mp<L, FULL>(z, v.re, a.P);
        mp_store(z, out, 2 * a.N, e);
        dd_to_mp<L, FULL>(z, v.im, a.P);
        mp_store(z, out, 2 * a.N, a.N + e);
    } else {
        double2 *p = dst + a.stride * b;
        p[e] = make_double2(v.re.hi, v.re.lo);
        p[a.N + e] = make_double2(v.im.hi, v.im.lo);
    }
}

// radix-2 pass without twiddles (the first pass when log2 N is odd)
template <int L, bool FULL, bool IN, bool OUT>
__global__ void __launch_bounds__(256)
dd_r2_kernel(FxArgs a, const double2 *src, double2 *dst)
{
    const uint32_t H = a.N >> 1, j = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (j >= H) return;
    const CDD v0 = dd_load<L, IN>(a, src, b, j), v1 = dd_load<L, IN>(a, src, b, j + H);
    dd_store<L, FULL, OUT>(a, dst, b, 2 * j, c_add(v0, v1));
    dd_store<L, FULL, OUT>(a, dst, b, 2 * j + 1, c_sub(v0, v1));
}

// Radix-4 Stockham pass of sub-transform size Ns: thread j (< N/4) reads
// v_r = x[j + r N/4], multiplies v_r by w^(r k N / (4 Ns)), k = j mod Ns
// (table entry toff + 3k + r - 1: re then im, each a double2 {hi, lo}), and
// writes y_q = sum_r v_r (-i)^(r q) to (j - k) 4 + k + q Ns.
template <int L, bool FULL, bool IN, bool OUT>
__global__ void __launch_bounds__(256)
dd_r4_kernel(FxArgs a, const double2 *src, double2 *dst, uint32_t Ns, uint32_t toff)
{
    const uint32_t Q = a.N >> 2, j = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (j >= Q) return;
    const uint32_t k = j & (Ns - 1);
    CDD v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = dd_load<L, IN>(a, src, b, j + r * Q);
    if (k) {
        const double2 *tw = reinterpret_cast<const double2 *>(a.tw);
#pragma unroll
        for (int r = 1; r < 4; r++) {
            const uint32_t t = toff + 3 * k + r - 1;
            const double2 wr = tw[2 * t], wi = tw[2 * t + 1];
            v[r] = c_mul(v[r], CDD{{wr.x, wr.y}, {wi.x, wi.y}});
        }
    }
    const CDD t0 = c_add(v[0], v[2]), t1 = c_sub(v[0], v[2]);
    const CDD t2 = c_add(v[1], v[3]), t3 = c_sub(v[1], v[3]);
    const uint32_t o = (j - k) * 4 + k;
    dd_store<L, FULL, OUT>(a, dst, b, o, c_add(t0, t2));
    dd_store<L, FULL, OUT>(a, dst, b, o + Ns, c_sub_ib(t1, t3));
    dd_store<L, FULL, OUT>(a, dst, b, o + 2 * Ns, c_sub(t0, t2));
    dd_store<L, FULL, OUT>(a, dst, b, o + 3 * Ns, c_add_ib(t1, t3));
}

// N = 1: the transform is the identity; the input split and rounded back
// (exact for P <= 106), with its range check
template <int L, bool FULL>
__global__ void __launch_bounds__(256)
dd_copy_kernel(FxArgs a)
{
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.batch) return;
    dd_store<L, FULL, true>(a, nullptr, b, 0, dd_load<L, true>(a, nullptr, b, 0));
}

template <int L, bool FULL, bool IN, bool OUT>
void launch_pass(const FxArgs &a, uint32_t R, const double2 *src, double2 *dst, uint32_t Ns,
                 uint32_t toff, hipStream_t st)
{
    const uint32_t threads = a.N / R;
    const dim3 grid((threads + 255) / 256, a.batch);
    if (R == 2)
        hipLaunchKernelGGL((dd_r2_kernel<L, FULL, IN, OUT>), grid, dim3(256), 0, st, a, src, dst);
    else
        hipLaunchKernelGGL((dd_r4_kernel<L, FULL, IN, OUT>), grid, dim3(256), 0, st, a, src, dst,
                           Ns, toff);
}

template <int L, bool FULL>
hipError_t launch_dd(const FxArgs &a, int s0, int s1, hipStream_t st)
{
    const uint32_t N = a.N;
    if (N == 1) {
        if (s0 <= 0 && s1 > 0)
            hipLaunchKernelGGL((dd_copy_kernel<L, FULL>), dim3((a.batch + 255) / 256), dim3(256),
                               0, st, a);
        return hipGetLastError();
    }
    uint32_t lg = 0;
    while ((1u << lg) < N) lg++;
    const int npass = fx_passes(N);
    double2 *src = reinterpret_cast<double2 *>(a.buf1), *dst = reinterpret_cast<double2 *>(a.buf0);
    uint32_t Ns = 1, toff = 0;
    for (int s = 0; s < npass; s++) {
        const uint32_t R = (s == 0 && (lg & 1)) ? 2u : 4u;
        if (s >= s0 && s < s1) {
            const bool first = s == 0, last = s == npass - 1;
            if (first && last) launch_pass<L, FULL, true, true>(a, R, src, dst, Ns, toff, st);
            else if (first) launch_pass<L, FULL, true, false>(a, R, src, dst, Ns, toff, st);
            else if (last) launch_pass<L, FULL, false, true>(a, R, src, dst, Ns, toff, st);
            else launch_pass<L, FULL, false, false>(a, R, src, dst, Ns, toff, st);
        }
        if (R == 4) toff += 3 * Ns;
        Ns *= R;
        std::swap(src, dst);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t dd_fft_launch(int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st)
{
    const bool full = (P == 32 * L);
    switch (L) {
    case 2: return full ? launch_dd<2, true>(a, s0, s1, st) : launch_dd<2, false>(a, s0, s1, st);
    case 3: return full ? launch_dd<3, true>(a, s0, s1, st) : launch_dd<3, false>(a, s0, s1, st);
    case 4: return full ? launch_dd<4, true>(a, s0, s1, st) : launch_dd<4, false>(a, s0, s1, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mpfft
