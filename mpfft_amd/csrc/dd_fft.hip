// Double-word and quad-word forms (MPFFT_FORM_DD, MPFFT_FORM_QD): the FFT on
// unevaluated sums of K = 2 or 4 binary64 numbers, for precisions of 53 to
// 106, resp. 212 bits.
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
// Quad-word numbers c0 + c1 + c2 + c3 follow the same principles (see
// below).
//
// Structure: Stockham passes (mw_plan in kernels.h), radix 16 on four waves
// per butterfly for both forms.  The first pass reads
// the SoA-limb input and splits every real into its words exactly; the last
// pass rounds the words' sum to P bits and writes SoA limbs (double-word:
// the correctly rounded sum of two binary64 numbers, dd_conv.h; quad-word:
// the limb adder of mp_arith.h).  In between, a real is K/2 double2.
#include "dd_conv.h"
#include "kernels.h"
#include "mp_arith.h"

namespace mpfft {

namespace {

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

// ---- quad-word numbers (K = 4) ---------------------------------------------
//
// c0 + c1 + c2 + c3, each word at most about half an ulp of the one before.
// Products and sums collect their partial terms by order (a term of order k
// is ~u^k times the largest), carry every rounding error of orders 0 .. 2
// exactly to the next order with TwoSum / the FMA residual, round the order-3
// sum, and renormalise: relative error of order u^4 ~ 2^-212 per operation
// (a few units of it), as for the usual quad-word algorithms.

struct QD {
    double c[4];
};

struct CQD {
    QD re, im;
};

// four overlapping words (roughly decreasing) -> four: a bottom-up TwoSum
// pass leaves the rounded total in c0 and each step's error below it, a
// top-down pass pushes what each word holds beyond half an ulp of the one
// above into the next (exact: the words' sum is unchanged)
__device__ __forceinline__ QD q_renorm(double c0, double c1, double c2, double c3)
{
    DD s = two_sum(c2, c3);
    c2 = s.hi; c3 = s.lo;
    s = two_sum(c1, c2);
    c1 = s.hi; c2 = s.lo;
    s = two_sum(c0, c1);
    c0 = s.hi; c1 = s.lo;
    s = two_sum(c1, c2);
    c1 = s.hi; c2 = s.lo;
    s = two_sum(c2, c3);
    c2 = s.hi; c3 = s.lo;
    return {{c0, c1, c2, c3}};
}

__device__ __forceinline__ QD q_add(const QD &a, const QD &b)
{
    const DD s0 = two_sum(a.c[0], b.c[0]);            // order 0 (+ its error, order 1)
    const DD s1 = two_sum(a.c[1], b.c[1]);            // order 1
    const DD s2 = two_sum(a.c[2], b.c[2]);            // order 2
    const DD t1 = two_sum(s1.hi, s0.lo);              // order 1 total
    const DD t2 = two_sum(s2.hi, s1.lo);              // order 2 ...
    const DD u2 = two_sum(t2.hi, t1.lo);              // ... total
    const double o3 = ((a.c[3] + b.c[3]) + s2.lo) + (t2.lo + u2.lo);
    return q_renorm(s0.hi, t1.hi, u2.hi, o3);
}

__device__ __forceinline__ QD q_neg(const QD &x) { return {{-x.c[0], -x.c[1], -x.c[2], -x.c[3]}}; }

__device__ __forceinline__ QD q_sub(const QD &a, const QD &b) { return q_add(a, q_neg(b)); }

__device__ __forceinline__ DD two_prod(double a, double b)
{
    const double p = a * b;
    return {p, __builtin_fma(a, b, -p)};
}

__device__ __forceinline__ QD q_mul(const QD &a, const QD &b)
{
    const DD p00 = two_prod(a.c[0], b.c[0]);
    const DD p01 = two_prod(a.c[0], b.c[1]), p10 = two_prod(a.c[1], b.c[0]);
    const DD p02 = two_prod(a.c[0], b.c[2]), p11 = two_prod(a.c[1], b.c[1]),
             p20 = two_prod(a.c[2], b.c[0]);
    // order 1: p00.lo + p01 + p10
    const DD s1 = two_sum(p01.hi, p10.hi);
    const DD s1b = two_sum(s1.hi, p00.lo);
    // order 2: p02 + p11 + p20 + the order-1 residuals, every addition exact
    // (a rounded sum of order-2 terms would lose an order-3 error)
    const DD s2 = two_sum(p02.hi, p11.hi);
    const DD s2b = two_sum(s2.hi, p20.hi);
    const DD s2c = two_sum(s2b.hi, p01.lo);
    const DD s2d = two_sum(s2c.hi, p10.lo);
    const DD s2e = two_sum(s2d.hi, s1.lo);
    const DD s2f = two_sum(s2e.hi, s1b.lo);
    // order 3: the remaining products and residuals, rounded
    double o3 = __builtin_fma(a.c[0], b.c[3], a.c[1] * b.c[2]);
    o3 = __builtin_fma(a.c[2], b.c[1], o3);
    o3 = __builtin_fma(a.c[3], b.c[0], o3);
    o3 += ((p02.lo + p11.lo) + (p20.lo + s2.lo)) + ((s2b.lo + s2c.lo) + (s2d.lo + s2e.lo)) +
          s2f.lo;
    return q_renorm(p00.hi, s1b.hi, s2f.hi, o3);
}

// ---- word-type traits: one real of K binary64 words ---------------------------

template <int K> struct Word;
template <> struct Word<2> {
    using R = DD;
    using C = CDD;
    static __device__ __forceinline__ R add(const R &x, const R &y) { return dd_add(x, y); }
    static __device__ __forceinline__ R sub(const R &x, const R &y) { return dd_sub(x, y); }
    static __device__ __forceinline__ R mul(const R &x, const R &y) { return dd_mul(x, y); }
    static __device__ __forceinline__ R neg(const R &x) { return dd_neg(x); }
    static __device__ __forceinline__ void get(R &x, const double2 *p) { x = {p[0].x, p[0].y}; }
    static __device__ __forceinline__ void put(double2 *p, const R &x)
    {
        p[0] = make_double2(x.hi, x.lo);
    }
};
template <> struct Word<4> {
    using R = QD;
    using C = CQD;
    static __device__ __forceinline__ R add(const R &x, const R &y) { return q_add(x, y); }
    static __device__ __forceinline__ R sub(const R &x, const R &y) { return q_sub(x, y); }
    static __device__ __forceinline__ R mul(const R &x, const R &y) { return q_mul(x, y); }
    static __device__ __forceinline__ R neg(const R &x) { return q_neg(x); }
    static __device__ __forceinline__ void get(R &x, const double2 *p)
    {
        const double2 u = p[0], v = p[1];
        x = {{u.x, u.y, v.x, v.y}};
    }
    static __device__ __forceinline__ void put(double2 *p, const R &x)
    {
        p[0] = make_double2(x.c[0], x.c[1]);
        p[1] = make_double2(x.c[2], x.c[3]);
    }
};

template <int K> using RW_ = typename Word<K>::R;
template <int K> using CW_ = typename Word<K>::C;

template <int K>
__device__ __forceinline__ CW_<K> w_add(const CW_<K> &a, const CW_<K> &b)
{
    return {Word<K>::add(a.re, b.re), Word<K>::add(a.im, b.im)};
}

template <int K>
__device__ __forceinline__ CW_<K> w_sub(const CW_<K> &a, const CW_<K> &b)
{
    return {Word<K>::sub(a.re, b.re), Word<K>::sub(a.im, b.im)};
}

template <int K>
__device__ __forceinline__ CW_<K> w_mul(const CW_<K> &a, const CW_<K> &w)
{
    return {Word<K>::sub(Word<K>::mul(a.re, w.re), Word<K>::mul(a.im, w.im)),
            Word<K>::add(Word<K>::mul(a.re, w.im), Word<K>::mul(a.im, w.re))};
}

template <int K>
__device__ __forceinline__ CW_<K> w_sub_ib(const CW_<K> &a, const CW_<K> &b)   // a - i b
{
    return {Word<K>::add(a.re, b.im), Word<K>::sub(a.im, b.re)};
}

template <int K>
__device__ __forceinline__ CW_<K> w_add_ib(const CW_<K> &a, const CW_<K> &b)   // a + i b
{
    return {Word<K>::sub(a.re, b.im), Word<K>::add(a.im, b.re)};
}

// P-bit limb value times 2^-eoff (the block exponent) -> quad-word, exact
// for P <= 212: the significand cut into four 53-bit pieces (all of the
// value's sign), then renormalised.  A scaled exponent below -lim (the last
// piece would leave the normal binary64 range; the value is below 2^-lim of
// the transform's largest input) is taken as zero; one above lim returns 1.
template <int L>
__device__ __forceinline__ uint32_t qd_from_mp(QD &z, const Mp<L> &x, int lim, int32_t eoff)
{
    z = {{0.0, 0.0, 0.0, 0.0}};
    const int32_t e = x.e - eoff;
    if (x.zero || e < -lim) {
        z.c[0] = x.neg ? -0.0 : 0.0;
        return 0;
    }
    if (e > lim) return 1;
    double c[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // 53 bits from significand bit 53 i on (bit 0 = the leading one)
        const int s = 53 * i, li = s >> 5, off = s & 31;
        const uint64_t w0 = li < L ? x.m[li] : 0u, w1 = li + 1 < L ? x.m[li + 1] : 0u,
                       w2 = li + 2 < L ? x.m[li + 2] : 0u;
        uint64_t top = (w0 << 32) | w1;
        if (off) top = (top << off) | (w2 >> (32 - off));
        const double v = __builtin_ldexp((double)(top >> 11), e - 53 * (i + 1));
        c[i] = x.neg ? -v : v;
    }
    z = q_renorm(c[0], c[1], c[2], c[3]);
    return 0;
}

// c0 + c1 + c2 + c3 rounded to P bits: the lower words summed at the full
// 32 L bits, then one rounding to P (P <= 212 < 32 L - 40 for L = 8; below,
// the partial sums lie ~2^-53 under the result, so their rounding cannot
// move the final one but at an exact tie)
template <int L, bool FULL>
__device__ __forceinline__ void qd_to_mp(Mp<L> &z, const QD &x, int P, int32_t eoff)
{
    Mp<L> a, b;
    mp_from_double(a, x.c[2]);
    mp_from_double(b, x.c[3]);
    mp_add<L, true>(b, a, 0u, b, 0u, 32 * L);
    mp_from_double(a, x.c[1]);
    mp_add<L, true>(b, a, 0u, b, 0u, 32 * L);
    mp_from_double(a, x.c[0]);
    mp_add<L, FULL>(z, a, 0u, b, 0u, P);
    if (!z.zero) z.e += eoff;
}

template <int K, int L>
__device__ __forceinline__ uint32_t w_from_mp(RW_<K> &z, const Mp<L> &x, int lim, int32_t eoff)
{
    if constexpr (K == 2) return dd_from_mp(z, x, lim, eoff);
    else return qd_from_mp(z, x, lim, eoff);
}

template <int K, int L, bool FULL>
__device__ __forceinline__ void w_to_mp(Mp<L> &z, const RW_<K> &x, int P, int32_t eoff)
{
    if constexpr (K == 2) dd_to_mp<L>(z, x, P, eoff);
    else qd_to_mp<L, FULL>(z, x, P, eoff);
}

// ---- passes ------------------------------------------------------------------
//
// Element e of transform b (re: e < N, im: N + e) is K/2 double2 at
// b * stride + (K/2) e.

// elements e[0..R) of transform b: from the limb input (IN) or a word
// buffer.  The limb input is loaded for all R elements before any is
// converted (a conversion branches on its value, and a load behind a branch
// waits for it), and an input beyond the range sets `bad`, which the caller
// reports once.
template <int K, int L, bool IN, int R>
__device__ __forceinline__ void w_load(const FxArgs &a, const double2 *src, uint32_t b,
                                       const uint32_t (&e)[R], CW_<K> (&v)[R], uint32_t &bad)
{
    if constexpr (IN) {
        const uint32_t *in = a.in + a.in_stride * b;
        const int32_t eo = a.eoff[b];
        Mp<L> xr[R], xi[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            mp_load(xr[r], in, 2 * a.N, e[r]);
            mp_load(xi[r], in, 2 * a.N, a.N + e[r]);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            bad |= w_from_mp<K, L>(v[r].re, xr[r], a.lim, eo);
            bad |= w_from_mp<K, L>(v[r].im, xi[r], a.lim, eo);
        }
    } else {
        const double2 *p = src + a.stride * b;
#pragma unroll
        for (int r = 0; r < R; r++) {
            Word<K>::get(v[r].re, p + (K / 2) * e[r]);
            Word<K>::get(v[r].im, p + (K / 2) * (a.N + e[r]));
        }
    }
}

template <int K, int L, bool FULL, bool OUT>
__device__ __forceinline__ void w_store(const FxArgs &a, double2 *dst, uint32_t b, uint32_t e,
                                        const CW_<K> &v)
{
    if constexpr (OUT) {
        uint32_t *out = a.out + a.out_stride * b;
        const int32_t eo = a.eoff[b];
        Mp<L> z;
        w_to_mp<K, L, FULL>(z, v.re, a.P, eo);
        mp_store(z, out, 2 * a.N, e);
        w_to_mp<K, L, FULL>(z, v.im, a.P, eo);
        mp_store(z, out, 2 * a.N, a.N + e);
    } else {
        double2 *p = dst + a.stride * b;
        Word<K>::put(p + (K / 2) * e, v.re);
        Word<K>::put(p + (K / 2) * (a.N + e), v.im);
    }
}

// radix-2 pass without twiddles (the first pass when log2 N is odd)
template <int K, int L, bool FULL, bool IN, bool OUT>
__global__ void __launch_bounds__(256)
mw_r2_kernel(FxArgs a, const double2 *src, double2 *dst)
{
    const uint32_t H = a.N >> 1, j = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (j >= H) return;
    uint32_t bad = 0;
    CW_<K> v[2];
    w_load<K, L, IN>(a, src, b, {j, j + H}, v, bad);
    if (bad) atomicOr(a.flag, 1u);
    w_store<K, L, FULL, OUT>(a, dst, b, 2 * j, w_add<K>(v[0], v[1]));
    w_store<K, L, FULL, OUT>(a, dst, b, 2 * j + 1, w_sub<K>(v[0], v[1]));
}

// Radix-4 Stockham pass of sub-transform size Ns: thread j (< N/4) reads
// v_r = x[j + r N/4], multiplies v_r by w^(r k N / (4 Ns)), k = j mod Ns
// (table entry t = toff + (r - 1) Ns + k: K double2, the real part's K/2 then
// the imaginary part's), and writes y_q = sum_r v_r (-i)^(r q) to
// (j - k) 4 + k + q Ns.
template <int K, int L, bool FULL, bool IN, bool OUT>
__global__ void __launch_bounds__(256)
mw_r4_kernel(FxArgs a, const double2 *src, double2 *dst, uint32_t Ns, uint32_t toff)
{
    const uint32_t Q = a.N >> 2, j = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (j >= Q) return;
    const uint32_t k = j & (Ns - 1);
    CW_<K> v[4];
    uint32_t bad = 0;
    w_load<K, L, IN>(a, src, b, {j, j + Q, j + 2 * Q, j + 3 * Q}, v, bad);
    if (bad) atomicOr(a.flag, 1u);
    if (k) {
        const double2 *tw = reinterpret_cast<const double2 *>(a.tw);
#pragma unroll
        for (int r = 1; r < 4; r++) {
            const double2 *t = tw + (size_t)K * (toff + (r - 1) * Ns + k);
            CW_<K> w;
            Word<K>::get(w.re, t);
            Word<K>::get(w.im, t + K / 2);
            v[r] = w_mul<K>(v[r], w);
        }
    }
    const CW_<K> t0 = w_add<K>(v[0], v[2]), t1 = w_sub<K>(v[0], v[2]);
    const CW_<K> t2 = w_add<K>(v[1], v[3]), t3 = w_sub<K>(v[1], v[3]);
    const uint32_t o = (j - k) * 4 + k;
    w_store<K, L, FULL, OUT>(a, dst, b, o, w_add<K>(t0, t2));
    w_store<K, L, FULL, OUT>(a, dst, b, o + Ns, w_sub_ib<K>(t1, t3));
    w_store<K, L, FULL, OUT>(a, dst, b, o + 2 * Ns, w_sub<K>(t0, t2));
    w_store<K, L, FULL, OUT>(a, dst, b, o + 3 * Ns, w_add_ib<K>(t1, t3));
}

template <int K>
__device__ __forceinline__ CW_<K> w_neg_i(const CW_<K> &a)        // -i a
{
    return {a.im, Word<K>::neg(a.re)};
}

// forward DFT4 in place: x_q <- sum_r x_r (-i)^(r q)
template <int K>
__device__ __forceinline__ void w_dft4(CW_<K> &x0, CW_<K> &x1, CW_<K> &x2, CW_<K> &x3)
{
    const CW_<K> t0 = w_add<K>(x0, x2), t1 = w_sub<K>(x0, x2);
    const CW_<K> t2 = w_add<K>(x1, x3), t3 = w_sub<K>(x1, x3);
    x0 = w_add<K>(t0, t2);
    x1 = w_sub_ib<K>(t1, t3);
    x2 = w_sub<K>(t0, t2);
    x3 = w_add_ib<K>(t1, t3);
}

// One complex value of K words in the exchange buffer of the radix-16 pass:
// K/2 double2 of the real part, then K/2 of the imaginary part, each a row
// of 64 lanes (a wave reads or writes 1 KB runs).
template <int K>
__device__ __forceinline__ void w_xput(double2 (*x)[64], uint32_t l, const CW_<K> &v)
{
    if constexpr (K == 2) {
        x[0][l] = make_double2(v.re.hi, v.re.lo);
        x[1][l] = make_double2(v.im.hi, v.im.lo);
    } else {
        x[0][l] = make_double2(v.re.c[0], v.re.c[1]);
        x[1][l] = make_double2(v.re.c[2], v.re.c[3]);
        x[2][l] = make_double2(v.im.c[0], v.im.c[1]);
        x[3][l] = make_double2(v.im.c[2], v.im.c[3]);
    }
}

template <int K>
__device__ __forceinline__ void w_xget(CW_<K> &v, double2 (*x)[64], uint32_t l)
{
    if constexpr (K == 2) {
        const double2 r = x[0][l], i = x[1][l];
        v.re = {r.x, r.y};
        v.im = {i.x, i.y};
    } else {
        const double2 r0 = x[0][l], r1 = x[1][l], i0 = x[2][l], i1 = x[3][l];
        v.re = {{r0.x, r0.y, r1.x, r1.y}};
        v.im = {{i0.x, i0.y, i1.x, i1.y}};
    }
}

// The same words in the output staging of the radix-16 pass: word pair h
// of element e at st[h 1024 + e]
template <int K>
__device__ __forceinline__ void w_sput(double2 *st, uint32_t e, const CW_<K> &v)
{
    if constexpr (K == 2) {
        st[e] = make_double2(v.re.hi, v.re.lo);
        st[1024 + e] = make_double2(v.im.hi, v.im.lo);
    } else {
        st[e] = make_double2(v.re.c[0], v.re.c[1]);
        st[1024 + e] = make_double2(v.re.c[2], v.re.c[3]);
        st[2048 + e] = make_double2(v.im.c[0], v.im.c[1]);
        st[3072 + e] = make_double2(v.im.c[2], v.im.c[3]);
    }
}

template <int K>
__device__ __forceinline__ void w_sget(CW_<K> &v, const double2 *st, uint32_t e)
{
    if constexpr (K == 2) {
        const double2 r = st[e], i = st[1024 + e];
        v.re = {r.x, r.y};
        v.im = {i.x, i.y};
    } else {
        const double2 r0 = st[e], r1 = st[1024 + e], i0 = st[2048 + e], i1 = st[3072 + e];
        v.re = {{r0.x, r0.y, r1.x, r1.y}};
        v.im = {{i0.x, i0.y, i1.x, i1.y}};
    }
}

// Radix-16 Stockham pass of sub-transform size Ns, one butterfly on four
// waves.  Butterfly g = 64 blockIdx + lane is j = g mod N/16 of transform
// g / (N/16) (small transforms share a block).  Wave w reads v_r = x[j + r
// N/16] for r = w + 4 r2 (r2 = 0 .. 3), multiplies v_r by w^(r k N / (16 Ns)),
// k = j mod Ns (entry toff + (r - 1) Ns + k), takes the DFT4 over r2 (the DFT16
// as 4 x 4: U_w(q1) = sum_r2 v_(w + 4 r2) (-i)^(r2 q1)) and the internal
// twiddles w16^(w q1) (entries cent + w q1; w16^4 = -i exactly), and puts
// U_w(q1) into the exchange buffer.  After the barrier wave w holds q1 = w:
// the DFT4 over r1 of U_r1(w) gives the outputs y_(w + 4 q2), written to
// (j - k) 16 + k + (w + 4 q2) Ns.  A thread holds four complex values, so
// the pass runs at several waves per SIMD where one thread per butterfly
// (sixteen values) ran at three, and the first and last passes convert
// from and to limbs in it.
template <int K, int L, bool FULL, bool IN, bool OUT>
__global__ void __launch_bounds__(256)
mw_r16_kernel(FxArgs a, const double2 *src, double2 *dst, uint32_t Ns, uint32_t toff,
              uint32_t cent, uint32_t lgS)
{
    __shared__ double2 xs[16][K][64];          // (r1, q1), then the words
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t g = (uint64_t)blockIdx.x * 64 + lane;
    const uint32_t S = a.N >> 4;
    const bool live = g < ((uint64_t)a.batch << lgS);
    const uint32_t b = live ? (uint32_t)(g >> lgS) : 0u, j = (uint32_t)g & (S - 1);
    const uint32_t k = j & (Ns - 1);
    const double2 *tw = reinterpret_cast<const double2 *>(a.tw);
    auto entry = [&](uint32_t t) {
        CW_<K> e;
        Word<K>::get(e.re, tw + (size_t)K * t);
        Word<K>::get(e.im, tw + (size_t)K * t + K / 2);
        return e;
    };
    CW_<K> v[4];
    if (live) {
        uint32_t bad = 0;
        w_load<K, L, IN>(a, src, b, {j + w * S, j + (w + 4) * S, j + (w + 8) * S, j + (w + 12) * S},
                         v, bad);
        if (bad) atomicOr(a.flag, 1u);
        if (k) {
#pragma unroll
            for (int r2 = 0; r2 < 4; r2++) {
                const uint32_t r = w + 4 * r2;
                if (r) v[r2] = w_mul<K>(v[r2], entry(toff + (r - 1) * Ns + k));
            }
        }
        w_dft4<K>(v[0], v[1], v[2], v[3]);     // v[q1] = U_w(q1)
#pragma unroll
        for (int q1 = 1; q1 < 4; q1++) {
            const uint32_t m = w * q1;
            if (m == 4) v[q1] = w_neg_i<K>(v[q1]);
            else if (m) v[q1] = w_mul<K>(v[q1], entry(cent + m));
        }
#pragma unroll
        for (int q1 = 0; q1 < 4; q1++) w_xput<K>(xs[4 * w + q1], lane, v[q1]);
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int r1 = 0; r1 < 4; r1++) w_xget<K>(v[r1], xs[4 * r1 + w], lane);
        w_dft4<K>(v[0], v[1], v[2], v[3]);     // v[q2] = y_(w + 4 q2)
    }
    if (Ns >= 16) {
        // a wave's outputs y_q lie on runs of at least 16 consecutive
        // positions (256 bytes and more)
        if (!live) return;
        const uint32_t o = (j - k) * 16 + k;
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++)
            w_store<K, L, FULL, OUT>(a, dst, b, o + (w + 4 * q2) * Ns, v[q2]);
        return;
    }
    // Ns < 16: the block's outputs fill 1024 consecutive positions of the
    // (transform, position) order, output q of lane l at e = 16 (l - k) + k
    // + q Ns, but a wave would store them in runs of Ns elements 16 Ns
    // apart: they go through LDS (row h of the words at h 1024, each group of
    // 8 elements rotated by e / 16 so neither side conflicts on the banks)
    // and leave as runs of 64 consecutive elements (the first pass at
    // N = 2^16, 106 bits: 79 -> 64 us)
    double2 *st = &xs[0][0][0];
    auto swz = [](uint32_t e) { return (e & ~7u) | ((e + (e >> 4)) & 7u); };
    __syncthreads();                           // the slots have been read
    if (live) {
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++) {
            const uint32_t e = swz(16 * (lane - k) + k + (w + 4 * q2) * Ns);
            w_sput<K>(st, e, v[q2]);
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
        const uint32_t bb = (uint32_t)(G >> lgN), pos = (uint32_t)G & (a.N - 1), es = swz(e);
        CW_<K> y;
        w_sget<K>(y, st, es);
        w_store<K, L, FULL, OUT>(a, dst, bb, pos, y);
    }
}

// N = 1: the transform is the identity; the input split and rounded back
// (exact), with its range check
template <int K, int L, bool FULL>
__global__ void __launch_bounds__(256)
mw_copy_kernel(FxArgs a)
{
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.batch) return;
    uint32_t bad = 0;
    CW_<K> v[1];
    w_load<K, L, true>(a, nullptr, b, {0u}, v, bad);
    if (bad) atomicOr(a.flag, 1u);
    w_store<K, L, FULL, true>(a, nullptr, b, 0, v[0]);
}

template <int K, int L, bool FULL, bool IN, bool OUT>
void launch_pass(const FxArgs &a, const MwPass &ps, uint32_t cent, const double2 *src,
                 double2 *dst, hipStream_t st)
{
    const uint32_t threads = a.N / ps.R;
    const dim3 grid((threads + 255) / 256, a.batch);
    if (ps.R == 2)
        hipLaunchKernelGGL((mw_r2_kernel<K, L, FULL, IN, OUT>), grid, dim3(256), 0, st, a, src,
                           dst);
    else if (ps.R == 4)
        hipLaunchKernelGGL((mw_r4_kernel<K, L, FULL, IN, OUT>), grid, dim3(256), 0, st, a, src,
                           dst, ps.Ns, ps.toff);
    else {
        // four waves per block, one butterfly per lane
        uint32_t lgS = 0;
        while ((2u << lgS) <= a.N >> 4) lgS++;
        const uint64_t bf = (uint64_t)a.batch << lgS;
        hipLaunchKernelGGL((mw_r16_kernel<K, L, FULL, IN, OUT>), dim3((uint32_t)((bf + 63) / 64)),
                           dim3(256), 0, st, a, src, dst, ps.Ns, ps.toff, cent, lgS);
    }
}

template <int K, int L, bool FULL>
hipError_t launch_mw(const FxArgs &a, int s0, int s1, hipStream_t st)
{
    const uint32_t N = a.N;
    if (N == 1) {
        if (s0 <= 0 && s1 > 0)
            hipLaunchKernelGGL((mw_copy_kernel<K, L, FULL>), dim3((a.batch + 255) / 256),
                               dim3(256), 0, st, a);
        return hipGetLastError();
    }
    MwPass ps[32];
    uint32_t cent = 0;
    const int npass = mw_plan(N, K, ps, &cent);
    double2 *src = reinterpret_cast<double2 *>(a.buf1), *dst = reinterpret_cast<double2 *>(a.buf0);
    for (int s = 0; s < npass; s++) {
        if (s >= s0 && s < s1) {
            const bool first = s == 0, last = s == npass - 1;
            if (first && last) launch_pass<K, L, FULL, true, true>(a, ps[s], cent, src, dst, st);
            else if (first) launch_pass<K, L, FULL, true, false>(a, ps[s], cent, src, dst, st);
            else if (last) launch_pass<K, L, FULL, false, true>(a, ps[s], cent, src, dst, st);
            else launch_pass<K, L, FULL, false, false>(a, ps[s], cent, src, dst, st);
        }
        std::swap(src, dst);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t dd_fft_launch(int K, int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st)
{
    const bool full = (P == 32 * L);
#define CALLW(Kv, Lv) (full ? launch_mw<Kv, Lv, true>(a, s0, s1, st) \
                            : launch_mw<Kv, Lv, false>(a, s0, s1, st))
    if (K == 2) {
        switch (L) {
        case 2: return CALLW(2, 2);
        case 3: return CALLW(2, 3);
        case 4: return CALLW(2, 4);
        default: return hipErrorInvalidValue;
        }
    }
    if (K == 4) {
        switch (L) {
        case 2: return CALLW(4, 2);
        case 3: return CALLW(4, 3);
        case 4: return CALLW(4, 4);
        case 6: return CALLW(4, 6);
        case 8: return CALLW(4, 8);
        default: return hipErrorInvalidValue;
        }
    }
#undef CALLW
    return hipErrorInvalidValue;
}

}  // namespace mpfft
