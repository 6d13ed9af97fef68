// Multiprecision split-radix conjugate-pair DIT FFT on gfx950 ("plain
// form"): the reference's recursion (lab/serial/mp/serial_mp_fft3.c:121-253)
// run breadth-first as one leaf launch plus one launch per node size
// n = 8, 16, ..., N.  Every element sees the reference's exact sequence of
// correctly rounded P-bit operations, so results are bit-identical to
// MPFR's.
//
// Data layout in HBM (per transform): buffer = hdr int32[2N] then limb
// uint32[L][2N]; element i < N is the real part of X[i], N+i its imaginary
// part.  Limb-major SoA: a wave reading limb j of 32 consecutive elements
// touches one 128-byte segment.  Twiddles (SoA): every level n has its own
// run of n/4 entries, entry k = LUT[k N/n], so the lanes of a wave (k
// consecutive) read consecutive entries; rows of E entries: for the 3m3a
// order of CODE 3, LUTr, LUTs, LUTd (LUTi is not used,
// serial_mp_fft3.c:79-80); for the 4m2a orders of CODE 1 and CODE 2
// (serial_mp_fft1.c, serial_mp_fft2.c), LUTr then LUTi.
//
// Work mapping: two lanes (a DPP quad pair) per butterfly.  The even lane
// computes m1 = Z+ * w^k, the odd lane m2 = Z- * conj(w^k), with the same
// instruction stream (operand signs are data); the pair then swaps m1/m2
// and each lane produces two of the four outputs.  Butterflies are laid
// out node by node, k fastest: a wave's accesses to U0, U1, Z+ and Z- are
// dense runs, at the price of some divergence where k = 0 (no product) and
// k = n/8 (two products) sit beside general k (three products).
#include "mp_arith.h"
#include "kernels.h"

namespace mpfft {

__device__ __forceinline__ uint32_t wrap_idx(int32_t sn, uint32_t s, uint32_t j, uint32_t N)
{
    // x[(sn + j*s) mod N]: only j == 0 can be negative (|sn| < s)
    int64_t v = (int64_t)sn + (int64_t)j * s;
    return (uint32_t)(v < 0 ? v + N : v);
}

// Blocks per CU for the leaf and level kernels: two up to 32 limbs; one at
// 64 limbs, where a lane's numbers need more than 256 registers and the
// level kernel's LDS parking slots (2 x 65 words x 256 threads, 133 KB) fill
// a CU's LDS alone.
template <int L>
constexpr int kMpBlocks = L > 32 ? 1 : 2;

// The kernels below run each of their additions (and the level kernel each
// of its multiplications) through one instance of mp_add / mp_mul inside a
// loop whose trip count is uniform across the wave; operands are selected
// per iteration.  Inlining every operation separately would give ~200 KB of
// code at 1024 bits, far more than the instruction cache holds.

// Element-major copy of the input for the leaf gathers: element i becomes
// the W = L+1 (rounded up to 4) consecutive words hdr, limb 0..L-1, so a
// gather of one element is W/4 16-byte loads from one place instead of
// L+1 loads from L+1 rows.  Reads are coalesced per limb row; each lane
// writes its own contiguous W words.
template <int L>
__global__ void __launch_bounds__(256)
mp_to_aos_kernel(LeafArgs a)
{
    constexpr int W = Aos<L>::W;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t M = 2 * a.N;
    if (i >= M) return;
    const uint32_t *src = a.in + a.in_stride * blockIdx.y;
    uint4 *dst = reinterpret_cast<uint4 *>(a.aos + a.aos_stride * blockIdx.y + (size_t)i * W);
    const uint32_t j = a.in_swap ? (i < a.N ? i + a.N : i - a.N) : i;
    uint32_t t[W];
#pragma unroll
    for (int k = 0; k < W; k++) t[k] = k <= L ? src[(size_t)k * M + j] : 0u;
#pragma unroll
    for (int k = 0; k < W / 4; k++)
        dst[k] = make_uint4(t[4 * k], t[4 * k + 1], t[4 * k + 2], t[4 * k + 3]);
}

// Leaves (serial_mp_fft3.c:136-177), one lane per output real: lane
// (h, p) computes part h (0 real, 1 imaginary) of out[p], so the lanes of
// a wave store 64 consecutive words of every limb row.  With r = p - o the
// position inside its leaf (offset o, size n, input stride s = N/n):
//   n = 2: out[o+r] = x0 +- x1                                  (1 add)
//   n = 4: y  = x0 +- x2        (y0 for even r, y1 for odd r)
//          a  = x1 + x3 (a1, same part h) for even r,
//               x1 - x3 (a2, the other part) for odd r
//          out = y +- a, the signs of :166-176               (3 adds)
// Each add is the reference's own correctly rounded operation on the same
// operands, so recomputing y and a per lane changes no result bit.
// Inputs are gathered as whole elements from the element-major copy.
template <int L, bool FULL>
__device__ __forceinline__ void mp_leaf_value(Mp<L> &res, const LeafArgs &a, uint32_t b,
                                              uint32_t h, uint32_t p)
{
    const uint32_t N = a.N;
    const uint32_t *in = a.aos + a.aos_stride * b;
    const int P = a.P;
    const int32_t sn = a.pos_sn[p];
    const uint32_t nr = a.pos_nr[p], n = nr & 0xffu, r = nr >> 8;
    const bool four = n == 4u;
    const uint32_t s = N / n;
    const uint32_t odd = r & 1u;
    const uint32_t hb = h * N, ha = (four && odd) ? (N - hb) : hb;   // part of a
    const int nit = four ? 3 : 1;
    Mp<L> u, v, y;
#pragma unroll 1
    for (int it = 0; it < nit; it++) {
        uint32_t fy;
        if (it == 0) {
            mp_load_aos(u, in, hb + wrap_idx(sn, s, 0, N));
            mp_load_aos(v, in, hb + wrap_idx(sn, s, four ? 2 : 1, N));
            fy = four ? odd : r;
        } else if (it == 1) {
            mp_load_aos(u, in, ha + wrap_idx(sn, s, 1, N));
            mp_load_aos(v, in, ha + wrap_idx(sn, s, 3, N));
            fy = odd;
        } else {
            u = y; v = res;
            fy = (r >> 1) ^ (odd & h);
        }
        mp_add<L, FULL>(res, u, 0, v, fy, P);
        if (it == 0) y = res;
    }
}

template <int L, bool FULL>
__global__ void __launch_bounds__(256, kMpBlocks<L>)
mp_leaf_kernel(LeafArgs a)
{
    const uint32_t N = a.N, M = 2 * N;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    const uint32_t h = t >= N ? 1u : 0u, p = t - h * N;
    uint32_t *out = a.data + a.stride * blockIdx.y;
    Mp<L> res;
    mp_leaf_value<L, FULL>(res, a, blockIdx.y, h, p);
    mp_store(res, out, M, t);
}

// One butterfly of a size-n node (serial_mp_fft3.c:184-248): the pair of
// lanes B = 0 (Z+) and B = 1 (Z-) of butterfly k of the node at offset o.
// The transform's reals sit in `data` as SoA rows of row stride M (hdr
// row, then limb rows), real parts at positions 0.., imaginary parts at Np
// + position: the HBM buffer (M = 2N, Np = N) or a subtree held in LDS (M =
// 2s, Np = s, positions relative to the subtree).  `lds` holds the two
// parking slots of each of the nt threads of the block.
template <int L, bool FULL, bool FOUR_M>
__device__ __forceinline__ void mp_butterfly(uint32_t *data, uint32_t M, uint32_t Np,
                                             uint32_t o, uint32_t k, uint32_t n, uint32_t B,
                                             int code, const uint32_t *lut, uint32_t lut_m,
                                             uint32_t lut_e, uint32_t lut_off, int P,
                                             uint32_t *lds, uint32_t nt, uint32_t lt)
{
    uint32_t *const out = data;
    const uint32_t q = n >> 2, N = Np;
    // CODE 1 (serial_mp_fft1.c:181-218) has no special butterflies
    const uint32_t kind = code == 1 ? 0u : (k == 0 ? 2u : (k == (n >> 3) ? 1u : 0u));
    const uint32_t zi = o + k + (B ? 3 * q : 2 * q);  // Z+ or Z-
    const uint32_t ui = o + k + (B ? q : 0);          // U0 or U1
    const uint32_t kt = lut_off + k;                  // this level's twiddle entry k

    // o1, o2: m1 = Z+ * w^k on the Z+ lane, m2 = Z- * conj(w^k) on the Z- lane.
    // Intermediates that must survive a multiplication wait in LDS slots
    // 0 and 1, so only the multiplication's own state is live across it.
    Mp<L> o1, o2;
    mp_load(o1, out, M, zi);
    mp_load(o2, out, M, N + zi);
    if (kind != 2) {
        // m4: the general butterfly in the 4m2a order of CODE 1 and CODE 2
        // (serial_mp_fft2.c:219-230, serial_mp_fft1.c:187-199); otherwise
        // the 3m3a order of CODE 3.
        const bool m4 = FOUR_M && kind == 0;
        // pre-additions.  general k, 3m3a (serial_mp_fft3.c:220-229):
        //   t = Zr +- Zi.  k = n/8 (:209-217): Zr +- Zi and -+Zr + Zi.
        Mp<L> r;
        const int npre = m4 ? 0 : (kind == 0 ? 1 : 2);
#pragma unroll 1
        for (int it = npre - 1; it >= 0; it--) {
            mp_add<L, FULL>(r, o1, it ? (B ^ 1u) : 0u, o2, it ? 0u : B, P);
            if (it) lds_put(lds, nt, lt, 1, r);        // second pre-sum -> slot 1
        }
        // products; every result goes to an LDS slot, so nothing but the
        // product's own state stays in registers across a multiplication.
        //   3m3a: T = t*LUTr -> 0; m.r = T -+ LUTd*Zi -> 1;
        //         m.i = T - LUTs*Zr (Z+ lane) or LUTs*Zr - T (Z- lane) -> 0.
        //   k = n/8: pre-sum j times LUTr[N/8] -> j.
        //   4m2a: A = LUTr*Zr -> 0; m.r = A +- LUTi*Zi -> 1;
        //         C = LUTr*Zi -> 0; m.i = C -+ LUTi*Zr -> 0.
        const int nmul = m4 ? 4 : (kind == 0 ? 3 : 2);
#pragma unroll 1
        for (int it = 0; it < nmul; it++) {
            Mp<L> f, w;
            uint32_t row;                              // twiddle row: r, s|i, d
            if (it == 0 && !m4) f = r;
            else if (kind == 1) lds_get(f, lds, nt, lt, 1);
            else mp_load(f, out, M, (m4 ? (it == 0 || it == 3) : it == 2) ? zi : N + zi);
            const uint32_t E = lut_e;
            if (kind == 1) row = 0;
            else if (m4) row = (it & 1) ? E : 0;
            else row = it == 0 ? 0 : (it == 1 ? 2 * E : E);
            mp_load(w, lut, lut_m, row + kt);
            mp_mul<L, FULL>(r, f, w, P);
            if (m4 ? (it & 1) : (kind == 0 && it > 0)) {
                Mp<L> T;
                lds_get(T, lds, nt, lt, 0);
                const uint32_t fx = (!m4 && it == 2) ? B : 0u;
                const uint32_t fy = (m4 && it == 1) ? B : (B ^ 1u);
                mp_add<L, FULL>(r, T, fx, r, fy, P);
            }
            lds_put(lds, nt, lt, it == 1 ? 1 : 0, r);
        }
        lds_get(o1, lds, nt, lt, kind == 1 ? 0 : 1);
        lds_get(o2, lds, nt, lt, kind == 1 ? 1 : 0);
    }
    // pair exchange: the Z+ lane gets m2, the Z- lane m1; then
    // Z+ lane: a1 = m1 + m2, Z- lane: a2 = m1 - m2 (serial_mp_fft3.c:234-238)
    Mp<L> v1, v2;
#pragma unroll 1
    for (int it = 0; it < 2; it++) {
        Mp<L> mine, other;
        mp_select(mine, (uint32_t)it, o2, o1);
        mp_swap_pair(other, mine);
        Mp<L> r;
        mp_add<L, FULL>(r, mine, B, other, 0, P);
        if (it == 0) v1 = r;
        else v2 = r;
    }
    // outputs (serial_mp_fft3.c:240-248):
    //   Z+ lane: X[k] = U0 + a1, X[k+n/2] = U0 - a1
    //   Z- lane: X[k+n/4] = U1 - i a2, X[k+3n/4] = U1 + i a2
    const uint32_t f0 = o + k + (B ? q : 0), s0 = f0 + 2 * q;
    Mp<L> U;
#pragma unroll 1
    for (int it = 0; it < 4; it++) {
        const uint32_t im = it >> 1;                 // 0: real parts, 1: imaginary parts
        if ((it & 1) == 0) mp_load(U, out, M, (im ? N : 0) + ui);
        Mp<L> x, r;
        // real: U.r +- (a1.r | a2.i); imaginary: U.i +- (a1.i | -a2.r)
        mp_select(x, B ^ im, v2, v1);
        const uint32_t fy = (uint32_t)(it & 1) ^ (im & B);
        mp_add<L, FULL>(r, U, 0, x, fy, P);
        mp_store(r, out, M, (im ? N : 0) + ((it & 1) ? s0 : f0));
    }
}

// Two waves per SIMD (256 registers per lane): one wave per SIMD leaves the
// dependent multiply chains nothing to overlap with.  Butterflies are laid
// out node-major: all k of a node on neighbouring lanes, so each load or
// store instruction covers dense runs of positions; the three kinds
// (general k, k = n/8, k = 0) diverge only around the products.
template <int L, bool FULL, bool FOUR_M>
__global__ void __launch_bounds__(256, kMpBlocks<L>)
mp_level_kernel(LevelArgs a)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t pair = tid >> 1, B = tid & 1u;   // B: 0 -> Z+ lane, 1 -> Z- lane
    const uint32_t q = a.n >> 2;
    const uint32_t node = pair / q;
    if (node >= a.cnt) return;
    extern __shared__ uint32_t lds[];
    mp_butterfly<L, FULL, FOUR_M>(a.data + a.stride * blockIdx.y, 2 * a.N, a.N,
                                  a.node_off[node], pair - node * q, a.n, B, a.code, a.lut,
                                  a.lut_m, a.lut_e, a.lut_off, a.P, lds, blockDim.x,
                                  threadIdx.x);
}

// A subtree finished in LDS (serial_mp_fft3.c:180-182 below node size s):
// workgroup (subtree, transform).  Its 2s reals live in LDS as SoA rows of
// 2s words (real parts, then imaginary parts, positions relative to the
// subtree); the leaves are computed from the element-major input copy
// straight into them, then each combine level n = 8 .. s runs on them with
// a barrier between levels (the same butterflies, operand for operand, as
// the level kernel), and the subtree is written back once.  The levels of a
// subtree are the node offsets of one subtree of its size (the tree below a
// node depends only on the node's size).
template <int L, bool FULL, bool FOUR_M>
__global__ void __launch_bounds__(256)
mp_sub_kernel(SubArgs a)
{
    extern __shared__ uint32_t lds[];
    const uint32_t nt = a.nthr, lt = threadIdx.x;
    uint32_t *const park = lds;                              // 2 (L+1) nt words
    uint32_t *const buf = lds + 2 * (L + 1) * nt;            // (L+1) x 2s words
    const uint32_t sd = a.sub[blockIdx.x];
    const uint32_t kind = sd & 1u, o = sd >> 1;
    const uint32_t lgs = a.lgs[kind], s = 1u << lgs, Ml = 2 * s;
    const uint32_t N = a.leaf.N, M = 2 * N;
    const int P = a.leaf.P;
    uint32_t *const out = a.leaf.data + a.leaf.stride * blockIdx.y;
    for (uint32_t idx = lt; idx < Ml; idx += nt) {
        const uint32_t h = idx >= s ? 1u : 0u;
        Mp<L> res;
        mp_leaf_value<L, FULL>(res, a.leaf, blockIdx.y, h, o + idx - h * s);
        lds_put(buf, Ml, idx, 0, res);
    }
    __syncthreads();
    for (uint32_t l = 3; l <= lgs; l++) {
        const uint32_t n = 1u << l, q = n >> 2;
        const uint32_t lanes = a.rel_cnt[kind][l] * (n >> 1);
        const uint32_t *const no = a.rel + a.rel_first[kind][l];
        for (uint32_t base = 0; base < lanes; base += nt) {
            // lanes come in pairs (nt and lanes are even): a pair is active
            // or idle as a whole, as the pair exchange needs
            const uint32_t tid = base + lt;
            if (tid < lanes) {
                const uint32_t pair = tid >> 1, node = pair / q;
                mp_butterfly<L, FULL, FOUR_M>(buf, Ml, s, no[node], pair - node * q, n,
                                              tid & 1u, a.code, a.lut, a.lut_m, a.lut_e,
                                              a.lut_off[l], P, park, nt, lt);
            }
        }
        __syncthreads();
    }
    for (uint32_t idx = lt; idx < Ml; idx += nt) {
        const uint32_t h = idx >= s ? 1u : 0u, g = h * N + o + idx - h * s;
#pragma unroll
        for (int j = 0; j <= L; j++) out[(size_t)j * M + g] = buf[j * Ml + idx];
    }
}

// Element-wise correctly rounded op on SoA arrays (device unit test of the
// arithmetic): op 0 add, 1 sub, 2 mul.
template <int L, bool FULL>
__global__ void __launch_bounds__(256)
mp_binop_kernel(int op, uint32_t count, const uint32_t *x, const uint32_t *y,
                uint32_t *z, int P)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    Mp<L> a, b, c;
    mp_load(a, x, count, i);
    mp_load(b, y, count, i);
    if (op == 2) mp_mul<L, FULL>(c, a, b, P);
    else mp_add<L, FULL>(c, a, 0, b, (uint32_t)op, P);
    mp_store(c, z, count, i);
}

// Closing step of the inverse transform by the swap trick
// (serial_mp_fft3.c:367-378 and :385): out.re = X.im / N, out.im = X.re / N
// for X the forward transform of the swapped input.  Division by N = 2^lgN
// only moves the exponent, so it is exact.
template <int L>
__global__ void __launch_bounds__(256)
mp_unswap_scale_kernel(uint32_t *data, uint64_t stride, uint32_t N, int32_t lgN)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint32_t *buf = data + stride * blockIdx.y;
    const uint32_t M = 2 * N;
    uint32_t ha = buf[i], hb = buf[N + i];
    auto scale = [lgN](uint32_t h) -> uint32_t {
        if (((int32_t)h & ~1) == kHdrZero) return h;
        return (uint32_t)((((int32_t)h >> 1) - lgN) * 2) | (h & 1u);
    };
    buf[i] = scale(hb);
    buf[N + i] = scale(ha);
#pragma unroll
    for (int j = 1; j <= L; j++) {
        uint32_t *row = buf + (size_t)j * M;
        const uint32_t a = row[i], b = row[N + i];
        row[i] = b;
        row[N + i] = a;
    }
}

// Error statistics of y against x, as the reference's test()
// (serial_mp_fft3.c:387-398): aE = x - y rounded at P bits, then per part
// (real, imaginary) sum x^2, sum aE^2 and max |aE|.  The sums are formed in
// double from the top 64 bits of each significand with a per-block common
// exponent (values near 2^-P*|x| are far below the double range), so they
// carry double rounding rather than the reference's P-bit rounding.  Per
// block and part: {sum x^2 mantissa, its exponent, sum aE^2 mantissa, its
// exponent, max log2|aE|}.
template <int L, bool FULL>
__global__ void __launch_bounds__(256)
mp_err_kernel(const uint32_t *x, const uint32_t *y, uint64_t stride, uint32_t N, int P,
              double *partial)
{
    __shared__ int32_t se[2][256];
    __shared__ double sv[2][256];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, M = 2 * N;
    const uint32_t lt = threadIdx.x;
    const uint32_t h = t >= N ? 1u : 0u;
    const bool live = t < M;
    int32_t ex = INT32_MIN, ed = INT32_MIN;
    double mx = 0.0, md = 0.0, lg = -INFINITY;
    if (live) {
        Mp<L> a, b, d;
        mp_load(a, x + stride * blockIdx.y, M, t);
        mp_load(b, y + stride * blockIdx.y, M, t);
        mp_add<L, FULL>(d, a, 0, b, 1, P);
        auto mant = [](const Mp<L> &v) {
            const double hi = (double)v.m[0], lo = (double)v.m[L > 1 ? 1 : 0] * (L > 1);
            return (hi * 4294967296.0 + lo) * 0x1p-64;         // in [0.5, 1)
        };
        if (!a.zero) { const double m = mant(a); mx = m * m; ex = 2 * a.e; }
        if (!d.zero) {
            const double m = mant(d);
            md = m * m; ed = 2 * d.e;
            lg = (double)d.e + log2(m);
        }
    }
    // per part: common exponent (max), then scaled sums; max of log2|aE|
    for (int q = 0; q < 2; q++) {
        const int32_t e = q ? ed : ex;
        const double v = q ? md : mx;
        for (uint32_t p = 0; p < 2; p++) se[p][lt] = (p == h) ? e : INT32_MIN;
        __syncthreads();
        for (uint32_t w = 128; w > 0; w >>= 1) {
            if (lt < w)
                for (int p = 0; p < 2; p++)
                    se[p][lt] = max(se[p][lt], se[p][lt + w]);
            __syncthreads();
        }
        const int32_t E0 = se[0][0], E1 = se[1][0];
        __syncthreads();
        for (uint32_t p = 0; p < 2; p++) {
            const int32_t E = p ? E1 : E0;
            sv[p][lt] = (p == h && e != INT32_MIN) ? ldexp(v, e - E) : 0.0;
        }
        __syncthreads();
        for (uint32_t w = 128; w > 0; w >>= 1) {
            if (lt < w)
                for (int p = 0; p < 2; p++) sv[p][lt] += sv[p][lt + w];
            __syncthreads();
        }
        if (lt == 0) {
            double *o = partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kErrPartial;
            for (int p = 0; p < 2; p++) {
                o[5 * p + 2 * q] = sv[p][0];
                o[5 * p + 2 * q + 1] = (double)(p ? E1 : E0);
            }
        }
        __syncthreads();
    }
    for (uint32_t p = 0; p < 2; p++) sv[p][lt] = (p == h) ? lg : -INFINITY;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (lt < w)
            for (int p = 0; p < 2; p++) sv[p][lt] = fmax(sv[p][lt], sv[p][lt + w]);
        __syncthreads();
    }
    if (lt == 0) {
        double *o = partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kErrPartial;
        o[4] = sv[0][0];
        o[9] = sv[1][0];
    }
}

// ---------------------------------------------------------------------------
// host-side launchers (one instantiation per supported limb count)

template <int L, bool FULL>
static hipError_t launch_fft(const LeafArgs &la, const LevelArgs *lv, int nlev,
                             uint32_t batch, hipStream_t st, const SubArgs *sub)
{
    if (la.npos) {
        dim3 grid((2 * la.N + 255) / 256, batch);
        hipLaunchKernelGGL((mp_to_aos_kernel<L>), grid, dim3(256), 0, st, la);
        if (L <= 16 && sub) {
            SubArgs sa = *sub;
            sa.leaf = la;
            const size_t sh = sizeof(uint32_t) * (L + 1) *
                              (2 * (size_t)sa.nthr + 2 * ((size_t)1 << sa.lgs[0]));
            if constexpr (L <= 16) {
                if (sa.code == 3)
                    hipLaunchKernelGGL((mp_sub_kernel<L, FULL, false>), dim3(sa.nsub, batch),
                                       dim3(sa.nthr), sh, st, sa);
                else
                    hipLaunchKernelGGL((mp_sub_kernel<L, FULL, true>), dim3(sa.nsub, batch),
                                       dim3(sa.nthr), sh, st, sa);
            }
        } else {
            hipLaunchKernelGGL((mp_leaf_kernel<L, FULL>), grid, dim3(256), 0, st, la);
        }
    }
    for (int i = 0; i < nlev; i++) {
        uint32_t thr = 2 * lv[i].cnt * (lv[i].n / 4);
        dim3 grid((thr + 255) / 256, batch);
        // two parking slots of L+1 words per thread
        const size_t sh = sizeof(uint32_t) * 2 * (L + 1) * 256;
        if (lv[i].code == 3)
            hipLaunchKernelGGL((mp_level_kernel<L, FULL, false>), grid, dim3(256), sh, st, lv[i]);
        else
            hipLaunchKernelGGL((mp_level_kernel<L, FULL, true>), grid, dim3(256), sh, st, lv[i]);
    }
    return hipGetLastError();
}

template <int L, bool FULL>
static hipError_t launch_binop(int op, uint32_t count, const uint32_t *x,
                               const uint32_t *y, uint32_t *z, int P, hipStream_t st)
{
    dim3 grid((count + 255) / 256);
    hipLaunchKernelGGL((mp_binop_kernel<L, FULL>), grid, dim3(256), 0, st,
                       op, count, x, y, z, P);
    return hipGetLastError();
}

// limb counts with compiled kernels; a precision uses the smallest one that
// holds it
// (64 limbs: the plain op orders and the fixed form on panels; the per-lane
// fixed form and the word forms stop below)
static const int kLimbs[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 64};

int limbs_for_precision(long P)
{
    long need = (P + 31) / 32;
    for (int L : kLimbs) if (L >= need) return L;
    return -1;
}

hipError_t mp_fft_launch(int L, int P, const LeafArgs &la, const LevelArgs *lv,
                         int nlev, uint32_t batch, hipStream_t st, const SubArgs *sub)
{
    const bool full = (P == 32 * L);
#define CALLF(Lv, F) launch_fft<Lv, F>(la, lv, nlev, batch, st, sub)
    switch (L) {
    case 1: return full ? CALLF(1, true) : CALLF(1, false);
    case 2: return full ? CALLF(2, true) : CALLF(2, false);
    case 3: return full ? CALLF(3, true) : CALLF(3, false);
    case 4: return full ? CALLF(4, true) : CALLF(4, false);
    case 6: return full ? CALLF(6, true) : CALLF(6, false);
    case 8: return full ? CALLF(8, true) : CALLF(8, false);
    case 12: return full ? CALLF(12, true) : CALLF(12, false);
    case 16: return full ? CALLF(16, true) : CALLF(16, false);
    case 24: return full ? CALLF(24, true) : CALLF(24, false);
    case 32: return full ? CALLF(32, true) : CALLF(32, false);
    case 64: return full ? CALLF(64, true) : CALLF(64, false);
    default: return hipErrorInvalidValue;
    }
#undef CALLF
}

template <int L>
static hipError_t launch_unswap(uint32_t *data, uint64_t stride, uint32_t N, int32_t lgN,
                                uint32_t batch, hipStream_t st)
{
    dim3 grid((N + 255) / 256, batch);
    hipLaunchKernelGGL((mp_unswap_scale_kernel<L>), grid, dim3(256), 0, st, data, stride, N, lgN);
    return hipGetLastError();
}

hipError_t mp_unswap_scale_launch(int L, uint32_t *data, uint64_t stride, uint32_t N,
                                  int32_t lgN, uint32_t batch, hipStream_t st)
{
#define CALLU(Lv) launch_unswap<Lv>(data, stride, N, lgN, batch, st)
    switch (L) {
    case 1: return CALLU(1);
    case 2: return CALLU(2);
    case 3: return CALLU(3);
    case 4: return CALLU(4);
    case 6: return CALLU(6);
    case 8: return CALLU(8);
    case 12: return CALLU(12);
    case 16: return CALLU(16);
    case 24: return CALLU(24);
    case 32: return CALLU(32);
    case 64: return CALLU(64);
    default: return hipErrorInvalidValue;
    }
#undef CALLU
}

template <int L, bool FULL>
static hipError_t launch_err(const uint32_t *x, const uint32_t *y, uint64_t stride, uint32_t N,
                             int P, double *partial, uint32_t batch, hipStream_t st)
{
    dim3 grid((2 * N + kErrBlock - 1) / kErrBlock, batch);
    hipLaunchKernelGGL((mp_err_kernel<L, FULL>), grid, dim3(kErrBlock), 0, st,
                       x, y, stride, N, P, partial);
    return hipGetLastError();
}

hipError_t mp_err_launch(int L, int P, const uint32_t *x, const uint32_t *y, uint64_t stride,
                         uint32_t N, double *partial, uint32_t batch, hipStream_t st)
{
    const bool full = (P == 32 * L);
#define CALLE(Lv, F) launch_err<Lv, F>(x, y, stride, N, P, partial, batch, st)
    switch (L) {
    case 1: return full ? CALLE(1, true) : CALLE(1, false);
    case 2: return full ? CALLE(2, true) : CALLE(2, false);
    case 3: return full ? CALLE(3, true) : CALLE(3, false);
    case 4: return full ? CALLE(4, true) : CALLE(4, false);
    case 6: return full ? CALLE(6, true) : CALLE(6, false);
    case 8: return full ? CALLE(8, true) : CALLE(8, false);
    case 12: return full ? CALLE(12, true) : CALLE(12, false);
    case 16: return full ? CALLE(16, true) : CALLE(16, false);
    case 24: return full ? CALLE(24, true) : CALLE(24, false);
    case 32: return full ? CALLE(32, true) : CALLE(32, false);
    case 64: return full ? CALLE(64, true) : CALLE(64, false);
    default: return hipErrorInvalidValue;
    }
#undef CALLE
}

hipError_t mp_binop_launch(int L, int P, int op, uint32_t count, const uint32_t *x,
                           const uint32_t *y, uint32_t *z, hipStream_t st)
{
    const bool full = (P == 32 * L);
#define CALLB(Lv, F) launch_binop<Lv, F>(op, count, x, y, z, P, st)
    switch (L) {
    case 1: return full ? CALLB(1, true) : CALLB(1, false);
    case 2: return full ? CALLB(2, true) : CALLB(2, false);
    case 3: return full ? CALLB(3, true) : CALLB(3, false);
    case 4: return full ? CALLB(4, true) : CALLB(4, false);
    case 6: return full ? CALLB(6, true) : CALLB(6, false);
    case 8: return full ? CALLB(8, true) : CALLB(8, false);
    case 12: return full ? CALLB(12, true) : CALLB(12, false);
    case 16: return full ? CALLB(16, true) : CALLB(16, false);
    case 24: return full ? CALLB(24, true) : CALLB(24, false);
    case 32: return full ? CALLB(32, true) : CALLB(32, false);
    case 64: return full ? CALLB(64, true) : CALLB(64, false);
    default: return hipErrorInvalidValue;
    }
#undef CALLB
}

}  // namespace mpfft

namespace mpfft {

// Chunk-major staging <-> SoA rows (one transform of N complex points, R =
// L + 1 rows).  Chunk c holds points [c cl, (c + 1) cl) as R rows of 2 cl
// words (real parts, then imaginary parts), so the host fills and uploads
// each chunk as one contiguous block while it converts the next.  Thread
// (row j, position e) moves one word; consecutive positions of one part of
// one chunk are consecutive on both sides.
__global__ void __launch_bounds__(256)
mp_chunk_relayout_kernel(const uint32_t *src, uint32_t *dst, uint32_t N, uint32_t R, uint32_t cl,
                         uint32_t to_soa)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (e >= 2 * N || j >= R) return;
    const uint32_t im = e >= N ? 1u : 0u, k = e - im * N;
    const uint32_t c = k / cl, loc = k - c * cl + im * cl;
    const size_t chunk = (size_t)c * R * 2 * cl + (size_t)j * 2 * cl + loc;
    const size_t soa = (size_t)j * 2 * N + e;
    if (to_soa) dst[soa] = src[chunk];
    else dst[chunk] = src[soa];
}

hipError_t mp_chunk_relayout_launch(const uint32_t *src, uint32_t *dst, uint32_t N, uint32_t R,
                                    uint32_t cl, bool to_soa, hipStream_t st)
{
    if (!N || !R || !cl || N % cl) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mp_chunk_relayout_kernel, dim3((2 * N + 255) / 256, R), dim3(256), 0, st,
                       src, dst, N, R, cl, to_soa ? 1u : 0u);
    return hipGetLastError();
}

}  // namespace mpfft
