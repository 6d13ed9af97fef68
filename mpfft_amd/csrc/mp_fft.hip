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
// touches one 128-byte segment.  Twiddles: hdr[3Q] + limb[L][3Q] with
// Q = N/4, LUTr at [0,Q), LUTs at [Q,2Q), LUTd at [2Q,3Q) (LUTi is not used
// by the kernel, serial_mp_fft3.c:79-80).
//
// Work mapping: two lanes (a DPP quad pair) per butterfly.  The even lane
// computes m1 = Z+ * w^k, the odd lane m2 = Z- * conj(w^k), with the same
// instruction stream (operand signs are data); the pair then swaps m1/m2
// and each lane produces two of the four outputs.  Butterflies of the three
// kinds (k = 0, k = n/8, general k) are laid out in separate 32-pair ranges
// so every wave runs one kind only.
#include "mp_arith.h"
#include "kernels.h"

namespace mpfft {

__device__ __forceinline__ uint32_t wrap_idx(int32_t sn, uint32_t s, uint32_t j, uint32_t N)
{
    // x[(sn + j*s) mod N]: only j == 0 can be negative (|sn| < s)
    int64_t v = (int64_t)sn + (int64_t)j * s;
    return (uint32_t)(v < 0 ? v + N : v);
}

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
    uint32_t t[W];
#pragma unroll
    for (int k = 0; k < W; k++) t[k] = k <= L ? src[(size_t)k * M + i] : 0u;
#pragma unroll
    for (int k = 0; k < W / 4; k++)
        dst[k] = make_uint4(t[4 * k], t[4 * k + 1], t[4 * k + 2], t[4 * k + 3]);
}

// Leaves, one pass, in output order with both sizes interleaved, so the
// stores of a wave cover a contiguous range of outputs.  Two lanes per
// leaf (real and imaginary part); every step loads two operands, adds and
// stores, with the values that outlive a step waiting in LDS slots:
//   leaf of 4 (serial_mp_fft3.c:154-174): y0 = x0 + x2, y1 = x0 - x2,
//     a1 = x1 + x3 (slot 0), a2 = x1 - x3; the real lane needs a2.im and
//     the imaginary lane a2.re, so a2 is swapped with the partner lane
//     (slot 1); then y2 = y0 - a1, y0 = y0 + a1, y3 = y1 -+ a2',
//     y1 = y1 +- a2'.
//   leaf of 2 (:141-145): the first two steps with x1 in place of x2.
// Inputs are gathered as whole elements from the element-major copy.
template <int L, bool FULL>
__global__ void __launch_bounds__(256, 2)
mp_leaf_kernel(LeafArgs a)
{
    extern __shared__ uint32_t lds[];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t pair = tid >> 1, B = tid & 1u;   // B: 0 real part, 1 imag part
    const uint32_t *in = a.aos + a.aos_stride * blockIdx.y;
    uint32_t *out = a.data + a.stride * blockIdx.y;
    const uint32_t N = a.N, M = 2 * N, base = B ? N : 0;
    const uint32_t lt = threadIdx.x, nt = blockDim.x;
    const int P = a.P;
    if (pair >= a.nleaf) return;
    const int32_t sn = a.leaf_sn[pair];
    const uint32_t o = a.leaf_off[pair];
    const bool four = a.leaf_n[pair] == 4u;
    const uint32_t s = four ? N >> 2 : N >> 1;
    const int nit = four ? 8 : 2;
    Mp<L> u, v, r;
#pragma unroll 1
    for (int it = 0; it < nit; it++) {
        // even steps load a new operand pair, odd steps reuse it
        if ((it & 1) == 0) {
            if (it < 4) {
                const uint32_t j = (uint32_t)it >> 1;          // x0,x2 (x0,x1) then x1,x3
                mp_load_aos(u, in, base + wrap_idx(sn, s, j, N));
                mp_load_aos(v, in, base + wrap_idx(sn, s, j ? 3 : (four ? 2 : 1), N));
            } else {
                const uint32_t k = (uint32_t)(it - 4) >> 1;    // y0 with a1, y1 with a2'
                mp_load(u, out, M, base + o + k);
                lds_get(v, lds, nt, lt, k);
            }
        }
        // sign of v and destination per step
        uint32_t fy, dst;
        switch (it) {
        case 0: fy = 0; dst = o; break;
        case 1: fy = 1; dst = o + 1; break;
        case 2: fy = 0; dst = 0; break;
        case 3: fy = 1; dst = 0; break;
        case 4: fy = 1; dst = o + 2; break;
        case 5: fy = 0; dst = o; break;
        case 6: fy = B ^ 1u; dst = o + 3; break;
        default: fy = B; dst = o + 1; break;
        }
        mp_add<L, FULL>(r, u, 0, v, fy, P);
        if (it == 2) {
            lds_put(lds, nt, lt, 0, r);
        } else if (it == 3) {
            mp_swap_pair(r, r);                     // word-wise, so in place is fine
            lds_put(lds, nt, lt, 1, r);
        } else {
            mp_store(r, out, M, base + dst);
        }
    }
}

// Two waves per SIMD (256 registers per lane): one wave per SIMD leaves the
// dependent multiply chains nothing to overlap with.
template <int L, bool FULL>
__global__ void __launch_bounds__(256, 2)
mp_level_kernel(LevelArgs a)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t pair = tid >> 1, B = tid & 1u;   // B: 0 -> Z+ lane, 1 -> Z- lane
    uint32_t *out = a.data + a.stride * blockIdx.y;
    const uint32_t N = a.N, n = a.n, q = n >> 2, M = 2 * N, Q = N >> 2, ML = 3 * Q;
    const int P = a.P;

    uint32_t kind, node, k;
    if (pair < a.rG) {
        if (pair >= a.nG) return;
        const uint32_t per = q - 2;                  // k in [1, n/4) \ {n/8}
        node = pair / per;
        k = pair - node * per + 1;
        k += (k >= (n >> 3));
        kind = 0;
    } else if (pair < a.rG + a.rS) {
        node = pair - a.rG;
        if (node >= a.nS) return;
        k = n >> 3;
        kind = 1;
    } else {
        node = pair - a.rG - a.rS;
        if (node >= a.nZ) return;
        k = 0;
        kind = 2;
    }
    const uint32_t o = a.node_off[node];
    const uint32_t zi = o + k + (B ? 3 * q : 2 * q);  // Z+ or Z-
    const uint32_t ui = o + k + (B ? q : 0);          // U0 or U1
    const uint32_t kn = k * (N / n);

    // o1, o2: m1 = Z+ * w^k on the Z+ lane, m2 = Z- * conj(w^k) on the Z- lane.
    // Intermediates that must survive a multiplication wait in LDS slots
    // 0 and 1, so only the multiplication's own state is live across it.
    extern __shared__ uint32_t lds[];
    const uint32_t lt = threadIdx.x, nt = blockDim.x;
    Mp<L> o1, o2;
    mp_load(o1, out, M, zi);
    mp_load(o2, out, M, N + zi);
    if (kind != 2) {
        // pre-additions.  general k (3m3a, serial_mp_fft3.c:220-229):
        //   t = Zr +- Zi.  k = n/8 (:209-217): Zr +- Zi and -+Zr + Zi.
        Mp<L> r;
        const int npre = kind == 0 ? 1 : 2;
#pragma unroll 1
        for (int it = npre - 1; it >= 0; it--) {
            mp_add<L, FULL>(r, o1, it ? (B ^ 1u) : 0u, o2, it ? 0u : B, P);
            if (it) lds_put(lds, nt, lt, 1, r);        // second pre-sum -> slot 1
        }
        // products.  general k: T = t*LUTr -> slot 0; m.r = T -+ LUTd*Zi ->
        //   slot 1; m.i = T - LUTs*Zr (Z+ lane) or LUTs*Zr - T (Z- lane).
        //   k = n/8: slot j = pre-sum j times LUTr[N/8].
        const int nmul = kind == 0 ? 3 : 2;
#pragma unroll 1
        for (int it = 0; it < nmul; it++) {
            Mp<L> f, w;
            uint32_t li = kn;
            if (it == 0) f = r;
            else if (kind == 0) mp_load(f, out, M, it == 1 ? N + zi : zi);
            else lds_get(f, lds, nt, lt, 1);
            if (kind == 0) li = it == 0 ? kn : (it == 1 ? 2 * Q + kn : Q + kn);
            mp_load(w, a.lut, ML, li);
            mp_mul<L, FULL>(r, f, w, P);
            if (kind == 0 && it > 0) {
                Mp<L> T;
                lds_get(T, lds, nt, lt, 0);
                mp_add<L, FULL>(o2, T, it == 2 ? B : 0u, r, B ^ 1u, P);
                if (it == 1) lds_put(lds, nt, lt, 1, o2);
            } else {
                lds_put(lds, nt, lt, it, r);
            }
        }
        lds_get(o1, lds, nt, lt, kind == 0 ? 1 : 0);
        if (kind == 1) lds_get(o2, lds, nt, lt, 1);
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

// ---------------------------------------------------------------------------
// host-side launchers (one instantiation per supported limb count)

template <int L, bool FULL>
static hipError_t launch_fft(const LeafArgs &la, const LevelArgs *lv, int nlev,
                             uint32_t batch, hipStream_t st)
{
    {
        uint32_t thr = 2 * la.nleaf;
        if (thr) {
            dim3 gridc((2 * la.N + 255) / 256, batch);
            hipLaunchKernelGGL((mp_to_aos_kernel<L>), gridc, dim3(256), 0, st, la);
            dim3 grid((thr + 255) / 256, batch);
            const size_t sh = sizeof(uint32_t) * 2 * (L + 1) * 256;
            hipLaunchKernelGGL((mp_leaf_kernel<L, FULL>), grid, dim3(256), sh, st, la);
        }
    }
    for (int i = 0; i < nlev; i++) {
        uint32_t thr = 2 * (lv[i].rG + lv[i].rS + lv[i].nZ);
        dim3 grid((thr + 255) / 256, batch);
        // two parking slots of L+1 words per thread
        const size_t sh = sizeof(uint32_t) * 2 * (L + 1) * 256;
        hipLaunchKernelGGL((mp_level_kernel<L, FULL>), grid, dim3(256), sh, st, lv[i]);
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
static const int kLimbs[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32};

int limbs_for_precision(long P)
{
    long need = (P + 31) / 32;
    for (int L : kLimbs) if (L >= need) return L;
    return -1;
}

hipError_t mp_fft_launch(int L, int P, const LeafArgs &la, const LevelArgs *lv,
                         int nlev, uint32_t batch, hipStream_t st)
{
    const bool full = (P == 32 * L);
#define CALLF(Lv, F) launch_fft<Lv, F>(la, lv, nlev, batch, st)
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
    default: return hipErrorInvalidValue;
    }
#undef CALLF
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
    default: return hipErrorInvalidValue;
    }
#undef CALLB
}

}  // namespace mpfft
