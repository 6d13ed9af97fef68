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

template <int L, bool FULL>
__global__ void __launch_bounds__(256)
mp_leaf_kernel(LeafArgs a)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t pair = tid >> 1, B = tid & 1u;   // B: 0 real part, 1 imag part
    const uint32_t *in = a.in + a.in_stride * blockIdx.y;
    uint32_t *out = a.data + a.stride * blockIdx.y;
    const uint32_t N = a.N, M = 2 * N, base = B ? N : 0;
    const int P = a.P;
    Mp<L> x0, x1, e0, e1;
    if (pair < a.r4) {
        // size-4 leaf (serial_mp_fft3.c:149-177), stride s = N/4
        if (pair >= a.n4) return;
        const int32_t sn = a.sn4[pair];
        const uint32_t o = a.off4[pair], s = N >> 2;
        mp_load(x0, in, M, base + wrap_idx(sn, s, 0, N));
        mp_load(x1, in, M, base + wrap_idx(sn, s, 2, N));
        mp_add<L, FULL>(e0, x0, 0, x1, 0, P);
        mp_add<L, FULL>(e1, x0, 0, x1, 1, P);
        mp_load(x0, in, M, base + wrap_idx(sn, s, 1, N));
        mp_load(x1, in, M, base + wrap_idx(sn, s, 3, N));
        Mp<L> a1, a2, pa2;
        mp_add<L, FULL>(a1, x0, 0, x1, 0, P);
        mp_add<L, FULL>(a2, x0, 0, x1, 1, P);
        // the real lane needs a2.im, the imaginary lane a2.re
        mp_swap_pair(pa2, a2);
        mp_add<L, FULL>(x0, e0, 0, a1, 1, P);
        mp_store(x0, out, M, base + o + 2);
        mp_add<L, FULL>(x0, e0, 0, a1, 0, P);
        mp_store(x0, out, M, base + o);
        mp_add<L, FULL>(x0, e1, 0, pa2, B ^ 1u, P);
        mp_store(x0, out, M, base + o + 3);
        mp_add<L, FULL>(x0, e1, 0, pa2, B, P);
        mp_store(x0, out, M, base + o + 1);
    } else {
        // size-2 leaf (serial_mp_fft3.c:136-148), stride s = N/2
        const uint32_t q = pair - a.r4;
        if (q >= a.n2) return;
        const int32_t sn = a.sn2[q];
        const uint32_t o = a.off2[q], s = N >> 1;
        mp_load(x0, in, M, base + wrap_idx(sn, s, 0, N));
        mp_load(x1, in, M, base + wrap_idx(sn, s, 1, N));
        mp_add<L, FULL>(e0, x0, 0, x1, 0, P);
        mp_store(e0, out, M, base + o);
        mp_add<L, FULL>(e1, x0, 0, x1, 1, P);
        mp_store(e1, out, M, base + o + 1);
    }
}

template <int L, bool FULL>
__global__ void __launch_bounds__(256)
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

    Mp<L> o1, o2;
    {
        Mp<L> X, Y;
        mp_load(X, out, M, zi);
        mp_load(Y, out, M, N + zi);
        if (kind == 0) {
            // 3m3a (serial_mp_fft3.c:220-229).  Z+ lane: m1 = Z+ * (wr - i wi);
            // Z- lane: m2 = Z- * (wr + i wi).
            Mp<L> t, T, W, Pr;
            mp_add<L, FULL>(t, X, 0, Y, B, P);          // Zr +- Zi
            mp_load(W, a.lut, ML, kn);                  // LUTr
            mp_mul<L, FULL>(T, t, W, P);
            mp_load(W, a.lut, ML, 2 * Q + kn);          // LUTd
            mp_mul<L, FULL>(Pr, W, Y, P);
            mp_add<L, FULL>(o1, T, 0, Pr, B ^ 1u, P);   // m1r = T - d Zi | m2r = T + d Zi
            mp_load(W, a.lut, ML, Q + kn);              // LUTs
            mp_mul<L, FULL>(Pr, W, X, P);
            mp_add<L, FULL>(o2, T, B, Pr, B ^ 1u, P);   // m1i = T - s Zr | m2i = s Zr - T
        } else if (kind == 1) {
            // k = n/8: w = (1 - i) c (serial_mp_fft3.c:209-217)
            Mp<L> t, W;
            mp_load(W, a.lut, ML, kn);
            mp_add<L, FULL>(t, X, 0, Y, B, P);          // Zr + Zi | Zr - Zi
            mp_mul<L, FULL>(o1, t, W, P);
            mp_add<L, FULL>(t, X, B ^ 1u, Y, 0, P);     // Zi - Zr | Zr + Zi
            mp_mul<L, FULL>(o2, t, W, P);
        } else {
            o1 = X;                                     // k = 0: m = Z
            o2 = Y;
        }
    }
    // pair exchange: Z+ lane gets m2, Z- lane gets m1
    Mp<L> p1, p2, v1, v2;
    mp_swap_pair(p1, o1);
    mp_swap_pair(p2, o2);
    // Z+ lane: a1 = m1 + m2; Z- lane: a2 = m1 - m2 (serial_mp_fft3.c:234-238)
    mp_add<L, FULL>(v1, o1, B, p1, 0, P);
    mp_add<L, FULL>(v2, o2, B, p2, 0, P);
    // outputs (serial_mp_fft3.c:240-248):
    //   Z+ lane: X[k] = U0 + a1, X[k+n/2] = U0 - a1
    //   Z- lane: X[k+n/4] = U1 - i a2, X[k+3n/4] = U1 + i a2
    Mp<L> U, R;
    const uint32_t f = o + k + (B ? q : 0), s = f + 2 * q;
    mp_load(U, out, M, ui);
    {
        Mp<L> al;
        mp_select(al, B, v2, v1);
        mp_add<L, FULL>(R, U, 0, al, 0, P);
        mp_store(R, out, M, f);
        mp_add<L, FULL>(R, U, 0, al, 1, P);
        mp_store(R, out, M, s);
    }
    mp_load(U, out, M, N + ui);
    {
        Mp<L> be;
        mp_select(be, B, v1, v2);
        mp_add<L, FULL>(R, U, 0, be, B, P);
        mp_store(R, out, M, N + f);
        mp_add<L, FULL>(R, U, 0, be, B ^ 1u, P);
        mp_store(R, out, M, N + s);
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
        uint32_t thr = 2 * (la.r4 + la.n2);
        if (thr) {
            dim3 grid((thr + 255) / 256, batch);
            hipLaunchKernelGGL((mp_leaf_kernel<L, FULL>), grid, dim3(256), 0, st, la);
        }
    }
    for (int i = 0; i < nlev; i++) {
        uint32_t thr = 2 * (lv[i].rG + lv[i].rS + lv[i].nZ);
        dim3 grid((thr + 255) / 256, batch);
        hipLaunchKernelGGL((mp_level_kernel<L, FULL>), grid, dim3(256), 0, st, lv[i]);
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
