// The fixed form of the FFT on panels of transforms, with the twiddle
// products on the int8 matrix cores (v_mfma_i32_32x32x32_i8).
//
// Same transform and the same fixed-point format as fx_fft.hip (radix-4
// Stockham passes, radix 2 first when log2 N is odd; W-word two's-complement
// reals with F = 32 (W - 1) fraction bits); what changes is where a product
// is computed.  In the per-lane kernels every lane multiplies its own operand
// by its own twiddle with 29-bit digit products (v_mad_u64_u32); here a wave
// holds the same element of 64 transforms, so the 64 operands of a wave
// share one twiddle, and the product becomes a matrix product:
//
//   x = sum_i X_i 256^i, t = sum_j T_j 256^j   (balanced digits in [-128, 127])
//   column c of x t:  S_c = sum_i T_(c-i) X_i = (Toeplitz(T) X)_c
//
// With the columns c in tiles of 32 (rows of an MFMA) and the operand digits
// i in tiles of 32 (its reduction axis), tile (p, q) is a 32 x 32 Toeplitz
// block of T that depends only on the diagonal p - q; the 32 MFMA columns are
// 32 transforms.  Column sums stay below 2^23 in int32, and a carry pass
// (7 VALU operations per output word) turns the columns at and above
// 2^(F - 32) into the W result words: the result is floor(x t / 2^F) up to
// the dropped columns, whose sum is below 2^(F - 17).
//
// A complex product (c + i d)(re + i im) takes two accumulating products per
// part: re' = c re + (-d) im, im' = d re + c im, so the twiddle table holds
// c, d and -d for each entry.
//
// Pass kernel: one workgroup = one j and 64 transforms; wave w holds
// v_r = x[j + r N/4] with r = (w + block) mod 4 (rotated so that the waves
// without products, r = 0, spread over the SIMDs).  After the products the
// four waves exchange through LDS and each forms one output of the DFT4.
#include "fx_arith.h"
#include "fx_lane.h"
#include "kernels.h"

// Compile-time tuning of the pass kernel: FXP_OCC blocks per CU (persistent
// grid), FXP_SETS LDS slot sets (2: alternate between items, one barrier per
// item; needs one block per CU), FXP_ACACHE 1 keeps a product's A tiles in
// registers (spills at two blocks per CU at W = 34).
#ifndef FXP_OCC
#define FXP_OCC 2
#endif
#ifndef FXP_SETS
#define FXP_SETS 1
#endif
#ifndef FXP_ACACHE
#define FXP_ACACHE 0
#endif
// radix-8 pass: up to this many words per real the next item's operands are
// fetched right after the products (more registers live through stages 1-3)
#ifndef FXP_R8_EARLY_MAXW
#define FXP_R8_EARLY_MAXW 10
#endif

namespace mpfft {

namespace {

// blocks per CU of the radix-4 pass kernel: reals wider than 34 words (up to
// 66, P = 2048) need the whole register file and 4 LDS slots of 34 KB
constexpr int fxp_occ(int W) { return W > 34 ? 1 : FXP_OCC; }

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// XCD-aware block order: consecutive logical blocks run on the same XCD
// (blocks are dealt to the 8 XCDs round robin), so blocks that share
// cache lines share an L2
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t nb = gridDim.x, bid = blockIdx.x;
    return (nb & 7u) == 0 ? (bid & 7u) * (nb >> 3) + (bid >> 3) : bid;
}

// Inside the W x 64 words of one element of a panel, word w of the panel's
// transform l sits at pan_word(w) + pan_lane(l): for even W as word pairs,
// (w / 2) 128 + 2 l + w % 2, so a lane moves its words 8 bytes at a time
// and a wave 512 contiguous bytes; for odd W word-major, 64 w + l.
template <int W> constexpr bool kPairs = W % 2 == 0;

template <int W>
__device__ __forceinline__ uint32_t pan_word(int w)
{
    return kPairs<W> ? (uint32_t)((w >> 1) * 2 * kPanel + (w & 1)) : (uint32_t)(w * kPanel);
}

template <int W>
__device__ __forceinline__ uint32_t pan_lane(uint32_t l)
{
    return kPairs<W> ? 2 * l : l;
}

// element e of panel grp: its W x 64 words start here
__device__ __forceinline__ size_t pan_elem(uint32_t grp, uint32_t e, uint32_t N, int W)
{
    return ((size_t)grp * (2 * (size_t)N) + e) * (size_t)W * kPanel;
}

// word w of transform b at element e in a panel buffer: pan_base + pan_word(w)
template <int W>
__device__ __forceinline__ size_t pan_base(uint32_t b, uint32_t e, uint32_t N)
{
    return pan_elem(b / kPanel, e, N, W) + pan_lane<W>(b % kPanel);
}

// the W words of one real (W x 64 words of the element, lane part included)
template <int W>
__device__ __forceinline__ void pan_get(Fx<W> &x, const uint32_t *p)
{
    if constexpr (kPairs<W>) {
#pragma unroll
        for (int j = 0; j < W / 2; j++) {
            const uint2 v = *reinterpret_cast<const uint2 *>(p + j * 2 * kPanel);
            x.w[2 * j] = v.x;
            x.w[2 * j + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < W; j++) x.w[j] = p[j * kPanel];
    }
}

template <int W>
__device__ __forceinline__ void pan_put(uint32_t *p, const Fx<W> &x)
{
    if constexpr (kPairs<W>) {
#pragma unroll
        for (int j = 0; j < W / 2; j++)
            *reinterpret_cast<uint2 *>(p + j * 2 * kPanel) = make_uint2(x.w[2 * j], x.w[2 * j + 1]);
    } else {
#pragma unroll
        for (int j = 0; j < W; j++) p[j * kPanel] = x.w[j];
    }
}

template <int W>
__device__ __forceinline__ void pan_load(Fx<W> &re, Fx<W> &im, const uint32_t *buf, uint32_t b,
                                         uint32_t e, uint32_t N)
{
    const uint32_t *p = buf + pan_base<W>(b, e, N);
    pan_get<W>(re, p);
    pan_get<W>(im, p + (size_t)N * W * kPanel);
}

template <int W>
__device__ __forceinline__ void pan_store(const Fx<W> &re, const Fx<W> &im, uint32_t *buf,
                                          uint32_t b, uint32_t e, uint32_t N)
{
    uint32_t *p = buf + pan_base<W>(b, e, N);
    pan_put<W>(p, re);
    pan_put<W>(p + (size_t)N * W * kPanel, im);
}

// The same element e of the 64 transforms of group grp (one panel): lane l
// holds transform 64 grp + l.  The element's words are one contiguous block
// of W x 256 bytes; every address is a wave-uniform base plus the lane part
// plus an immediate offset.
template <int W>
__device__ __forceinline__ void row_load(Fx<W> &re, Fx<W> &im, const uint32_t *buf, uint32_t grp,
                                         uint32_t e, uint32_t N, uint32_t lane)
{
    const uint32_t *p = buf + pan_elem(grp, e, N, W);
    pan_get<W>(re, p + pan_lane<W>(lane));
    pan_get<W>(im, p + (size_t)N * W * kPanel + pan_lane<W>(lane));
}

template <int W>
__device__ __forceinline__ void row_store(const Fx<W> &re, const Fx<W> &im, uint32_t *buf,
                                          uint32_t grp, uint32_t e, uint32_t N, uint32_t lane)
{
    uint32_t *p = buf + pan_elem(grp, e, N, W);
    pan_put<W>(p + pan_lane<W>(lane), re);
    pan_put<W>(p + (size_t)N * W * kPanel + pan_lane<W>(lane), im);
}

// Output R of the DFT4 over the four slots: with a = v0 +- v2, c = v1 +- v3
// (+ for R even), y0 = a + c, y2 = a - c, y1 = a - i c, y3 = a + i c.
// vs[s] points at lane's word 0 of v_s in the LDS slots.
template <bool SUB>
__device__ __forceinline__ uint32_t addsub_c(uint32_t a, uint32_t b, uint32_t &c)
{
    return SUB ? sub_bb(a, b, c) : add_cc(a, b, c);
}

template <int W, int R>
__device__ __forceinline__ void dft4_out(Fx<W> &yr, Fx<W> &yi, const uint32_t *const (&vs)[4])
{
    constexpr bool odd = R & 1;
    constexpr bool sre = (R == 2 || R == 3), sim = (R == 1 || R == 2);
    uint32_t ca_r = 0, ca_i = 0, cc_r = 0, cc_i = 0, cy_r = 0, cy_i = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t ar = addsub_c<odd>(vs[0][i * 64], vs[2][i * 64], ca_r);
        const uint32_t ai = addsub_c<odd>(vs[0][(W + i) * 64], vs[2][(W + i) * 64], ca_i);
        const uint32_t cr = addsub_c<odd>(vs[1][i * 64], vs[3][i * 64], cc_r);
        const uint32_t ci = addsub_c<odd>(vs[1][(W + i) * 64], vs[3][(W + i) * 64], cc_i);
        yr.w[i] = addsub_c<sre>(ar, odd ? ci : cr, cy_r);
        yi.w[i] = addsub_c<sim>(ai, odd ? cr : ci, cy_i);
    }
}

// Balanced base-256 digits of x as MFMA B operands.  Digit i of x is byte i
// of (x + 0x8080...80) minus 128, i.e. that byte XOR 0x80 read as int8.
// K tile q covers digits [32 q, 32 q + 32) = words [8 q, 8 q + 8); in an MFMA
// the lanes n and n + 32 supply column n (digits 16 g + [0, 16) from lane
// n + 32 g), so group 0 (the operands of lanes 0..31) needs words 8q + 4..7
// of lane n in lane n + 32, and group 1 words 8q..8q+3 of lane n + 32 in lane
// n: one permlane32 swap per word pair.
template <int W, int KT>
__device__ __forceinline__ void mf_bdigits(const Fx<W> &x, v4i (&B0)[KT], v4i (&B1)[KT])
{
    uint32_t d[8 * KT];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8 * KT; i++)
        d[i] = i < W ? (add_cc(x.w[i], 0x80808080u, c) ^ 0x80808080u) : 0u;
#pragma unroll
    for (int q = 0; q < KT; q++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (8 * q + i >= W) {            // both halves zero
                B0[q][i] = 0;
                B1[q][i] = 0;
                continue;
            }
            const auto s = __builtin_amdgcn_permlane32_swap(d[8 * q + i], d[8 * q + 4 + i], false,
                                                            false);
            B0[q][i] = (int)s[0];
            B1[q][i] = (int)s[1];
        }
    }
}

// A operand of diagonal tile delta: lane (m = lane % 32, g = lane / 32) holds
// T[C0 + 32 delta + m - 16 g - b] in byte b, read from the twiddle real's
// digit table R (byte u = digit U - u) at byte offset off.
__device__ __forceinline__ v4i mf_atile(const uint32_t *R, uint32_t off)
{
    const uint32_t b4 = off >> 2, sh = off & 3u;
    const uint32_t d0 = R[b4], d1 = R[b4 + 1], d2 = R[b4 + 2], d3 = R[b4 + 3], d4 = R[b4 + 4];
    v4i a;
    a[0] = (int)__builtin_amdgcn_alignbyte(d1, d0, sh);
    a[1] = (int)__builtin_amdgcn_alignbyte(d2, d1, sh);
    a[2] = (int)__builtin_amdgcn_alignbyte(d3, d2, sh);
    a[3] = (int)__builtin_amdgcn_alignbyte(d4, d3, sh);
    return a;
}

// The B operands of a complex operand (re, im): its digit tiles for both
// 32-lane groups.
template <int W>
struct MfOperand {
    static constexpr int KT = mf_geom(W).KT;
    v4i r0[KT], r1[KT], i0[KT], i1[KT];
};

template <int W>
__device__ __forceinline__ void mf_operand(MfOperand<W> &B, const Fx<W> &re, const Fx<W> &im)
{
    mf_bdigits<W, MfOperand<W>::KT>(re, B.r0, B.r1);
    mf_bdigits<W, MfOperand<W>::KT>(im, B.i0, B.i1);
}

// (re, im) *= (c + i d) for the 64 operands of the wave (one per lane), given
// as B operands; the wave's twiddle reals c, d, -d are digit tables in LDS.
// REAL: (re, im) *= c (a real twiddle: one product per part, Rd and Rnd unused).
// TWO: the operands of lanes 32..63 take another twiddle, tables Rc2, Rd2,
// Rnd2 (the MFMAs of the two 32-lane groups, D0 and D1, get their own A
// tiles; the B operands are laid out as before).
// AM, how the A tiles are sourced for reals of up to 34 words (where the
// output tiles are unrolled): 0 built from LDS at each use; 1 built once per
// diagonal, at the first output tile that needs it, and kept in registers
// while later tiles use it (W = 34: six builds per table instead of
// nineteen); 2 built at each use, but the LDS reads for the next live tile
// are issued before the current tile's matrix products, so their latency
// overlaps the products and the carry pass (registers for one tile's reads
// instead of a diagonal cache).
// Results go to the wave's LDS slot: word j of re' at slot[j 64 + lane], of
// im' at slot[(W + j) 64 + lane].  The twiddle tables may lie in the im' part
// of the slot: all of them are read before the first word of im' is written.
// first output tile p whose product needs diagonal d (tile (p, p - d) live)
constexpr int mf_first(int W, int d)
{
    const MfGeom g = mf_geom(W);
    for (int p = 0; p < g.MT; p++)
        if (p - d >= 0 && p - d < g.KT && mf_live(W, p, p - d)) return p;
    return -1;
}

// the live tile after (p, q) in the order p, then q, as p KT + q (-1: none)
constexpr int mf_next(int W, int p, int q)
{
    const MfGeom g = mf_geom(W);
    for (int i = p * g.KT + q + 1; i < g.MT * g.KT; i++)
        if (mf_live(W, i / g.KT, i % g.KT)) return i;
    return -1;
}

template <int W, bool REAL = false, bool TWO = false, int AM = 0>
__device__ __forceinline__ void mf_cmul(const MfOperand<W> &B, const uint32_t *Rc,
                                        const uint32_t *Rd, const uint32_t *Rnd, uint32_t lane,
                                        uint32_t *slot, const uint32_t *Rc2 = nullptr,
                                        const uint32_t *Rd2 = nullptr,
                                        const uint32_t *Rnd2 = nullptr)
{
    static_assert(!(REAL && TWO), "one real twiddle per product");
    constexpr MfGeom G = mf_geom(W);
    constexpr int KT = G.KT, MT = G.MT, ND = G.DHI - G.DLO + 1;
    // |S| <= 2 (4W - 3) 2^14 (two products of balanced digits): the 32-bit
    // form of the carry step below holds up to W = 34; wider reals take
    // 64-bit sums (the output-tile loop is unrolled at every width: at
    // W = 66 the radix-4 pass takes 1.62 ms against 2.08 ms rolled, N = 2^16,
    // batch 64, with 256 VGPRs and about 200 AGPRs and no spill)
    constexpr bool kNarrow = 257LL * 2 * (4 * W - 3) * 16384 + (1LL << 16) < (1LL << 31);
    constexpr int kAM = kNarrow ? AM : 0;
    constexpr int NTAB = REAL ? 1 : TWO ? 4 : 2;
    const uint32_t m = lane & 31u, g = lane >> 5;
    // byte offset of diagonal delta: U - (C0 + 32 delta + m - 16 g)
    const uint32_t off0 = (uint32_t)(G.U - G.C0 - 32 * G.DLO) - m + 16 * g;
    const uint32_t sh = off0 & 3u, b0i = off0 >> 2;
    // the two parts one after the other (rolled: overlapping them holds two
    // sets of accumulators and carries)
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
        // h = 0: re' = c re + (-d) im;  h = 1: im' = d re + c im
        // (REAL: h = 0: re' = c re;  h = 1: im' = c im)
        const uint32_t *R1 = REAL ? Rc : (h ? Rd : Rc);
        const uint32_t *R2 = h ? Rc : Rnd;
        const uint32_t *R1b = TWO ? (h ? Rd2 : Rc2) : R1;
        const uint32_t *R2b = TWO ? (h ? Rc2 : Rnd2) : R2;
        const uint32_t *tab[4] = {R1, R2, R1b, R2b};
        // the lane's table offset, opaque here: addresses derived from it
        // stay inside this loop (a constant offset from one register) rather
        // than being computed once per tile and kept live across the items
        uint32_t b0 = b0i;
        if constexpr (kAM != 0) asm volatile("" : "+v"(b0));
#if FXP_ACACHE
        v4i A1[ND], A2[ND];
#pragma unroll
        for (int dd = 0; dd < ND; dd++) {
            A1[dd] = mf_atile(R1, off0 - 32 * dd);
            A2[dd] = REAL ? A1[dd] : mf_atile(R2, off0 - 32 * dd);
        }
#endif
        int32_t carry = 0;
        uint32_t *out = slot + (uint32_t)h * W * 64 + lane;
        // kAM 1: tiles by diagonal (lower half's tables 0, 1; upper half's
        // 2, 3 with TWO); kAM 2: the reads of the current and the next tile
        v4i Wt[kAM == 1 ? NTAB : 1][kAM == 1 ? ND : 1];
        uint32_t rc[kAM == 2 ? NTAB : 1][5], rn[kAM == 2 ? NTAB : 1][5];
        auto fetch = [&](uint32_t (&r)[kAM == 2 ? NTAB : 1][5], int dd) {
#pragma unroll
            for (int t = 0; t < (kAM == 2 ? NTAB : 1); t++)
#pragma unroll
                for (int i = 0; i < 5; i++) r[t][i] = tab[t][b0 - 8 * dd + i];
        };
        auto align = [&](const uint32_t (&r)[5]) {
            v4i a;
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = (int)__builtin_amdgcn_alignbyte(r[i + 1], r[i], sh);
            return a;
        };
        if constexpr (kAM == 2) {
            constexpr int f = mf_next(W, 0, -1);
            fetch(rc, f / KT - f % KT - G.DLO);
        }
#pragma unroll (MT)
        for (int p = 0; p < MT; p++) {
            if constexpr (kAM == 1) {
#pragma unroll
                for (int q = KT - 1; q >= 0; q--) {
                    if (!mf_live(W, p, q) || mf_first(W, p - q) != p) continue;
                    const int dd = p - q - G.DLO;
#pragma unroll
                    for (int t = 0; t < NTAB; t++) Wt[t][dd] = mf_atile(tab[t], off0 - 32 * dd);
                }
            }
            v16i D0 = {}, D1 = {};
#pragma unroll
            for (int q = 0; q < KT; q++) {
                if (!mf_live(W, p, q)) continue;
                const int dd = p - q - G.DLO;
                v4i at[NTAB];
                if constexpr (kAM == 1) {
#pragma unroll
                    for (int t = 0; t < NTAB; t++) at[t] = Wt[t][dd];
                } else if constexpr (kAM == 2) {
                    // the next live tile's reads go out before this tile's products
                    const int nx = mf_next(W, p, q);
                    if (nx >= 0) fetch(rn, nx / KT - nx % KT - G.DLO);
#pragma unroll
                    for (int t = 0; t < NTAB; t++) at[t] = align(rc[t]);
                    if (nx >= 0) {
#pragma unroll
                        for (int t = 0; t < NTAB; t++)
#pragma unroll
                            for (int i = 0; i < 5; i++) rc[t][i] = rn[t][i];
                    }
                } else if constexpr (kAM == 3) {
                    // all tables' reads for this tile go out together: one
                    // LDS round trip per tile instead of one per table
                    uint32_t r3[NTAB][5];
#pragma unroll
                    for (int t = 0; t < NTAB; t++)
#pragma unroll
                        for (int i = 0; i < 5; i++) r3[t][i] = tab[t][b0 - 8 * dd + i];
#pragma unroll
                    for (int t = 0; t < NTAB; t++) at[t] = align(r3[t]);
                } else {
#if FXP_ACACHE
                    at[0] = A1[dd];
                    if (!REAL) at[1] = A2[dd];
#else
                    at[0] = mf_atile(R1, off0 - 32 * dd);
                    if (!REAL) at[REAL ? 0 : 1] = mf_atile(R2, off0 - 32 * dd);
                    if (TWO) {
                        at[TWO ? 2 : 0] = mf_atile(R1b, off0 - 32 * dd);
                        at[TWO ? 3 : 0] = mf_atile(R2b, off0 - 32 * dd);
                    }
#endif
                }
                const v4i a1 = at[0];
                if (REAL) {
                    D0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, h ? B.i0[q] : B.r0[q], D0, 0, 0, 0);
                    D1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, h ? B.i1[q] : B.r1[q], D1, 0, 0, 0);
                    continue;
                }
                const v4i a2 = at[REAL ? 0 : 1];
                const v4i b1 = TWO ? at[TWO ? 2 : 0] : a1;
                const v4i b2 = TWO ? at[TWO ? 3 : 0] : a2;
                D0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, B.r0[q], D0, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(b1, B.r1[q], D1, 0, 0, 0);
                D0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, B.i0[q], D0, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(b2, B.i1[q], D1, 0, 0, 0);
                // keep the pipeline's order: at most one tile's reads ahead
                if constexpr (kAM >= 2) __builtin_amdgcn_sched_barrier(0);
            }
            // D0 lane l: rows 8 (v/4) + 4 (l/32) + v%4 of operand l % 32,
            // D1 the same of operand l % 32 + 32.  After one swap per v the
            // lane holds rows 8 (v/4) + v%4 (S0) and those + 4 (S1) of its
            // own operand.
            int32_t S0[16], S1[16];
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int row = 8 * (v / 4) + v % 4;
                if (32 * p + row >= G.NCOL) { S0[v] = 0; S1[v] = 0; continue; }
                const auto s = __builtin_amdgcn_permlane32_swap((uint32_t)D0[v], (uint32_t)D1[v],
                                                                false, false);
                S0[v] = (int32_t)s[0];
                S1[v] = (int32_t)s[1];
            }
            // rows 4u..4u+3 = output word t = 8 p + u - 1 (t = -1: the four
            // columns below 2^-F, which only carry)
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int t = 8 * p + u - 1;
                if (t >= W) break;
                const int32_t *S = (u & 1) ? &S1[4 * (u >> 1)] : &S0[4 * (u >> 1)];
                if constexpr (!kNarrow) {
                    const int64_t V = (int64_t)S[0] + ((int64_t)S[1] << 8) +
                                      ((int64_t)S[2] << 16) + ((int64_t)S[3] << 24) + carry;
                    if (t >= 0) out[(uint32_t)t * 64] = (uint32_t)V;
                    carry = (int32_t)(V >> 32);
                    continue;
                }
                // V = S0 + 2^8 S1 + 2^16 S2 + 2^24 S3 + carry = P0 + 2^16 P1
                const int32_t P0 = (int32_t)((uint32_t)S[0] + ((uint32_t)S[1] << 8)) + carry;
                const int32_t P1 = (int32_t)((uint32_t)S[2] + ((uint32_t)S[3] << 8));
                if (t >= 0) out[(uint32_t)t * 64] = (uint32_t)P0 + ((uint32_t)P1 << 16);
                carry = (P1 + (P0 >> 16)) >> 16;
            }
        }
    }
}

}  // namespace

// Input conversion: float SoA -> fixed point in the panel layout, through an
// LDS transpose.  A block covers 32 elements (of the 2N re/im positions) of
// 16 transforms, a thread converts two reals: the limb side reads 128-byte
// runs (32 elements), the panel side writes 128-byte runs (16 transforms'
// word pairs).  Blocks are ordered transform group fastest, so the four
// blocks that fill one panel row run next to each other on one XCD.
template <int L, int W>
__global__ void __launch_bounds__(256)
fxp_in_kernel(FxArgs a, uint32_t *dst)
{
    __shared__ uint32_t tile[W * 32 * 17];
    const uint32_t N = a.N, M = 2 * N, t = threadIdx.x;
    const uint32_t lb = xcd_block(), ng = 4 * fxp_groups(a.batch);
    const uint32_t g16 = lb % ng, eb = lb / ng;
    {
        // lane = (element t % 32, transform t / 32 and that + 8); the padding
        // transforms of a last, partial panel are zeros
        const uint32_t el = t & 31u;
#pragma unroll 1
        for (uint32_t bl = t >> 5; bl < 16; bl += 8) {
            const uint32_t b = g16 * 16 + bl;
            Fx<W> z;
            if (b < a.batch) {
                Mp<L> x;
                mp_load(x, a.in + a.in_stride * b, M, eb * 32 + el);
                if (fx_from_mp<L, W>(z, x, a.lim, a.eoff[b])) atomicOr(a.flag, 1u);
            } else {
#pragma unroll
                for (int w = 0; w < W; w++) z.w[w] = 0u;
            }
#pragma unroll
            for (int w = 0; w < W; w++) tile[(w * 32 + el) * 17 + bl] = z.w[w];
        }
    }
    __syncthreads();
    // lane = (transform t % 16, element t / 16 and that + 16)
    const uint32_t bl = t & 15u;
#pragma unroll 1
    for (uint32_t el = t >> 4; el < 32; el += 16) {
        uint32_t *p = dst + pan_base<W>(g16 * 16 + bl, eb * 32 + el, N);
        Fx<W> z;
#pragma unroll
        for (int w = 0; w < W; w++) z.w[w] = tile[(w * 32 + el) * 17 + bl];
        pan_put<W>(p, z);
    }
}

// A radix-2 pass without twiddles (the first pass when log2 N is odd): a
// lane holds the same j of one transform for both halves.
template <int W>
__global__ void __launch_bounds__(256)
fxp_r2_kernel(FxArgs a, const uint32_t *src, uint32_t *dst)
{
    const uint32_t N = a.N, H = N >> 1;
    const uint32_t wid = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t j = wid % H, grp = wid / H, b = grp * 64 + lane;
    Fx<W> r0, i0, r1, i1, yr, yi;
    pan_load(r0, i0, src, b, j, N);
    pan_load(r1, i1, src, b, j + H, N);
    fx_add(yr, r0, r1);
    fx_add(yi, i0, i1);
    pan_store(yr, yi, dst, b, 2 * j, N);
    fx_sub(yr, r0, r1);
    fx_sub(yi, i0, i1);
    pan_store(yr, yi, dst, b, 2 * j + 1, N);
}

// A radix-4 pass with twiddles, panel layout in and out (see the top of the
// file).  Persistent: a block runs items it = block, block + grid, ...
// (item = one j and 64 transforms) and loads the next item's operands and
// twiddle while it multiplies the current one, so the HBM traffic of a pass
// overlaps its matrix-core and VALU work even at one wave per SIMD.  LDS:
// two sets (alternating items) of four slots of 2 W words x 64 lanes; one
// barrier per item.
template <int W>
__global__ void __launch_bounds__(256, fxp_occ(W))
fxp_r4_kernel(FxArgs a, const uint32_t *src, uint32_t *dst, uint32_t Ns, uint32_t toff,
              uint32_t nitems)
{
    extern __shared__ uint32_t lds[];
    constexpr MfGeom G = mf_geom(W);
    constexpr uint32_t SLOT = 2 * W * 64, TWN = 3 * G.RW, TWL = (TWN + 63) / 64;
    const uint32_t N = a.N, Q = N >> 2;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    uint32_t it = blockIdx.x;
    if (it >= nitems) return;
    // operands and twiddle words of item it (prefetched one item ahead)
    auto item = [&](uint32_t i, uint32_t &j, uint32_t &grp, uint32_t &rot, uint32_t &r,
                    uint32_t &k) {
        j = i % Q; grp = i / Q; rot = i & 3u; r = (wv + rot) & 3u; k = j & (Ns - 1);
    };
    auto fetch = [&](uint32_t i, Fx<W> &re, Fx<W> &im, uint32_t (&tw)[TWL]) {
        uint32_t j, grp, rot, r, k;
        item(i, j, grp, rot, r, k);
        if (r != 0 && k != 0) {
            const uint32_t *t = a.tw + (size_t)(toff + 3 * k + r - 1) * TWN;
#pragma unroll
            for (uint32_t q = 0; q < TWL; q++)
                tw[q] = (lane + 64 * q < TWN) ? (t + 64 * q)[lane] : 0u;
        }
        row_load(re, im, src, grp, j + r * Q, N, lane);
    };
    Fx<W> re, im;
    uint32_t tw[TWL];
    fetch(it, re, im, tw);
    for (uint32_t n = 0;; n++) {
        uint32_t j, grp, rot, r, k;
        item(it, j, grp, rot, r, k);
        const uint32_t nx = it + gridDim.x;
        uint32_t *set = lds + (FXP_SETS == 2 ? (n & 1u) : 0u) * 4 * SLOT;
        if (FXP_SETS == 1 && n > 0) __syncthreads();   // the slots of the last item are read
        uint32_t *slot = set + wv * SLOT;
        // twiddle tables at the end of the slot, or (wide reals) after the slots
        uint32_t *R = mf_tail_fits(W) ? slot + SLOT - TWN : lds + FXP_SETS * 4 * SLOT + wv * TWN;
        const bool prod = r != 0 && k != 0;
        // consume this item's operands before the next item's loads are
        // issued: loads and stores share one in-order counter (at most 63
        // outstanding), so a wait for these operands placed after the next
        // item's loads would wait for those too
        MfOperand<W> B;
        if (prod) {
#pragma unroll
            for (uint32_t q = 0; q < TWL; q++)
                if (lane + 64 * q < TWN) R[lane + 64 * q] = tw[q];
            mf_operand<W>(B, re, im);
        } else {
#pragma unroll
            for (int i = 0; i < W; i++) slot[i * 64 + lane] = re.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slot[(W + i) * 64 + lane] = im.w[i];
        }
        // wide reals (W > 34) hold no prefetched operands across the product
        constexpr bool kPrefetch = W <= 34;
        if (kPrefetch && nx < nitems) fetch(nx, re, im, tw);
        if (prod) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            mf_cmul<W>(B, R, R + G.RW, R + 2 * G.RW, lane, slot);
        }
        if (!kPrefetch && nx < nitems) fetch(nx, re, im, tw);
        __syncthreads();
        // DFT4, output y_r of this wave; v_s sits in the slot of wave
        // (s - rot) mod 4
        const uint32_t *const vs[4] = {set + ((0u - rot) & 3u) * SLOT + lane,
                                       set + ((1u - rot) & 3u) * SLOT + lane,
                                       set + ((2u - rot) & 3u) * SLOT + lane,
                                       set + ((3u - rot) & 3u) * SLOT + lane};
        Fx<W> yr, yi;
        switch (r) {
        case 0: dft4_out<W, 0>(yr, yi, vs); break;
        case 1: dft4_out<W, 1>(yr, yi, vs); break;
        case 2: dft4_out<W, 2>(yr, yi, vs); break;
        default: dft4_out<W, 3>(yr, yi, vs); break;
        }
        row_store(yr, yi, dst, grp, (j - k) * 4 + k + r * Ns, N, lane);
        if (nx >= nitems) break;
        it = nx;
    }
}

// A radix-8 pass, panel layout in and out.  One item = one j and 64
// transforms; the eight waves of a block hold v_r = x[j + r N/8] (r rotated
// by the item index) and multiply it by w^(r k) on the matrix cores.  The
// butterfly y_q = sum_r v_r w8^(r q) runs in three LDS-exchange stages:
//   u_(r1, q0) = v_r1 + (-1)^q0 v_(r1+4)            (wave of input r1 + 4 q0)
//   u'_(r1, 1) = w8^r1 u_(r1, 1)                     (r1 = 1, 3: two real products
//                                                     by 1/sqrt(2), matrix cores)
//   y_(q0 + 2 q1) = sum_r1 (-i)^(r1 q1) u'_(r1, q0)  (DFT4 of four slots)
// LDS: two sets of eight slots of 2 W words x 64 lanes (148 KB at W = 18,
// one block per CU, two waves per SIMD), the table of 1/sqrt(2) and, where a
// product's twiddle tables do not fit at the end of a slot (W = 8), one
// table area per wave.
template <int W>
__global__ void __launch_bounds__(512, 1)
fxp_r8_kernel(FxArgs a, const uint32_t *src, uint32_t *dst, uint32_t Ns, uint32_t toff,
              uint32_t nitems, uint32_t c8e)
{
    extern __shared__ uint32_t lds[];
    constexpr MfGeom G = mf_geom(W);
    constexpr uint32_t SLOT = 2 * W * 64, TWN = 3 * G.RW, TWL = (TWN + 63) / 64;
    const uint32_t N = a.N, E8 = N >> 3;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    // two sets of eight slots: the products and stage 1 use set A, stage 2
    // writes and stage 3 reads set B, so an item needs two barriers
    constexpr uint32_t SETB = 8 * SLOT;
    uint32_t *c8t = lds + 16 * SLOT;
    constexpr bool kEarlyFetch = W <= FXP_R8_EARLY_MAXW;
    for (uint32_t i = threadIdx.x; i < (uint32_t)G.RW; i += blockDim.x)
        c8t[i] = a.tw[(size_t)c8e * TWN + i];
    __syncthreads();                                   // c8t staged
    uint32_t it = blockIdx.x;
    if (it >= nitems) return;
    auto item = [&](uint32_t i, uint32_t &j, uint32_t &grp, uint32_t &rot, uint32_t &r,
                    uint32_t &k) {
        j = i % E8; grp = i / E8; rot = i & 7u; r = (wv + rot) & 7u; k = j & (Ns - 1);
    };
    auto fetch = [&](uint32_t i, Fx<W> &re, Fx<W> &im, uint32_t (&tw)[TWL]) {
        uint32_t j, grp, rot, r, k;
        item(i, j, grp, rot, r, k);
        if (r != 0 && k != 0) {
            const uint32_t *t = a.tw + (size_t)(toff + 7 * k + r - 1) * TWN;
#pragma unroll
            for (uint32_t q = 0; q < TWL; q++)
                tw[q] = (lane + 64 * q < TWN) ? (t + 64 * q)[lane] : 0u;
        }
        row_load(re, im, src, grp, j + r * E8, N, lane);
    };
    Fx<W> re, im;
    uint32_t tw[TWL];
    fetch(it, re, im, tw);
    for (uint32_t n = 0;; n++) {
        uint32_t j, grp, rot, r, k;
        item(it, j, grp, rot, r, k);
        const uint32_t nx = it + gridDim.x;
        // slot of the wave holding input s: wave (s - rot) mod 8.  No barrier
        // is needed before this item's writes: set A was last read in the
        // previous item's stage 1, and a wave gets here only after that
        // item's second barrier, which every wave reaches after its stage 1;
        // set B (read in the previous item's stage 3) is written only after
        // this item's first barrier, which every wave reaches after its
        // stage 3.
        auto sl = [&](uint32_t s_) { return lds + SETB + ((s_ - rot) & 7u) * SLOT + lane; };
        auto sla = [&](uint32_t s_) { return lds + ((s_ - rot) & 7u) * SLOT + lane; };
        uint32_t *slot = lds + wv * SLOT, *slotb = lds + SETB + wv * SLOT;
        // twiddle tables at the end of the slot, or after the slots and c8t
        uint32_t *R = mf_tail_fits(W) ? slot + SLOT - TWN : lds + 16 * SLOT + G.RW + wv * TWN;
        const bool prod = r != 0 && k != 0;
        MfOperand<W> B;
        if (prod) {
#pragma unroll
            for (uint32_t q = 0; q < TWL; q++)
                if (lane + 64 * q < TWN) R[lane + 64 * q] = tw[q];
            mf_operand<W>(B, re, im);
        } else {
#pragma unroll
            for (int i = 0; i < W; i++) slot[i * 64 + lane] = re.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slot[(W + i) * 64 + lane] = im.w[i];
        }
        if (prod) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            mf_cmul<W>(B, R, R + G.RW, R + 2 * G.RW, lane, slot);
        }
        // re, im and tw are consumed: up to FXP_R8_EARLY_MAXW words per real
        // (registers to spare) the next item's loads start here and overlap
        // both barriers and stages 1-3
        if (kEarlyFetch && nx < nitems) fetch(nx, re, im, tw);
        __syncthreads();
        // stage 1: u = v_r + v_(r+4) (r < 4) or v_(r-4) - v_r (r >= 4)
        Fx<W> ur, ui;
        {
            const uint32_t *lo = sla(r & 3u), *hi = sla((r & 3u) + 4);
            uint32_t cr = 0, ci = 0;
            if (r < 4) {
#pragma unroll
                for (int i = 0; i < W; i++) {
                    ur.w[i] = add_cc(lo[i * 64], hi[i * 64], cr);
                    ui.w[i] = add_cc(lo[(W + i) * 64], hi[(W + i) * 64], ci);
                }
            } else {
#pragma unroll
                for (int i = 0; i < W; i++) {
                    ur.w[i] = sub_bb(lo[i * 64], hi[i * 64], cr);
                    ui.w[i] = sub_bb(lo[(W + i) * 64], hi[(W + i) * 64], ci);
                }
            }
        }
        // stage 2 (into set B): u' = w8^r1 u for the odd half (r = 4 + r1)
        if (r == 5 || r == 7) {
            // w8 (a + ib) = ((a + b) + i (b - a)) / sqrt2; w8^3 (a + ib) = ((b - a) - i (a + b)) / sqrt2
            Fx<W> s_, d_;
            fx_add(s_, ur, ui);
            fx_sub(d_, ui, ur);
            if (r == 7) {
                Fx<W> z;
#pragma unroll
                for (int i = 0; i < W; i++) z.w[i] = 0;
                fx_sub(ur, z, s_);            // -(a + b)
                s_ = d_;
                d_ = ur;
            }
            MfOperand<W> Bq;
            mf_operand<W>(Bq, s_, d_);
            mf_cmul<W, true>(Bq, c8t, c8t, c8t, lane, slotb);
        } else {
            if (r == 6) {                     // -i (a + ib) = b - ia
                Fx<W> z, t_;
#pragma unroll
                for (int i = 0; i < W; i++) z.w[i] = 0;
                fx_sub(t_, z, ur);
                ur = ui;
                ui = t_;
            }
#pragma unroll
            for (int i = 0; i < W; i++) slotb[i * 64 + lane] = ur.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slotb[(W + i) * 64 + lane] = ui.w[i];
        }
        // wider reals: the next item's operands load during stage 3 and the
        // stores (issued here, after both products, to keep them out of the
        // products' registers)
        if (!kEarlyFetch && nx < nitems) fetch(nx, re, im, tw);
        __syncthreads();
        // stage 3: output q = q0 + 2 q1 with q0 = r / 4, q1 = r % 4: the DFT4 of
        // the slots of inputs 4 q0 + r1, r1 = 0..3
        const uint32_t q0 = r >> 2, q1 = r & 3u;
        const uint32_t *const vs[4] = {sl(4 * q0), sl(4 * q0 + 1), sl(4 * q0 + 2), sl(4 * q0 + 3)};
        Fx<W> yr, yi;
        switch (q1) {
        case 0: dft4_out<W, 0>(yr, yi, vs); break;
        case 1: dft4_out<W, 1>(yr, yi, vs); break;
        case 2: dft4_out<W, 2>(yr, yi, vs); break;
        default: dft4_out<W, 3>(yr, yi, vs); break;
        }
        row_store(yr, yi, dst, grp, (j - k) * 8 + k + (q0 + 2 * q1) * Ns, N, lane);
        if (nx >= nitems) break;
        it = nx;
    }
}

// Output R of the DFT4 over four LDS operands, R chosen per lane at run time
// (the two 32-lane halves of a wave may compute different outputs): the
// same sums as dft4_out, subtraction as addition of the complement.  v_s
// points at the lane's word 0 of operand s (word i at i 64, imaginary words
// at (W + i) 64).
template <int W>
__device__ __forceinline__ void dft4_rt(Fx<W> &yr, Fx<W> &yi, const uint32_t *v0,
                                        const uint32_t *v1, const uint32_t *v2,
                                        const uint32_t *v3, uint32_t R)
{
    const uint32_t odd = R & 1u;
    const uint32_t mo = 0u - odd;                                  // a = v0 -+ v2, c = v1 -+ v3
    const uint32_t sre = (R == 2u || R == 3u) ? 1u : 0u;           // yr = ar -+ (odd ? ci : cr)
    const uint32_t sim = (R == 1u || R == 2u) ? 1u : 0u;           // yi = ai -+ (odd ? cr : ci)
    const uint32_t mre = 0u - sre, mim = 0u - sim;
    const uint32_t cx = odd ? (uint32_t)W * 64u : 0u;              // part of c added into yr
    const uint32_t cy = odd ? 0u : (uint32_t)W * 64u;              // part of c added into yi
    uint32_t ca_r = odd, ca_i = odd, cc_x = odd, cc_y = odd, cy_r = sre, cy_i = sim;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t o = (uint32_t)i * 64u, oi = (uint32_t)(W + i) * 64u;
        const uint32_t ar = add_cc(v0[o], v2[o] ^ mo, ca_r);
        const uint32_t ai = add_cc(v0[oi], v2[oi] ^ mo, ca_i);
        const uint32_t c1 = add_cc(v1[cx + o], v3[cx + o] ^ mo, cc_x);
        const uint32_t c2 = add_cc(v1[cy + o], v3[cy + o] ^ mo, cc_y);
        yr.w[i] = add_cc(ar, c1 ^ mre, cy_r);
        yi.w[i] = add_cc(ai, c2 ^ mim, cy_i);
    }
}

// Radix-16 pass layout.  Lane-half h = 2 w + (lane >> 5) of wave w (eight
// waves) holds input position r = w + 8 (lane >> 5) of 32 transforms, and
// the DFT16 runs as 4 x 4 (r = r2 + 4 r1, q = m1 + 4 m2):
//   stage A:  Y_(r2, m1) = sum_r1 v_(r2 + 4 r1) (-i)^(r1 m1)       (DFT4)
//   internal: Y'_(r2, m1) = w16^(r2 m1) Y_(r2, m1)
//   stage B:  y_(m1 + 4 m2) = sum_r2 Y'_(r2, m1) (-i)^(r2 m2)      (DFT4)
// Lane-half h computes Y_(kR2[h], kM1[h]) in stage A and y_q with m1 =
// kM1[h], m2 = kR2[h] in stage B.  The pairs are placed so that the eight
// nontrivial internal twiddles fall in waves 0..3, both halves of a wave
// under one twiddle (one matrix operand A per tile, as in the radix-4
// pass): wave 0 w16^3 (pairs (1,3), (3,1)), wave 1 w16^1 (pairs (1,1) and
// (3,3), whose twiddle w16^9 = -w16^1: the upper half's operand is negated
// first), wave 2 w16^2 ((1,2), (2,1)), wave 3 w16^6 ((2,3), (3,2)); the
// trivial ones in waves 4..7 (1, and -i in wave 7's upper half).
__constant__ const uint8_t kR2[16] = {1, 3, 1, 3, 1, 2, 2, 3, 0, 0, 0, 0, 1, 2, 3, 2};
__constant__ const uint8_t kM1[16] = {3, 1, 1, 3, 2, 1, 3, 2, 0, 1, 2, 3, 0, 0, 0, 2};

// internal twiddle table of wave w < 4, an index into the tables w16^1,
// w16^3, w16^9, w16^2, w16^6
__device__ __forceinline__ uint32_t r16_int_table(uint32_t w)
{
    return w == 0 ? 1u : w == 1 ? 0u : w + 1;
}

// blocks per CU of the radix-16 pass: two up to FXP_R16_OCC2_MAXW words
// per real (one slot set each, 128 registers per wave, four waves per SIMD),
// else one.  Measured at N = 2^16, batch 64: at 6 words (128 bits) the pass
// with products takes 109-111 us against 133 us at one block, the first
// 84-86 against 99; at 10 words the product spills 35 registers at 128 and
// the pass takes 312 us against 249
#ifndef FXP_R16_OCC2_MAXW
#define FXP_R16_OCC2_MAXW 8
#endif
constexpr int kR16Occ(int W) { return W <= FXP_R16_OCC2_MAXW ? 2 : 1; }

constexpr int kR16Sets(int W) { return (W <= 18 && kR16Occ(W) == 1) ? 2 : 1; }

// A-tile sourcing of the radix-16 products (mf_cmul's AM): the diagonal
// cache up to 18 words per real (measured 6 % faster passes at W = 18); at
// wider reals it does not fit beside the operands in 256 registers (the
// spills cost more than the LDS reads saved), nor do the other forms
constexpr int r16_am(int W) { return (W <= 18 && kR16Occ(W) == 1) ? 1 : 0; }

// LDS words of the radix-16 pass: the slot sets (eight slots of 2 W words
// x 64 lanes each), two external twiddle tables per wave, and the five
// internal twiddle tables
template <int W>
constexpr uint32_t r16_lds_words()
{
    return (uint32_t)(kR16Sets(W) * 8 * 2 * W * 64 + (16 + 5) * 3 * mf_geom(W).RW);
}

// A radix-16 pass, panel layout in and out.  One item = one j and half a
// panel (32 transforms); persistent over items it = block, block + grid, ...
// (it = 2 (grp N/16 + j) + half, so the two halves of one j run side by
// side).  Twiddle table: the pass's entries toff + 15 k + r - 1 (k = j mod
// Ns, r = 1..15); c16e: the entry w^0 = 1, then w^(N/16), w^(3N/16),
// w^(9N/16), w^(N/8), w^(3N/8) (w16^1, w16^3, w16^9, w16^2, w16^6).  The
// external product runs for every item (k = 0 and position 0 take the
// entry 1: exact), so every wave runs the same code; without EXT (the first
// pass, Ns = 1: every twiddle is 1) it is skipped.  One set of slots (W > 18):
// four barriers per item (products write, stage A reads, internal products
// write, stage B reads, all in the one set); two sets: two barriers (stage A
// reads set A, the internal products write set B, stage B reads set B).
template <int W, bool EXT>
__global__ void __launch_bounds__(512, 2 * kR16Occ(W))
fxp_r16_kernel(FxArgs a, const uint32_t *src, uint32_t *dst, uint32_t Ns, uint32_t toff,
               uint32_t nitems, uint32_t c16e)
{
    extern __shared__ uint32_t lds[];
    constexpr MfGeom G = mf_geom(W);
    constexpr uint32_t SLOT = 2 * W * 64, TWN = 3 * G.RW, TWL = (TWN + 63) / 64;
    constexpr int SETS = kR16Sets(W);
    const uint32_t N = a.N, J = N >> 4;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t hl = lane >> 5, col = lane & 31u, h = 2 * wv + hl;
    uint32_t *const setA = lds, *const setB = lds + (SETS - 1) * 8 * SLOT;
    uint32_t *const tabE = lds + SETS * 8 * SLOT + wv * 2 * TWN;     // this wave's two tables
    uint32_t *const w16t = lds + SETS * 8 * SLOT + 16 * TWN;         // the internal tables
    for (uint32_t i = threadIdx.x; i < 5 * TWN; i += blockDim.x)
        w16t[i] = a.tw[(size_t)(c16e + 1) * TWN + i];
    __syncthreads();                                   // constant tables staged
    uint32_t it = blockIdx.x;
    if (it >= nitems) return;
    // per lane-half, fixed over items: stage A reads v_(r2 + 4 s) (position r
    // lives in the slot of wave r % 8, lane half r / 8), stage B reads
    // Y'_(s, m1), which lane-half g with kR2[g] = s, kM1[g] = m1 holds
    const uint32_t r2 = kR2[h], m1 = kM1[h];
    uint32_t sa[4], sb[4];                             // word offsets in lds
#pragma unroll
    for (uint32_t s_ = 0; s_ < 4; s_++) {
        const uint32_t r = r2 + 4 * s_;
        sa[s_] = (r & 7u) * SLOT + 32 * (r >> 3) + col;
        uint32_t g = 0;
#pragma unroll
        for (uint32_t x = 0; x < 16; x++)
            if (kR2[x] == s_ && kM1[x] == m1) g = x;
        sb[s_] = (uint32_t)(SETS - 1) * 8 * SLOT + (g >> 1) * SLOT + 32 * (g & 1u) + col;
    }
    const uint32_t q = m1 + 4 * r2;                    // stage B output of this lane-half
    const uint32_t pos = wv + 8 * hl;                  // input position of this lane-half
    auto item = [&](uint32_t i, uint32_t &j, uint32_t &grp, uint32_t &half, uint32_t &k) {
        half = i & 1u;
        const uint32_t jj = i >> 1;
        j = jj % J; grp = jj / J; k = j & (Ns - 1);
    };
    // operands of item i for this lane, and the wave's two twiddle tables:
    // positions wv and wv + 8 (the entry "1" for position 0 and for k = 0).
    // In two parts (the real words and the tables, then the imaginary
    // words), so that at most one part is held in registers across a product.
    auto operand = [&](uint32_t i) {
        uint32_t j, grp, half, k;
        item(i, j, grp, half, k);
        return src + pan_elem(grp, j + pos * J, N, W) + pan_lane<W>(32 * half + col);
    };
    auto fetch_re = [&](uint32_t i, Fx<W> &re) { pan_get<W>(re, operand(i)); };
    auto fetch_tw = [&](uint32_t i, uint32_t (&ta)[TWL], uint32_t (&tb)[TWL]) {
        uint32_t j, grp, half, k;
        item(i, j, grp, half, k);
        const uint32_t ea = (wv == 0 || k == 0) ? c16e : toff + 15 * k + wv - 1;
        const uint32_t eb = k == 0 ? c16e : toff + 15 * k + wv + 7;
        const uint32_t *t0 = a.tw + (size_t)ea * TWN, *t1 = a.tw + (size_t)eb * TWN;
#pragma unroll
        for (uint32_t u = 0; u < TWL; u++) {
            ta[u] = (lane + 64 * u < TWN) ? (t0 + 64 * u)[lane] : 0u;
            tb[u] = (lane + 64 * u < TWN) ? (t1 + 64 * u)[lane] : 0u;
        }
    };
    auto fetch_im = [&](uint32_t i, Fx<W> &im) {
        pan_get<W>(im, operand(i) + (size_t)N * W * kPanel);
    };
    Fx<W> re, im;
    uint32_t ta[TWL], tb[TWL];
    if (EXT) fetch_tw(it, ta, tb);
    fetch_re(it, re);
    fetch_im(it, im);
    for (uint32_t n = 0;; n++) {
        uint32_t j, grp, half, k;
        item(it, j, grp, half, k);
        const uint32_t nx = it + gridDim.x;
        uint32_t *const slotA = setA + wv * SLOT, *const slotB = setB + wv * SLOT;
        // one set: the slots were last read in the previous item's stage B
        if (SETS == 1 && n > 0) __syncthreads();
        // external twiddles w^(r k), r = this lane-half's position
        if (!EXT) {
#pragma unroll
            for (int i = 0; i < W; i++) slotA[i * 64 + lane] = re.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slotA[(W + i) * 64 + lane] = im.w[i];
        } else {
#pragma unroll
            for (uint32_t u = 0; u < TWL; u++)
                if (lane + 64 * u < TWN) {
                    tabE[lane + 64 * u] = ta[u];
                    tabE[TWN + lane + 64 * u] = tb[u];
                }
            MfOperand<W> Bx;
            mf_operand<W>(Bx, re, im);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            mf_cmul<W, false, true, r16_am(W)>(Bx, tabE, tabE + G.RW, tabE + 2 * G.RW, lane, slotA,
                                          tabE + TWN, tabE + TWN + G.RW, tabE + TWN + 2 * G.RW);
        }
        __syncthreads();
        // stage A
        Fx<W> yr, yi;
        dft4_rt<W>(yr, yi, lds + sa[0], lds + sa[1], lds + sa[2], lds + sa[3], m1);
        // internal twiddles (wave-uniform kind).  With one slot set they
        // write where stage A reads: after a barrier, and the operands are
        // converted after it too (no matrix operand is held across it)
        if (wv == 7 && hl == 1) {
            // -i (a + ib) = b - ia
            Fx<W> z, t_;
#pragma unroll
            for (int i = 0; i < W; i++) z.w[i] = 0;
            fx_sub(t_, z, yr);
            yr = yi;
            yi = t_;
        }
        if (SETS == 1) __syncthreads();               // stage A's reads of the one set
        if (wv < 4) {
            if (wv == 1 && hl == 1) {                  // w16^9 Y = w16^1 (-Y)
                Fx<W> z;
#pragma unroll
                for (int i = 0; i < W; i++) z.w[i] = 0;
                fx_sub(yr, z, yr);
                fx_sub(yi, z, yi);
            }
            MfOperand<W> B;
            mf_operand<W>(B, yr, yi);
            const uint32_t *t0 = w16t + r16_int_table(wv) * TWN;
            mf_cmul<W, false, false, r16_am(W)>(B, t0, t0 + G.RW, t0 + 2 * G.RW, lane, slotB);
        } else {
#pragma unroll
            for (int i = 0; i < W; i++) slotB[i * 64 + lane] = yr.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slotB[(W + i) * 64 + lane] = yi.w[i];
        }
        // the next item's operands and twiddle tables load during stage B,
        // where few registers are live (none is held across a product)
        if (nx < nitems) {
            fetch_re(nx, re);
            fetch_im(nx, im);
            if (EXT) fetch_tw(nx, ta, tb);
        }
        __syncthreads();
        // stage B: output q = m1 + 4 m2 of this lane-half
        dft4_rt<W>(yr, yi, lds + sb[0], lds + sb[1], lds + sb[2], lds + sb[3], r2);
        uint32_t *p = dst + pan_elem(grp, (j - k) * 16 + k + q * Ns, N, W) +
                      pan_lane<W>(32 * half + col);
        pan_put<W>(p, yr);
        pan_put<W>(p + (size_t)N * W * kPanel, yi);
        if (nx >= nitems) break;
        it = nx;
    }
}

// Output conversion: panel layout -> float SoA rounded to P bits, through
// an LDS transpose.  A block covers 32 elements of 16 transforms (block order
// as fxp_in_kernel), a thread converts two reals: the panel side reads
// 128-byte runs (16 transforms' word pairs), the limb side writes 128-byte
// runs (32 elements; with 16 elements, 64-byte runs measured slower).
template <int W, int L, bool FULL>
__global__ void __launch_bounds__(256)
fxp_out_kernel(FxArgs a, const uint32_t *src)
{
    __shared__ uint32_t tile[(L + 1) * 32 * 17];
    const uint32_t N = a.N, M = 2 * N, t = threadIdx.x;
    const uint32_t lb = xcd_block(), ng = (a.batch + 15) / 16;
    const uint32_t g16 = lb % ng, eb = lb / ng;
    {
        // lane = (transform t % 16, element t / 16 and that + 16)
        const uint32_t bl = t & 15u, b = g16 * 16 + bl;
        const int32_t eo = b < a.batch ? a.eoff[b] : 0;     // padding transforms: none
#pragma unroll 1
        for (uint32_t el = t >> 4; el < 32; el += 16) {
            const uint32_t *p = src + pan_base<W>(b, eb * 32 + el, N);
            Fx<W> x;
            pan_get<W>(x, p);
            Mp<L> z;
            fx_to_mp<W, L, FULL>(z, x, a.P, eo);
            tile[el * 17 + bl] =
                z.zero ? ((uint32_t)kHdrZero | z.neg) : (((uint32_t)z.e << 1) | z.neg);
#pragma unroll
            for (int j = 0; j < L; j++) tile[((j + 1) * 32 + el) * 17 + bl] = z.zero ? 0u : z.m[j];
        }
    }
    __syncthreads();
    // lane = (element t % 32, transform t / 32 and that + 8)
    const uint32_t el = t & 31u;
#pragma unroll 1
    for (uint32_t bl = t >> 5; bl < 16; bl += 8) {
        if (g16 * 16 + bl >= a.batch) break;          // padding transforms
        uint32_t *o = a.out + a.out_stride * (g16 * 16 + bl) + eb * 32 + el;
#pragma unroll
        for (int j = 0; j <= L; j++) o[(size_t)j * M] = tile[(j * 32 + el) * 17 + bl];
    }
}

// ---- four-step plan (kernels.h, fxs_plan) -----------------------------------

// The tiled buffers between the four-step passes (kernels.h, FxsPass): word
// pair k (k < W/2 real words 2k, 2k+1; then the imaginary ones) of element E
// of transform t.  A tile of shape (Tt, Te, Se) holds Tt transforms x Te
// elements Se apart (E = row + Se m, m < Te); slice k of a tile is Tt Te
// consecutive pairs, transform fastest.  Returns the word offset of pair 0;
// pair k is k << (lgTt + lgTe + 1) words further.
template <uint32_t lgTt, uint32_t lgTe, uint32_t lgSe>
__device__ __forceinline__ size_t tile_at(uint32_t t, uint32_t E, uint32_t N, int W)
{
    const uint32_t rowE = ((E >> (lgSe + lgTe)) << lgSe) | (E & ((1u << lgSe) - 1u));
    const uint32_t m = (E >> lgSe) & ((1u << lgTe) - 1u);
    const size_t R = (size_t)(t >> lgTt) * (N >> lgTe) + rowE;
    return ((R * (size_t)W) << (lgTt + lgTe + 1)) +
           ((size_t)((m << lgTt) | (t & ((1u << lgTt) - 1u))) << 1);
}

// The shapes of a four-step pass (kernels.h, fxs_plan), fixed per kind:
// T48 marks a step-1 pass that writes the 4 x 8 tiles the twist pass reads.
template <uint32_t MODE, bool T48>
struct FxsShape {
    static constexpr bool step1 = MODE == kFxsFirst || MODE == kFxsShared;
    // columns: 2^lgCt transforms x 2^(5 - lgCt) column indices
    static constexpr uint32_t lgCt = MODE == kFxsTwist ? 5 : (MODE == kFxsLast || T48) ? 2 : 0;
    static constexpr uint32_t iTt = step1 ? 0 : 2, iTe = step1 ? 5 : 3;
    static constexpr uint32_t iSe = MODE == kFxsLast ? 8 : 0;
    static constexpr uint32_t oTt = (MODE == kFxsTwist || T48) ? 2 : 0;
    static constexpr uint32_t oTe = (MODE == kFxsTwist || T48) ? 3 : 5;
    static constexpr uint32_t oSe = MODE == kFxsTwist ? 8 : 0;
    // element strides of the column index u and the position (N2 = 256);
    // the last pass's output stride of its position is N1 (run time)
    static constexpr uint32_t ics = step1 ? 1 : 256, ies = step1 ? 256 : 1;
    static constexpr uint32_t ocs = MODE == kFxsTwist ? 256 : 1;
    static constexpr uint32_t oes = step1 ? 256 : 1;
};

// every tile shape of the plan holds Tt Te = 32 pairs per slice: slices are
// 64 words apart
constexpr uint32_t kTileSlice = 64;

template <int W>
__device__ __forceinline__ void tile_get(Fx<W> &re, Fx<W> &im, const uint32_t *p)
{
    static_assert(W % 2 == 0, "tiled buffers hold word pairs");
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p + k * kTileSlice);
        if (k < W / 2) { re.w[2 * k] = v.x; re.w[2 * k + 1] = v.y; }
        else { im.w[2 * k - W] = v.x; im.w[2 * k - W + 1] = v.y; }
    }
}

template <int W>
__device__ __forceinline__ void tile_put(uint32_t *p, const Fx<W> &re, const Fx<W> &im)
{
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint2 v = k < W / 2 ? make_uint2(re.w[2 * k], re.w[2 * k + 1])
                                  : make_uint2(im.w[2 * k - W], im.w[2 * k - W + 1]);
        *reinterpret_cast<uint2 *>(p + k * kTileSlice) = v;
    }
}

// A radix-16 pass of the four-step plan.  The DFT16, the LDS slots and the
// products are those of fxp_r16_kernel (one item = one butterfly index j of
// 32 columns; eight waves hold the 16 positions, position w + 8 h in lane
// half h of wave w; stage A, internal twiddles, stage B); what the mode sets
// is which butterflies form an item, where their elements are and which
// twiddle each position takes (kernels.h, FxsPass).  Columns beyond the
// batch load zeros and store nothing.  first: the loads convert the limb
// input (block exponent a.eoff); last: the stores round into the limb
// output (and the next item's loads follow them, so the rounding does not
// share the registers with them).
template <int W, int L, uint32_t MODE, bool FULL, bool T48>
__global__ void __launch_bounds__(512, 1)
fxs_r16_kernel(FxArgs a, const uint32_t *src, uint32_t *dst, FxsPass ps, uint32_t c16e)
{
    using S = FxsShape<MODE, T48>;
    extern __shared__ uint32_t lds[];
    constexpr bool EXT = MODE != kFxsFirst;
    constexpr MfGeom G = mf_geom(W);
    constexpr uint32_t SLOT = 2 * W * 64, TWN = 3 * G.RW, TWL = (TWN + 63) / 64;
    constexpr int SETS = kR16Sets(W);
    const uint32_t N = a.N, M2 = 2 * N, J = 1u << ps.lgJ;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t hl = lane >> 5, col = lane & 31u, h = 2 * wv + hl;
    uint32_t *const setA = lds, *const setB = lds + (SETS - 1) * 8 * SLOT;
    uint32_t *const tabE = lds + SETS * 8 * SLOT + wv * 2 * TWN;     // this wave's two tables
    uint32_t *const w16t = lds + SETS * 8 * SLOT + 16 * TWN;         // the internal tables
    for (uint32_t i = threadIdx.x; i < 5 * TWN; i += blockDim.x)
        w16t[i] = a.tw[(size_t)(c16e + 1) * TWN + i];
    __syncthreads();                                   // constant tables staged
    uint32_t it = blockIdx.x;
    const uint32_t nitems = ps.nitems;
    if (it >= nitems) return;
    const uint32_t r2 = kR2[h], m1 = kM1[h];
    uint32_t sa[4], sb[4];                             // word offsets in lds
#pragma unroll
    for (uint32_t s_ = 0; s_ < 4; s_++) {
        const uint32_t r = r2 + 4 * s_;
        sa[s_] = (r & 7u) * SLOT + 32 * (r >> 3) + col;
        uint32_t g = 0;
#pragma unroll
        for (uint32_t x = 0; x < 16; x++)
            if (kR2[x] == s_ && kM1[x] == m1) g = x;
        sb[s_] = (uint32_t)(SETS - 1) * 8 * SLOT + (g >> 1) * SLOT + 32 * (g & 1u) + col;
    }
    const uint32_t q = m1 + 4 * r2;                    // stage B output of this lane-half
    const uint32_t pos = wv + 8 * hl;                  // input position of this lane-half
    // blocks of 32 columns: Ct transforms x (32 / Ct) column indices u
    constexpr uint32_t lgCe = 5 - S::lgCt;
    const uint32_t lgNub = ps.lgU - lgCe, nub = 1u << lgNub;
    const uint32_t ntg = (a.batch + (1u << S::lgCt) - 1) >> S::lgCt;
    // item i: butterfly j, column block blk and (twist) the k1 of the item.
    // Twist order: 32-transform blocks fastest, then j and k1 in 8 x 8 blocks
    // (the items that read one input tile, j..j+7, and those that write one
    // output tile, k1..k1+7, run side by side)
    // Slot i of the persistent loop (i = block + grid n) takes item
    // xcd_item(i): the grid's blocks are dealt to the 8 XCDs round robin, so
    // the slots of one XCD in a round take consecutive items, and items that
    // share cache lines share an L2 (the last, partial round keeps i)
    const uint32_t ng = gridDim.x, full = (nitems / ng) * ng;
    auto xcd_item = [&](uint32_t i) {
        if ((ng & 7u) != 0 || i >= full) return i;
        const uint32_t b = i % ng;
        return i - b + (b & 7u) * (ng >> 3) + (b >> 3);
    };
    auto item = [&](uint32_t i0, uint32_t &j, uint32_t &blk, uint32_t &ki) {
        const uint32_t i = xcd_item(i0);
        if constexpr (MODE == kFxsTwist) {
            // 4 x 4 blocks of (j, k1) (a 128-byte line of a 4 x 8 tile holds
            // four j of the input, four k1 of the output), in 8 x 8 blocks
            blk = i % ntg;
            const uint32_t r1 = i / ntg;
            const uint32_t jl = (r1 & 3u) | ((r1 >> 2) & 4u), kl = ((r1 >> 2) & 3u) | ((r1 >> 3) & 4u);
            const uint32_t r3 = r1 >> 6, nj8 = J >> 3;
            j = (r3 % nj8) * 8 + jl;
            ki = (r3 / nj8) * 8 + kl;
        } else {
            // column blocks of one transform group fastest, then j: the
            // items that run side by side sweep whole sub-transforms of a
            // few transforms (long runs in every buffer)
            ki = 0;
            const uint32_t nb = i & (nub - 1u), rest = i >> lgNub;
            j = rest & (J - 1u);
            blk = ((rest >> ps.lgJ) << lgNub) | nb;
        }
    };
    // this lane's transform t and column index u; false beyond the batch
    auto column = [&](uint32_t blk, uint32_t ki, uint32_t &t, uint32_t &u) {
        if constexpr (MODE == kFxsTwist) {
            t = 32 * blk + col;
            u = ki;
        } else {
            const uint32_t tg = blk >> lgNub, ub = blk & (nub - 1u);
            t = (tg << S::lgCt) + (col & ((1u << S::lgCt) - 1u));
            u = (ub << lgCe) + (col >> S::lgCt);
        }
        return t < a.batch;
    };
    auto fetch = [&](uint32_t i, Fx<W> &re, Fx<W> &im) {
        uint32_t j, blk, ki, t, u;
        item(i, j, blk, ki);
        if (!column(blk, ki, t, u)) {
#pragma unroll
            for (int w = 0; w < W; w++) { re.w[w] = 0u; im.w[w] = 0u; }
            return;
        }
        const uint32_t E = u * S::ics + (j + pos * J) * S::ies;
        if constexpr (MODE == kFxsFirst) {
            const uint32_t *b = a.in + a.in_stride * t;
            const int32_t eo = a.eoff[t];
            Mp<L> x;
            mp_load(x, b, M2, E);
            uint32_t bad = fx_from_mp<L, W>(re, x, a.lim, eo);
            mp_load(x, b, M2, N + E);
            bad |= fx_from_mp<L, W>(im, x, a.lim, eo);
            if (bad) atomicOr(a.flag, 1u);
        } else {
            tile_get<W>(re, im, src + tile_at<S::iTt, S::iTe, S::iSe>(t, E, N, W));
        }
    };
    // the wave's two twiddle tables: positions wv and wv + 8
    auto fetch_tw = [&](uint32_t i, uint32_t (&ta)[TWL], uint32_t (&tb)[TWL]) {
        uint32_t j, blk, ki;
        item(i, j, blk, ki);
        uint32_t ea, eb;
        if constexpr (MODE == kFxsTwist) {
            // w^(k1 n2), n2 = j + r J: entry toff + (k1 J + j) 16 + r (the
            // table in the order the items read it)
            ea = ps.toff + ((ki * J + j) << 4) + wv;
            eb = ea + 8;
        } else {
            const uint32_t k = j & (ps.Ns - 1u);
            ea = (wv == 0 || k == 0) ? c16e : ps.toff + 15 * k + wv - 1;
            eb = k == 0 ? c16e : ps.toff + 15 * k + wv + 7;
        }
        const uint32_t *t0 = a.tw + (size_t)ea * TWN, *t1 = a.tw + (size_t)eb * TWN;
#pragma unroll
        for (uint32_t u = 0; u < TWL; u++) {
            ta[u] = (lane + 64 * u < TWN) ? (t0 + 64 * u)[lane] : 0u;
            tb[u] = (lane + 64 * u < TWN) ? (t1 + 64 * u)[lane] : 0u;
        }
    };
    Fx<W> re, im;
    uint32_t ta[TWL], tb[TWL];
    if constexpr (EXT) fetch_tw(it, ta, tb);
    fetch(it, re, im);
    for (uint32_t n = 0;; n++) {
        uint32_t j, blk, ki;
        item(it, j, blk, ki);
        const uint32_t nx = it + gridDim.x;
        uint32_t *const slotA = setA + wv * SLOT, *const slotB = setB + wv * SLOT;
        // one set: the slots were last read in the previous item's stage B
        if (SETS == 1 && n > 0) __syncthreads();
        if constexpr (!EXT) {
#pragma unroll
            for (int i = 0; i < W; i++) slotA[i * 64 + lane] = re.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slotA[(W + i) * 64 + lane] = im.w[i];
        } else {
#pragma unroll
            for (uint32_t u = 0; u < TWL; u++)
                if (lane + 64 * u < TWN) {
                    tabE[lane + 64 * u] = ta[u];
                    tabE[TWN + lane + 64 * u] = tb[u];
                }
            MfOperand<W> Bx;
            mf_operand<W>(Bx, re, im);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            mf_cmul<W, false, true, r16_am(W)>(Bx, tabE, tabE + G.RW, tabE + 2 * G.RW, lane, slotA,
                                          tabE + TWN, tabE + TWN + G.RW, tabE + TWN + 2 * G.RW);
        }
        __syncthreads();
        // stage A
        Fx<W> yr, yi;
        dft4_rt<W>(yr, yi, lds + sa[0], lds + sa[1], lds + sa[2], lds + sa[3], m1);
        if (wv == 7 && hl == 1) {
            // -i (a + ib) = b - ia
            Fx<W> z, t_;
#pragma unroll
            for (int i = 0; i < W; i++) z.w[i] = 0;
            fx_sub(t_, z, yr);
            yr = yi;
            yi = t_;
        }
        if (SETS == 1) __syncthreads();               // stage A's reads of the one set
        if (wv < 4) {
            if (wv == 1 && hl == 1) {                  // w16^9 Y = w16^1 (-Y)
                Fx<W> z;
#pragma unroll
                for (int i = 0; i < W; i++) z.w[i] = 0;
                fx_sub(yr, z, yr);
                fx_sub(yi, z, yi);
            }
            MfOperand<W> B;
            mf_operand<W>(B, yr, yi);
            const uint32_t *t0 = w16t + r16_int_table(wv) * TWN;
            mf_cmul<W, false, false, r16_am(W)>(B, t0, t0 + G.RW, t0 + 2 * G.RW, lane, slotB);
        } else {
#pragma unroll
            for (int i = 0; i < W; i++) slotB[i * 64 + lane] = yr.w[i];
#pragma unroll
            for (int i = 0; i < W; i++) slotB[(W + i) * 64 + lane] = yi.w[i];
        }
        // the next item's operands and twiddle tables load during stage B
        // (the last pass: after its stores)
        if (nx < nitems) {
            if constexpr (MODE != kFxsLast) fetch(nx, re, im);
            if constexpr (MODE != kFxsLast && EXT) fetch_tw(nx, ta, tb);
        }
        __syncthreads();
        // stage B: output q = m1 + 4 m2 of this lane-half
        dft4_rt<W>(yr, yi, lds + sb[0], lds + sb[1], lds + sb[2], lds + sb[3], r2);
        {
            uint32_t t, u;
            if (column(blk, ki, t, u)) {
                const uint32_t k = j & (ps.Ns - 1u);
                const uint32_t oes = MODE == kFxsLast ? (N >> 8) : S::oes;
                const uint32_t E = u * S::ocs + ((j - k) * 16 + k + q * ps.Ns) * oes;
                if constexpr (MODE == kFxsLast) {
                    uint32_t *b = a.out + a.out_stride * t;
                    const int32_t eo = a.eoff[t];
                    Mp<L> z;
                    fx_to_mp<W, L, FULL>(z, yr, a.P, eo);
                    mp_store(z, b, M2, E);
                    fx_to_mp<W, L, FULL>(z, yi, a.P, eo);
                    mp_store(z, b, M2, N + E);
                } else {
                    tile_put<W>(dst + tile_at<S::oTt, S::oTe, S::oSe>(t, E, N, W), yr, yi);
                }
            }
        }
        if (nx >= nitems) break;
        if constexpr (MODE == kFxsLast) {
            fetch(nx, re, im);
            fetch_tw(nx, ta, tb);
        }
        it = nx;
    }
}

template <int L, int W, bool FULL>
static hipError_t launch_fxs(const FxArgs &a, const FxsPass *ps, int np, uint32_t c16e, int s0,
                             int s1, hipStream_t st)
{
    if constexpr (W > kPanelR16MaxW || W % 2 != 0) {
        return hipErrorInvalidValue;              // fxs_eligible: even W up to 34
    } else {
        static_assert(sizeof(uint32_t) * r16_lds_words<W>() <= 160 * 1024,
                      "radix-16 pass exceeds the LDS of a CU");
        const size_t sh = sizeof(uint32_t) * r16_lds_words<W>();
        // buf1 and buf0 alternate between the passes; the first reads a.in,
        // the last writes a.out
        for (int i = 0; i < np; i++) {
            if (i < s0 || i >= s1) continue;
            const FxsPass &P = ps[i];
            const uint32_t *src = (i & 1) ? a.buf0 : a.buf1;
            uint32_t *dst = (i & 1) ? a.buf1 : a.buf0;
            const uint32_t nblk = P.nitems < a.ncu ? P.nitems : a.ncu;
            if (P.nitems == 0) continue;
            // the pass must have the shapes its kernel is compiled for
            const bool t48 = P.tout == fxs_tile(2, 3, 0);
            auto shapes_ok = [&](auto shape) {
                using Sh = decltype(shape);
                return P.lgCt == Sh::lgCt && P.ics == Sh::ics && P.ies == Sh::ies &&
                       P.ocs == Sh::ocs && (P.mode == kFxsLast || P.oes == Sh::oes) &&
                       (P.mode == kFxsFirst || P.tin == fxs_tile(Sh::iTt, Sh::iTe, Sh::iSe)) &&
                       (P.mode == kFxsLast || P.tout == fxs_tile(Sh::oTt, Sh::oTe, Sh::oSe)) &&
                       (P.mode != kFxsLast || P.oes == (a.N >> 8));
            };
#define FXS_GO(MD, T)                                                                         \
    do {                                                                                      \
        if (!shapes_ok(FxsShape<MD, T>{})) return hipErrorInvalidValue;                       \
        hipLaunchKernelGGL((fxs_r16_kernel<W, L, MD, FULL, T>), dim3(nblk), dim3(512), sh, st, \
                           a, src, dst, P, c16e);                                             \
    } while (0)
            switch (P.mode) {
            case kFxsFirst:
                if (t48) FXS_GO(kFxsFirst, true);
                else FXS_GO(kFxsFirst, false);
                break;
            case kFxsShared:
                if (t48) FXS_GO(kFxsShared, true);
                else FXS_GO(kFxsShared, false);
                break;
            case kFxsTwist: FXS_GO(kFxsTwist, false); break;
            default: FXS_GO(kFxsLast, false); break;
            }
#undef FXS_GO
        }
        return hipGetLastError();
    }
}

hipError_t fxs_fft_launch(int L, int P, const FxArgs &a, const FxsPass *ps, int np,
                          uint32_t c16e, int s0, int s1, hipStream_t st)
{
    if (!fxs_eligible(a.N, L + 2) || np < 2 || np > kFxsMaxPasses || ps[0].mode != kFxsFirst ||
        ps[np - 1].mode != kFxsLast)
        return hipErrorInvalidValue;
    const bool full = (P == 32 * L);
#define CALLS(Lv) (full ? launch_fxs<Lv, Lv + 2, true>(a, ps, np, c16e, s0, s1, st) \
                        : launch_fxs<Lv, Lv + 2, false>(a, ps, np, c16e, s0, s1, st))
    switch (L) {
    case 2: return CALLS(2);
    case 4: return CALLS(4);
    case 6: return CALLS(6);
    case 8: return CALLS(8);
    case 12: return CALLS(12);
    case 16: return CALLS(16);
    case 24: return CALLS(24);
    case 32: return CALLS(32);
    default: return hipErrorInvalidValue;
    }
#undef CALLS
}

// ---------------------------------------------------------------------------

template <int L, int W, bool FULL>
static hipError_t launch_fxp(const FxArgs &a, int s0, int s1, hipStream_t st)
{
    const uint32_t N = a.N;
    auto in_range = [&](int s) { return s >= s0 && s < s1; };
    uint32_t *src = a.buf1, *dst = a.buf0;
    const uint32_t NG = fxp_groups(a.batch);          // panels, the last maybe partial
    // stage 0: input conversion into buf1 (every panel row, padding included)
    if (in_range(0))
        hipLaunchKernelGGL((fxp_in_kernel<L, W>), dim3((2 * N / 32) * 4 * NG), dim3(256), 0, st,
                           a, src);
    FxpPass ps[32];
    uint32_t nent = 0;
    const int np = fxp_plan(N, W, a.prb, ps, &nent);
    constexpr MfGeom G = mf_geom(W);
    int stage = 1;
    for (int i = 0; i < np; i++, stage++) {
        const FxpPass &P = ps[i];
        if (in_range(stage)) {
            if (P.R == 2) {
                hipLaunchKernelGGL((fxp_r2_kernel<W>), dim3((N / 2) * NG / 4),
                                   dim3(256), 0, st, a, (const uint32_t *)src, dst);
            } else if (P.R == 4) {
                const size_t sh = sizeof(uint32_t) * (FXP_SETS * 4 * 2 * W * 64 +
                                                      4 * mf_table_words(W));
                const uint32_t nitems = (N / 4) * NG;
                const uint32_t occ = fxp_occ(W);
                const uint32_t nblk = nitems < occ * a.ncu ? nitems : occ * a.ncu;
                hipLaunchKernelGGL((fxp_r4_kernel<W>), dim3(nblk), dim3(256), sh, st, a,
                                   (const uint32_t *)src, dst, P.Ns, P.toff, nitems);
            } else if (P.R == 16) {
                if constexpr (W <= kPanelR16MaxW) {
                    static_assert(sizeof(uint32_t) * r16_lds_words<W>() <= 160 * 1024,
                                  "radix-16 pass exceeds the LDS of a CU");
                    const size_t sh = sizeof(uint32_t) * r16_lds_words<W>();
                    const uint32_t nitems = (N / 16) * NG * 2;
                    const uint32_t cap = kR16Occ(W) * a.ncu;
                    const uint32_t nblk = nitems < cap ? nitems : cap;
                    if (P.Ns == 1)
                        hipLaunchKernelGGL((fxp_r16_kernel<W, false>), dim3(nblk), dim3(512), sh,
                                           st, a, (const uint32_t *)src, dst, P.Ns, P.toff,
                                           nitems, nent + 1);
                    else
                        hipLaunchKernelGGL((fxp_r16_kernel<W, true>), dim3(nblk), dim3(512), sh,
                                           st, a, (const uint32_t *)src, dst, P.Ns, P.toff,
                                           nitems, nent + 1);
                } else {
                    return hipErrorInvalidValue;    // no radix-16 pass at this width
                }
            } else if constexpr (W <= kPanelR8MaxW) {
                const size_t sh =
                    sizeof(uint32_t) * (16 * 2 * W * 64 + G.RW + 8 * mf_table_words(W));
                const uint32_t nitems = (N / 8) * NG;
                const uint32_t nblk = nitems < a.ncu ? nitems : a.ncu;
                hipLaunchKernelGGL((fxp_r8_kernel<W>), dim3(nblk), dim3(512), sh, st, a,
                                   (const uint32_t *)src, dst, P.Ns, P.toff, nitems, nent);
            } else {
                return hipErrorInvalidValue;        // no radix-8 pass at this width
            }
        }
        std::swap(src, dst);
    }
    // the last stage: output conversion
    if (in_range(stage))
        hipLaunchKernelGGL((fxp_out_kernel<W, L, FULL>), dim3((2 * N / 32) * ((a.batch + 15) / 16)),
                           dim3(256), 0, st, a, (const uint32_t *)src);
    return hipGetLastError();
}

hipError_t fxp_fft_launch(int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st)
{
    if (!fxp_eligible(a.N, a.batch)) return hipErrorInvalidValue;
    const bool full = (P == 32 * L);
#define CALLP(Lv) (full ? launch_fxp<Lv, Lv + 2, true>(a, s0, s1, st) \
                        : launch_fxp<Lv, Lv + 2, false>(a, s0, s1, st))
    switch (L) {
    case 1: return CALLP(1);
    case 2: return CALLP(2);
    case 3: return CALLP(3);
    case 4: return CALLP(4);
    case 6: return CALLP(6);
    case 8: return CALLP(8);
    case 12: return CALLP(12);
    case 16: return CALLP(16);
    case 24: return CALLP(24);
    case 32: return CALLP(32);
    case 64: return CALLP(64);
    default: return hipErrorInvalidValue;
    }
#undef CALLP
}

}  // namespace mpfft
