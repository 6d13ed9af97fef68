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

// Subtrees finished in LDS (plain forms): the maximal subtrees of node size
// <= S (nodes of size S or S/2 whose parent is larger; the whole transform
// when N <= S), one workgroup each: its leaves, then every combine level
// inside it, on the subtree held in LDS, then one write back.
constexpr int kSubMaxLevel = 16;
struct SubArgs {
    LeafArgs leaf;              // inputs (element-major copy), leaf positions, out
    const uint32_t *lut;        // twiddles as in LevelArgs
    uint32_t lut_m, lut_e;
    int code;
    const uint32_t *sub;        // per subtree: offset << 1 | kind (0: size 2^lgs[0],
                                // 1: size 2^lgs[1])
    uint32_t nsub;
    const uint32_t *rel;        // node offsets inside a subtree, per kind and level
    uint32_t rel_first[2][kSubMaxLevel], rel_cnt[2][kSubMaxLevel];
    uint32_t lut_off[kSubMaxLevel];
    uint32_t lgs[2];
    uint32_t nthr;              // threads per workgroup
};

// kernels exist for these limb counts; a precision P uses the smallest
// count >= ceil(P/32)
int limbs_for_precision(long P);

// Default subtree size finished in LDS (0: none; MPFFT_SUB_SIZE sets one).
// Measured at N = 2^16, batch 64, against the level kernels on HBM: equal
// at 64 and 128 bits (S = 512: 69.0 k vs 68.9 k FFT/s; S = 128: 46.4 k vs
// 45.9 k), slower at 256 bits (best S = 128: 25.5 k vs 26.6 k) and 512 bits
// (best S = 128: 11.0 k vs 13.7 k).  A subtree held in LDS costs 4 (L+1)
// words per lane beside the 2 (L+1) parking words, which caps a CU near 12
// waves where the level kernels run about 24, and at these precisions a
// level's HBM sweep costs no more than its arithmetic: so none by default.
inline uint32_t mp_sub_size(int) { return 0u; }

// leaves and levels; with sub != nullptr the leaves and the levels of node
// size <= 2^lgs[0] run as the subtree kernel first
hipError_t mp_fft_launch(int L, int P, const LeafArgs &la, const LevelArgs *lv,
                         int nlev, uint32_t batch, hipStream_t st,
                         const SubArgs *sub = nullptr);
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

// One transform of N points between chunk-major staging (chunk c = points
// [c cl, (c + 1) cl): R rows of 2 cl words, real parts then imaginary parts)
// and SoA rows (R rows of 2N words); to_soa picks the direction; cl divides N.
hipError_t mp_chunk_relayout_launch(const uint32_t *src, uint32_t *dst, uint32_t N, uint32_t R,
                                    uint32_t cl, bool to_soa, hipStream_t st);

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
    int lim;                    // scaled inputs x 2^-eoff must have exponent <= lim
                                // (|x 2^-eoff| < 2^lim); the word forms: >= -lim
                                // too, smaller ones are taken as zero
    int P;
    uint32_t ncu;               // compute units of the device (persistent grids)
    // block exponent: per transform b, eoff[b] = the largest exponent of its
    // 2N input reals (0 when all are zero); the conversions divide the input
    // by 2^eoff[b] and multiply the output by it (exact shifts), so the
    // fixed-point and word forms see |x| < 1 whatever the input's scale
    int32_t *eoff;              // batch entries
    int32_t *eoff_part;         // batch x eoff_nblk partial maxima of the scan
    uint32_t eoff_nblk;         // blocks of the scan per transform (fx_eoff_blocks)
    int prb;                    // panel kernels: log2 of the main radix (fxp_plan)
};

// Exponent range of the limb format's inputs: |exponent| < kInputExpLimit
// (MPFR's default range is 1 - 2^30 .. 2^30 - 1; an FFT of up to 2^24
// points keeps every intermediate and output inside it from there).
constexpr int32_t kInputExpLimit = 1 << 29;

// blocks per transform of the block-exponent scan: one per 8192 header words
inline uint32_t fx_eoff_blocks(uint32_t N)
{
    const uint32_t b = (2 * N + 8191) / 8192;
    return b < 1 ? 1u : b;
}

// The block-exponent scan of the fixed and word forms: eoff[b] for every
// transform of a.in (two launches, one where 2N <= 8192; no memset: it can
// be captured in a graph).  Flags (a.flag) a transform whose largest exponent is outside
// the limb format's input range.
hipError_t fx_eoff_launch(const FxArgs &a, hipStream_t st);

// fixed-point words per real for limb count L (F = 32 (L + 1) fraction bits)
constexpr int fx_words(int L) { return L + 2; }
// digit width of the per-lane twiddle product for reals of W words (the
// host cuts the per-lane twiddle tables into digits of this width): 29 bits
// while a real has at most 63 digits (W <= 57), 28 beyond
constexpr int fx_lane_digit_bits(int W) { return (32 * W + 28) / 29 <= 63 ? 29 : 28; }

// stages [s0, s1) of the fixed form
hipError_t fx_fft_launch(int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st);

// ---- fixed form, panel kernels (fx_panel.hip) ----
//
// Transforms are interleaved in panels of kPanel: element e of the 64
// transforms of panel g is a contiguous block of W x 256 bytes at
// (g 2N + e) W 64 words (inside it, word pairs per transform for even W:
// fx_panel.hip, pan_word).  A twiddle
// product runs on the int8 matrix cores: the twiddle t (|t| <= 1) and the
// operand x are written in balanced base-256 digits (each in [-128, 127]),
// and the columns of x t are a Toeplitz matrix of t's digits times the
// digit vectors of 32 operands that share t (the same element of 32
// transforms), one v_mfma_i32_32x32x32_i8 per 32 x 32 tile.

constexpr uint32_t kPanel = 64;        // transforms per panel

// panels of a batch; a last, partial panel is padded with zero transforms
__host__ __device__ constexpr uint32_t fxp_groups(uint32_t batch)
{
    return (batch + kPanel - 1) / kPanel;
}

struct MfGeom {
    int NXB;       // digits of an operand: 4 W
    int NTB;       // digits of a twiddle real: 4 W - 3 (|t| <= 1 at F bits)
    int C0;        // lowest product column kept: F / 8 - 4
    int NCOL;      // columns kept: C0 .. C0 + NCOL - 1 (4 below 2^-F, 4 W above)
    int KT, MT;    // 32-wide tiles over the operand digits / the kept columns
    int DLO, DHI;  // tile diagonals p - q with a nonzero twiddle digit
    int U;         // digit table of a twiddle real: byte u holds digit U - u
    int RW;        // 32-bit words of that table (with 4 bytes of over-read)
};

// tile (p, q) holds some twiddle digit T[C0 + 32 (p - q) + m - k], m < 32
// rows (kept columns), k < 32 operand digits
constexpr bool mf_live(int W, int p, int q)
{
    const int NXB = 4 * W, NTB = 4 * W - 3, C0 = 4 * W - 8, NCOL = 4 * W + 4;
    const int mmax = (NCOL - 32 * p < 32 ? NCOL - 32 * p : 32) - 1;
    const int kmax = (NXB - 32 * q < 32 ? NXB - 32 * q : 32) - 1;
    const int lo = C0 + 32 * (p - q) - kmax, hi = C0 + 32 * (p - q) + mmax;
    return mmax >= 0 && kmax >= 0 && hi >= 0 && lo <= NTB - 1;
}

constexpr MfGeom mf_geom(int W)
{
    MfGeom g{};
    g.NXB = 4 * W;
    g.NTB = 4 * W - 3;
    g.C0 = 4 * W - 8;
    g.NCOL = 4 * W + 4;
    g.KT = (g.NXB + 31) / 32;
    g.MT = (g.NCOL + 31) / 32;
    g.DLO = 1 << 20;
    g.DHI = -(1 << 20);
    for (int p = 0; p < g.MT; p++)
        for (int q = 0; q < g.KT; q++)
            if (mf_live(W, p, q)) {
                if (p - q < g.DLO) g.DLO = p - q;
                if (p - q > g.DHI) g.DHI = p - q;
            }
    g.U = g.C0 + 32 * g.DHI + 31;
    g.RW = (32 * (g.DHI - g.DLO) + 67 + 3) / 4;
    return g;
}

// A product's three twiddle digit tables (c, d, -d: 3 RW words) may sit at
// the end of the wave's LDS slot when they overlap only im' words that the
// last product tile writes (im' word t at slot words (W + t) 64 ..; tile p
// writes words t = 8p - 1 .. 8p + 6), i.e. up to W = 34; wider reals keep
// them in a per-wave area after the slots.
constexpr bool mf_tail_fits(int W)
{
    return 3 * mf_geom(W).RW <= (W - 8 * (mf_geom(W).MT - 1) + 1) * 64;
}

constexpr int mf_table_words(int W) { return mf_tail_fits(W) ? 0 : 3 * mf_geom(W).RW; }

// the panel kernels run any batch at N >= 16
inline bool fxp_eligible(uint32_t N, uint32_t batch)
{
    return N >= 16 && batch > 0;
}

// The per-lane kernels run radix-16 passes up to this many words per real
// (fxl_plan).  Beyond, the exchange of a butterfly's 16 values (16 x 2 W
// words x 64 lanes of LDS) leaves one block per CU: at 14 and 18 words the
// radix-16 lanes measured 1.4-2.5 times the radix-4 lanes' time and 2.0-2.7
// times the panels' at N = 2^10 .. 2^16, batch 64.
constexpr int kFxl16MaxW = 10;

// Where those beat the panels at any batch: measured on MI355X, time per
// transform of the per-lane radix-16 passes over the panels at batch 64:
// 0.64-0.92 for N >= 2^14 at 3 to 10 words per real, 0.81-1.06 at N = 2^12,
// and for N <= 2^10 0.81-1.07 at 3 and 4 words but 0.92-1.6 at 5 to 10
// (N = 2^16, 256 bits: 63.4 k against 47.2 k FFT/s).
inline bool fxl_preferred(uint32_t N, int W)
{
    return W <= kFxl16MaxW && (N >= 4096 || W <= 4);
}

// MPFFT_FORM_FIXED takes the panels (reals of W words) unless fxl_preferred,
// when the padding of the last panel is at most a quarter of the batch, and
// at W >= 34 for N <= 1024 at any batch.  Measured on MI355X, time per
// transform of the per-lane radix-4 kernels over the panels at N = 2^6 ..
// 2^14, P = 128 .. 1024, batches 16 .. 128: 1.1-1.3 at full panels, falling
// with the padding to ~1 near a quarter of the batch for N >= 2^12; at W =
// 34 and N <= 2^10 above 1 at every batch (the per-lane kernels run few
// waves of very wide reals there).
inline bool fxp_preferred(uint32_t N, uint32_t batch, int W)
{
    if (!fxp_eligible(N, batch) || fxl_preferred(N, W)) return false;
    if (W >= 34 && N <= 1024) return true;
    return 4 * (fxp_groups(batch) * kPanel - batch) <= batch;
}

// Passes of the fixed form on panels, for a main radix 2^rb: rb = 4 (radix
// 16: eight waves hold 16 positions of 32 transforms and run the DFT16 as
// 4 x 4 through LDS; W <= kPanelR16MaxW), rb = 3 (radix 8: 8 positions of
// 64 transforms; W <= kPanelR8MaxW) or rb = 2 (radix 4, any width).  When
// log2 N is not a multiple of rb the first pass has radix 2^(log2 N mod rb)
// (radix 8 there needs W <= kPanelR8MaxW; wider: radix 2 then radix 4).
// Pass i has radix R and sub-transform size Ns; its twiddles w^(r k) (k <
// Ns, r = 1 .. R-1) are table entries toff + (R - 1) k + r - 1.  Seven more
// entries follow at index *nent: w^(N/8) (its real part 1/sqrt(2) is the
// constant of the radix-8 butterfly), w^0 = 1, then the radix-16 internal
// twiddles w16^1, w16^3, w16^9, w16^2, w16^6 = w^(N/16), w^(3N/16),
// w^(9N/16), w^(N/8), w^(3N/8).
constexpr int kPanelR8MaxW = 18;
constexpr int kPanelR16MaxW = 34;
constexpr int kPanelConstEntries = 7;

struct FxpPass {
    uint32_t R, Ns, toff;
};

// the main radixes a width allows
inline bool fxp_radix_ok(int W, int rb)
{
    return rb == 2 || (rb == 3 && W <= kPanelR8MaxW) || (rb == 4 && W <= kPanelR16MaxW);
}

// The default main radix per width: radix 16 up to kPanelR16MaxW words per
// real, radix 4 beyond.  Measured on MI355X at N = 2^16, batch 64 (FFT/s
// with the radix-16 plan against the best other): 81.0k / 70.1k (radix 8) at
// 128 bits, 46.2k / 42.1k at 256, 28.2k / 25.4k at 512, 18.8k / 16.6k
// (radix 4) at 768, 13.6k / 12.4k at 1024.
inline int fxp_radix_default(int W) { return W <= kPanelR16MaxW ? 4 : 2; }

inline int fxp_plan(uint32_t N, int W, int rb, FxpPass (&p)[32], uint32_t *nent)
{
    int lg = 0;
    while ((1u << lg) < N) lg++;
    if (!fxp_radix_ok(W, rb)) rb = 2;
    int n = 0;
    uint32_t Ns = 1, toff = 0;
    auto add = [&](uint32_t R) {
        p[n++] = {R, Ns, toff};
        toff += (R - 1) * Ns;
        Ns *= R;
    };
    const int rem = lg % rb;
    if (rem == 3 && W > kPanelR8MaxW) {
        add(2);
        add(4);
    } else if (rem) {
        add(1u << rem);
    }
    while (Ns < N) add(1u << rb);
    *nent = toff;
    return n;
}

// stages [s0, s1) of the fixed form on panels: the input conversion, one
// stage per pass, the output conversion (rounding back to floats)
hipError_t fxp_fft_launch(int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st);

// ---- fixed form on panels, four-step plan (fx_panel.hip, fxs_*) ----
//
// For N = 16^m (m >= 3) and reals of at most kPanelR16MaxW words: N = N1 N2
// with N2 = 256, and with x[n1 N2 + n2] the input
//   step 1: A[k1, n2] = sum_n1 x[n1 N2 + n2] w_N1^(n1 k1)   (held at k1 N2 + n2)
//   step 2: B[k1, n2] = A[k1, n2] w_N^(k1 n2)
//   step 3: X[k1 + N1 k2] = sum_n2 B[k1, n2] w_N2^(n2 k2)
// Every pass is a radix-16 Stockham DIT pass (the radix-16 panel kernel's
// DFT16) over the sub-transforms of one step.  A work item is one
// butterfly index j of 32 "columns" that share its twiddles:
//   first  (step 1, Ns = 1): no twiddles; columns 32 consecutive
//          (transform, n2); reads the limb input and converts it;
//   shared (step 1 or 3, Ns > 1): twiddles w_N^(r k N / (16 Ns)), k = j mod
//          Ns, the same for every column: columns 32 consecutive
//          (transform, n2) in step 1, (transform, k1) in step 3;
//   twist  (step 3, Ns = 1): the twiddles of step 2, w_N^(k1 n2) with
//          n2 = j + r N2/16, one per (k1, n2): columns 32 transforms;
//   last   (step 3's last pass): as shared, and rounds the output into the
//          limb layout (X[k1 + N1 k2]: 32 consecutive k1 are 32 consecutive
//          elements).
// So the input and output conversions are the first and last passes' loads
// and stores.  Columns are Ct consecutive transforms times 32/Ct consecutive
// column indices u (n2 in step 1, k1 in step 3; twist: 32 transforms of one
// k1), and the buffers between passes are tiled to match: a tile holds
// Tt transforms x Te elements that are Se apart, as W word-pair slices of
// Tt Te = 32 pairs (256 bytes), transform fastest inside a slice.
// A half-wave whose columns match the tile moves whole 256-byte slices; the
// two transposing transitions (step 1 -> twist -> last) use 4 x 8 tiles, so
// the twist pass moves eight 32-byte runs per slice on either side:
//   first, shared (step 1): columns 1 x 32, tiles 1 x 32 (the pass before
//     the twist: columns 4 x 8, writing 4 x 8 tiles over n2);
//   twist: columns 32 x 1, reading 4 x 8 tiles over n2, writing 4 x 8 tiles
//     over k1 (Se = N2); its items in 8 x 8 blocks of (j, k1), so the items
//     that share a tile run side by side;
//   last: columns 4 x 8 (8 consecutive k1: 32-byte runs of the limb output).
// Element addresses: input u ics + i ies, output u ocs + i' oes, with i =
// j + r M/16 the position read and i' = (j - k) 16 + k + q Ns the one written.
enum : uint32_t { kFxsFirst = 0, kFxsShared = 1, kFxsTwist = 2, kFxsLast = 3 };

// a tile shape: log2 Tt | log2 Te << 8 | log2 Se << 16
constexpr uint32_t fxs_tile(uint32_t lgTt, uint32_t lgTe, uint32_t lgSe)
{
    return lgTt | lgTe << 8 | lgSe << 16;
}

struct FxsPass {
    uint32_t mode;
    uint32_t Ns;                // sub-transform size before the pass (1: none)
    uint32_t toff;              // shared, last: entry of (k, r) = toff + 15 k + r - 1;
                                // twist: entry of (k1, j, r) = toff + (16 k1 + j) 16 + r,
                                // w^(k1 (j + 16 r))
    uint32_t lgJ;               // log2 of the butterflies per sub-transform (M / 16)
    uint32_t lgU;               // log2 of the column indices u per transform (N2 or N1)
    uint32_t lgCt;              // log2 of the transforms in a block of 32 columns
    uint32_t ics, ies, ocs, oes;
    uint32_t tin, tout;         // tile shapes of the input and output buffers
    uint32_t order;             // 0: column block fastest, then j; 1: the twist order
    uint32_t nitems;
};

constexpr int kFxsMaxPasses = 8;

inline bool fxs_eligible(uint32_t N, int W)
{
    int lg = 0;
    while ((1u << lg) < N) lg++;
    return W <= kPanelR16MaxW && W % 2 == 0 && (1u << lg) == N && lg % 4 == 0 && lg >= 12 &&
           lg <= 24;
}

// The four-step plan for N (fxs_eligible) and a batch: the passes and the
// table entries they use (*nent; the kPanelConstEntries constants follow)
inline int fxs_plan(uint32_t N, uint32_t batch, FxsPass (&p)[kFxsMaxPasses], uint32_t *nent)
{
    const uint32_t xt = 2, xe = 3;           // the transposing tiles: 4 x 8
    int lg = 0;
    while ((1u << lg) < N) lg++;
    const uint32_t N2 = 256, N1 = N / N2, lg1 = (uint32_t)lg - 8;
    int n = 0;
    uint32_t toff = 0;
    auto blocks = [&](uint32_t lgCt, uint32_t lgU) {     // blocks of 32 columns
        const uint32_t Ct = 1u << lgCt;
        return ((batch + Ct - 1) / Ct) * ((1u << lgU) / (32 / Ct));
    };
    // step 1: N1 = 16^a points over n1, columns (transform, n2)
    for (uint32_t Ns = 1; Ns < N1; Ns *= 16) {
        const bool before_twist = Ns * 16 == N1;
        FxsPass q{};
        q.mode = Ns == 1 ? kFxsFirst : kFxsShared;
        q.Ns = Ns; q.toff = toff; q.lgJ = lg1 - 4; q.lgU = 8;
        q.lgCt = before_twist ? 2 : 0;
        q.ics = 1; q.ies = N2; q.ocs = 1; q.oes = N2;
        q.tin = Ns == 1 ? 0u : fxs_tile(0, 5, 0);
        q.tout = before_twist ? fxs_tile(xt, xe, 0) : fxs_tile(0, 5, 0);
        q.order = 0;
        q.nitems = (N1 / 16) * blocks(q.lgCt, q.lgU);
        if (Ns > 1) toff += 15 * Ns;
        p[n++] = q;
    }
    // step 3 with the step-2 twiddles in its first pass: 256 points over n2
    {
        FxsPass q{};
        q.mode = kFxsTwist;
        q.Ns = 1; q.toff = toff; q.lgJ = 4; q.lgU = lg1; q.lgCt = 5;
        q.ics = N2; q.ies = 1; q.ocs = N2; q.oes = 1;
        q.tin = fxs_tile(xt, xe, 0); q.tout = fxs_tile(xt, xe, 8);
        q.order = 1;
        q.nitems = ((batch + 31) / 32) * N1 * 16;
        toff += N;
        p[n++] = q;
    }
    {
        FxsPass q{};
        q.mode = kFxsLast;
        q.Ns = 16; q.toff = toff; q.lgJ = 4; q.lgU = lg1; q.lgCt = 2;
        q.ics = N2; q.ies = 1; q.ocs = 1; q.oes = N1;
        q.tin = fxs_tile(xt, xe, 8); q.tout = 0;
        q.order = 0;
        q.nitems = 16 * blocks(q.lgCt, q.lgU);
        toff += 15 * 16;
        p[n++] = q;
    }
    *nent = toff;
    return n;
}

// passes [s0, s1) of the four-step plan (pass 0 reads a.in and needs the
// block exponents of fx_eoff_launch; the last writes a.out); c16e: the
// table entry of w^0 = 1, the five internal twiddles of the DFT16 after it
hipError_t fxs_fft_launch(int L, int P, const FxArgs &a, const FxsPass *ps, int np,
                          uint32_t c16e, int s0, int s1, hipStream_t st);

// ---- double- and quad-word forms (dd_fft.hip) ----
//
// Same arguments as the fixed form, for K = 2 or 4 binary64 words per real:
// buf0 / buf1 hold K/2 double2 per real (stride = N K double2 per
// transform), tw holds ntw entries of 2K doubles (the real part's K words,
// then the imaginary part's) in the per-lane fixed form's order, lim bounds
// |exponent| of the inputs.  Stages: one per pass (the first splits the limb
// input, the last rounds to P bits); N = 1: one.
hipError_t dd_fft_launch(int K, int L, int P, const FxArgs &a, int s0, int s1, hipStream_t st);

// Passes of the word forms (double and quad words alike): radix 16, four
// threads per butterfly, with a radix-2 pass first where log2 N is odd and a
// radix-4 pass after the first radix-16 one where bit 1 of log2 N is set;
// radix 2 and 4 alone below N = 16.  (A radix-16 pass with one thread per
// butterfly ran at three waves per SIMD for double words and converting at
// one; the quad-word form ran radix 4 throughout, 16 % slower at N = 2^16.)
// The first pass also splits the limb input into words and the last rounds
// back to limbs; the radix-2 kernel has no twiddles, so it runs only as a
// first pass (Ns = 1).
// Pass i has radix R and sub-transform size Ns; its twiddles w^(r k N / (R
// Ns)) (k < Ns, r = 1 .. R-1) are table entries toff + (r - 1) Ns + k (k
// fastest: the lanes of a wave, consecutive k, read consecutive entries).
// With a radix-16 pass, ten entries from *nent on hold w16^m = w^(m N / 16),
// m = 0 .. 9, the internal twiddles of the 4 x 4 butterfly.
struct MwPass {
    uint32_t R, Ns, toff;
};

inline int mw_plan(uint32_t N, int K, MwPass (&p)[32], uint32_t *nent)
{
    int lg = 0;
    while ((1u << lg) < N) lg++;
    int n = 0;
    uint32_t Ns = 1, toff = 0;
    auto add = [&](uint32_t R) {
        p[n++] = {R, Ns, toff};
        toff += (R - 1) * Ns;
        Ns *= R;
    };
    (void)K;                                         // the same plan for K = 2 and 4
    if (lg >= 4) {
        if (lg & 1) add(2);
        add(16);
        if (lg & 2) add(4);
        for (int i = 1; i < lg / 4; i++) add(16);
    } else {
        if (lg & 1) add(2);
        while (Ns < N) add(4);
    }
    *nent = toff;
    return n;
}

// Passes of the fixed form on lanes (fx_fft.hip).  Up to kFxl16MaxW words
// per real (P <= 256) and from N = 16 on, the word forms' plan: radix 16
// (four waves per butterfly, a lane holding four positions), a radix-2 pass
// first where log2 N is odd and a radix-4 pass after the first radix-16 one
// where bit 1 of log2 N is set; wider reals (four positions per lane would
// not fit the registers) and smaller N: a radix-2 pass first where log2 N
// is odd, then radix 4.  Twiddle entries of a pass from toff: radix 16
// r-major, (r - 1) Ns + k (the lanes of a wave hold consecutive k); radix 4
// and 2 k-major, (R - 1) k + r - 1 (a DPP quad holds the r of one k); with a
// radix-16 pass, ten entries from *nent hold w16^m, m = 0 .. 9.
inline int fxl_plan(uint32_t N, int W, MwPass (&p)[32], uint32_t *nent)
{
    if (W <= kFxl16MaxW) return mw_plan(N, 2, p, nent);
    int lg = 0;
    while ((1u << lg) < N) lg++;
    int n = 0;
    uint32_t Ns = 1, toff = 0;
    auto add = [&](uint32_t R) {
        p[n++] = {R, Ns, toff};
        toff += (R - 1) * Ns;
        Ns *= R;
    };
    if (lg & 1) add(2);
    while (Ns < N) add(4);
    *nent = toff;
    return n;
}

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
