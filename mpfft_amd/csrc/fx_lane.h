// Cross-lane butterflies of the fixed form: the radix-4 and radix-2 DFTs over
// the lanes of a DPP quad (or pair), each lane holding one complex fixed-point
// value.  Shared by the per-lane kernels (fx_fft.hip) and the first pass of
// the panel kernels (fx_panel.hip).
#pragma once
#include "fx_arith.h"

namespace mpfft {

template <int W>
__device__ __forceinline__ void fx_dpp(Fx<W> &o, const Fx<W> &x, int ctrl)
{
    // quad permutation of every word; ctrl is a compile-time constant at
    // each call site (0x4E: [2,3,0,1], 0xB1: [1,0,3,2])
#pragma unroll
    for (int j = 0; j < W; j++) {
        o.w[j] = ctrl == 0x4E
                     ? (uint32_t)__builtin_amdgcn_mov_dpp((int)x.w[j], 0x4E, 0xF, 0xF, false)
                     : (uint32_t)__builtin_amdgcn_mov_dpp((int)x.w[j], 0xB1, 0xF, 0xF, false);
    }
}

template <int W>
__device__ __forceinline__ void fx_select(Fx<W> &o, uint32_t c, const Fx<W> &a, const Fx<W> &b)
{
#pragma unroll
    for (int j = 0; j < W; j++) o.w[j] = c ? a.w[j] : b.w[j];
}

// z = A + (-1)^s B with (A, B) = (p, x) when sw, else (x, p)
template <int W>
__device__ __forceinline__ void fx_pair(Fx<W> &z, const Fx<W> &x, const Fx<W> &p, uint32_t sw,
                                        uint32_t s)
{
    const uint32_t m = s ? 0xffffffffu : 0u;
    uint32_t c = s;
#pragma unroll
    for (int j = 0; j < W; j++) {
        const uint32_t a = sw ? p.w[j] : x.w[j], b = sw ? x.w[j] : p.w[j];
        z.w[j] = add_cc(a, b ^ m, c);
    }
}

// DFT4 over a quad: lane r holds v_r on entry and y_{rout(r)} on exit,
// y_q = sum_r v_r (-i)^(r q), rout = (0, 2, 1, 3).
//   Stage 1 (partner r ^ 2): lanes 0, 1 form t0 = v0 + v2, t2 = v1 + v3;
//                            lanes 2, 3 form t1 = v0 - v2, t3 = v1 - v3.
//   Stage 2 (partner r ^ 1): lane 0: y0 = t0 + t2, lane 1: y2 = t0 - t2,
//                            lane 2: y1 = t1 - i t3, lane 3: y3 = t1 + i t3.
__device__ __forceinline__ uint32_t fx_dft4_rout(uint32_t r)
{
    return (r == 1) ? 2u : (r == 2 ? 1u : r);
}

template <int W>
__device__ __forceinline__ void fx_dft4_quad(Fx<W> &re, Fx<W> &im, uint32_t r)
{
    {
        const uint32_t hi = r >> 1;
        Fx<W> p;
        fx_dpp(p, re, 0x4E);
        fx_pair(re, re, p, hi, hi);
        fx_dpp(p, im, 0x4E);
        fx_pair(im, im, p, hi, hi);
    }
    //   lanes 0, 1: re = own.re +- p.re, im likewise (lane 1: p - own)
    //   lane 2 (own t1, p t3): re = own.re + p.im, im = own.im - p.re
    //   lane 3 (own t3, p t1): re = p.re - own.im, im = p.im + own.re
    {
        const uint32_t odd = r & 1u, cross = r >> 1;
        Fx<W> pre, pim, nre;
        fx_dpp(pre, re, 0xB1);
        fx_dpp(pim, im, 0xB1);
        {
            Fx<W> x, y;
            fx_select(x, cross, odd ? pre : re, re);
            fx_select(y, cross, odd ? im : pim, pre);
            fx_pair(nre, x, y, cross ? 0u : odd, odd);
        }
        {
            Fx<W> x, y;
            fx_select(x, cross, odd ? pim : im, im);
            fx_select(y, cross, odd ? re : pre, pim);
            fx_pair(im, x, y, cross ? 0u : odd, cross ? (odd ^ 1u) : odd);
        }
        re = nre;
    }
}

// DFT2 over a lane pair: lane 0 gets v0 + v1, lane 1 gets v0 - v1.
template <int W>
__device__ __forceinline__ void fx_dft2_pair(Fx<W> &re, Fx<W> &im, uint32_t r)
{
    Fx<W> pre, pim, x, y;
    fx_dpp(pre, re, 0xB1);
    fx_dpp(pim, im, 0xB1);
    fx_select(x, r, pre, re);
    fx_select(y, r, re, pre);
    fx_addsub(re, x, y, r);
    fx_select(x, r, pim, im);
    fx_select(y, r, im, pim);
    fx_addsub(im, x, y, r);
}

}  // namespace mpfft
