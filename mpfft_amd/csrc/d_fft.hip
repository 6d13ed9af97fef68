// Double-precision split-radix conjugate-pair DIT FFT on gfx950, in the
// reference's operation order (built with -ffp-contract=off so each
// product and sum rounds separately, as gcc -O3 on x86-64 does):
//   code 5: lab/serial/double/serial_d_fft5.c:93-229 (three butterflies,
//           leaves n <= 4)
//   code 6: lab/serial/double/serial_d_fft6.c:93-429 (one butterfly for every
//           k, hard-coded 8- and 16-point leaves); code 7 (serial_d_fft7.c)
//           performs the same arithmetic.
//
// Layout: split real/imaginary arrays of N doubles (re at [0,N), im at
// [N,2N) of one buffer per transform).  Twiddles: wr[N/4], wi[N/4] and five
// fixed constants (w8, w16_1..4), computed on the host with the reference's
// own libm expressions (serial_d_fft5.c:52-62).
//
// Small transforms (2N doubles fit in LDS) run as one workgroup per
// transform: leaves, then every level from LDS with a barrier in between.
// Larger ones run one launch per level on HBM.
#include "kernels.h"

namespace mpfft {

__device__ __forceinline__ uint32_t d_idx(int32_t sn, uint32_t s, uint32_t j, uint32_t N)
{
    int64_t v = (int64_t)sn + (int64_t)j * s;
    return (uint32_t)(v < 0 ? v + N : v);
}

// leaves n = 1, 2, 4 (serial_d_fft5.c:97-147, identical in serial_d_fft6.c)
__device__ void d_leaf_small(uint32_t n, int32_t sn, uint32_t N, const double *xr,
                             const double *xi, double *yr, double *yi)
{
    const uint32_t s = N / n;
    const uint32_t i0 = d_idx(sn, s, 0, N);
    if (n == 1) { yr[0] = xr[i0]; yi[0] = xi[i0]; return; }
    const uint32_t i1 = d_idx(sn, s, 1, N);
    if (n == 2) {
        yr[0] = xr[i0] + xr[i1]; yr[1] = xr[i0] - xr[i1];
        yi[0] = xi[i0] + xi[i1]; yi[1] = xi[i0] - xi[i1];
        return;
    }
    const uint32_t i2 = d_idx(sn, s, 2, N), i3 = d_idx(sn, s, 3, N);
    const double er0 = xr[i0] + xr[i2], er1 = xr[i0] - xr[i2];
    const double or0 = xr[i1] + xr[i3], or1 = xr[i1] - xr[i3];
    const double ei0 = xi[i0] + xi[i2], ei1 = xi[i0] - xi[i2];
    const double oi0 = xi[i1] + xi[i3], oi1 = xi[i1] - xi[i3];
    yr[0] = er0 + or0; yr[1] = er1 + oi1; yr[2] = er0 - or0; yr[3] = er1 - oi1;
    yi[0] = ei0 + oi0; yi[1] = ei1 - or1; yi[2] = ei0 - oi0; yi[3] = ei1 + or1;
}

// 8-point leaf of code 6 (serial_d_fft6.c:148-212)
__device__ void d_leaf8(int32_t sn, uint32_t N, const double *xr, const double *xi,
                        double *yr, double *yi, double w8)
{
    const uint32_t s = N / 8;
    double R[8], I[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t g = d_idx(sn, s, j, N);
        R[j] = xr[g]; I[j] = xi[g];
    }
    // input pairs (0,4) (2,6) (1,7) (3,5): sum / difference
    const double s0r = R[0] + R[4], d0r = R[0] - R[4], s1r = R[2] + R[6], d1r = R[2] - R[6];
    const double s2r = R[1] + R[7], d2r = R[1] - R[7], s3r = R[3] + R[5], d3r = R[3] - R[5];
    const double s0i = I[0] + I[4], d0i = I[0] - I[4], s1i = I[2] + I[6], d1i = I[2] - I[6];
    const double s2i = I[1] + I[7], d2i = I[1] - I[7], s3i = I[3] + I[5], d3i = I[3] - I[5];
    const double m1r = w8 * (d2r + d3r), m1i = w8 * (d2i + d3i);
    const double m2r = w8 * (s2r - s3r), m2i = w8 * (s2i - s3i);
    const double b1r = s2r + s3r, b1i = s2i + s3i;
    const double b2r = s0r + s1r, b2i = s0i + s1i;
    const double b3r = s0r - s1r, b3i = s0i - s1i;
    const double b4r = d2r - d3r, b4i = d2i - d3i;
    const double b5r = d0r + m2r, b5i = d0i + m2i;
    const double b6r = d0r - m2r, b6i = d0i - m2i;
    const double b7r = d1r + m1r, b7i = d1i + m1i;
    const double b8r = d1r - m1r, b8i = d1i - m1i;
    yr[0] = b2r + b1r; yr[1] = b5r + b7i; yr[2] = b3r + b4i; yr[3] = b6r - b8i;
    yr[4] = b2r - b1r; yr[5] = b6r + b8i; yr[6] = b3r - b4i; yr[7] = b5r - b7i;
    yi[0] = b2i + b1i; yi[1] = b5i - b7r; yi[2] = b3i - b4r; yi[3] = b6i + b8r;
    yi[4] = b2i - b1i; yi[5] = b6i - b8r; yi[6] = b3i + b4r; yi[7] = b5i + b7r;
}

// 16-point Winograd leaf of code 6 (serial_d_fft6.c:213-385)
__device__ void d_leaf16(int32_t sn, uint32_t N, const double *xr, const double *xi,
                         double *yr, double *yi, const double *c)
{
    const double w8 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], c4 = c[4];
    const uint32_t s = N / 16;
    double uR[8], vR[8], uI[8], vI[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t g0 = d_idx(sn, s, j, N), g1 = d_idx(sn, s, j + 8, N);
        uR[j] = xr[g0] + xr[g1]; vR[j] = xr[g0] - xr[g1];
        uI[j] = xi[g0] + xi[g1]; vI[j] = xi[g0] - xi[g1];
    }
    // the reference forms this one as -x4 + x12 (same value as x12 - x4)
    vI[4] = -xi[d_idx(sn, s, 4, N)] + xi[d_idx(sn, s, 12, N)];
    const double b1r = uR[0] - uR[4], b2r = uR[0] + uR[4];
    const double b3r = uR[2] - uR[6], b4r = uR[2] + uR[6];
    const double b5r = b2r + b4r, b6r = b2r - b4r;
    const double b7r = uR[1] - uR[5], b8r = uR[1] + uR[5];
    const double b9r = uR[3] - uR[7], b10r = uR[3] + uR[7];
    const double b11r = b8r + b10r, b12r = b8r - b10r;
    const double b13r = vR[1] + vR[7], b14r = vR[1] - vR[7];
    const double b15r = vR[3] + vR[5], b16r = vR[5] - vR[3];
    const double b1i = uI[0] - uI[4], b2i = uI[0] + uI[4];
    const double b3i = -uI[2] + uI[6], b4i = uI[2] + uI[6];
    const double b5i = b2i + b4i, b6i = b2i - b4i;
    const double b7i = uI[1] - uI[5], b8i = uI[1] + uI[5];
    const double b9i = uI[3] - uI[7], b10i = uI[3] + uI[7];
    const double b11i = b8i + b10i, b12i = -b8i + b10i;
    const double b13i = vI[1] + vI[7], b14i = vI[1] - vI[7];
    const double b15i = vI[3] + vI[5], b16i = vI[5] - vI[3];
    const double m1r = w8 * (vR[2] - vR[6]), m2r = w8 * (vR[2] + vR[6]);
    const double m3r = w8 * (b7r - b9r), m4r = w8 * (b7r + b9r);
    const double m5r = c1 * (b13r + b15r), m6r = -c2 * b13r, m7r = c4 * b15r;
    const double m8r = c3 * (b14r + b16r), m9r = c4 * b14r, m10r = -c2 * b16r;
    const double m1i = w8 * (vI[2] - vI[6]), m2i = -w8 * (vI[2] + vI[6]);
    const double m3i = w8 * (b7i - b9i), m4i = -w8 * (b7i + b9i);
    const double m5i = -c1 * (b13i + b15i), m6i = c2 * b13i, m7i = -c4 * b15i;
    const double m8i = c3 * (b14i + b16i), m9i = c4 * b14i, m10i = -c2 * b16i;
    const double c1r = b1r + m3r, c2r = b1r - m3r, c3r = b3r + m4r, c4r = m4r - b3r;
    const double c5r = vR[0] + m1r, c6r = vR[0] - m1r, c7r = m9r - m8r, c8r = m10r - m8r;
    const double c9r = c5r + c7r, c10r = c5r - c7r, c11r = c6r + c8r, c12r = c6r - c8r;
    const double c13r = vR[4] + m2r, c14r = vR[4] - m2r, c15r = m5r + m6r, c16r = m5r - m7r;
    const double c17r = c13r + c15r, c18r = c13r - c15r, c19r = c14r + c16r, c20r = c14r - c16r;
    const double c1i = b1i + m3i, c2i = b1i - m3i, c3i = b3i + m4i, c4i = m4i - b3i;
    const double c5i = vI[0] + m1i, c6i = vI[0] - m1i, c7i = m9i - m8i, c8i = m10i - m8i;
    const double c9i = c5i + c7i, c10i = c5i - c7i, c11i = c6i + c8i, c12i = c6i - c8i;
    const double c13i = vI[4] + m2i, c14i = vI[4] - m2i, c15i = m5i + m6i, c16i = m5i - m7i;
    const double c17i = c13i + c15i, c18i = c13i - c15i, c19i = c14i + c16i, c20i = c14i - c16i;
    yr[0] = b5r + b11r; yr[1] = c9r - c17i; yr[2] = c1r - c3i; yr[3] = c12r + c20i;
    yr[4] = b6r - b12i; yr[5] = c11r - c19i; yr[6] = c2r - c4i; yr[7] = c10r + c18i;
    yr[8] = b5r - b11r; yr[9] = c10r - c18i; yr[10] = c2r + c4i; yr[11] = c11r + c19i;
    yr[12] = b6r + b12i; yr[13] = c12r - c20i; yr[14] = c1r + c3i; yr[15] = c9r + c17i;
    yi[0] = b5i + b11i; yi[1] = c9i - c17r; yi[2] = c1i - c3r; yi[3] = c12i + c20r;
    yi[4] = b6i - b12r; yi[5] = c11i - c19r; yi[6] = c2i - c4r; yi[7] = c10i + c18r;
    yi[8] = b5i - b11i; yi[9] = c10i - c18r; yi[10] = c2i + c4r; yi[11] = c11i + c19r;
    yi[12] = b6i + b12r; yi[13] = c12i - c20r; yi[14] = c1i + c3r; yi[15] = c9i + c17r;
}

// one butterfly k of a size-n node whose outputs start at y (4m2a;
// serial_d_fft5.c:154-227 for code 5, serial_d_fft6.c:392-427 for code 6)
__device__ __forceinline__ void d_butterfly(int code, uint32_t N, uint32_t n, uint32_t k,
                                            const double *wrt, const double *wit,
                                            double *yr, double *yi)
{
    const uint32_t q = n >> 2;
    const uint32_t i0 = k, i1 = k + q, i2 = k + 2 * q, i3 = k + 3 * q;
    const double U0r = yr[i0], U1r = yr[i1], Zpr = yr[i2], Zmr = yr[i3];
    const double U0i = yi[i0], U1i = yi[i1], Zpi = yi[i2], Zmi = yi[i3];
    double m1r, m1i, m2r, m2i;
    if (code == 5 && k == 0) {
        m1r = Zpr; m1i = Zpi; m2r = Zmr; m2i = Zmi;
    } else {
        const uint32_t kn = k * (N / n);
        const double wr = wrt[kn], wi = wit[kn];
        if (code == 5 && k == (n >> 3)) {
            m1r = wr * (Zpr + Zpi); m1i = wr * (Zpi - Zpr);
            m2r = wr * (Zmr - Zmi); m2i = wr * (Zmi + Zmr);
        } else {
            m1r = wr * Zpr + wi * Zpi; m1i = wr * Zpi - wi * Zpr;
            m2r = wr * Zmr - wi * Zmi; m2i = wr * Zmi + wi * Zmr;
        }
    }
    const double a1r = m1r + m2r, a1i = m1i + m2i, a2r = m1r - m2r, a2i = m1i - m2i;
    yr[i0] = U0r + a1r; yr[i2] = U0r - a1r; yr[i1] = U1r + a2i; yr[i3] = U1r - a2i;
    yi[i0] = U0i + a1i; yi[i2] = U0i - a1i; yi[i1] = U1i - a2r; yi[i3] = U1i + a2r;
}

__device__ __forceinline__ void d_leaf(const DTables &t, uint32_t i, const double *xr,
                                       const double *xi, double *yr, double *yi)
{
    const uint32_t n = t.leaf_n[i], o = t.leaf_off[i];
    const int32_t sn = t.leaf_sn[i];
    if (n <= 4) d_leaf_small(n, sn, t.N, xr, xi, yr + o, yi + o);
    else if (n == 8) d_leaf8(sn, t.N, xr, xi, yr + o, yi + o, t.c[0]);
    else d_leaf16(sn, t.N, xr, xi, yr + o, yi + o, t.c);
}

// whole transform in one workgroup, working set in LDS
__global__ void __launch_bounds__(256)
d_fft_lds_kernel(DTables t, const double *in, double *out, uint64_t stride)
{
    extern __shared__ double sm[];
    const double *xr = in + stride * blockIdx.x, *xi = xr + t.N;
    double *yr = sm, *yi = sm + t.N;
    for (uint32_t i = threadIdx.x; i < t.nleaf; i += blockDim.x)
        d_leaf(t, i, xr, xi, yr, yi);
    __syncthreads();
    for (uint32_t l = t.lmin; l <= t.lgN; l++) {
        const uint32_t n = 1u << l, q = n >> 2;
        const uint32_t b0 = t.lvl_begin[l], cnt = t.lvl_begin[l + 1] - b0;
        for (uint32_t w = threadIdx.x; w < cnt * q; w += blockDim.x) {
            const uint32_t node = w / q, k = w - node * q, o = t.lvl_off[b0 + node];
            d_butterfly(t.code, t.N, n, k, t.wr, t.wi, yr + o, yi + o);
        }
        __syncthreads();
    }
    double *dst = out + stride * blockIdx.x;
    for (uint32_t i = threadIdx.x; i < 2 * t.N; i += blockDim.x) dst[i] = sm[i];
}

__global__ void __launch_bounds__(256)
d_leaf_kernel(DTables t, const double *in, double *out, uint64_t stride)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.nleaf) return;
    const double *xr = in + stride * blockIdx.y;
    double *yr = out + stride * blockIdx.y;
    d_leaf(t, i, xr, xr + t.N, yr, yr + t.N);
}

__global__ void __launch_bounds__(256)
d_level_kernel(DTables t, uint32_t l, double *out, uint64_t stride)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << l, q = n >> 2;
    const uint32_t b0 = t.lvl_begin[l], cnt = t.lvl_begin[l + 1] - b0;
    if (w >= cnt * q) return;
    const uint32_t node = w / q, k = w - node * q, o = t.lvl_off[b0 + node];
    double *yr = out + stride * blockIdx.y;
    d_butterfly(t.code, t.N, n, k, t.wr, t.wi, yr + o, yr + t.N + o);
}

hipError_t d_fft_launch(const DTables &t, const double *in, double *out, uint64_t stride,
                        uint32_t batch, hipStream_t st)
{
    if (t.N <= kDLdsMaxN) {
        size_t sh = sizeof(double) * 2 * t.N;
        hipLaunchKernelGGL(d_fft_lds_kernel, dim3(batch), dim3(256), sh, st, t, in, out, stride);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(d_leaf_kernel, dim3((t.nleaf + 255) / 256, batch), dim3(256), 0, st,
                       t, in, out, stride);
    for (uint32_t l = t.lmin; l <= t.lgN; l++) {
        uint32_t work = (t.N >> 2);   // upper bound of cnt * n/4 per level
        hipLaunchKernelGGL(d_level_kernel, dim3((work + 255) / 256, batch), dim3(256), 0, st,
                           t, l, out, stride);
    }
    return hipGetLastError();
}

}  // namespace mpfft
