// VALU issue-rate probe for the roofline of the limb-arithmetic kernels:
// lane-operations per second of the multiply forms the kernels can use
// (v_mad_u64_u32 with and without a carry chain, v_mul_lo_u32, plain 32-bit
// integer ops, v_fma_f64), each as 8 independent chains per lane over a grid
// that fills every CU.  Built by mpfft_amd/build.py as mpfft_amd/valu_probe.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#define ITERS 2048
#define CH 8
__global__ void k_mad64(uint64_t *out, uint32_t m) {
  uint64_t acc[CH];
  for (int k = 0; k < CH; k++) acc[k] = threadIdx.x * 7 + k + 1;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int k = 0; k < CH; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * m + acc[k];
  }
  uint64_t s = 0; for (int k = 0; k < CH; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
__global__ void k_madc(uint64_t *out, uint32_t m) {
  uint64_t acc[CH]; uint32_t c2[CH];
  for (int k = 0; k < CH; k++) { acc[k] = threadIdx.x * 7 + k + 1; c2[k] = 0; }
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int k = 0; k < CH; k++) {
      uint64_t cc;
      uint32_t a = (uint32_t)acc[k] ^ it;
      asm volatile("v_mad_u64_u32 %0, %2, %3, %4, %0\n\tv_addc_co_u32_e64 %1, %2, 0, %1, %2"
                   : "+v"(acc[k]), "+v"(c2[k]), "=&s"(cc) : "v"(a), "v"(m));
    }
  }
  uint64_t s = 0; for (int k = 0; k < CH; k++) s ^= acc[k] + c2[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
__global__ void k_mullo(uint64_t *out, uint32_t m) {
  uint32_t acc[CH];
  for (int k = 0; k < CH; k++) acc[k] = threadIdx.x * 7 + k + 1;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int k = 0; k < CH; k++) acc[k] = acc[k] * m + 1;
  }
  uint64_t s = 0; for (int k = 0; k < CH; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
__global__ void k_add(uint64_t *out, uint32_t m) {
  uint32_t acc[CH];
  for (int k = 0; k < CH; k++) acc[k] = threadIdx.x * 7 + k + 1;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int k = 0; k < CH; k++) acc[k] = (acc[k] ^ m) + k;
  }
  uint64_t s = 0; for (int k = 0; k < CH; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
__global__ void k_fma64(uint64_t *out, uint32_t m) {
  double acc[CH]; double f = 1.0 + m * 1e-9;
  for (int k = 0; k < CH; k++) acc[k] = threadIdx.x * 7 + k + 1;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int k = 0; k < CH; k++) acc[k] = __builtin_fma(acc[k], f, 0.5);
  }
  double s = 0; for (int k = 0; k < CH; k++) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint64_t)s;
}
typedef void (*K)(uint64_t*, uint32_t);
int main() {
  const int blocks = 256 * 8, threads = 256;
  uint64_t *out; hipMalloc(&out, sizeof(uint64_t) * blocks * threads);
  struct { const char *name; K k; int insts_per_op; } ks[] = {
    {"v_mad_u64_u32 (C)", k_mad64, 1}, {"mad_u64_u32+addc (asm)", k_madc, 2},
    {"v_mul_lo_u32+add", k_mullo, 1}, {"v_xor+v_add (2 ops)", k_add, 2}, {"v_fma_f64", k_fma64, 1}};
  hipEvent_t a, b; hipEventCreate(&a); hipEventCreate(&b);
  for (auto &e : ks) {
    hipLaunchKernelGGL(e.k, dim3(blocks), dim3(threads), 0, 0, out, 12345u);
    hipDeviceSynchronize();
    hipEventRecord(a);
    for (int r = 0; r < 5; r++) hipLaunchKernelGGL(e.k, dim3(blocks), dim3(threads), 0, 0, out, 12345u + r);
    hipEventRecord(b); hipEventSynchronize(b);
    float ms; hipEventElapsedTime(&ms, a, b);
    double ops = 5.0 * blocks * threads * (double)ITERS * CH;
    printf("%-26s %8.3f Tops/s (lane-ops of the named form)\n", e.name, ops / (ms * 1e-3) / 1e12);
  }
  hipError_t err = hipGetLastError(); printf("status %s\n", hipGetErrorString(err));
  hipFree(out);
  return err == hipSuccess ? 0 : 1;
}
