// Host-only stand-in for the part of the HIP runtime API the host library
// uses (declarations only; tests/host_stub/stub.cpp implements them), so the
// host logic (plans, batched form, tables) runs under AddressSanitizer on a
// CPU (tests/test_host_logic.py).
#pragma once
#include <cstddef>
#include <cstdint>
#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __align__(n) alignas(n)
// a symbol's "device address" on the stand-in device is its host address
#define HIP_SYMBOL(x) (&(x))
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotReady = 600, hipErrorUnknown = 999 };
struct ihipStream_t; typedef ihipStream_t *hipStream_t;
struct ihipEvent_t; typedef ihipEvent_t *hipEvent_t;
struct ihipGraph; typedef ihipGraph *hipGraph_t;
struct hipGraphExec; typedef hipGraphExec *hipGraphExec_t;
typedef void *hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
#define hipEventDisableTiming 2
#define hipHostMallocDefault 0
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
hipError_t hipMalloc(void **p, size_t b);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t b, unsigned f);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t b, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t b, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipDeviceSynchronize();
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int d);
hipError_t hipGetLastError();
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *n, char *log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
