// Host-only stand-in for the subset of the HIP runtime API that
// one-synchronization-bcgs_amd/csrc/orth_host.cpp calls, so that the host
// schedules (including the batch scheduler and its graph capture) can be
// compiled with g++ under AddressSanitizer / UndefinedBehaviorSanitizer and
// exercised on a machine without a GPU.  "Device" memory is host memory,
// copies are memcpy, kernels (hip_stub.cpp) do nothing, and streams, events
// and graphs are bookkeeping objects whose misuse (capture on a capturing
// stream, ending a capture with unjoined forks, double destroys) is reported.
#pragma once

#include <stddef.h>
#include <stdint.h>

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorStreamCaptureUnjoined = 904,
    hipErrorStreamCaptureInvalidated = 901,
    hipErrorIllegalState = 401,
} hipError_t;

typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
typedef struct ihipGraph *hipGraph_t;
typedef struct hipGraphExec *hipGraphExec_t;
typedef struct hipGraphNode *hipGraphNode_t;

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
} hipMemcpyKind;

typedef enum hipMemoryType {
    hipMemoryTypeUnregistered = 0,
    hipMemoryTypeHost = 1,
    hipMemoryTypeDevice = 2,
    hipMemoryTypeManaged = 3
} hipMemoryType;

typedef struct hipPointerAttribute_t {
    hipMemoryType type;
    int device;
    void *devicePointer;
    void *hostPointer;
} hipPointerAttribute_t;

typedef enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 } hipDeviceAttribute_t;

typedef struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
} hipDeviceProp_t;

typedef enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2
} hipStreamCaptureMode;

typedef enum hipStreamCaptureStatus {
    hipStreamCaptureStatusNone = 0,
    hipStreamCaptureStatusActive = 1,
    hipStreamCaptureStatusInvalidated = 2
} hipStreamCaptureStatus;

#define hipStreamNonBlocking 0x01
#define hipEventDisableTiming 0x02
#define hipHostMallocDefault 0x0

hipError_t hipMalloc(void **p, size_t n);
template <class T>
static inline hipError_t hipMalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, n);
}
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
template <class T>
static inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags)
{
    return hipHostMalloc((void **)p, n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy2DAsync(void *d, size_t dpitch, const void *s, size_t spitch, size_t width,
                            size_t height, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *err, char *log,
                               size_t logn);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGetLastError(void);
hipError_t hipGetDevice(int *d);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int dev);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int dev);
hipError_t hipDeviceSynchronize(void);
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *p);
const char *hipGetErrorString(hipError_t e);
