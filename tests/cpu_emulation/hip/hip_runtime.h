// Host stand-in for <hip/hip_runtime.h> used only to run the reactor's device code
// serially on the CPU (one lane at a time) in tests; built with clang++ (the device
// code uses clang vector extensions).  Cross-lane operations assume a
// single active lane, so only the one-thread-per-cell schedule is meaningful here.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(x)

using std::isfinite;

typedef int hipError_t;
typedef void* hipStream_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotSupported = 801 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
typedef void* hipEvent_t;
inline hipError_t hipEventCreate(hipEvent_t*) { return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }

// graph capture is not emulated (the serial runs issue trips from the host)
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
// device queries: an emulated device of 4 compute units, one resident wave each
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount };
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 4; return hipSuccess; }
template <class F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F, int, size_t) { *n = 1; return hipSuccess; }

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct Idx3 { unsigned x = 0, y = 0, z = 0; };
extern thread_local Idx3 threadIdx, blockIdx, blockDim, gridDim;

template <class T> inline T max(T a, T b) { return a > b ? a : b; }
template <class T> inline T min(T a, T b) { return a < b ? a : b; }
inline int __lane_id() { return 0; }   // every emulated lane reduces on its own
template <class T> inline T __shfl_down(T v, int, int = 64) { return T(0); }
template <class T> inline T __shfl(T v, int, int = 64) { return v; }
template <class T> inline T __shfl_xor(T v, int, int = 64) { return v; }
inline void __syncthreads() {}
// matrix-core products need all 64 lanes at once: kernels that use them (the learned
// predictor) are compiled here but never launched
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, x, y, z) (c)
// one lane at a time cannot run a matrix-core product (mi.lu_form = mfma is not launched here)
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) (c)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) (c)
#define __shared__ static
inline unsigned long long __ballot(int p) { return p ? (1ull << __lane_id()) : 0ull; }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline long long __double_as_longlong(double x) { long long r; __builtin_memcpy(&r, &x, 8); return r; }
inline double __longlong_as_double(long long x) { double r; __builtin_memcpy(&r, &x, 8); return r; }
#define __builtin_amdgcn_readlane(v, l) (v)
// one lane at a time: a DPP move reads this lane's own value
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) (src)
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned int __float_as_uint(float x) {
  unsigned int u;
  __builtin_memcpy(&u, &x, 4);
  return u;
}
template <class T> inline T atomicAdd(T* p, T v) { T o = *p; *p += v; return o; }
template <class T> inline T atomicMax(T* p, T v) { T o = *p; if (v > o) *p = v; return o; }

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)              \
  do {                                                                          \
    const dim3 g_ = (grid), b_ = (block);                                       \
    blockDim.x = b_.x;                                                          \
    gridDim.x = g_.x;                                                           \
    for (unsigned bx_ = 0; bx_ < g_.x; ++bx_)                                   \
      for (unsigned tx_ = 0; tx_ < b_.x; ++tx_) {                               \
        blockIdx.x = bx_;                                                       \
        threadIdx.x = tx_;                                                      \
        kernel(__VA_ARGS__);                                                    \
      }                                                                         \
  } while (0)
