// CPU emulation of the HIP subset used by prtcl_amd/csrc: kernels run on the
// host, one fiber per thread, with wave (64-lane) shuffles/ballots and block
// barriers resolved by a scheduler. Built with AddressSanitizer it turns
// out-of-bounds device accesses into host reports. Debug tool only.
#pragma once
#include <math.h>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>
#include <memory>
#include <new>
#include <limits>
#include <array>
#include <map>
#include <climits>
#include <cfloat>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define __restrict__ __restrict
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3_ { unsigned x, y, z; };

struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct alignas(16) float4 { float x, y, z, w; };
struct double2 { double x, y; };
struct alignas(32) double4 { double x, y, z, w; };
struct alignas(8) uint2 { unsigned x, y; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct alignas(16) int4 { int x, y, z, w; };
struct alignas(8) int2 { int x, y; };
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }
inline uint2 make_uint2(unsigned a, unsigned b) { return uint2{a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
inline double4 make_double4(double a, double b, double c, double d) { return double4{a, b, c, d}; }

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorInvalidConfiguration = 9,
  hipErrorNotReady = 600,
  hipErrorLaunchTimeOut = 702,
  hipErrorUnknown = 999,
} hipError_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u
#define hipHostMallocMapped 2u
#define hipHostMallocCoherent 0x40000000u
#ifndef __HIP_MEMORY_SCOPE_SYSTEM
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#endif

namespace emu {
struct Stream;
struct Event;
struct Graph;
}  // namespace emu
typedef emu::Stream *hipStream_t;
typedef emu::Event *hipEvent_t;
typedef emu::Graph *hipGraph_t;
typedef emu::Graph *hipGraphExec_t;
typedef void *hipGraphNode_t;

namespace emu {
struct ThreadState {
  uint3_ idx;
};
extern ThreadState *g_cur;
extern uint3_ g_bidx;
extern dim3 g_bdim, g_gdim;
enum WaveOp { kOpShfl = 1, kOpShflUp, kOpShflDown, kOpShflXor, kOpBallot };
uint64_t wave_op(int op, uint64_t bits, int arg, int width, int line, const char *file);
void barrier(int line, const char *file);
unsigned lane_id();
void submit(hipStream_t s, const char *name, dim3 g, dim3 b, std::function<void()> fn);
// reports 8-byte words of a kernel argument that look like host addresses
// outside every device, pinned or symbol allocation
void check_arg(const char *kernel, int index, const void *p, size_t n);

template <class T>
inline uint64_t to_bits(T v) {
  static_assert(sizeof(T) <= 8, "shuffle of > 8 bytes");
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T from_bits(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}
template <class... P, class... A>
void launch(const char *name, void (*k)(P...), dim3 g, dim3 b, size_t, hipStream_t s, A &&...a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  std::tuple<std::decay_t<P>...> args(static_cast<std::decay_t<P>>(std::forward<A>(a))...);
  std::apply([name](const auto &...x) {
    int i = 0;
    (check_arg(name, i++, &x, sizeof(x)), ...);
  }, args);
  submit(s, name, g, b, [k, args]() { std::apply(k, args); });
}
}  // namespace emu

#define threadIdx (::emu::g_cur->idx)
#define blockIdx (::emu::g_bidx)
#define blockDim (::emu::g_bdim)
#define gridDim (::emu::g_gdim)
#define warpSize 64

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  ::emu::launch(#kernel, kernel, dim3(grid), dim3(block), static_cast<size_t>(shmem), stream, ##__VA_ARGS__)

#define __syncthreads() ::emu::barrier(__LINE__, __FILE__)
#define __lane_id() ::emu::lane_id()
#define __shfl(v, s, ...) ::emu::shfl_(v, s, __LINE__, __FILE__, ##__VA_ARGS__)
#define __shfl_up(v, d, ...) ::emu::shfl_up_(v, d, __LINE__, __FILE__, ##__VA_ARGS__)
#define __shfl_down(v, d, ...) ::emu::shfl_down_(v, d, __LINE__, __FILE__, ##__VA_ARGS__)
#define __shfl_xor(v, m, ...) ::emu::shfl_xor_(v, m, __LINE__, __FILE__, ##__VA_ARGS__)
#define __ballot(p) ::emu::wave_op(::emu::kOpBallot, (p) ? 1u : 0u, 0, 64, __LINE__, __FILE__)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
// a wavefront barrier: every live lane of the wave reaches it before any goes
// on (a ballot rendezvous); fences order nothing on one host thread
#define __builtin_amdgcn_wave_barrier() \
  ((void)::emu::wave_op(::emu::kOpBallot, 0u, 0, 64, __LINE__, __FILE__))
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_rsqf(x) (1.0f / std::sqrt(static_cast<float>(x)))
#define __builtin_amdgcn_rcpf(x) (1.0f / static_cast<float>(x))

namespace emu {
template <class T>
inline T shfl_(T v, int src, int line, const char *file, int width = 64) {
  return from_bits<T>(wave_op(kOpShfl, to_bits(v), src, width, line, file));
}
template <class T>
inline T shfl_up_(T v, unsigned d, int line, const char *file, int width = 64) {
  return from_bits<T>(wave_op(kOpShflUp, to_bits(v), static_cast<int>(d), width, line, file));
}
template <class T>
inline T shfl_down_(T v, unsigned d, int line, const char *file, int width = 64) {
  return from_bits<T>(wave_op(kOpShflDown, to_bits(v), static_cast<int>(d), width, line, file));
}
template <class T>
inline T shfl_xor_(T v, int m, int line, const char *file, int width = 64) {
  return from_bits<T>(wave_op(kOpShflXor, to_bits(v), m, width, line, file));
}
}  // namespace emu

inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll(static_cast<long long>(x)); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned long long __umul64hi(unsigned long long a, unsigned long long b) { return static_cast<unsigned long long>((static_cast<unsigned __int128>(a) * b) >> 64); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll(static_cast<unsigned long long>(x)); }
inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(unsigned u) { float f; std::memcpy(&f, &u, 4); return f; }
inline long long __double_as_longlong(double d) { long long u; std::memcpy(&u, &d, 8); return u; }
inline double __longlong_as_double(long long u) { double d; std::memcpy(&d, &u, 8); return d; }
inline float rsqrt(float x) { return 1.0f / std::sqrt(x); }
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
using std::min;
using std::max;
inline unsigned min(unsigned a, int b) { return a < static_cast<unsigned>(b) ? a : static_cast<unsigned>(b); }
inline long min(long a, int b) { return a < b ? a : b; }

template <class T, class U>
inline T atomicAdd(T *p, U v) { T o = *p; *p = static_cast<T>(o + v); return o; }
template <class T, class U>
inline T atomicExch(T *p, U v) { T o = *p; *p = static_cast<T>(v); return o; }
template <class T, class U>
inline T atomicMax(T *p, U v) { T o = *p; if (static_cast<T>(v) > o) *p = static_cast<T>(v); return o; }
// the GPU threads of the emulation share one address space and run on one
// host thread at a time: a device-scope fence has nothing to order
inline void __threadfence() {}
template <class T, class U>
inline T atomicMin(T *p, U v) { T o = *p; if (static_cast<T>(v) < o) *p = static_cast<T>(v); return o; }

// runtime API
hipError_t hipMalloc(void **p, size_t n);
template <class T>
inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc(reinterpret_cast<void **>(p), n); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
template <class T>
inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags = 0) {
  return hipHostMalloc(reinterpret_cast<void **>(p), n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipHostGetDevicePointer(void **d, void *h, unsigned flags);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamQuery(hipStream_t s);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipDeviceSynchronize();
hipError_t hipGetSymbolAddress(void **p, const void *sym);
