// CPU emulation of the HIP subset used by rpd-statmc-denoiser_amd/csrc, for
// tests only: the engine's own kernels compiled with g++ and run on the host,
// so that every global and dynamic-LDS access is checked by AddressSanitizer
// and static arrays by UBSan bounds checks.
//
// Execution model: a launch runs its blocks one after another on the calling
// thread; the threads of a block are fibers (ucontext). __syncthreads() and
// the wave collectives (__ballot, __shfl, __shfl_xor over 64-lane waves) are
// cooperative barriers; a barrier that can never complete (divergent
// collectives, a hang) aborts with a report. Device memory is host memory;
// streams are synchronous; stream capture records the launches and copies
// into a graph that hipGraphLaunch replays.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <tuple>
#include <type_traits>
#include <utility>

// ---- qualifiers -----------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline)) inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __constant__
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) alignas(n)

// ---- vector types -----------------------------------------------------------------
struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3 { unsigned x, y, z; };
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
struct uint2 { unsigned x, y; };
struct uint4 { unsigned x, y, z, w; };
struct double2 { double x, y; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int4 make_int4(int a, int b, int c, int d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return {a, b, c, d}; }
inline double2 make_double2(double a, double b) { return {a, b}; }

// ---- runtime types ------------------------------------------------------------------
enum hipError_t {
    hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorInvalidConfiguration = 9,
    hipErrorStreamCaptureUnsupported = 900, hipErrorStreamCaptureImplicit = 906, hipErrorUnknown = 999
};
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
#define hipStreamNonBlocking 1u
#define hipStreamDefault 0u
struct ihipStream_t;
struct ihipEvent_t;
struct ihipGraph;
struct hipGraphExec;
typedef ihipStream_t* hipStream_t;
typedef ihipEvent_t* hipEvent_t;
typedef ihipGraph* hipGraph_t;
typedef hipGraphExec* hipGraphExec_t;
typedef void* hipGraphNode_t;
struct hipDeviceProp_t {
    char name[256];
    size_t totalGlobalMem;
    int multiProcessorCount;
    int warpSize;
    int maxThreadsPerBlock;
    size_t sharedMemPerBlock;
};

// ---- runtime API ----------------------------------------------------------------------
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int* d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int d);
hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b);
hipError_t hipMalloc(void** p, size_t bytes);
template <class T> inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc(reinterpret_cast<void**>(p), bytes); }
hipError_t hipFree(void* p);
// a file-scope device array is ordinary host memory here: its address is itself
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) { *p = const_cast<void*>(symbol); return hipSuccess; }
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemset(void* dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
#define hipEventDisableTiming 2u
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* err_node, char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void* f, int block, size_t lds);
template <class F> inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F f, int block, size_t lds) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, reinterpret_cast<const void*>(f), block, lds);
}

namespace rpd_emu {

// ---- execution of one launch ----------------------------------------------------------
// Runs `body` as every thread of every block of the grid (or records it when
// the stream is being captured). Validates the launch shape like the device.
hipError_t launch(dim3 grid, dim3 block, size_t dyn_lds, hipStream_t s, std::function<void()> body);

// the calling fiber's coordinates and its block's dynamic LDS
uint3 thread_idx();
uint3 block_idx();
dim3 block_dim();
dim3 grid_dim();
void* dynamic_lds();

// cooperative barriers
void block_barrier();
__attribute__((noinline)) uint64_t wave_exchange_u64(uint64_t v, int mode, int arg);  // mode 0: ballot of v != 0, 1: shfl from lane arg, 2: shfl_xor arg

}  // namespace rpd_emu

#define threadIdx (::rpd_emu::thread_idx())
#define blockIdx (::rpd_emu::block_idx())
#define blockDim (::rpd_emu::block_dim())
#define gridDim (::rpd_emu::grid_dim())

inline void __syncthreads() { ::rpd_emu::block_barrier(); }
inline unsigned long long __ballot(int pred) { return ::rpd_emu::wave_exchange_u64(pred != 0, 0, 0); }
template <class T> inline T __shfl(T v, int src, int width = 64) {
    static_assert(sizeof(T) <= 8, "shfl of at most 64 bits");
    (void)width;
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    bits = ::rpd_emu::wave_exchange_u64(bits, 1, src);
    T r;
    std::memcpy(&r, &bits, sizeof(T));
    return r;
}
template <class T> inline T __shfl_xor(T v, int mask, int width = 64) {
    (void)width;
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    bits = ::rpd_emu::wave_exchange_u64(bits, 2, mask);
    T r;
    std::memcpy(&r, &bits, sizeof(T));
    return r;
}
// v of lane (lane - delta); lanes below delta keep their own v (HIP's semantics)
template <class T> inline T __shfl_up(T v, unsigned delta, int width = 64) {
    (void)width;
    const int lane = int(threadIdx.x) & 63;
    return __shfl(v, lane >= int(delta) ? lane - int(delta) : lane);
}
// v of lane `lane` (a wave-uniform index, as the device instruction requires)
inline int __builtin_amdgcn_readlane(int v, int lane) { return __shfl(v, lane); }
// issue priority of the wave: a scheduling hint with no effect on results
inline void __builtin_amdgcn_s_setprio(short) {}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __clz(unsigned v) { return v ? __builtin_clz(v) : 32; }
inline int __clz(int v) { return __clz(unsigned(v)); }
// fibers of one process share memory; atomics are sequentially consistent
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __ffsll(unsigned long long v) { return __builtin_ffsll(static_cast<long long>(v)); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline float __uint_as_float(unsigned v) { float f; std::memcpy(&f, &v, 4); return f; }
inline unsigned __float_as_uint(float f) { unsigned v; std::memcpy(&v, &f, 4); return v; }
inline int __float_as_int(float f) { int v; std::memcpy(&v, &f, 4); return v; }
inline float __int_as_float(int v) { float f; std::memcpy(&f, &v, 4); return f; }
inline long long __double_as_longlong(double d) { long long v; std::memcpy(&v, &d, 8); return v; }
inline double __longlong_as_double(long long v) { double d; std::memcpy(&d, &v, 8); return d; }

template <class T, class U> inline T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, T(v), __ATOMIC_SEQ_CST); }
template <class F> inline F atomic_add_fp(F* p, F v) {
    F old;
    __atomic_load(p, &old, __ATOMIC_SEQ_CST);
    F want = old + v;
    while (!__atomic_compare_exchange(p, &old, &want, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) want = old + v;
    return old;
}
inline float atomicAdd(float* p, float v) { return atomic_add_fp(p, v); }
inline double atomicAdd(double* p, double v) { return atomic_add_fp(p, v); }
template <class T, class U> inline T atomicMin(T* p, U v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (T(v) < old && !__atomic_compare_exchange_n(p, &old, T(v), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}
template <class T, class U> inline T atomicMax(T* p, U v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (T(v) > old && !__atomic_compare_exchange_n(p, &old, T(v), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}
template <class T, class U> inline T atomicExch(T* p, U v) { return __atomic_exchange_n(p, T(v), __ATOMIC_SEQ_CST); }

// device min/max overloads as HIP declares them (same-type pairs; mixed pairs
// promote like the device headers' integer overloads)
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline unsigned long min(unsigned long a, unsigned long b) { return a < b ? a : b; }
inline unsigned long max(unsigned long a, unsigned long b) { return a > b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
inline float min(float a, float b) { return fminf(a, b); }
inline float max(float a, float b) { return fmaxf(a, b); }
inline double min(double a, double b) { return fmin(a, b); }
inline double max(double a, double b) { return fmax(a, b); }
inline long min(int a, long b) { return min(long(a), b); }
inline long min(long a, int b) { return min(a, long(b)); }
inline long max(int a, long b) { return max(long(a), b); }
inline long max(long a, int b) { return max(a, long(b)); }
// device math without a host libm counterpart: cos(pi x), sin(pi x)
inline float cospif(float x) { return float(std::cos(3.14159265358979323846 * double(x))); }
inline float sinpif(float x) { return float(std::sin(3.14159265358979323846 * double(x))); }
inline double cospi(double x) { return std::cos(3.14159265358979323846 * x); }
inline double sinpi(double x) { return std::sin(3.14159265358979323846 * x); }
inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }
using std::isfinite;
using std::isnan;
using std::isinf;

// kernel launch: converts the arguments to the kernel's parameter types once,
// then runs (or records) the kernel as every thread of the grid
template <class... P, class... A>
inline hipError_t hipLaunchKernelGGL(void (*k)(P...), dim3 grid, dim3 block, size_t dyn_lds, hipStream_t s, A&&... a) {
    auto args = std::make_shared<std::tuple<std::decay_t<P>...>>(std::forward<A>(a)...);
    return ::rpd_emu::launch(grid, block, dyn_lds, s, [k, args]() { std::apply(k, *args); });
}
