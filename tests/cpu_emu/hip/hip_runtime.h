// CPU stand-in for the parts of the HIP runtime and device intrinsics that the
// library uses: kernels run on fibres (one per GPU thread), device memory sits
// against guard pages so that any overrun faults at the access.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3e { unsigned x, y, z; };

struct double4 { double x, y, z, w; };
struct double2 { double x, y; };
struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct int4 { int x, y, z, w; };
struct int2 { int x, y; };
inline int2 make_int2(int a, int b) { return int2{a, b}; }
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
inline double4 make_double4(double a, double b, double c, double d) { return double4{a, b, c, d}; }

namespace emu {
struct Fiber;
extern Fiber* g_cur;
const uint3e& tidx();
const uint3e& bidx();
const uint3e& bdim();
const uint3e& gdim();
void syncthreads();
long long wave_op(int kind, long long v, int arg, int width);  // 0 shfl, 1 xor, 2 up, 3 ballot, 4 readlane
void* dyn_smem();
void begin_kernel(const char* name, dim3 g, dim3 b, size_t smem);
void run_block(void (*entry)(void*), void* arg, unsigned bx, unsigned by, unsigned bz);
void end_kernel();
template <class T> long long to_bits(T v) { long long r = 0; std::memcpy(&r, &v, sizeof(T)); return r; }
template <class T> T from_bits(long long b) { T r; std::memcpy(&r, &b, sizeof(T)); return r; }
}  // namespace emu

#define threadIdx (emu::tidx())
#define blockIdx (emu::bidx())
#define blockDim (emu::bdim())
#define gridDim (emu::gdim())

inline void __syncthreads() { emu::syncthreads(); }
inline void __threadfence() {}
template <class T> inline T __shfl(T v, int src, int width = 64) {
    return emu::from_bits<T>(emu::wave_op(0, emu::to_bits(v), src, width));
}
template <class T> inline T __shfl_xor(T v, int mask, int width = 64) {
    return emu::from_bits<T>(emu::wave_op(1, emu::to_bits(v), mask, width));
}
template <class T> inline T __shfl_up(T v, unsigned delta, int width = 64) {
    return emu::from_bits<T>(emu::wave_op(2, emu::to_bits(v), (int)delta, width));
}
inline unsigned long long __ballot(int pred) { return (unsigned long long)emu::wave_op(3, pred ? 1 : 0, 0, 64); }
inline int __builtin_amdgcn_readlane(int v, int lane) { return (int)emu::wave_op(4, v, lane, 64); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
template <class T> inline T atomicAdd(T* p, T v) { T o = *p; *p = o + v; return o; }
inline int atomicMax(int* p, int v) { int o = *p; if (v > o) *p = v; return o; }
inline int atomicExch(int* p, int v) { int o = *p; *p = v; return o; }
inline int atomicOr(int* p, int v) { int o = *p; *p = o | v; return o; }
template <class A, class B> inline auto min(A a, B b) -> typename std::common_type<A, B>::type { return a < b ? a : b; }
template <class A, class B> inline auto max(A a, B b) -> typename std::common_type<A, B>::type { return a > b ? a : b; }

// ------------------------------------------------------------------ runtime
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorUnknown = 999, hipErrorNotSupported = 801, hipErrorInvalidValue = 1, hipErrorNotReady = 600 };
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph_exec* hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock = 1 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 1 };

hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t s);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFreeAsync(void* p, hipStream_t s);
hipError_t hipFree(void* p);
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t s);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned f);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned f);
hipError_t hipHostFree(void* p);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetDevice(int* d);
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int d);
hipError_t hipFuncSetAttribute(const void* f, hipFuncAttribute a, int v);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);

namespace emu {
template <class... KA, class... A>
void launch(const char* name, void (*k)(KA...), dim3 g, dim3 b, size_t smem, A&&... a) {
    using Tup = std::tuple<typename std::decay<KA>::type...>;
    struct Pack { void (*k)(KA...); Tup args; };
    Pack pk{k, Tup(std::forward<A>(a)...)};
    auto entry = [](void* p) {
        Pack* q = static_cast<Pack*>(p);
        std::apply(q->k, q->args);
    };
    begin_kernel(name, g, b, smem);
    for (unsigned z = 0; z < g.z; ++z)
        for (unsigned y = 0; y < g.y; ++y)
            for (unsigned x = 0; x < g.x; ++x) run_block(entry, &pk, x, y, z);
    end_kernel();
}
}  // namespace emu
#define hipLaunchKernelGGL(F, G, B, S, ST, ...) emu::launch(#F, F, dim3(G), dim3(B), (size_t)(S), ##__VA_ARGS__)
