// CPU emulation of the HIP subset that beat-box_amd's kernels use, so that the
// device code (beat-box_amd/csrc/bb_step.hip, unchanged) can be compiled by
// g++ with AddressSanitizer and its indexing checked on the host.  Each
// workgroup of a launch runs on its own OS thread (so a team of workgroups
// really runs concurrently and its global barrier works; EMU_MAX_BLOCKS=K
// runs at most K workgroups at a time, as a GPU that cannot hold a whole
// team resident would); each work-item is a
// fiber of that thread; __syncthreads, wave operations (shuffles, ballots)
// and s_sleep switch fibers.  Wave operations are rendezvous of the lanes of
// a width-aligned segment.  Atomics are sequentially consistent, so the
// emulation checks indexing and team logic, not the GPU's memory model or its
// compiler.  Device memory is host memory.
#pragma once
#include <stdint.h>
#include <string.h>
#include <stddef.h>
#include <math.h>
#include <chrono>
#include <functional>
#include <utility>
#include <tuple>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(x) alignas(x)
#define HIP_DYNAMIC_SHARED(type, var) type* var = (type*)emu::dyn_lds();

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};

#define EMU_VEC4(T, N)                                                  \
  struct N##4 { T x, y, z, w; };                                        \
  struct N##3 { T x, y, z; };                                           \
  struct N##2 { T x, y; };                                              \
  static inline N##4 make_##N##4(T a, T b, T c, T d) { return N##4{a, b, c, d}; } \
  static inline N##3 make_##N##3(T a, T b, T c) { return N##3{a, b, c}; } \
  static inline N##2 make_##N##2(T a, T b) { return N##2{a, b}; }
EMU_VEC4(float, float)
EMU_VEC4(int, int)
EMU_VEC4(unsigned int, uint)
EMU_VEC4(double, double)

namespace emu {
struct Fiber;
Fiber* cur();
const dim3& tid_of(Fiber*);
extern thread_local dim3 t_block, t_grid, t_bdim;
void barrier();
void yield();
[[noreturn]] void end_program();
uint64_t wave_op(int kind, uint64_t v, int arg, int width);
void* dyn_lds();
void launch(dim3 grid, dim3 block, size_t shmem, std::function<void()> fn);
}  // namespace emu

#define threadIdx (emu::tid_of(emu::cur()))
#define blockIdx (emu::t_block)
#define gridDim (emu::t_grid)
#define blockDim (emu::t_bdim)

static inline void __syncthreads() { emu::barrier(); }
// the device's constant-rate clock, here nanoseconds / 10 (100 MHz)
static inline unsigned long long wall_clock64() {
  return (unsigned long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch()).count() / 10ull;
}
static inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
static inline void __threadfence_block() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
#define __builtin_amdgcn_s_sleep(x) emu::yield()
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)
// s_endpgm: the work-item's fiber ends here
#define __builtin_amdgcn_endpgm() emu::end_program()

enum { EMU_SHFL = 0, EMU_SHFL_UP = 1, EMU_SHFL_DOWN = 2, EMU_BALLOT = 3 };
template <class T>
static inline uint64_t emu_bits(T v) {
  uint64_t b = 0;
  memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
static inline T emu_from(uint64_t b) {
  T v;
  memcpy(&v, &b, sizeof(T));
  return v;
}
template <class T>
static inline T __shfl(T v, int src, int width = 64) {
  return emu_from<T>(emu::wave_op(EMU_SHFL, emu_bits(v), src, width));
}
template <class T>
static inline T __shfl_up(T v, unsigned d, int width = 64) {
  return emu_from<T>(emu::wave_op(EMU_SHFL_UP, emu_bits(v), (int)d, width));
}
template <class T>
static inline T __shfl_down(T v, unsigned d, int width = 64) {
  return emu_from<T>(emu::wave_op(EMU_SHFL_DOWN, emu_bits(v), (int)d, width));
}
static inline uint64_t __ballot(int p) { return emu::wave_op(EMU_BALLOT, p ? 1 : 0, 0, 64); }

static inline int __popc(uint32_t x) { return __builtin_popcount(x); }
static inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
static inline int __ffs(uint32_t x) { return __builtin_ffs((int)x); }
static inline int __ffs(int x) { return __builtin_ffs(x); }
static inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
static inline int __ffsll(long long x) { return __builtin_ffsll(x); }
static inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((uint32_t)x); }
static inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((uint64_t)x); }
static inline int __float_as_int(float f) { return emu_from<int>(emu_bits(f)); }
static inline uint32_t __float_as_uint(float f) { return emu_from<uint32_t>(emu_bits(f)); }
static inline float __int_as_float(int i) { return emu_from<float>(emu_bits(i)); }
static inline float __uint_as_float(uint32_t i) { return emu_from<float>(emu_bits(i)); }

template <class T>
static inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
static inline float atomicAdd(float* p, float v) {
  float o = *p, n;
  do { n = o + v; } while (!__atomic_compare_exchange(p, &o, &n, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
  return o;
}
template <class T>
static inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T>
static inline T atomicAnd(T* p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST); }
template <class T>
static inline T atomicExch(T* p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T>
static inline T atomicMax(T* p, T v) {
  T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (o < v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return o;
}
template <class T>
static inline T atomicMin(T* p, T v) {
  T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (o > v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return o;
}
template <class T>
static inline T atomicCAS(T* p, T cmp, T v) {
  __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return cmp;
}
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0
#define __hip_atomic_load(p, o, s) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, o, s) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_add(p, v, o, s) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)

// ------------------------------------------------------------- host API
typedef enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 } hipError_t;
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph* hipGraphExec_t;
typedef struct emu_node* hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1, hipDeviceAttributeWallClockRate = 2 };
#define hipStreamNonBlocking 1u
#define hipStreamDefault 0u
#define hipHostMallocDefault 0u
#define hipEventDefault 0u
#define hipEventDisableTiming 2u

hipError_t hipMalloc(void** p, size_t n);
template <class T>
static inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags = 0);
template <class T>
static inline hipError_t hipHostMalloc(T** p, size_t n, unsigned flags = 0) { return hipHostMalloc((void**)p, n, flags); }
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemset(void* d, int v, size_t n);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* n, char* log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned f);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int dev);
template <class K>
static inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) {
  *n = 1;
  return hipSuccess;
}

namespace emu {
// Records fn into the stream's capture, or runs it now.
void submit(hipStream_t st, std::function<void()> fn);
}
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...)                                       \
  emu::submit((ST), [=]() { emu::launch(dim3(G), dim3(B), (SH), [=]() { K(__VA_ARGS__); }); })
