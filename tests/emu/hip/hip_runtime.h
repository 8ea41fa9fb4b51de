// Host emulation of the HIP execution model (test infrastructure only).
//
// The library's .hip sources are compiled for x86 with this header in place of ROCm's
// <hip/hip_runtime.h>, so that their kernels run on the CPU under the CPU test suite: every
// work-group is one host thread running its work-items as fibers (work-group barriers switch
// fibers), wavefront operations (shuffles, DPP moves, lane reads, ballots, MFMA) gather the 64
// lanes of the wavefront and compute the instruction's result from the documented lane layouts,
// and up to kResident work-groups of a launch run concurrently in ascending order, as the GPU
// dispatches them (persistent and cooperative grids stay co-resident).  Device memory is host
// memory; every launch and copy completes before the call returns (all streams are serialised).
// Nothing here is part of the product: the shipped library is built by hipcc for gfx950.
#pragma once

#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

// ---- qualifiers ---------------------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
// work-group shared memory: one copy per host thread, and every work-group runs on one thread
#define __shared__ static thread_local
// AMDGPU function attributes become a harmless host attribute
#define amdgpu_waves_per_eu(...) unused
#define amdgpu_flat_work_group_size(...) unused

// ---- vector types and launch geometry ----------------------------------------------------------
struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

// HIP's short vector types (the members the kernels use)
#define EMU_VEC2(T, N)                                                      \
    struct N {                                                               \
        T x, y;                                                              \
    };                                                                       \
    inline N make_##N(T x, T y) { return N{x, y}; }
#define EMU_VEC4(T, N)                                                      \
    struct N {                                                               \
        T x, y, z, w;                                                        \
    };                                                                       \
    inline N make_##N(T x, T y, T z, T w) { return N{x, y, z, w}; }
EMU_VEC2(int, int2)
EMU_VEC2(unsigned, uint2)
EMU_VEC2(float, float2)
EMU_VEC2(double, double2)
EMU_VEC2(long long, longlong2)
EMU_VEC4(int, int4)
EMU_VEC4(float, float4)
EMU_VEC4(double, double4)

// ---- runtime types -------------------------------------------------------------------------
typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorInvalidConfiguration = 9,
    hipErrorInvalidDevice = 101,
    hipErrorNoDevice = 100,
    hipErrorLaunchFailure = 719,
    hipErrorCooperativeLaunchTooLarge = 720,
    hipErrorInvalidDeviceFunction = 98,
} hipError_t;

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
} hipMemcpyKind;

struct emu_stream;
struct emu_event;
typedef emu_stream *hipStream_t;
typedef emu_event *hipEvent_t;

#define hipStreamDefault 0x0
#define hipStreamNonBlocking 0x1
#define hipEventDefault 0x0
#define hipEventBlockingSync 0x1
#define hipEventDisableTiming 0x2

struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    size_t totalGlobalMem;
    size_t sharedMemPerBlock;
    int regsPerBlock;
    int warpSize;
    int maxThreadsPerBlock;
    int maxThreadsDim[3];
    int maxGridSize[3];
    int clockRate;
    int multiProcessorCount;
    int cooperativeLaunch;
    int cooperativeMultiDeviceLaunch;
    int maxThreadsPerMultiProcessor;
    int l2CacheSize;
    int major, minor;
};

// ---- emulator interface ----------------------------------------------------------------------
namespace emu {

constexpr int kWave = 64;
constexpr int kResident = 32;        // work-groups resident at once (emulated CUs x occupancy)
constexpr int kCUs = 8;              // compute units reported to the library
constexpr int kBlocksPerCU = 2;      // occupancy reported for every kernel

const dim3 &thread_idx();
const dim3 &block_idx();
const dim3 &block_dim();
const dim3 &grid_dim();

void block_barrier();
int block_reduce(int pred, int op);   // 0 and, 1 or, 2 count (after a barrier)
void nap();                            // s_sleep: let the other work-items run
// Gathers `bytes` (<= 32) of every lane of the calling wavefront that executes the operation at
// `site`: *all points to 64 x 32 bytes, valid until the lane's next wavefront operation or
// barrier, and the result is the mask of the lanes that took part (the other lanes, inactive or exited,
// contribute zeros); own lane index in *lane.  A wavefront's operation proceeds once none of its
// lanes can run any further (the others wait at a barrier, at another operation, or exited), so
// the lanes of a divergent branch form the operation's execution mask as on the hardware.
uint64_t wave_gather(uint64_t site, const void *mine, size_t bytes, const void **all, int *lane);
inline uint64_t site_id(const char *file, int line) { return (uint64_t)(uintptr_t)file * 0x9E3779B97F4A7C15ull ^ (uint64_t)line; }

// one grid of blocks; body runs once per work-item
void run_grid(const char *what, dim3 grid, dim3 block, bool cooperative, const std::function<void()> &body);
void set_launch_error(hipError_t e);

// kernels launched by address (cooperative launches): an invoker per registered kernel
using Invoker = std::function<void(void **)>;
void register_kernel(const void *f, Invoker inv);
const Invoker *find_kernel(const void *f);

template <class... P, size_t... I>
inline void invoke_with(void (*k)(P...), void **a, std::index_sequence<I...>) {
    k(*static_cast<std::remove_reference_t<P> *>(a[I])...);
}
template <class... P>
inline void register_kernel(void (*k)(P...)) {
    register_kernel(reinterpret_cast<const void *>(k),
                    [k](void **a) { invoke_with(k, a, std::index_sequence_for<P...>()); });
}

template <class... P, class... A>
inline void launch(const char *what, void (*k)(P...), dim3 g, dim3 b, size_t, hipStream_t, A &&...a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<std::decay_t<P>...> args(static_cast<std::decay_t<P>>(a)...);
    run_grid(what, g, b, false, [&]() { std::apply(k, args); });
}

template <class T>
inline T bits_to(uint64_t v) {
    T t;
    std::memcpy(&t, &v, sizeof(T));
    return t;
}
template <class T>
inline uint64_t to_bits(T v) {
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}

// value of lane src of every lane's v (lanes that left the kernel read 0)
template <class T>
inline T lane_value(uint64_t site, T v, int (*src)(int lane, int width, int arg), int width, int arg) {
    static_assert(sizeof(T) <= 8, "lane value");
    const void *all;
    int lane = 0;
    wave_gather(site, &v, sizeof(T), &all, &lane);
    const int s = src(lane, width, arg) & (kWave - 1);
    T out;
    std::memcpy(&out, static_cast<const char *>(all) + 32 * s, sizeof(T));
    return out;
}

inline int src_xor(int lane, int width, int mask) {
    const int idx = lane ^ mask;
    return idx >= ((lane + width) & ~(width - 1)) ? lane : idx;
}
inline int src_idx(int lane, int width, int s) { return s + (lane & ~(width - 1)); }
inline int src_down(int lane, int width, int d) { return (int)((lane & (width - 1)) + d) >= width ? lane : lane + d; }
inline int src_up(int lane, int width, int d) { return (int)(lane & (width - 1)) - d < 0 ? lane : lane - d; }

int readfirstlane(uint64_t site, int v);
int readlane(uint64_t site, int v, int l);
int dpp(uint64_t site, int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl, bool keep_old_if_masked);
int ds_swizzle(uint64_t site, int v, int pattern);
uint64_t ballot(uint64_t site, int pred);
template <class T> inline T shfl_xor(uint64_t site, T v, int mask, int width = 64) { return lane_value(site, v, src_xor, width, mask); }
template <class T> inline T shfl(uint64_t site, T v, int src, int width = 64) { return lane_value(site, v, src_idx, width, src); }
template <class T> inline T shfl_down(uint64_t site, T v, unsigned d, int width = 64) { return lane_value(site, v, src_down, width, (int)d); }
template <class T> inline T shfl_up(uint64_t site, T v, unsigned d, int width = 64) { return lane_value(site, v, src_up, width, (int)d); }

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
f32x4 mfma_f32_16x16x4(uint64_t site, float a, float b, f32x4 c);
f32x16 mfma_f32_32x32x2(uint64_t site, float a, float b, f32x16 c);
f64x4 mfma_f64_16x16x4(uint64_t site, double a, double b, f64x4 c);

}  // namespace emu

#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::block_idx())
#define blockDim (::emu::block_dim())
#define gridDim (::emu::grid_dim())
constexpr int warpSize = 64;

#define hipLaunchKernelGGL(k, g, b, sm, st, ...) ::emu::launch(#k, k, dim3(g), dim3(b), (sm), (st), ##__VA_ARGS__)

// ---- work-group and wavefront intrinsics ----------------------------------------------------------
inline void __syncthreads() { ::emu::block_barrier(); }
inline int __syncthreads_and(int p) { return ::emu::block_reduce(p, 0); }
inline int __syncthreads_or(int p) { return ::emu::block_reduce(p, 1); }
inline int __syncthreads_count(int p) { return ::emu::block_reduce(p, 2); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }

#define EMU_SITE (::emu::site_id(__FILE__, __LINE__))
#define __shfl_xor(...) ::emu::shfl_xor(EMU_SITE, __VA_ARGS__)
#define __shfl(...) ::emu::shfl(EMU_SITE, __VA_ARGS__)
#define __shfl_down(...) ::emu::shfl_down(EMU_SITE, __VA_ARGS__)
#define __shfl_up(...) ::emu::shfl_up(EMU_SITE, __VA_ARGS__)
#define __ballot(p) ((unsigned long long)::emu::ballot(EMU_SITE, (int)(p)))
inline unsigned int __lane_id() {
    return (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)) & (warpSize - 1);
}

#define __builtin_amdgcn_readfirstlane(v) ::emu::readfirstlane(EMU_SITE, (int)(v))
#define __builtin_amdgcn_readlane(v, l) ::emu::readlane(EMU_SITE, (int)(v), (int)(l))
#define __builtin_amdgcn_mov_dpp(src, ctrl, rm, bm, bc) ::emu::dpp(EMU_SITE, 0, (int)(src), (ctrl), (rm), (bm), (bc), false)
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) ::emu::dpp(EMU_SITE, (int)(old), (int)(src), (ctrl), (rm), (bm), (bc), true)
#define __builtin_amdgcn_ds_swizzle(v, p) ::emu::ds_swizzle(EMU_SITE, (int)(v), (p))
#define __builtin_amdgcn_s_sleep(n) ::emu::nap()
#define __builtin_amdgcn_s_waitcnt(n) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) ::emu::mfma_f32_16x16x4(EMU_SITE, (a), (b), (c))
#define __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, x, y, z) ::emu::mfma_f32_32x32x2(EMU_SITE, (a), (b), (c))
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) ::emu::mfma_f64_16x16x4(EMU_SITE, (a), (b), (c))

// ---- atomics ---------------------------------------------------------------------------------
#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)

template <class T>
inline T atomicAdd(T *p, T v) {
    if constexpr (std::is_integral<T>::value) {
        return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
    } else {
        T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
        while (!__atomic_compare_exchange_n(p, &old, old + v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
        }
        return old;
    }
}
template <class T> inline T atomicSub(T *p, T v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicMax(T *p, T v) { return __atomic_fetch_max(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicMin(T *p, T v) { return __atomic_fetch_min(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicCAS(T *p, T cmp, T v) {
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}

// ---- bit casts and integer intrinsics -----------------------------------------------------------
inline long long __double_as_longlong(double x) { long long r; std::memcpy(&r, &x, 8); return r; }
inline double __longlong_as_double(long long x) { double r; std::memcpy(&r, &x, 8); return r; }
inline int __float_as_int(float x) { int r; std::memcpy(&r, &x, 4); return r; }
inline float __int_as_float(int x) { float r; std::memcpy(&r, &x, 4); return r; }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffs(int x) { return __builtin_ffs(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
inline unsigned __brev(unsigned x) { return __builtin_bitreverse32(x); }
inline unsigned long long __brevll(unsigned long long x) { return __builtin_bitreverse64(x); }

// device min / max of mixed arithmetic types, as the HIP headers overload them
template <class A, class B, class = std::enable_if_t<std::is_arithmetic<A>::value && std::is_arithmetic<B>::value>>
inline std::common_type_t<A, B> min(A a, B b) { return b < a ? (std::common_type_t<A, B>)b : (std::common_type_t<A, B>)a; }
template <class A, class B, class = std::enable_if_t<std::is_arithmetic<A>::value && std::is_arithmetic<B>::value>>
inline std::common_type_t<A, B> max(A a, B b) { return a < b ? (std::common_type_t<A, B>)b : (std::common_type_t<A, B>)a; }
using std::fabs;
using std::sqrt;
using std::fma;
using std::fmax;
using std::fmin;
using std::exp;
using std::sin;
using std::cos;
using std::floor;
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }

// ---- runtime API -------------------------------------------------------------------------------
hipError_t hipMalloc(void **p, size_t bytes);
template <class T> inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc(reinterpret_cast<void **>(p), bytes); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags = 0);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemcpy2D(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind kind);
hipError_t hipMemcpy2DAsync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemset(void *p, int v, size_t bytes);
hipError_t hipMemsetAsync(void *p, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipDeviceSynchronize();
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *f, int block, size_t shmem);
template <class F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F f, int block, size_t shmem) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, reinterpret_cast<const void *>(f), block, shmem);
}
hipError_t hipLaunchCooperativeKernel(const void *f, dim3 grid, dim3 block, void **args, unsigned shmem, hipStream_t s);
