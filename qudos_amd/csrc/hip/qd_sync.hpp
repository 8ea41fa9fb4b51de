// Communication between work-groups of one running kernel on MI355X (8 XCDs, one L2 each).
//
// Data and flags shared across work-groups go through agent-coherent loads and stores: relaxed
// atomics at agent scope, which the compiler emits as sc1 accesses that no XCD's L2 keeps stale.
// They need no cache-wide write-back or invalidation (an acquire/release fence at agent scope
// emits buffer_wbl2 / buffer_inv on the whole L2, which is what polling loops must avoid).
// Protocol: a producer issues its coherent stores, waits until they completed (vmcnt 0), passes
// a work-group barrier, and one thread publishes the flag; a consumer thread polls the flag with
// coherent loads, and the work-group loads the data after a barrier.
#pragma once

#include "qd_device.hpp"

namespace qd {

constexpr long kSpinLimit = 1L << 22;   // polls before a wait gives up (seconds of waiting)
constexpr int kFlagStride = 32;         // ints between progress flags: one 128-byte line each

__device__ inline c64 ld_coh(const c64 *p) {
    const unsigned long long *q = reinterpret_cast<const unsigned long long *>(p);
    const unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return mk(__longlong_as_double((long long)a), __longlong_as_double((long long)b));
}

__device__ inline void st_coh(c64 *p, c64 v) {
    unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
    __hip_atomic_store(q, (unsigned long long)__double_as_longlong(v.re), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, (unsigned long long)__double_as_longlong(v.im), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st_coh_d(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sum of ld_coh(base[i * stride]) for i = first, first + step, ... < count, with four
// independent accumulators so that several coherent loads are in flight at once
__device__ inline c64 sum_coh(const c64 *base, long stride, int first, int step, int count) {
    c64 a0 = mk(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
    int i = first;
    for (; i + 3 * step < count; i += 4 * step) {
        const c64 x0 = ld_coh(base + (long)i * stride), x1 = ld_coh(base + (long)(i + step) * stride);
        const c64 x2 = ld_coh(base + (long)(i + 2 * step) * stride), x3 = ld_coh(base + (long)(i + 3 * step) * stride);
        a0 = a0 + x0;
        a1 = a1 + x1;
        a2 = a2 + x2;
        a3 = a3 + x3;
    }
    for (; i < count; i += step) a0 = a0 + ld_coh(base + (long)i * stride);
    return (a0 + a1) + (a2 + a3);
}

// The same sum with every load issued before the first add (one memory round trip for
// count <= first + MAXQ * step): indices are clamped into range and the surplus terms masked, so
// no load sits in a branch.  Larger counts continue in further rounds of MAXQ loads.
template <int MAXQ>
__device__ inline c64 sum_coh_wide(const c64 *base, long stride, int first, int step, int count) {
    c64 acc = mk(0.0, 0.0);
    for (int i0 = first; i0 < count; i0 += MAXQ * step) {
        c64 x[MAXQ];
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) x[q] = ld_coh(base + (long)min(i0 + q * step, count - 1) * stride);
#pragma unroll
        for (int q = 0; q < MAXQ; ++q)
            if (i0 + q * step < count) acc = acc + x[q];
    }
    return acc;
}

// s_waitcnt vmcnt(0): every vector memory access of this wavefront (stores included) completed
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ inline void publish(int *flag, int value) {
    __hip_atomic_store(flag, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread-local wait until *flag >= target.  sync[1] is the kernel's abort flag, sync[2] a sticky
// error flag read by the host: a wait that exceeds kSpinLimit polls sets both and returns false,
// and every other waiting work-group then returns false at its next poll.  *seen receives the
// last value read.  max_nap bounds the back-off between polls (units of s_sleep 2, ~128 cycles):
// waits on a latency-critical chain with few pollers use 1.
__device__ inline bool wait_at_least(int *flag, int target, int *sync, int *seen = nullptr, int max_nap = 16) {
    long spins = 0;
    int v;
    while ((v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < target) {
        if ((++spins & 15) == 0 && __hip_atomic_load(&sync[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (spins > kSpinLimit) {
            __hip_atomic_store(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sync[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        // back off (up to ~2k cycles between polls) so waiting work-groups do not flood the
        // memory system with coherent loads
        const int nap = min(max_nap, spins < 4 ? 1 : (spins < 16 ? 4 : 16));
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(2);
    }
    if (seen) *seen = v;
    return true;
}

// Grid-barrier state of a co-resident (cooperative) grid of G <= 256 work-groups (ints, zero before
// the launch), laid out for grid_sync and grid_flags_sync below.
constexpr int kGridBarrierInts = (17 + 256) * kFlagStride;

// Grid-wide barrier of a co-resident (cooperative) grid of G <= 256 work-groups, two levels deep
// so that no line sees more than 16 atomic arrivals: work-group w adds to the counter of its group
// of 16, the last arrival of a group adds to the top counter, and the last arrival overall
// releases every work-group through its own flag (one line each, so the polls do not contend).
// gb (ints, zero before the launch): group counters at g * kFlagStride (g < 16), top counter at
// 16 * kFlagStride, release flag of work-group w at (17 + w) * kFlagStride.  epoch numbers the
// barriers of the launch from 1.  s_flag, s_ok: work-group shared ints.  False if aborted.
__device__ inline bool grid_sync(int *gb, int epoch, int *sync, int *s_flag, int *s_ok) {
    const int G = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
    const int ngroups = (G + 15) >> 4;
    wait_vmem();
    __syncthreads();
    if (tid == 0) {
        const int g = w >> 4;
        const int gsize = min(16, G - g * 16);
        const int old = atomicAdd(&gb[g * kFlagStride], 1);
        int rel = 0;
        if (old + 1 == epoch * gsize) {
            const int t = atomicAdd(&gb[16 * kFlagStride], 1);
            rel = (t + 1 == epoch * ngroups);
        }
        *s_flag = rel;
    }
    __syncthreads();
    if (*s_flag)
        for (int i = tid; i < G; i += blockDim.x) publish(&gb[(17 + i) * kFlagStride], epoch);
    if (tid == 0) *s_ok = wait_at_least(&gb[(17 + w) * kFlagStride], epoch, sync, nullptr, 1);
    __syncthreads();
    return *s_ok != 0;
}

// All-to-all arrival without atomics, for a co-resident grid of G <= 256 work-groups: work-group
// w publishes the epoch at its own flag (int (17 + w) * kFlagStride of gb, zero before the
// launch) once its coherent stores completed, then waits until every work-group's flag reached
// the epoch (thread t polls flags t, t + blockDim.x, ...).  One flag store and one polling round
// trip instead of grid_sync's two levels of returning atomics, for G times the polling traffic
// (grid_sync suits kernels that share the GPU with others); the flags keep the last epoch, so
// the same state serves every step of the launch.  False if aborted.
__device__ inline bool grid_flags_sync(int *gb, int epoch, int *sync) {
    const int G = gridDim.x, tid = threadIdx.x;
    wait_vmem();
    __syncthreads();
    if (tid == 0) publish(&gb[(17 + blockIdx.x) * kFlagStride], epoch);
    bool ok = true;
    for (int w = tid; w < G; w += blockDim.x) ok = wait_at_least(&gb[(17 + w) * kFlagStride], epoch, sync, nullptr, 1) && ok;
    return __syncthreads_and(ok) != 0;
}

}  // namespace qd
