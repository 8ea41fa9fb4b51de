// MFMA and wavefront helpers shared by the dense kernels (gfx950, 64-lane wavefronts).
#pragma once

#include "qd_device.hpp"

namespace qd {

template <class T> struct VecT;
template <> struct VecT<float> { typedef float v4 __attribute__((ext_vector_type(4))); };
template <> struct VecT<double> { typedef double v4 __attribute__((ext_vector_type(4))); };

// 16x16x4 MFMA: lane (li = lane % 16, lk = lane / 16) supplies A(li, lk) and B(lk, li) of one
// 4-deep slice; accumulator element e of lane (li, lk) holds C(acc_row(lk, e), li).  The f32
// instruction stores rows 4*lk + e, the f64 instruction rows lk + 4*e (probed on gfx950).
template <class T> __device__ inline int acc_row(int lk, int e);
template <> __device__ inline int acc_row<float>(int lk, int e) { return 4 * lk + e; }
template <> __device__ inline int acc_row<double>(int lk, int e) { return lk + 4 * e; }

__device__ inline VecT<float>::v4 mfma(float a, float b, VecT<float>::v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline VecT<double>::v4 mfma(double a, double b, VecT<double>::v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 32x32x2 FP32 MFMA: lane l supplies A(i = l % 32, k = l / 32) and B(k = l / 32, j = l % 32);
// accumulator element e of lane l holds C(8 (e / 4) + 4 (l / 32) + e % 4, l % 32).
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ inline int acc_row32(int l, int e) { return 8 * (e >> 2) + 4 * (l >> 5) + (e & 3); }
__device__ inline f32x16 mfma32(float a, float b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum over aligned groups of `width` lanes (width a power of two <= 64)
template <class T>
__device__ inline T group_sum(T v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
__device__ inline c64 ld64(const cpx<T> *p) { return cast<double>(*p); }

// Lane reductions with DPP moves (register to register, no LDS round trip).  DPP controls:
// quad_perm [1,0,3,2] = 0xB1 (lanes 1 apart), quad_perm [2,3,0,1] = 0x4E (lanes 2 apart),
// row_half_mirror = 0x141 (lane i <-> 7-i within 8), row_mirror = 0x140 (i <-> 15-i within 16).
template <int CTRL>
__device__ inline double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum over the 4 lanes of a quad (lanes 4q .. 4q+3); every lane receives it
__device__ inline double quad_sum(double v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    return v;
}
// sum over aligned groups of 8 lanes
__device__ inline double sum8(double v) {
    v = quad_sum(v);
    v += dpp_mov<0x141>(v);
    return v;
}
// sum over aligned groups of 16 lanes
__device__ inline double sum16(double v) {
    v = sum8(v);
    v += dpp_mov<0x140>(v);
    return v;
}
// DPP move with a row mask: lanes of the rows selected by ROWMASK receive CTRL's source lane,
// the other lanes 0.  row_bcast15 = 0x142 (lane 15 of a row to the next row), row_bcast31 =
// 0x143 (lane 31 to rows 2 and 3).
template <int CTRL, int ROWMASK>
__device__ inline double dpp_rows(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum over the 64 lanes, returned to every lane (all lanes active): 16-lane DPP sums, the four
// row sums gathered into lane 63 by two row broadcasts, then read from lane 63
__device__ inline double wave_sum_dpp(double v) {
    v = sum16(v);
    v += dpp_rows<0x142, 0xA>(v);
    v += dpp_rows<0x143, 0xC>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// One level of a transposed butterfly over the 2N values v[0 .. 2N): lanes whose bit BIT is set
// keep the upper half, the others the lower half, and each adds the half it keeps from the DPP
// partner (CTRL), whose bit BIT differs and whose kept range so far is the same.
template <int N, int CTRL, int BIT>
__device__ __attribute__((always_inline)) inline void bfly_level(double (&v)[32]) {
    const bool hi = (threadIdx.x >> BIT) & 1;
#pragma unroll
    for (int t = 0; t < N; ++t) {
        const double send = hi ? v[t] : v[t + N];
        const double keep = hi ? v[t + N] : v[t];
        v[t] = keep + dpp_mov<CTRL>(send);
    }
}
// Sums over the 64 lanes of 32 values per lane (all lanes active), scattered over the lanes:
// lane l receives the total of value j(l) = 16 b3 + 8 b2 + 4 b1 + 2 b0 + b4 (b_k: bit k of l;
// lanes l and l ^ 32 receive the same).  Levels: row mirror (bit 3), half-row mirror (bit 2),
// lanes 2 apart (bit 1), lanes 1 apart (bit 0), lanes 16 apart (bit 4, ds_swizzle), lanes 32
// apart -- 16 + 8 + 4 + 2 + 1 + 1 exchanges for the 32 sums instead of 6 per sum.
__device__ inline int scatter32_index(int lane) {
    return 16 * ((lane >> 3) & 1) + 8 * ((lane >> 2) & 1) + 4 * ((lane >> 1) & 1) + 2 * (lane & 1) + ((lane >> 4) & 1);
}
__device__ inline double wave_sum32_scatter(double (&v)[32]) {
    bfly_level<16, 0x140, 3>(v);
    bfly_level<8, 0x141, 2>(v);
    bfly_level<4, 0x4E, 1>(v);
    bfly_level<2, 0xB1, 0>(v);
    const bool hi = (threadIdx.x >> 4) & 1;
    const double send = hi ? v[0] : v[1], keep = hi ? v[1] : v[0];
    const long long b = __double_as_longlong(send);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), 0x401F);   // lane ^ 16 within 32
    const int up = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
    const double w = keep + __longlong_as_double(((long long)up << 32) | (unsigned int)lo);
    return w + __shfl_xor(w, 32, 64);
}
// value of lane 0 (all lanes active)
__device__ inline double lane0(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

}  // namespace qd
