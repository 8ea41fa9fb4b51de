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

}  // namespace qd
