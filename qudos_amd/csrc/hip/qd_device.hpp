// Device-side building blocks shared by the MI355X kernels: complex arithmetic on plain
// (re, im) pairs, launch-error checking, and the per-element 8-band k.p formula that both the
// GPU Hamiltonian kernels and host-side tests evaluate.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "qudos_gpu.h"

namespace qd {

template <class T>
struct cpx {
    T re, im;
};

template <class T> __host__ __device__ inline cpx<T> mk(T r, T i) { return cpx<T>{r, i}; }
template <class T> __host__ __device__ inline cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> __host__ __device__ inline cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> __host__ __device__ inline cpx<T> operator-(cpx<T> a) { return {-a.re, -a.im}; }
template <class T> __host__ __device__ inline cpx<T> operator*(cpx<T> a, cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> __host__ __device__ inline cpx<T> operator*(T s, cpx<T> a) { return {s * a.re, s * a.im}; }
template <class T> __host__ __device__ inline cpx<T> conj(cpx<T> a) { return {a.re, -a.im}; }
template <class T> __host__ __device__ inline T norm2(cpx<T> a) { return a.re * a.re + a.im * a.im; }
// a * conj(b)
template <class T> __host__ __device__ inline cpx<T> mulc(cpx<T> a, cpx<T> b) {
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// conj(a) * b
template <class T> __host__ __device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// acc += a * b
template <class T> __host__ __device__ inline void fma_(cpx<T> &acc, cpx<T> a, cpx<T> b) {
    acc.re = std::fma(a.re, b.re, std::fma(-a.im, b.im, acc.re));
    acc.im = std::fma(a.re, b.im, std::fma(a.im, b.re, acc.im));
}
template <class T> __host__ __device__ inline cpx<T> div_(cpx<T> a, cpx<T> b) {
    const T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <class To, class From> __host__ __device__ inline cpx<To> cast(cpx<From> a) {
    return {(To)a.re, (To)a.im};
}

using c64 = cpx<double>;
using c32 = cpx<float>;

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string &m, int c) : std::runtime_error(m), code(c) {}
};

#define QD_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::qd::HipError(std::string(#call) + ": " + hipGetErrorString(e_), QD_E_DEVICE); \
    } while (0)

#define QD_STR2(x) #x
#define QD_STR(x) QD_STR2(x)
#define QD_LAUNCH_CHECK()                                                                     \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            throw ::qd::HipError(std::string("kernel launch at " __FILE__ ":" QD_STR(__LINE__) ": ") + \
                                     hipGetErrorString(e_), QD_E_DEVICE);                     \
    } while (0)

// ---------------------------------------------------------------------------------------------
// Per-element Hamiltonian.  Row index r = b_r*N_G + g_r, g = (ix*Ny + iy)*Nz + iz
// (hamiltonian_8_band.c:39-52).  The Delta-G table index uses d = m_row - m_col + NT/2 per axis.
// The lower block table (row band >= col band) is hamiltonian_8_band.c:78-492; the elements
// E_CB ... U are the sums of hamiltonian_8_band_matrix_elements.c:27-969 with every per-shape
// sum folded into the QD_T_* tables.
// ---------------------------------------------------------------------------------------------
struct HGeom {
    int N[3], NT[3], NG;
    double K[3];
    const double *G[3];      // device or host pointers to the NT-long axes
};

__host__ __device__ inline void g_split(const HGeom &g, int gi, int &ix, int &iy, int &iz) {
    const int nyz = g.N[1] * g.N[2];
    ix = gi / nyz;
    const int rem = gi - ix * nyz;
    iy = rem / g.N[2];
    iz = rem - iy * g.N[2];
}

// A basis index r split into its band b and plane-wave indices (x, y, z), with k = G + K.
struct PWIndex {
    int b, x, y, z;
    double kx, ky, kz;
};
__host__ __device__ inline void pw_k(const HGeom &g, PWIndex &p) {
    p.kx = g.G[0][g.NT[0] / 2 - g.N[0] / 2 + p.x] + g.K[0];
    p.ky = g.G[1][g.NT[1] / 2 - g.N[1] / 2 + p.y] + g.K[1];
    p.kz = g.G[2][g.NT[2] / 2 - g.N[2] / 2 + p.z] + g.K[2];
}
__host__ __device__ inline PWIndex pw_index(const HGeom &g, int r) {
    PWIndex p;
    p.b = r / g.NG;
    g_split(g, r - p.b * g.NG, p.x, p.y, p.z);
    pw_k(g, p);
    return p;
}
// the index r + s (s >= 0) by carries instead of divisions
__host__ __device__ inline void pw_advance(const HGeom &g, PWIndex &p, int s) {
    p.z += s;
    while (p.z >= g.N[2]) {
        p.z -= g.N[2];
        if (++p.y == g.N[1]) {
            p.y = 0;
            if (++p.x == g.N[0]) {
                p.x = 0;
                ++p.b;
            }
        }
    }
    pw_k(g, p);
}

// Element (r, c) for r >= c of the 8x8-band Hermitian H, from the split indices of r and c.  The
// formula is linear in the tables; tab(k, d) returns table k at Delta-G index d (the device
// tables, or a one-hot selector when the adjoint needs the coefficient of one table).
template <class TabAt>
__host__ __device__ inline c64 h_lower_pw(const HGeom &g, TabAt tab, const PWIndex &R, const PWIndex &Cc) {
    const int br = R.b, bc = Cc.b;
    const int rx = R.x, ry = R.y, rz = R.z, cx = Cc.x, cy = Cc.y, cz = Cc.z;
    const long d = ((long)(rx - cx + g.NT[0] / 2) * g.NT[1] + (ry - cy + g.NT[1] / 2)) * g.NT[2] + (rz - cz + g.NT[2] / 2);
    // k = G + K of row / column plane waves
    const double k1x = R.kx, k1y = R.ky, k1z = R.kz;
    const double k2x = Cc.kx, k2y = Cc.ky, k2z = Cc.kz;
    auto T = [&](int k) { return tab(k, d); };
    const double hb = 7.61996;
    const double kxy = k1x * k2x + k1y * k2y, kz = k1z * k2z;
    const double sx = k1x + k2x, sy = k1y + k2y, sz = k1z + k2z;
    const double r2 = 1.4142135623730951, r3 = 1.7320508075688772, rh = 0.7071067811865476;
    const double r6 = 2.449489742783178;
    const c64 I1 = mk(0.0, 1.0);
    // element kinds
    auto ECB = [&]() { return T(QD_T_ECB) + (0.5 * hb * (kxy + kz)) * T(QD_T_SC); };
    auto EHH = [&]() {
        const c64 g1 = T(QD_T_G1), g2 = T(QD_T_G2);
        return T(QD_T_EHH) - (0.5 * hb * kxy) * (g1 + g2) - (0.5 * hb * kz) * (g1 - 2.0 * g2);
    };
    auto ELH = [&]() {
        const c64 g1 = T(QD_T_G1), g2 = T(QD_T_G2);
        return T(QD_T_ELH) - (0.5 * hb * kxy) * (g1 - g2) - (0.5 * hb * kz) * (g1 + 2.0 * g2);
    };
    auto ESO = [&]() { return T(QD_T_ESO) - (0.5 * hb * (kxy + kz)) * T(QD_T_G1); };
    auto Qe = [&]() { return (-rh * hb * kxy + r2 * hb * kz) * T(QD_T_G2) + T(QD_T_Q); };
    // R (sgn = -1) and R* (sgn = +1)
    auto Re_ = [&](double sgn) {
        const c64 a = (0.5 * r3 * hb * (k1x * k2x - k1y * k2y)) * T(QD_T_G2);
        const c64 b = (0.5 * r3 * hb * (k1x * k2y + k1y * k2x)) * T(QD_T_G3);
        return a + sgn * (I1 * b) + T(QD_T_RB) - sgn * (I1 * T(QD_T_RD));
    };
    // S (sgn = -1) and S* (sgn = +1)
    auto Se = [&](double sgn) {
        const double c = 0.5 * 1.224744871391589 * hb;   // sqrt(3/2)/2 * hbar^2/m0
        const c64 g3 = T(QD_T_G3);
        const c64 a = (c * (k1x * k2z + k1z * k2x)) * g3;
        const c64 b = (c * (k1y * k2z + k1z * k2y)) * g3;
        return a + sgn * (I1 * b) - rh * (T(QD_T_SXZ) + sgn * (I1 * T(QD_T_SYZ)));
    };
    // T+ (sgn = +1) and T- (sgn = -1)
    auto Te = [&](double sgn) {
        const c64 P = T(QD_T_P);
        const c64 kin = (1.0 / (2.0 * r6)) * (sx * P + sgn * (I1 * (sy * P)));
        const c64 Pxx = T(QD_T_PXX), Pyy = T(QD_T_PYY), Pxy = T(QD_T_PXY), Pxz = T(QD_T_PXZ), Pyz = T(QD_T_PYZ);
        const c64 st = sx * (Pxx + sgn * (I1 * Pxy)) + sy * (Pxy + sgn * (I1 * Pyy)) + sz * (Pxz + sgn * (I1 * Pyz));
        return kin - (0.5 / r6) * st;
    };
    auto Ue = [&]() {
        return (sz / (2.0 * r3)) * T(QD_T_P) - (0.5 / r3) * (sx * T(QD_T_PXZ) + sy * T(QD_T_PYZ) + sz * T(QD_T_PZZ));
    };
    const c64 zero = mk(0.0, 0.0);
    switch (bc * 8 + br) {
    case 0 * 8 + 0: return ECB();
    case 0 * 8 + 1: return -r3 * Te(-1.0);
    case 0 * 8 + 2: return r2 * Ue();
    case 0 * 8 + 3: return -Ue();
    case 0 * 8 + 6: return -Te(1.0);
    case 0 * 8 + 7: return -r2 * Te(1.0);
    case 1 * 8 + 1: return EHH();
    case 1 * 8 + 2: return r2 * Se(1.0);
    case 1 * 8 + 3: return -Se(1.0);
    case 1 * 8 + 6: return -Re_(1.0);
    case 1 * 8 + 7: return -r2 * Re_(1.0);
    case 2 * 8 + 2: return ELH();
    case 2 * 8 + 3: return Qe();
    case 2 * 8 + 4: return Te(1.0);
    case 2 * 8 + 5: return Re_(1.0);
    case 2 * 8 + 7: return r3 * Se(1.0);
    case 3 * 8 + 3: return ESO();
    case 3 * 8 + 4: return r2 * Te(1.0);
    case 3 * 8 + 5: return r2 * Re_(1.0);
    case 3 * 8 + 6: return -r3 * Se(1.0);
    case 4 * 8 + 4: return ECB();
    case 4 * 8 + 5: return -r3 * Te(1.0);
    case 4 * 8 + 6: return r2 * Ue();
    case 4 * 8 + 7: return -Ue();
    case 5 * 8 + 5: return EHH();
    case 5 * 8 + 6: return r2 * Se(-1.0);
    case 5 * 8 + 7: return -Se(-1.0);
    case 6 * 8 + 6: return ELH();
    case 6 * 8 + 7: return Qe();
    case 7 * 8 + 7: return ESO();
    default: return zero;
    }
}

template <class TabAt>
__host__ __device__ inline c64 h_lower_t(const HGeom &g, TabAt tab, int r, int c) {
    return h_lower_pw(g, tab, pw_index(g, r), pw_index(g, c));
}

// Element (r, c), r >= c, from QD_NTAB x nT complex tables.
__host__ __device__ inline c64 h_lower(const HGeom &g, const c64 *__restrict__ tab, long nT, int r, int c) {
    return h_lower_t(g, [&](int k, long d) { return tab[(long)k * nT + d]; }, r, c);
}

// Full Hermitian element (upper triangle by conjugation of the lower formula).
__host__ __device__ inline c64 h_full(const HGeom &g, const c64 *__restrict__ tab, long nT, int r, int c) {
    if (r >= c) {
        c64 v = h_lower(g, tab, nT, r, c);
        if (r == c) v.im = 0.0;
        return v;
    }
    return conj(h_lower(g, tab, nT, c, r));
}

}  // namespace qd
