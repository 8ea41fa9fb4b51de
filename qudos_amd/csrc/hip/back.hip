// Kernel set (b), part 3: application of Q = H(0) H(1) ... (the Householder reflectors stored
// column by column: one-stage reduction, or stage 1 of the two-stage reduction, where the
// reflector of column c starts at row c + off) to a block of k vectors, in blocks of kBT
// reflectors written in compact-WY form Q_b = I - V T V^H (LAPACK zlarft 'F','C' semantics).
//   k_gram    : split-K partials of V^H V for one block            (once per solve)
//   k_tfactor : triangular T of one block from V^H V and tau        (once per solve)
//   k_vhx     : split-K partials of Y = V^H X
//   k_ty      : Y' = T Y (Q X) or T^H Y (Q^H X)
//   k_xupd    : X -= V Y'
// Used for the eigenvector back-transformation (X = Q Z_T) and, in the refinement, for
// applying the FP32 factorisation Q (T - lambda)^{-1} Q^H.
#include "qd_internal.hpp"

namespace qd {
namespace {

constexpr int kBT = 64;      // reflectors per compact-WY block
constexpr int kCH = 512;     // rows per split-K chunk
constexpr int kSUB = 32;     // rows staged in LDS at a time
constexpr int kMAXK = 64;    // max vectors

template <class T>
__device__ inline c64 vref(const cpx<T> *A, long ld, int n, int off, int c, int r) {
    if (r < c + off || r >= n || c >= n - off) return mk(0.0, 0.0);
    return cast<double>(A[(long)c * ld + r]);
}

// partial G = V^H V of block b (columns b*kBT ...), rows chunk
template <class T>
__global__ __launch_bounds__(256) void k_gram(const cpx<T> *A, long ld, int n, int off, int cb, int nb, c64 *part) {
    __shared__ c64 sV[kSUB][kBT + 1];
    const int tid = threadIdx.x;
    const int r0 = cb + off + blockIdx.x * kCH;
    c64 acc[16];
    for (int m = 0; m < 16; ++m) acc[m] = mk(0.0, 0.0);
    const int qa = tid & 63, qb0 = tid >> 6;   // output (qa, qb0 + 4m)
    for (int s = 0; s < kCH; s += kSUB) {
        for (int idx = tid; idx < kSUB * kBT; idx += 256) {
            const int rr = idx % kSUB, q = idx / kSUB;
            sV[rr][q] = q < nb ? vref(A, ld, n, off, cb + q, r0 + s + rr) : mk(0.0, 0.0);
        }
        __syncthreads();
        for (int rr = 0; rr < kSUB; ++rr) {
            const c64 a = conj(sV[rr][qa]);
#pragma unroll
            for (int m = 0; m < 16; ++m) fma_(acc[m], a, sV[rr][qb0 + 4 * m]);
        }
        __syncthreads();
    }
    c64 *dst = part + (long)blockIdx.x * kBT * kBT;
    for (int m = 0; m < 16; ++m) dst[(qb0 + 4 * m) * kBT + qa] = acc[m];   // column-major G(qa, qb)
}

template <class T>
__global__ __launch_bounds__(256) void k_tfactor(const c64 *part, int nchunk, const c64 *tau, int cb, int nb, c64 *Tout) {
    __shared__ c64 G[kBT][kBT + 1];
    __shared__ c64 Tm[kBT][kBT + 1];   // Tm[row][col]
    const int tid = threadIdx.x;
    for (int idx = tid; idx < kBT * kBT; idx += 256) {
        c64 s = mk(0.0, 0.0);
        for (int ch = 0; ch < nchunk; ++ch) s = s + part[(long)ch * kBT * kBT + idx];
        const int a = idx % kBT, b = idx / kBT;
        G[a][b] = s;
        Tm[a][b] = mk(0.0, 0.0);
    }
    __syncthreads();
    for (int i = 0; i < nb; ++i) {
        const c64 ti = tau[cb + i];
        // z = G(0:i, i); Tm(0:i, i) = -tau_i Tm(0:i,0:i) z
        if (tid < i) {
            c64 s = mk(0.0, 0.0);
            for (int q = tid; q < i; ++q) s = s + Tm[tid][q] * G[q][i];
            Tm[tid][i] = -(ti * s);
        }
        if (tid == 0) Tm[i][i] = ti;
        __syncthreads();
    }
    for (int idx = tid; idx < kBT * kBT; idx += 256) {
        const int a = idx % kBT, b = idx / kBT;
        Tout[idx] = Tm[a][b];
    }
}

// split-K partials of Y = V^H X over 64-row chunks; thread (q, l0) owns outputs (q, l0 + 4m)
template <class T>
__global__ __launch_bounds__(256) void k_vhx(const cpx<T> *A, long ld, int n, int off, int cb, int nb, const c64 *X, long ldx,
                                             int k, c64 *part) {
    __shared__ c64 sV[64][kBT + 1];
    __shared__ c64 sX[64][kMAXK + 1];
    const int tid = threadIdx.x;
    const int r0 = cb + off + blockIdx.x * 64;
    for (int idx = tid; idx < 64 * kBT; idx += 256) {
        const int rr = idx & 63, qq = idx >> 6;
        sV[rr][qq] = qq < nb ? vref(A, ld, n, off, cb + qq, r0 + rr) : mk(0.0, 0.0);
    }
    for (int idx = tid; idx < 64 * k; idx += 256) {
        const int rr = idx & 63, l = idx >> 6;
        const int r = r0 + rr;
        sX[rr][l] = r < n ? X[(long)l * ldx + r] : mk(0.0, 0.0);
    }
    __syncthreads();
    const int q = tid & 63, l0 = tid >> 6;
    c64 acc[kMAXK / 4];
#pragma unroll
    for (int m = 0; m < kMAXK / 4; ++m) acc[m] = mk(0.0, 0.0);
    for (int rr = 0; rr < 64; ++rr) {
        const c64 v = conj(sV[rr][q]);
#pragma unroll
        for (int m = 0; m < kMAXK / 4; ++m)
            if (l0 + 4 * m < k) fma_(acc[m], v, sX[rr][l0 + 4 * m]);
    }
    c64 *dst = part + (long)blockIdx.x * kBT * kMAXK;
#pragma unroll
    for (int m = 0; m < kMAXK / 4; ++m)
        if (l0 + 4 * m < k) dst[(l0 + 4 * m) * kBT + q] = acc[m];
}

// Y'(:, l) = T Y(:, l) (or T^H Y(:, l)); one work-group per vector l reduces the chunk partials.
__global__ __launch_bounds__(256) void k_ty(const c64 *part, int nchunk, const c64 *Tb, int nb, int k, bool adjoint,
                                            c64 *Yout) {
    __shared__ c64 s4[4][kBT];
    __shared__ c64 Y[kBT];
    const int tid = threadIdx.x, l = blockIdx.x;
    const int q = tid & 63, pr = tid >> 6;
    c64 a8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a8[u] = mk(0.0, 0.0);
    for (int c0 = pr * 8; c0 < nchunk; c0 += 32) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c0 + u < nchunk) a8[u] = a8[u] + part[(long)(c0 + u) * kBT * kMAXK + l * kBT + q];
    }
    c64 acc = mk(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + a8[u];
    s4[pr][q] = acc;
    __syncthreads();
    if (tid < kBT) Y[tid] = s4[0][tid] + s4[1][tid] + s4[2][tid] + s4[3][tid];
    __syncthreads();
    if (tid < kBT) {
        c64 sum = mk(0.0, 0.0);
        if (q < nb) {
            if (!adjoint) {   // (T Y)(q) = sum_{p >= q} T(q,p) Y(p)
                for (int p = q; p < nb; ++p) sum = sum + Tb[p * kBT + q] * Y[p];
            } else {          // (T^H Y)(q) = sum_{p <= q} conj(T(p,q)) Y(p)
                for (int p = 0; p <= q; ++p) sum = sum + conj(Tb[q * kBT + p]) * Y[p];
            }
        }
        Yout[l * kBT + q] = sum;
    }
}

// X(r, l) -= sum_q V(r, q) Y'(q, l): thread per (row, vector), 64 rows x 4 vectors per work-group
template <class T>
__global__ __launch_bounds__(256) void k_xupd(const cpx<T> *A, long ld, int n, int off, int cb, int nb, c64 *X, long ldx, int k,
                                              const c64 *Y) {
    __shared__ c64 sY[kBT][4];
    const int tid = threadIdx.x;
    const int lb = blockIdx.y * 4;
    for (int idx = tid; idx < kBT * 4; idx += 256) {
        const int q = idx % kBT, ll = idx / kBT;
        sY[q][ll] = (lb + ll < k) ? Y[(lb + ll) * kBT + q] : mk(0.0, 0.0);
    }
    __syncthreads();
    const int r = cb + off + blockIdx.x * 64 + (tid & 63);
    const int ll = tid >> 6, l = lb + ll;
    if (r >= n || l >= k) return;
    const int qmax = min(nb, r - cb - off + 1);   // v(r, q) = 0 unless r >= cb + q + off
    c64 s = mk(0.0, 0.0);
#pragma unroll 8
    for (int q = 0; q < qmax; ++q) fma_(s, cast<double>(A[(long)(cb + q) * ld + r]), sY[q][ll]);
    X[(long)l * ldx + r] = X[(long)l * ldx + r] - s;
}

}  // namespace

template <class T>
void back_alloc(BackWork<T> &b, int n, int k) {
    b.n = n;
    b.k = k;
    b.npan = (n + kBT - 1) / kBT;
    const long nchunk = (n + 63) / 64 + 1;   // 64-row chunks of the split-K products
    QD_HIP(hipMalloc(&b.Tm, sizeof(c64) * (long)b.npan * kBT * kBT));
    QD_HIP(hipMalloc(&b.part, sizeof(c64) * nchunk * kBT * kMAXK));
    QD_HIP(hipMalloc(&b.Y, sizeof(c64) * kBT * kMAXK));
}

template <class T>
void back_free(BackWork<T> &b) {
    hipFree(b.Tm); hipFree(b.part); hipFree(b.Y);
    b = BackWork<T>();
}

template <class T>
void back_factors(BackWork<T> &b, const Reflectors<T> &r, hipStream_t s) {
    const int n = r.n;
    for (int p = 0; p < b.npan; ++p) {
        const int cb = p * kBT;
        const int nb = std::min(kBT, n - r.off - cb);
        if (nb <= 0) continue;
        const int rows = n - cb - r.off;
        const int nchunk = (rows + kCH - 1) / kCH;
        hipLaunchKernelGGL(k_gram<T>, dim3(nchunk), dim3(256), 0, s, r.A, r.ld, n, r.off, cb, nb, b.part);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_tfactor<T>, dim3(1), dim3(256), 0, s, b.part, nchunk, r.tau, cb, nb,
                           b.Tm + (long)p * kBT * kBT);
        QD_LAUNCH_CHECK();
    }
}

template <class T>
void back_apply(BackWork<T> &b, const Reflectors<T> &r, c64 *X, int k, bool adjoint, hipStream_t s) {
    if (k > kMAXK) throw HipError("too many vectors for the back-transformation", QD_E_UNSUPPORTED);
    const int n = r.n;
    for (int i = 0; i < b.npan; ++i) {
        const int p = adjoint ? i : b.npan - 1 - i;
        const int cb = p * kBT;
        const int nb = std::min(kBT, n - r.off - cb);
        if (nb <= 0) continue;
        const int rows = n - cb - r.off;
        const int nch = (rows + 63) / 64;
        hipLaunchKernelGGL(k_vhx<T>, dim3(nch), dim3(256), 0, s, r.A, r.ld, n, r.off, cb, nb, X, r.ld, k, b.part);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_ty, dim3(k), dim3(256), 0, s, b.part, nch, b.Tm + (long)p * kBT * kBT, nb, k, adjoint, b.Y);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_xupd<T>, dim3(nch, (k + 3) / 4), dim3(256), 0, s, r.A, r.ld, n, r.off, cb, nb, X, r.ld, k, b.Y);
        QD_LAUNCH_CHECK();
    }
}

template void back_alloc<float>(BackWork<float> &, int, int);
template void back_alloc<double>(BackWork<double> &, int, int);
template void back_free<float>(BackWork<float> &);
template void back_free<double>(BackWork<double> &);
template void back_factors<float>(BackWork<float> &, const Reflectors<float> &, hipStream_t);
template void back_factors<double>(BackWork<double> &, const Reflectors<double> &, hipStream_t);
template void back_apply<float>(BackWork<float> &, const Reflectors<float> &, c64 *, int, bool, hipStream_t);
template void back_apply<double>(BackWork<double> &, const Reflectors<double> &, c64 *, int, bool, hipStream_t);

}  // namespace qd
