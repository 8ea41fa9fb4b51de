// Kernel set (a): 8-band k.p plane-wave Hamiltonian assembly on MI355X.
//
// Each work-group owns one 64x64 tile of H.  Lanes map to consecutive rows of one column, so a
// wavefront stores 64 contiguous complex values (512 B in FP32, 1 KiB in FP64) of the
// column-major matrix.  The Delta-G tables (QD_NTAB complex FP64 arrays, ~5-16 MB) are read
// through L2 / Infinity Cache; H itself is write-once, so the kernel is bound by HBM writes:
// algorithmic bytes = N_H(N_H+1)/2 * sizeof(element) for the lower triangle.
#include "qd_internal.hpp"

namespace qd {

namespace {

// lower-triangular tile index t -> (I, J), I >= J
__device__ inline void tri_tile(long t, int &I, int &J) {
    long i = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    I = (int)i;
    J = (int)(t - i * (i + 1) / 2);
}

// The row index of a thread is split once; the column (c0 + 4 k, the same for the whole wavefront:
// readfirstlane makes that visible to the compiler) is advanced by carries, so no element pays the
// integer divisions of the index split.
template <class T>
__global__ __launch_bounds__(256) void k_build_lower(HGeom g, const c64 *__restrict__ tab, long nT, int n,
                                                     long ld, cpx<T> *__restrict__ H) {
    int I, J;
    tri_tile(blockIdx.x, I, J);
    const int r = min(I * 64 + (int)(threadIdx.x & 63), n - 1);
    const int c0 = __builtin_amdgcn_readfirstlane(J * 64 + (int)(threadIdx.x >> 6));
    if (c0 >= n) return;
    const bool row_in = I * 64 + (int)(threadIdx.x & 63) < n;
    const PWIndex R = pw_index(g, r);
    PWIndex C = pw_index(g, c0);
    auto at = [&](int k, long d) { return tab[(long)k * nT + d]; };
    for (int k = 0; k < 16; ++k) {
        const int c = c0 + 4 * k;
        if (c >= n) break;
        if (k > 0) pw_advance(g, C, 4);
        if (row_in && c <= r) {
            c64 v = h_lower_pw(g, at, R, C);
            if (r == c) v.im = 0.0;
            H[(long)c * ld + r] = cast<T>(v);
        }
    }
}

// Full Hermitian FP64 matrix (used by the FP64 residuals of the refinement).
__global__ __launch_bounds__(256) void k_build_full(HGeom g, const c64 *__restrict__ tab, long nT, int n, long ld,
                                                    c64 *__restrict__ H) {
    const int r = min((int)(blockIdx.x * 64 + (threadIdx.x & 63)), n - 1);
    const int c0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * 64 + (threadIdx.x >> 6)));
    if (c0 >= n) return;
    const bool row_in = (int)(blockIdx.x * 64 + (threadIdx.x & 63)) < n;
    const PWIndex R = pw_index(g, r);
    PWIndex C = pw_index(g, c0);
    auto at = [&](int k, long d) { return tab[(long)k * nT + d]; };
    for (int k = 0; k < 16; ++k) {
        const int c = c0 + 4 * k;
        if (c >= n) break;
        if (k > 0) pw_advance(g, C, 4);
        if (!row_in) continue;
        c64 v;
        if (r >= c) {
            v = h_lower_pw(g, at, R, C);
            if (r == c) v.im = 0.0;
        } else {
            v = conj(h_lower_pw(g, at, C, R));
        }
        H[(long)c * ld + r] = v;
    }
}

}  // namespace

template <class T>
void launch_build_lower(const DeviceTables &t, int n, long ld, cpx<T> *H, hipStream_t s) {
    const long nt = (n + 63) / 64;
    const long tiles = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_build_lower<T>, dim3((unsigned)tiles), dim3(256), 0, s, t.geom, t.tab, t.nT, n, ld, H);
    QD_LAUNCH_CHECK();
}
template void launch_build_lower<float>(const DeviceTables &, int, long, cpx<float> *, hipStream_t);
template void launch_build_lower<double>(const DeviceTables &, int, long, cpx<double> *, hipStream_t);

void launch_build_full(const DeviceTables &t, int n, long ld, c64 *H, hipStream_t s) {
    const unsigned nt = (n + 63) / 64;
    hipLaunchKernelGGL(k_build_full, dim3(nt, nt), dim3(256), 0, s, t.geom, t.tab, t.nT, n, ld, H);
    QD_LAUNCH_CHECK();
}

}  // namespace qd
