// Momentum matrix elements p_a(v, c) = <psi_v| dH/dk_a |psi_c> between valence states v and
// conduction states c of the resident window (momentum_matrix_element.c:31-103 over the dH/dk
// matrices of hamiltonian_derivative_8_band.c:23-500).
//
// Every element of H is bilinear in the row and column wave vectors k = G + K (kinetic terms),
// linear in them (Kane P terms) or independent of K (band edges, strain, piezoelectric and field
// terms), so dH/dK_a = [H(K + s e_a) - H(K - s e_a)] / (2 s) holds for any step s: the s^2 terms
// cancel.  The kernel evaluates the element formula at K +- s e_a on the fly, so no N_H x N_H
// matrix is stored; each lane owns one row, each wavefront one column at a time, and the
// conduction states are read as wavefront-uniform values.  Work per element: two element
// evaluations plus nc complex multiply-adds (compute bound, like the H assembly).
#include <functional>
#include <vector>

#include "qd_internal.hpp"

namespace qd {

namespace {

constexpr int kMomMaxCols = 16;   // conduction states per pass
// the step only scales an exact difference; 0.1 / Angstrom is the size of the plane-wave cut-offs,
// so both evaluations stay at the magnitude of the elements themselves
constexpr double kStep = 0.1;

__host__ __device__ inline void shift_k(PWIndex &p, int axis, double s) {
    if (axis == 0) p.kx += s;
    else if (axis == 1) p.ky += s;
    else p.kz += s;
}

// dH/dK_axis (r, c) from the element formula at the shifted wave vectors (Rp, Cp at K + s e_axis,
// Rm, Cm at K - s e_axis); upper-triangle elements by conjugation of the lower formula.
template <class TabAt>
__host__ __device__ inline c64 dk_element(const HGeom &g, TabAt at, const PWIndex &Rp, const PWIndex &Rm,
                                          const PWIndex &Cp, const PWIndex &Cm, int r, int c, double inv2s) {
    if (r >= c) {
        const c64 hp = h_lower_pw(g, at, Rp, Cp), hm = h_lower_pw(g, at, Rm, Cm);
        return mk((hp.re - hm.re) * inv2s, r == c ? 0.0 : (hp.im - hm.im) * inv2s);
    }
    const c64 hp = h_lower_pw(g, at, Cp, Rp), hm = h_lower_pw(g, at, Cm, Rm);
    return mk((hp.re - hm.re) * inv2s, -(hp.im - hm.im) * inv2s);
}

// Ypart[seg][j][r] = sum over columns c of segment seg of dH/dK_axis(r, c) X[j][c], j < nc
__global__ __launch_bounds__(256) void k_dk_apply(HGeom g, const c64 *__restrict__ tab, long nT, int n, int axis,
                                                  double step, const c64 *__restrict__ X, long ldx, int nc,
                                                  int cseg, c64 *__restrict__ Ypart) {
    const int r = min((int)(blockIdx.x * 64 + (threadIdx.x & 63)), n - 1);
    const bool row_in = (int)(blockIdx.x * 64 + (threadIdx.x & 63)) < n;
    const int seg = blockIdx.y;
    const int cbeg = seg * cseg, cend = min(n, cbeg + cseg);
    auto at = [&](int k, long d) { return tab[(long)k * nT + d]; };
    PWIndex Rp = pw_index(g, r), Rm = Rp;
    shift_k(Rp, axis, step);
    shift_k(Rm, axis, -step);
    c64 acc[kMomMaxCols];
#pragma unroll
    for (int j = 0; j < kMomMaxCols; ++j) acc[j] = mk(0.0, 0.0);
    const double inv2s = 0.5 / step;
    for (int c = cbeg + (int)(threadIdx.x >> 6); c < cend; c += 4) {
        const int cu = __builtin_amdgcn_readfirstlane(c);
        PWIndex Cp = pw_index(g, cu), Cm = Cp;
        shift_k(Cp, axis, step);
        shift_k(Cm, axis, -step);
        const c64 d = dk_element(g, at, Rp, Rm, Cp, Cm, r, cu, inv2s);
#pragma unroll
        for (int j = 0; j < kMomMaxCols; ++j)
            if (j < nc) fma_(acc[j], d, X[(long)j * ldx + cu]);
    }
    // the four wavefronts of the work-group cover interleaved columns: add them in LDS
    __shared__ c64 red[3][kMomMaxCols][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wv > 0) {
#pragma unroll
        for (int j = 0; j < kMomMaxCols; ++j) red[wv - 1][j][lane] = acc[j];
    }
    __syncthreads();
    if (wv == 0 && row_in) {
#pragma unroll
        for (int j = 0; j < kMomMaxCols; ++j)
            if (j < nc) Ypart[((long)seg * nc + j) * n + r] = acc[j] + red[0][j][lane] + red[1][j][lane] + red[2][j][lane];
    }
}

// P[j * nv + v] = sum_r conj(V[v][r]) sum_seg Ypart[seg][j][r]; one work-group per (v, j)
__global__ __launch_bounds__(256) void k_mom_dot(const c64 *__restrict__ V, long ldv, int nv, const c64 *__restrict__ Ypart,
                                                 int nseg, int nc, int n, c64 *__restrict__ P) {
    const int v = blockIdx.x % nv, j = blockIdx.x / nv;
    c64 acc = mk(0.0, 0.0);
    for (int r = threadIdx.x; r < n; r += 256) {
        c64 y = mk(0.0, 0.0);
        for (int sg = 0; sg < nseg; ++sg) y = y + Ypart[((long)sg * nc + j) * n + r];
        fma_(acc, conj(V[(long)v * ldv + r]), y);
    }
    __shared__ double sr[256], si[256];
    sr[threadIdx.x] = acc.re;
    si[threadIdx.x] = acc.im;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sr[threadIdx.x] += sr[threadIdx.x + w];
            si[threadIdx.x] += si[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) P[(long)j * nv + v] = mk(sr[0], si[0]);
}

}  // namespace

size_t momentum_scratch_bytes(int n, int nc) {
    const int rb = (n + 63) / 64;
    const int nseg = std::max(1, std::min((2048 + rb - 1) / rb, (n + 255) / 256));
    return sizeof(c64) * ((size_t)nseg * std::min(nc, kMomMaxCols) * n + (size_t)kMomMaxCols * 64);
}

// P (device, nc x nv per axis, P[(axis * nc + c) * nv + v]) of the nv valence states V and the nc
// conduction states C; the scratch holds momentum_scratch_bytes(n, nc).
void momentum_device(const DeviceTables &t, int n, const c64 *V, int nv, const c64 *C, int nc, long ld,
                     unsigned char *scratch, c64 *P, hipStream_t s) {
    const int rb = (n + 63) / 64;
    const int nseg = std::max(1, std::min((2048 + rb - 1) / rb, (n + 255) / 256));
    const int cseg = (((n + nseg - 1) / nseg) + 3) & ~3;
    c64 *Ypart = reinterpret_cast<c64 *>(scratch);
    for (int axis = 0; axis < 3; ++axis) {
        for (int c0 = 0; c0 < nc; c0 += kMomMaxCols) {
            const int ncp = std::min(kMomMaxCols, nc - c0);
            hipLaunchKernelGGL(k_dk_apply, dim3(rb, nseg), dim3(256), 0, s, t.geom, t.tab, t.nT, n, axis, kStep,
                               C + (long)c0 * ld, ld, ncp, cseg, Ypart);
            QD_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_mom_dot, dim3(nv * ncp), dim3(256), 0, s, V, ld, nv, Ypart, nseg, ncp, n,
                               P + ((long)axis * nc + c0) * nv);
            QD_LAUNCH_CHECK();
        }
    }
}

void parallel_for(long n, const std::function<void(long, long)> &body);   // host/shapes.cpp

// Host emulation (test support): the same element code, P in momentum_device's layout.
void momentum_host(const HGeom &g, const c64 *tab, long nT, int n, const c64 *V, int nv, const c64 *C, int nc,
                   long ld, c64 *P) {
    auto at = [&](int k, long d) { return tab[(long)k * nT + d]; };
    const double inv2s = 0.5 / kStep;
    std::vector<c64> Y((size_t)nc * n);
    for (int axis = 0; axis < 3; ++axis) {
        parallel_for(n, [&](long lo, long hi) {
            std::vector<c64> acc(nc);
            for (long r = lo; r < hi; ++r) {
                PWIndex Rp = pw_index(g, (int)r), Rm = Rp;
                shift_k(Rp, axis, kStep);
                shift_k(Rm, axis, -kStep);
                for (int j = 0; j < nc; ++j) acc[j] = mk(0.0, 0.0);
                for (int c = 0; c < n; ++c) {
                    PWIndex Cp = pw_index(g, c), Cm = Cp;
                    shift_k(Cp, axis, kStep);
                    shift_k(Cm, axis, -kStep);
                    const c64 d = dk_element(g, at, Rp, Rm, Cp, Cm, (int)r, c, inv2s);
                    for (int j = 0; j < nc; ++j) fma_(acc[j], d, C[(long)j * ld + c]);
                }
                for (int j = 0; j < nc; ++j) Y[(size_t)j * n + r] = acc[j];
            }
        });
        for (int j = 0; j < nc; ++j)
            for (int v = 0; v < nv; ++v) {
                c64 acc = mk(0.0, 0.0);
                for (int r = 0; r < n; ++r) fma_(acc, conj(V[(long)v * ld + r]), Y[(size_t)j * n + r]);
                P[((long)axis * nc + j) * nv + v] = acc;
            }
    }
}

// p[(a nc + c) nv + v] (interleaved complex) with v = 0 the highest valence state, from P of
// momentum_device / momentum_host (valence columns in ascending order).
void momentum_order(const c64 *P, int nv, int nc, double *p) {
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < nc; ++c)
            for (int v = 0; v < nv; ++v) {
                const c64 x = P[((size_t)a * nc + c) * nv + (nv - 1 - v)];
                p[2 * (((size_t)a * nc + c) * nv + v)] = x.re;
                p[2 * (((size_t)a * nc + c) * nv + v) + 1] = x.im;
            }
}

}  // namespace qd
