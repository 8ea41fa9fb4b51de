// Kernel set (b), two-stage reduction, stage 1: a Hermitian matrix (lower triangle, column
// major) is reduced to band form with bandwidth kBW by blocked Householder QR of panels and
// two-sided trailing updates on MFMA (LAPACK zhetrd_he2hb semantics, UPLO = 'L').
//
// Panel p covers columns p0 .. p0 + kBW - 1; its QR factorises A[t0:n, p0:t0] (t0 = p0 + kBW),
// Q_p = H_0 ... H_{nb-1} = I - V T V^H, and the trailing matrix A22 = A[t0:n, t0:n] becomes
// Q_p^H A22 Q_p = A22 - W V^H - V W^H with X = A22 V T, M = T^H V^H X, W = X - V M / 2.
//   k_pqr_step   : applies the previous reflector to a work-group's 256 panel rows and forms the
//                  group's partial sums for the next column: |x|^2, x^H a_k (later columns) and
//                  v_q^H x (earlier reflectors, for T).  One pass over the rows per column.
//   k_pqr_reduce : one work-group turns the partials into beta, tau, 1/(alpha - beta) (zlarfg),
//                  v^H a_k for the next step and column i of T (zlarft 'F', 'C').
//   k_pqr_finish : explicit V (zero above the unit diagonal) for the trailing update.
//   k_hemm       : X = A22 V on MFMA tiles staged through LDS.  The lower triangle is the only
//                  stored part: tiles above the diagonal are read conjugate-transposed and the
//                  diagonal tile is completed by symmetry.  Split over column segments.
//   k_xfin       : X <- (sum of segments) T, and per row block the partials of V^H X.
//   k_msum       : M = T^H V^H X.        k_wform : W = X - V M / 2.
//   k_her2k_band : A22 -= [V W] [W V]^H on MFMA (lower tiles only).
//   k_hemm_f32, k_her2k_f32 : the FP32 forms of the two trailing kernels on 32x32x2 MFMAs.
// The panel's reflectors stay in A below the band (unit entry row c + kBW of column c), the
// band (R of each panel) above them; band_extract copies the band for stage 2 and then writes
// the unit entries for the back-transformation.
#include <algorithm>
#include <array>
#include <type_traits>

#include "qd_internal.hpp"
#include "qd_mfma.hpp"
#include "qd_sync.hpp"

namespace qd {
namespace {

constexpr int B = kBW;          // band width = panel width
constexpr int CPW = B / 4;      // panel columns per wavefront in k_pqr_step

// ------------------------------------------------------------------------------------------
// Panel QR
// ------------------------------------------------------------------------------------------
// sc layout: [0] 1/(alpha - beta), [1] tau, [2] beta, [4 + k] v^H a_k of the last reflector.
template <class T>
__global__ __launch_bounds__(256) void k_pqr_step(cpx<T> *P, long ld, int m, int nbp, int i, const c64 *sc,
                                                  c64 *part, c64 *piv) {
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int rbase = blockIdx.x * 256;
    const int q = i - 1;
    const c64 one = mk(1.0, 0.0);
    __shared__ c64 sx[256];
    c64 s_q = mk(0.0, 0.0), tau_q = mk(0.0, 0.0), beta_q = mk(0.0, 0.0);
    c64 vha[CPW];
    if (i > 0) {
        s_q = sc[0];
        tau_q = sc[1];
        beta_q = sc[2];
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) vha[kk] = sc[4 + wv * CPW + kk];
    }
    // phase 1: loads (all of them before any store: column q is read by every wavefront)
    c64 a[4][CPW], xq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        xq[u] = mk(0.0, 0.0);
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) a[u][kk] = mk(0.0, 0.0);
        if (r < m) {
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) a[u][kk] = ld64(&P[(long)(wv * CPW + kk) * ld + r]);
            if (i > 0 && r > q) xq[u] = ld64(&P[(long)q * ld + r]);
        }
    }
    __syncthreads();
    // phase 2: apply H_q^H to the columns after q (rows >= q); store v_q (or beta at the pivot)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        if (r < m && i > 0 && r >= q) {
            const c64 vq = (r == q) ? one : s_q * xq[u];
            const c64 ct = conj(tau_q) * vq;
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) {
                const int k = wv * CPW + kk;
                if (k > q) {
                    a[u][kk] = a[u][kk] - ct * vha[kk];
                    P[(long)k * ld + r] = cast<T>(a[u][kk]);
                } else if (k == q) {
                    a[u][kk] = (r == q) ? beta_q : vq;
                    P[(long)k * ld + r] = cast<T>(a[u][kk]);
                }
            }
        }
        if (i < nbp && wv == i / CPW) {
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk)
                if (kk == i % CPW) sx[lane + 64 * u] = a[u][kk];
        }
    }
    if (i >= nbp) return;
    __syncthreads();
    // phase 3: partial sums for column i over rows r > i; pivot row values from row i
    c64 acc[CPW];
#pragma unroll
    for (int kk = 0; kk < CPW; ++kk) acc[kk] = mk(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        if (r >= m) continue;
        if (r == i) {
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) piv[wv * CPW + kk] = a[u][kk];
            continue;
        }
        if (r < i) continue;
        const c64 x = sx[lane + 64 * u];
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) {
            const int k = wv * CPW + kk;
            if (k > i) acc[kk] = acc[kk] + cmul(x, a[u][kk]);
            else if (k == i) acc[kk].re += norm2(x);
            else acc[kk] = acc[kk] + cmul(a[u][kk], x);
        }
    }
#pragma unroll
    for (int kk = 0; kk < CPW; ++kk) {
        const double re = wave_sum_dpp(acc[kk].re), im = wave_sum_dpp(acc[kk].im);
        if (lane == 0) part[(long)blockIdx.x * B + wv * CPW + kk] = mk(re, im);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_pqr_reduce(const c64 *part, int nblk, const c64 *piv, int i, c64 *sc,
                                                    c64 *Tp, c64 *tau_out) {
    __shared__ c64 red[8][B];
    __shared__ c64 ssum[B], sg[B];
    __shared__ c64 s_s, s_tau;
    const int tid = threadIdx.x, k = tid & (B - 1), g = tid >> 5;
    c64 acc = mk(0.0, 0.0);
    for (int blk = g; blk < nblk; blk += 8) acc = acc + part[(long)blk * B + k];
    red[g][k] = acc;
    __syncthreads();
    if (tid < B) {
        c64 s = mk(0.0, 0.0);
#pragma unroll
        for (int gg = 0; gg < 8; ++gg) s = s + red[gg][tid];
        ssum[tid] = s;
    }
    __syncthreads();
    if (tid == 0) {
        const c64 alpha = piv[i];
        const double xn2 = ssum[i].re;
        c64 tau, scl;
        double beta;
        if (xn2 == 0.0 && alpha.im == 0.0) {
            tau = mk(0.0, 0.0);
            beta = alpha.re;
            scl = mk(0.0, 0.0);
        } else {
            const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
            beta = alpha.re >= 0.0 ? -nrm : nrm;
            tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
            scl = div_(mk(1.0, 0.0), alpha - mk(beta, 0.0));
        }
        sc[0] = scl;
        sc[1] = tau;
        sc[2] = mk(beta, 0.0);
        *tau_out = tau;
        s_s = scl;
        s_tau = tau;
    }
    __syncthreads();
    const c64 scl = s_s, tau = s_tau;
    if (tid < B) {
        sc[4 + tid] = tid > i ? piv[tid] + conj(scl) * ssum[tid] : mk(0.0, 0.0);
        if (tid < i) sg[tid] = conj(piv[tid]) + scl * ssum[tid];   // V(:, tid)^H v_i
    }
    __syncthreads();
    if (tid < B) {
        c64 t = mk(0.0, 0.0);
        if (tid < i) {
            c64 s = mk(0.0, 0.0);
            for (int qq = tid; qq < i; ++qq) s = s + Tp[qq * B + tid] * sg[qq];
            t = -(tau * s);
        } else if (tid == i) {
            t = tau;
        }
        Tp[i * B + tid] = t;
    }
}

// Panel QR for the FP64 path, one column per grid-wide barrier.  The FP64 path runs the panel QR
// concurrently with the trailing update (look-ahead), where the footprint matters more than the
// speed: a cooperative launch waits until every work-group is resident, and this kernel stays
// within 256 registers per lane (a wavefront fits beside one of k_her2k_band<double> on a SIMD)
// and polls one flag per work-group (two-level barrier).  Every work-group keeps its 256 panel
// rows in registers for all columns.  Needs all work-groups resident (cooperative launch).
// part / pivb are double-buffered by column parity: a work-group can only write the buffers of
// column i+1 after every work-group passed the barrier of column i, i.e. finished reading the
// buffers of column i-1.  gb: grid-barrier state (zero at launch).
template <class T>
__global__ __launch_bounds__(256) void k_pqr_panel1(cpx<T> *P, long ld, int m, int nbp, c64 *part, c64 *pivb, c64 *gpart,
                                                   c64 *tau, int *sync, int *gb) {
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int blk = blockIdx.x, nblk = gridDim.x;
    const int rbase = blk * 256;
    __shared__ c64 sx[256];
    __shared__ c64 red[8][B];
    __shared__ c64 ssum[B], spiv[B], svha[B], sg[B];
    __shared__ c64 sV[64][B + 1];
    __shared__ c64 s_scl, s_tau;
    __shared__ double s_beta;
    __shared__ int s_ok, s_flag;
    c64 a[4][CPW];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) a[u][kk] = r < m ? ld64(&P[(long)(wv * CPW + kk) * ld + r]) : mk(0.0, 0.0);
    }
    if (wv == 0)
#pragma unroll
        for (int u = 0; u < 4; ++u) sx[lane + 64 * u] = a[u][0];
    __syncthreads();
    for (int i = 0; i < nbp; ++i) {
        c64 *pt = part + (long)(i & 1) * nblk * B;
        c64 *pv = pivb + (i & 1) * B;
        // partials of column i over rows r > i; the pivot row (row i, in work-group 0) is published
        c64 acc[CPW];
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) acc[kk] = mk(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = rbase + lane + 64 * u;
            if (r >= m || r < i) continue;
            if (r == i) {
#pragma unroll
                for (int kk = 0; kk < CPW; ++kk) st_coh(&pv[wv * CPW + kk], a[u][kk]);
                continue;
            }
            const c64 x = sx[lane + 64 * u];
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) {
                const int k = wv * CPW + kk;
                if (k > i) acc[kk] = acc[kk] + cmul(x, a[u][kk]);
                else if (k == i) acc[kk].re += norm2(x);
                else acc[kk] = acc[kk] + cmul(a[u][kk], x);
            }
        }
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) {
            const double re = wave_sum_dpp(acc[kk].re), im = wave_sum_dpp(acc[kk].im);
            if (lane == 0) st_coh(&pt[(long)blk * B + wv * CPW + kk], mk(re, im));
        }
        if (!grid_sync(gb, i + 1, sync, &s_flag, &s_ok)) return;
        {
            const int k = tid & (B - 1), g = tid >> 5;
            const c64 pivv = ld_coh(&pv[k]);   // issued first: its latency overlaps the partial sums
            red[g][k] = sum_coh_wide<16>(pt + k, B, g, 8, nblk);
            if (tid < B) spiv[tid] = pivv;
        }
        __syncthreads();
        if (tid < B) {
            c64 t = mk(0.0, 0.0);
#pragma unroll
            for (int g = 0; g < 8; ++g) t = t + red[g][tid];
            ssum[tid] = t;
        }
        __syncthreads();
        if (tid == 0) {
            const c64 alpha = spiv[i];
            const double xn2 = ssum[i].re;
            c64 tq, scl;
            double beta;
            if (xn2 == 0.0 && alpha.im == 0.0) {
                tq = mk(0.0, 0.0);
                beta = alpha.re;
                scl = mk(0.0, 0.0);
            } else {
                const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
                beta = alpha.re >= 0.0 ? -nrm : nrm;
                tq = mk((beta - alpha.re) / beta, -alpha.im / beta);
                scl = div_(mk(1.0, 0.0), alpha - mk(beta, 0.0));
            }
            s_scl = scl;
            s_tau = tq;
            s_beta = beta;
            if (blk == 0) tau[i] = tq;
        }
        __syncthreads();
        const c64 scl = s_scl, tq = s_tau;
        if (tid < B) {
            svha[tid] = tid > i ? spiv[tid] + conj(scl) * ssum[tid] : mk(0.0, 0.0);
            sg[tid] = tid < i ? conj(spiv[tid]) + scl * ssum[tid] : mk(0.0, 0.0);   // V(:, tid)^H v_i
        }
        __syncthreads();
        // apply H_i^H to the work-group's rows (rows >= i); column i becomes beta / v_i
        const c64 bq = mk((double)s_beta, 0.0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = rbase + lane + 64 * u;
            if (r >= m || r < i) continue;
            const c64 vq = (r == i) ? mk(1.0, 0.0) : scl * sx[lane + 64 * u];
            const c64 ct = conj(tq) * vq;
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) {
                const int k = wv * CPW + kk;
                if (k > i) a[u][kk] = a[u][kk] - ct * svha[k];
                else if (k == i) a[u][kk] = (r == i) ? bq : cast<double>(cast<T>(vq));   // v as stored
            }
        }
        __syncthreads();
        if (i + 1 < nbp && wv == (i + 1) / CPW) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int kk = 0; kk < CPW; ++kk)
                    if (kk == (i + 1) % CPW) sx[lane + 64 * u] = a[u][kk];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        if (r < m)
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) P[(long)(wv * CPW + kk) * ld + r] = cast<T>(a[u][kk]);
    }
    if (blk == 0 && tid < B && tid >= nbp) tau[tid] = mk(0.0, 0.0);
    // Gram partials V^H V of the stored reflectors (unit diagonal, zero above) over this
    // work-group's rows, 64 rows at a time: thread (k1 = tid % 32, k2 = tid / 32 + 8 v)
    c64 g[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) g[v] = mk(0.0, 0.0);
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) {
            const int k = wv * CPW + kk;
            c64 vv = mk(0.0, 0.0);
            if (r < m && k < nbp) vv = r == k ? mk(1.0, 0.0) : (r > k ? a[u][kk] : mk(0.0, 0.0));
            sV[lane][k] = vv;
        }
        __syncthreads();
        const int k1 = tid & (B - 1);
        for (int rr = 0; rr < 64; ++rr) {
            const c64 v1 = conj(sV[rr][k1]);
#pragma unroll
            for (int v = 0; v < 4; ++v) fma_(g[v], v1, sV[rr][(tid >> 5) + 8 * v]);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) gpart[(long)blk * B * B + ((tid >> 5) + 8 * v) * B + (tid & (B - 1))] = g[v];
}

// zlarfg on (alpha, x) given alpha and |x|^2: H^H (alpha; x) = (beta; 0) with H = I - tau v v^H,
// v = (1; scl x), beta real with the sign opposite to Re(alpha)
__device__ inline void larfg(c64 alpha, double xn2, c64 &tq, c64 &scl, double &beta) {
    if (xn2 == 0.0 && alpha.im == 0.0) {
        tq = mk(0.0, 0.0);
        beta = alpha.re;
        scl = mk(0.0, 0.0);
    } else {
        const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
        beta = alpha.re >= 0.0 ? -nrm : nrm;
        tq = mk((beta - alpha.re) / beta, -alpha.im / beta);
        scl = div_(mk(1.0, 0.0), alpha - mk(beta, 0.0));
    }
}

// Persistent panel QR: every work-group keeps its 256 panel rows in registers for all columns.
// Columns are factorised two at a time with one grid-wide barrier per pair.  Besides the partial
// sums of column i (its squared norm below the pivot and its inner products with the later
// columns), every work-group publishes its share of the Gram row of column i+1 (inner products of
// column i+1 with the later columns over rows > i), taken before H_i is applied, and the pivot rows
// i and i+1 are published as well.  H_i is unitary on rows >= i, so that Gram row over rows >= i is
// the same after H_i; subtracting the contributions of rows i and i+1 after H_i gives column i+1's
// squared norm below its pivot and its inner products with the later columns below row i+1: all
// that zlarfg and the update need for H_{i+1}.  The downdate cancels when column i+1 has little
// weight below row i+1: a pair whose downdated norm falls below kPairGuard times the Gram diagonal
// factorises column i alone (the next step starts at column i+1), so the relative error of the
// downdated norm stays within 1/kPairGuard ulps (1e4 FP64 ulps of the accumulated Gram sums, far
// below the rounding of the FP32-stored panel).  FP32 only: FP64 panels use k_pqr_panel1.  Needs all
// work-groups resident (cooperative launch).
// part / pivb are double-buffered by step parity: a work-group can only write the buffers of step
// e+1 after every work-group passed the barrier of step e, i.e. finished reading those of step e-1.
// gb: grid-barrier state (zero at launch).  part: 2 x nblk x 2B, pivb: 2 x 2B.
template <class T>
__global__ __launch_bounds__(256) void k_pqr_panel(cpx<T> *P, long ld, int m, int nbp, c64 *part, c64 *pivb, c64 *gpart,
                                                   c64 *tau, int *sync, int *gb) {
    static_assert(std::is_same<T, float>::value, "the paired panel step is the FP32 path's");
    constexpr double kPairGuard = 1e-4;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int blk = blockIdx.x, nblk = gridDim.x;
    const int rbase = blk * 256;
    __shared__ c64 sx[256], sy[256];   // columns i and i+1 before the step
    __shared__ c64 red[4][2 * B];
    __shared__ c64 ssum[2 * B], spiv[2 * B], svha[B], svha2[B], sri[B], sr1[B];
    __shared__ c64 sV[64][B + 1];
    __shared__ c64 s_scl, s_tau, s_scl2, s_tau2;
    __shared__ double s_beta, s_beta2;
    __shared__ int s_two;
    c64 a[4][CPW];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) a[u][kk] = r < m ? ld64(&P[(long)(wv * CPW + kk) * ld + r]) : mk(0.0, 0.0);
    }
    if (wv == 0)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sx[lane + 64 * u] = a[u][0];
            sy[lane + 64 * u] = a[u][1];
        }
    __syncthreads();
    // Column tests are made on wave-uniform scalars (this wave's first column, the step's first
    // column, the pair flag) so that they compile to scalar branches; row cases are selects.
    const int kb = __builtin_amdgcn_readfirstlane(wv * CPW);
    int i = 0;
    for (int ep = 1; i < nbp; ++ep) {
        const bool two = i + 1 < nbp;
        c64 *pt = part + (long)(ep & 1) * nblk * 2 * B;
        c64 *pv = pivb + (ep & 1) * 2 * B;
        // pivot rows i and i+1 (lanes i, i+1 of work-group 0) are published
        if (blk == 0 && (lane == i || lane == i + 1) && lane < m)
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) st_coh(&pv[(lane - i) * B + kb + kk], a[0][kk]);
        // partials over rows r > i of column i and column i+1 with every column (only k >= i and,
        // for column i+1, k > i are used; the k == i term of column i is its squared norm)
        c64 acc[CPW], acc2[CPW];
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) acc[kk] = acc2[kk] = mk(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = rbase + lane + 64 * u;
            const bool act = r > i && r < m;
            const c64 x = act ? sx[lane + 64 * u] : mk(0.0, 0.0), y = act ? sy[lane + 64 * u] : mk(0.0, 0.0);
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) {
                acc[kk] = acc[kk] + cmul(x, a[u][kk]);
                acc2[kk] = acc2[kk] + cmul(y, a[u][kk]);
            }
        }
        {
            // the 32 wave sums (re, im of acc and acc2), one per lane, stored by lanes 0..31
            double vs[32];
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) {
                vs[2 * kk] = acc[kk].re;
                vs[2 * kk + 1] = acc[kk].im;
                vs[16 + 2 * kk] = acc2[kk].re;
                vs[17 + 2 * kk] = acc2[kk].im;
            }
            const double tot = wave_sum32_scatter(vs);
            const int j = scatter32_index(lane), c = j >> 1;
            if (lane < 32 && (c < CPW || two))
                st_coh_d(reinterpret_cast<double *>(&pt[(long)blk * 2 * B + (c < CPW ? kb + c : B + kb + c - CPW)]) + (j & 1), tot);
        }
        if (!grid_flags_sync(gb, ep, sync)) return;
        {
            const int k2 = tid & (2 * B - 1), g = tid >> 6;
            const c64 pivv = ld_coh(&pv[k2]);   // issued first: its latency overlaps the partial sums
            red[g][k2] = (two || k2 < B) ? sum_coh_wide<24>(pt + k2, 2 * B, g, 4, nblk) : mk(0.0, 0.0);
            if (tid < 2 * B) spiv[tid] = pivv;
        }
        __syncthreads();
        if (tid < 2 * B) ssum[tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        __syncthreads();
        if (tid == 0) {
            c64 tq, scl;
            double beta;
            larfg(spiv[i], ssum[i].re, tq, scl, beta);
            s_scl = scl;
            s_tau = tq;
            s_beta = beta;
            if (blk == 0) tau[i] = tq;
        }
        __syncthreads();
        const c64 scl = s_scl, tq = s_tau;
        if (tid < B) {
            const c64 vh = tid > i ? spiv[tid] + conj(scl) * ssum[tid] : mk(0.0, 0.0);
            svha[tid] = vh;
            // rows i and i+1 after H_i
            sri[tid] = spiv[tid] - conj(tq) * vh;
            sr1[tid] = spiv[B + tid] - conj(tq) * (scl * spiv[B + i]) * vh;
        }
        __syncthreads();
        if (tid == 0) {
            bool pr = false;
            c64 tq2 = mk(0.0, 0.0), scl2 = mk(0.0, 0.0);
            double beta2 = 0.0;
            if (two) {
                // |column i+1|^2 over rows >= i (unchanged by H_i), less rows i and i+1 after H_i
                const c64 ri = sri[i + 1], alpha2 = sr1[i + 1];
                const double G = ssum[B + i + 1].re + norm2(spiv[i + 1]);
                const double xn2 = G - norm2(ri) - norm2(alpha2);
                pr = xn2 >= kPairGuard * G;
                if (pr) {
                    larfg(alpha2, xn2, tq2, scl2, beta2);
                    if (blk == 0) tau[i + 1] = tq2;
                }
            }
            s_two = pr;
            s_scl2 = scl2;
            s_tau2 = tq2;
            s_beta2 = beta2;
        }
        __syncthreads();
        const bool pair = __builtin_amdgcn_readfirstlane(s_two) != 0;
        const c64 scl2 = s_scl2, tq2 = s_tau2;
        if (tid < B) {
            c64 vh2 = mk(0.0, 0.0);
            if (pair && tid > i + 1) {
                // Gram row of column i+1 over rows > i+1 after H_i
                const c64 ri = sri[i + 1], alpha2 = sr1[i + 1];
                const c64 gk = ssum[B + tid] + conj(spiv[i + 1]) * spiv[tid];
                const c64 s2 = gk - conj(ri) * sri[tid] - conj(alpha2) * sr1[tid];
                vh2 = sr1[tid] + conj(scl2) * s2;
            }
            svha2[tid] = vh2;
        }
        __syncthreads();
        // apply H_i^H (rows >= i) and H_{i+1}^H (rows >= i+1): per row the coefficients ct, ct2 (zero
        // on rows the reflector does not touch); column i becomes beta / v_i and, for a pair,
        // column i+1 becomes (R(i, i+1); beta_2 / v_{i+1})
        const c64 bq = mk((double)s_beta, 0.0), bq2 = mk((double)s_beta2, 0.0);
        const c64 vh1 = svha[i + 1 < B ? i + 1 : i];
        const int nxt = i + (pair ? 2 : 1);
        c64 ct[4], ct2[4], vq[4], vq2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = rbase + lane + 64 * u;
            const bool act = r >= i && r < m;
            vq[u] = (r == i) ? mk(1.0, 0.0) : scl * sx[lane + 64 * u];
            ct[u] = act ? conj(tq) * vq[u] : mk(0.0, 0.0);
            vq2[u] = (r == i + 1) ? mk(1.0, 0.0) : scl2 * (sy[lane + 64 * u] - ct[u] * vh1);
            if (!(pair && r > i && r < m)) vq2[u] = mk(0.0, 0.0);
            ct2[u] = conj(tq2) * vq2[u];
        }
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) {
            const int k = kb + kk;
            if (k >= nxt) {
                const c64 h = svha[k], h2 = svha2[k];
#pragma unroll
                for (int u = 0; u < 4; ++u) a[u][kk] = a[u][kk] - ct[u] * h - ct2[u] * h2;
            } else if (k == i) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = rbase + lane + 64 * u;
                    a[u][kk] = (r == i) ? bq : ((r > i && r < m) ? cast<double>(cast<T>(scl * sx[lane + 64 * u])) : a[u][kk]);
                }
            } else if (k == i + 1) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = rbase + lane + 64 * u;
                    const c64 y = a[u][kk] - ct[u] * vh1;
                    a[u][kk] = (r == i + 1) ? bq2 : ((r > i + 1 && r < m) ? cast<double>(cast<T>(scl2 * y)) : y);
                }
            }
        }
        __syncthreads();
        i = nxt;
        if (i < nbp) {
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) {
                const int k = kb + kk;
                if (k == i) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) sx[lane + 64 * u] = a[u][kk];
                } else if (k == i + 1) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) sy[lane + 64 * u] = a[u][kk];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        if (r < m)
#pragma unroll
            for (int kk = 0; kk < CPW; ++kk) P[(long)(wv * CPW + kk) * ld + r] = cast<T>(a[u][kk]);
    }
    if (blk == 0 && tid < B && tid >= nbp) tau[tid] = mk(0.0, 0.0);
    // Gram partials V^H V of the stored reflectors (unit diagonal, zero above) over this
    // work-group's rows, 64 rows at a time: thread (k1 = tid % 32, k2 = tid / 32 + 8 v)
    c64 g[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) g[v] = mk(0.0, 0.0);
    for (int u = 0; u < 4; ++u) {
        const int r = rbase + lane + 64 * u;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CPW; ++kk) {
            const int k = wv * CPW + kk;
            c64 vv = mk(0.0, 0.0);
            if (r < m && k < nbp) vv = r == k ? mk(1.0, 0.0) : (r > k ? a[u][kk] : mk(0.0, 0.0));
            sV[lane][k] = vv;
        }
        __syncthreads();
        const int k1 = tid & (B - 1);
        for (int rr = 0; rr < 64; ++rr) {
            const c64 v1 = conj(sV[rr][k1]);
#pragma unroll
            for (int v = 0; v < 4; ++v) fma_(g[v], v1, sV[rr][(tid >> 5) + 8 * v]);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) gpart[(long)blk * B * B + ((tid >> 5) + 8 * v) * B + (tid & (B - 1))] = g[v];
}

// The panel kernel of each precision: FP64 uses the single-column kernel sized to run beside the
// trailing update, FP32 the paired one.
template <class T>
const void *panel_kernel() {
    if constexpr (std::is_same<T, double>::value) return (const void *)k_pqr_panel1<T>;
    else return (const void *)k_pqr_panel<T>;
}

// T factor of the panel (zlarft 'F', 'C') from the summed Gram partials of the stored V.
// Column i: T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).  The tau are staged in LDS first (no global load
// inside the column loop); row r of the mat-vec is split over 8 lanes (tid = 8 r + segment) and
// combined with an 8-lane DPP sum.
__global__ __launch_bounds__(256) void k_panel_tfactor(const c64 *gpart, int nblk, const c64 *tau, int nbp, c64 *Tp) {
    static_assert(B == 32, "32 rows x 8 segments");
    __shared__ c64 G[B][B + 1], Tm[B][B + 1], st[B];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < B * B; idx += 256) {
        c64 sacc = mk(0.0, 0.0);
        for (int b2 = 0; b2 < nblk; ++b2) sacc = sacc + gpart[(long)b2 * B * B + idx];
        G[idx % B][idx / B] = sacc;
        Tm[idx % B][idx / B] = mk(0.0, 0.0);
    }
    if (tid < B) st[tid] = tau[tid];
    __syncthreads();
    const int r = tid >> 3, sg = tid & 7;
    for (int i = 0; i < nbp; ++i) {
        c64 sm = mk(0.0, 0.0);
        for (int q2 = r + sg; q2 < i; q2 += 8) sm = sm + Tm[r][q2] * G[q2][i];
        sm.re = sum8(sm.re);
        sm.im = sum8(sm.im);
        if (sg == 0 && r < i) Tm[r][i] = -(st[i] * sm);
        if (tid == 0) Tm[i][i] = st[i];
        __syncthreads();
    }
    for (int idx = tid; idx < B * B; idx += 256) Tp[idx] = Tm[idx % B][idx / B];
}

template <class T>
__global__ __launch_bounds__(256) void k_pqr_finish(const cpx<T> *P, long ld, int m, int nbp, int t0, cpx<T> *V,
                                                    c64 *tau, c64 *Tp) {
    const int tid = threadIdx.x;
    const int r = blockIdx.x * 256 + tid;
    if (r < m) {
        // all 32 columns of this row loaded first (the stored entries at every row, masked after)
        cpx<T> pv[B];
#pragma unroll
        for (int k = 0; k < B; ++k) pv[k] = P[(long)k * ld + r];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            cpx<T> v = mk<T>(0, 0);
            if (k < nbp) v = r == k ? mk<T>(1, 0) : (r > k ? pv[k] : mk<T>(0, 0));
            V[(long)k * ld + t0 + r] = v;
        }
    }
    if (blockIdx.x == 0) {
        if (tid < B && tid >= nbp) tau[tid] = mk(0.0, 0.0);
        for (int idx = tid; idx < B * B; idx += 256)
            if (idx / B >= nbp) Tp[idx] = mk(0.0, 0.0);
    }
}

// ------------------------------------------------------------------------------------------
// Trailing update
// ------------------------------------------------------------------------------------------
// Xpart[seg] (rows of block I) = sum over J in the segment of A22(I, J) V(J).  The columns of
// the J tiles (the K dimension) stream through double-buffered LDS in chunks of kKC; the MFMAs
// compute X^T, so a lane's accumulator entries are 16 consecutive rows and the stores coalesce.
constexpr int kKC = 16;
template <class T>
__global__ __launch_bounds__(256) void k_hemm(const cpx<T> *A, long ld, int n, int t0, const cpx<T> *V,
                                              cpx<T> *Xpart, int per) {
    typedef typename VecT<T>::v4 v4;
    constexpr int CPT = 64 / kKC;               // chunks per tile
    const int I0 = t0 >> 6, ntb = (n + 63) >> 6;
    const int I = I0 + blockIdx.x, seg = blockIdx.y;
    const int Jb = I0 + seg * per, Je = min(ntb, Jb + per);
    const int nch = (Je - Jb) * CPT;
    __shared__ T sAr[2][64][kKC + 1], sAi[2][64][kKC + 1];   // [buf][row r of I][k]
    __shared__ T sVr[2][B][kKC + 1], sVi[2][B][kKC + 1];     // [buf][column c][k]
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    cpx<T> ra[4], rv[2];
    // entry u of this thread in chunk ch: (r, k) of the A22 tile and (c, k) of V
    auto a_index = [&](int J, int u, int &r, int &k) {
        const int idx = tid + 256 * u;
        if (I < J) { k = idx & (kKC - 1); r = idx / kKC; }   // read along the stored rows of tile (J, I)
        else { r = idx & 63; k = idx >> 6; }
    };
    // every load is issued unconditionally (addresses clamped to element 0 outside the matrix):
    // a load inside a divergent branch would be waited for inside that branch.  Conjugation and
    // masking happen when the chunk is stored to LDS.
    auto load_chunk = [&](int ch) {
        const int J = Jb + ch / CPT, kc = (ch % CPT) * kKC;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int r, k;
            a_index(J, u, r, k);
            const int row = I * 64 + r, col = J * 64 + kc + k;
            long off = row >= col ? (long)col * ld + row : (long)row * ld + col;
            ra[u] = A[(row < n && col < n) ? off : 0];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int idx = tid + 256 * u, k = idx & (kKC - 1), c = idx / kKC;
            rv[u] = V[(long)c * ld + min(J * 64 + kc + k, n - 1)];
        }
    };
    auto store_chunk = [&](int buf, int ch) {
        const int J = Jb + ch / CPT, kc = (ch % CPT) * kKC;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int r, k;
            a_index(J, u, r, k);
            const int row = I * 64 + r, col = J * 64 + kc + k;
            const bool ok = row < n && col < n;
            const T im = row > col ? ra[u].im : (row < col ? -ra[u].im : (T)0);
            sAr[buf][r][k] = ok ? ra[u].re : (T)0;
            sAi[buf][r][k] = ok ? im : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int idx = tid + 256 * u, k = idx & (kKC - 1), c = idx / kKC;
            const int row = J * 64 + kc + k;
            const bool ok = row >= t0 && row < n;
            sVr[buf][c][k] = ok ? rv[u].re : (T)0;
            sVi[buf][c][k] = ok ? rv[u].im : (T)0;
        }
    };
    v4 accR[2], accI[2];
    // FP32 accumulators are flushed into FP64 every kFlush chunks (one 64-column tile of K),
    // which keeps the rounding error of the long K sums at the level of a 64-term sum
    constexpr int kFlush = 64 / kKC;
    double dR[2][4], dI[2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        accR[m] = v4{0, 0, 0, 0};
        accI[m] = v4{0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 4; ++e) dR[m][e] = dI[m][e] = 0.0;
    }
    if (nch > 0) {
        load_chunk(0);
        store_chunk(0, 0);
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int b = ch & 1;
        if (ch + 1 < nch) load_chunk(ch + 1);
#pragma unroll
        for (int kk = 0; kk < kKC; kk += 4) {
            const T ar = sAr[b][16 * wv + li][kk + lk], ai = sAi[b][16 * wv + li][kk + lk];   // B operand (k x r)
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const T vr = sVr[b][16 * m + li][kk + lk], vi = sVi[b][16 * m + li][kk + lk];  // A operand (c x k)
                accR[m] = mfma(vr, ar, accR[m]);
                accR[m] = mfma(-vi, ai, accR[m]);
                accI[m] = mfma(vr, ai, accI[m]);
                accI[m] = mfma(vi, ar, accI[m]);
            }
        }
        if (ch + 1 < nch) store_chunk(b ^ 1, ch + 1);
        if ((ch + 1) % kFlush == 0) {
#pragma unroll
            for (int m = 0; m < 2; ++m) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    dR[m][e] += (double)accR[m][e];
                    dI[m][e] += (double)accI[m][e];
                }
                accR[m] = v4{0, 0, 0, 0};
                accI[m] = v4{0, 0, 0, 0};
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = 16 * m + acc_row<T>(lk, e);
            const long row = (long)I * 64 + 16 * wv + li;
            Xpart[((long)seg * B + c) * ld + row] = mk<T>((T)(dR[m][e] + (double)accR[m][e]), (T)(dI[m][e] + (double)accI[m][e]));
        }
}

// FP32 variant of k_hemm on 32x32x2 MFMAs (the 16x16x4 FP32 form issues at most ~85 % of the
// matrix-core rate).  Work-group: rows of block I x the kBW columns of V; wavefront w owns rows
// 32 (w & 1) .. +31 and the K half (w >> 1) of every chunk; the two halves are added at the end.
// MFMA step j of an 8-column group: lane half h supplies k = 8 g + 4 h + j, so each lane reads its
// four k values of a row as one float4 (LDS rows padded by 4 floats).  Staging: element u of a
// thread is a VGPR base offset plus u times a uniform stride; interior tiles take an unchecked
// path, the diagonal tile and tiles at n the checked one.  Accuracy: the FP32 accumulators hold
// 64-term blocks of the K sum (per wavefront), added into FP32 running sums after each block and
// converted to FP64 for the final sum of the halves (error bound ~ (64 + K / 64) eps).
constexpr int kKC32 = 16;              // K columns per LDS chunk
constexpr int kLDK32 = kKC32 + 4;      // padded LDS row
constexpr int kBlk32 = 64;             // K terms per wavefront between FP32 block sums
__global__ __launch_bounds__(256) void k_hemm_f32(const c32 *A, long ld, int n, int t0, const c32 *V,
                                                 c32 *Xpart, int per) {
    typedef float T;
    constexpr int KC = kKC32, LDK = kLDK32, CPT = 64 / KC;
    constexpr int NA = 64 * KC / 256, NV = B * KC / 256;
    constexpr int KH = KC / 2;                  // k per wavefront per chunk
    constexpr int RU = 256 / KC;                // rows advanced per staging element (transposed)
    const int I0 = t0 >> 6, ntb = (n + 63) >> 6;
    const int I = I0 + blockIdx.x, seg = blockIdx.y;
    const int Jb = I0 + seg * per, Je = min(ntb, Jb + per);
    const int nch = (Je - Jb) * CPT;
    constexpr int kTiles = 2 * 2 * 64 * LDK + 2 * 2 * B * LDK, kRed = 2 * 16 * 2 * 64 * 2;   // floats
    __shared__ __attribute__((aligned(16))) T smem[kTiles > kRed ? kTiles : kRed];
    constexpr int oAr = 0, oAi = 2 * 64 * LDK, oVr = 4 * 64 * LDK, oVi = 4 * 64 * LDK + 2 * B * LDK;
    constexpr int sA = 64 * LDK, sV = B * LDK;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, l32 = lane & 31, h = lane >> 5;
    const int rt = wv & 1, kh = wv >> 1;
    const bool nfull = (n & 63) == 0;
    const int ldi = (int)ld;
    // staging offsets: stored orientation (tile below the diagonal): r = lane, k = 4 wv + u, written
    // to LDS as one float4 per part (conflict-free: consecutive rows are LDK apart);
    // transposed (tile above) and V: k = tid % KC, r (or c) = tid / KC + RU u
    static_assert(NA == 4, "one float4 of k per thread and part");
    const int oN = (4 * wv) * ldi + lane, lN = lane * LDK + 4 * wv;
    const int oT = (tid / KC) * ldi + (tid & (KC - 1)), lT = (tid / KC) * LDK + (tid & (KC - 1));
    c32 ra[NA], rv[NV];
    auto load_chunk = [&](int ch) {
        const int J = Jb + ch / CPT, kc = (ch % CPT) * KC;
        const bool edge = !nfull && (I == ntb - 1 || J == ntb - 1);
        if (I > J && !edge) {
            const c32 *base = A + (long)(J * 64 + kc) * ld + I * 64 + oN;
#pragma unroll
            for (int u = 0; u < NA; ++u) ra[u] = base[(long)u * ld];
        } else if (I < J && !edge) {
            const c32 *base = A + (long)(I * 64) * ld + J * 64 + kc + oT;
#pragma unroll
            for (int u = 0; u < NA; ++u) ra[u] = base[(long)(RU * u) * ld];
        } else {
            const bool tr = I < J;
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                const int idx = tid + 256 * u;
                const int r = tr ? idx / KC : (idx & 63), k = tr ? (idx & (KC - 1)) : (idx >> 6);
                const int row = I * 64 + r, col = J * 64 + kc + k;
                const long off = row >= col ? (long)col * ld + row : (long)row * ld + col;
                ra[u] = A[(row < n && col < n) ? off : 0];
            }
        }
        const int r0 = J * 64 + kc;
        if (r0 >= t0 && r0 + KC <= n) {
            const c32 *base = V + r0 + oT;
#pragma unroll
            for (int u = 0; u < NV; ++u) rv[u] = base[(long)(RU * u) * ld];
        } else {
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const int idx = tid + 256 * u, k = idx & (KC - 1), c = idx / KC;
                rv[u] = V[(long)c * ld + min(r0 + k, n - 1)];
            }
        }
    };
    auto store_chunk = [&](int buf, int ch) {
        const int J = Jb + ch / CPT, kc = (ch % CPT) * KC;
        const bool edge = !nfull && (I == ntb - 1 || J == ntb - 1);
        T *pAr = smem + oAr + buf * sA, *pAi = smem + oAi + buf * sA;
        if (I > J && !edge) {
            *reinterpret_cast<f32x4 *>(pAr + lN) = f32x4{ra[0].re, ra[1].re, ra[2].re, ra[3].re};
            *reinterpret_cast<f32x4 *>(pAi + lN) = f32x4{ra[0].im, ra[1].im, ra[2].im, ra[3].im};
        } else if (I < J && !edge) {
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                pAr[lT + RU * LDK * u] = ra[u].re;
                pAi[lT + RU * LDK * u] = -ra[u].im;
            }
        } else {
            const bool tr = I < J;
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                const int idx = tid + 256 * u;
                const int r = tr ? idx / KC : (idx & 63), k = tr ? (idx & (KC - 1)) : (idx >> 6);
                const int row = I * 64 + r, col = J * 64 + kc + k;
                const bool ok = row < n && col < n;
                const T im = row > col ? ra[u].im : (row < col ? -ra[u].im : (T)0);
                pAr[r * LDK + k] = ok ? ra[u].re : (T)0;
                pAi[r * LDK + k] = ok ? im : (T)0;
            }
        }
        T *pVr = smem + oVr + buf * sV, *pVi = smem + oVi + buf * sV;
        const int r0 = J * 64 + kc;
        if (r0 >= t0 && r0 + KC <= n) {
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                pVr[lT + RU * LDK * u] = rv[u].re;
                pVi[lT + RU * LDK * u] = rv[u].im;
            }
        } else {
            const int k = tid & (KC - 1);
            const bool ok = r0 + k >= t0 && r0 + k < n;
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                pVr[lT + RU * LDK * u] = ok ? rv[u].re : (T)0;
                pVi[lT + RU * LDK * u] = ok ? rv[u].im : (T)0;
            }
        }
    };
    // the MFMAs compute X^T: A operand V(k, c = l % 32), B operand A22(r, k).  Complex products in
    // three real ones (Gauss): t1 = vr ar, t2 = vi ai, t3 = (vr + vi)(ar + ai); re = t1 - t2,
    // im = t3 - t1 - t2 (norm-wise error within a small factor of the four-product form)
    f32x16 t1, t2, t3;
    float sR[16], sI[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) { t1[e] = t2[e] = t3[e] = 0; sR[e] = sI[e] = 0; }
    const int aoff = (32 * rt + l32) * LDK + kh * KH + 4 * h, voff = l32 * LDK + kh * KH + 4 * h;
    if (nch > 0) {
        load_chunk(0);
        store_chunk(0, 0);
    }
    __syncthreads();
    constexpr int FP = kBlk32 / KH;    // chunks per block sum
    // nested loops: with the block sum inside the chunk loop the compiler moves the accumulators
    // between register files on every iteration
    for (int cb = 0; cb < nch; cb += FP) {
        const int ce = min(nch, cb + FP);
        for (int ch = cb; ch < ce; ++ch) {
            const int b = ch & 1;
            load_chunk(min(ch + 1, nch - 1));   // unconditional (the last chunk again at the end): no join before the wait
            const T *pAr = smem + oAr + b * sA + aoff, *pAi = smem + oAi + b * sA + aoff;
            const T *pVr = smem + oVr + b * sV + voff, *pVi = smem + oVi + b * sV + voff;
#pragma unroll
            for (int g = 0; g < KH / 8; ++g) {
                const f32x4 ar = *reinterpret_cast<const f32x4 *>(pAr + 8 * g);
                const f32x4 ai = *reinterpret_cast<const f32x4 *>(pAi + 8 * g);
                const f32x4 vr = *reinterpret_cast<const f32x4 *>(pVr + 8 * g);
                const f32x4 vi = *reinterpret_cast<const f32x4 *>(pVi + 8 * g);
                const f32x4 as = ar + ai, vs = vr + vi;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    t1 = mfma32(vr[j], ar[j], t1);
                    t2 = mfma32(vi[j], ai[j], t2);
                    t3 = mfma32(vs[j], as[j], t3);
                }
            }
            if (ch + 1 < nch) store_chunk(b ^ 1, ch + 1);
            __syncthreads();
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            sR[e] += t1[e] - t2[e];
            sI[e] += t3[e] - t1[e] - t2[e];
            t1[e] = t2[e] = t3[e] = 0;
        }
    }
    // K half 1 hands its sums to K half 0 through LDS (FP64)
    double *red = reinterpret_cast<double *>(smem);   // [rt][e][re/im][lane]
    if (kh == 1) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            red[((rt * 16 + e) * 2 + 0) * 64 + lane] = (double)sR[e];
            red[((rt * 16 + e) * 2 + 1) * 64 + lane] = (double)sI[e];
        }
    }
    __syncthreads();
    if (kh == 0) {
        const long row = (long)I * 64 + 32 * rt + l32;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const double xr = (double)sR[e] + red[((rt * 16 + e) * 2 + 0) * 64 + lane];
            const double xi = (double)sI[e] + red[((rt * 16 + e) * 2 + 1) * 64 + lane];
            Xpart[((long)seg * B + acc_row32(lane, e)) * ld + row] = mk<T>((T)xr, (T)xi);
        }
    }
}

// X = (sum_seg Xpart) T on rows of block I (zero outside t0 .. n-1); pm(I) = V_I^H X_I
template <class T>
__global__ __launch_bounds__(256) void k_xfin(const cpx<T> *Xpart, int S, long ld, int n, int t0, const c64 *Tp,
                                              const cpx<T> *V, cpx<T> *X, c64 *pm) {
    __shared__ c64 sX[64][B + 1];
    __shared__ c64 sT[B][B + 1];
    __shared__ cpx<T> sV[64][B + 1];   // V tile in its storage precision
    const int tid = threadIdx.x;
    const int I = (t0 >> 6) + blockIdx.x;
    constexpr int NE = 64 * B / 256;   // tile entries per thread: row idx & 63, column idx >> 6
    // the V tile and T are loaded first (unconditional, clamped, masked when stored), so that their
    // latency passes under the segment sums; the segments are summed four at a time (all their loads
    // in flight together)
    cpx<T> rv[NE];
    c64 rt[B * B / 256];
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const int idx = tid + 256 * u;
        rv[u] = V[(long)(idx >> 6) * ld + min(I * 64 + (idx & 63), n - 1)];
    }
#pragma unroll
    for (int u = 0; u < B * B / 256; ++u) rt[u] = Tp[tid + 256 * u];
    c64 acc[NE];
#pragma unroll
    for (int u = 0; u < NE; ++u) acc[u] = mk(0.0, 0.0);
    for (int sg0 = 0; sg0 < S; sg0 += 4) {
        cpx<T> xp[4][NE];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int u = 0; u < NE; ++u) {
                const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6;
                xp[q][u] = Xpart[((long)min(sg0 + q, S - 1) * B + cc) * ld + (long)I * 64 + rr];
            }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (sg0 + q < S)
#pragma unroll
                for (int u = 0; u < NE; ++u) acc[u] = acc[u] + cast<double>(xp[q][u]);
    }
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const int idx = tid + 256 * u, rw = I * 64 + (idx & 63);
        sX[idx & 63][idx >> 6] = acc[u];
        sV[idx & 63][idx >> 6] = (rw >= t0 && rw < n) ? rv[u] : mk<T>(0, 0);
    }
#pragma unroll
    for (int u = 0; u < B * B / 256; ++u) sT[(tid + 256 * u) % B][(tid + 256 * u) / B] = rt[u];
    __syncthreads();
    const int rr = tid & 63, cg = tid >> 6;
    const int row = I * 64 + rr;
    const bool valid = row >= t0 && row < n;
    c64 y[B / 4];
#pragma unroll
    for (int u = 0; u < B / 4; ++u) {
        const int c = cg + 4 * u;
        c64 s = mk(0.0, 0.0);
        for (int qq = 0; qq <= c; ++qq) fma_(s, sX[rr][qq], sT[qq][c]);   // T upper triangular
        y[u] = valid ? s : mk(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < B / 4; ++u) {
        const int c = cg + 4 * u;
        sX[rr][c] = y[u];
        if (row < (int)ld) X[(long)c * ld + row] = cast<T>(y[u]);
    }
    __syncthreads();
    // V^H X partial of this row block: thread (qv, c0 + 8 u)
    const int qv = tid & (B - 1), c0 = tid >> 5;
    c64 pa[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pa[u] = mk(0.0, 0.0);
    for (int r2 = 0; r2 < 64; ++r2) {
        const c64 v = cast<double>(sV[r2][qv]);
#pragma unroll
        for (int u = 0; u < 4; ++u) pa[u] = pa[u] + cmul(v, sX[r2][c0 + 8 * u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) pm[(long)blockIdx.x * B * B + (c0 + 8 * u) * B + qv] = pa[u];
}

// out = sum over count of the kBW x kBW blocks parts[b] (b < count).  Grid kBW work-groups:
// work-group g sums entries kBW g .. kBW g + kBW - 1, eight threads per entry with four
// independent accumulators each, then an LDS reduction.
__global__ __launch_bounds__(256) void k_sum_blocks(const c64 *parts, int count, c64 *out) {
    __shared__ c64 red[8][B];
    const int tid = threadIdx.x, e = blockIdx.x * B + (tid & (B - 1)), g = tid >> 5;
    c64 a0 = mk(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
    int b = g;
    for (; b + 24 < count; b += 32) {
        const c64 x0 = parts[(long)b * B * B + e], x1 = parts[(long)(b + 8) * B * B + e];
        const c64 x2 = parts[(long)(b + 16) * B * B + e], x3 = parts[(long)(b + 24) * B * B + e];
        a0 = a0 + x0;
        a1 = a1 + x1;
        a2 = a2 + x2;
        a3 = a3 + x3;
    }
    for (; b < count; b += 8) a0 = a0 + parts[(long)b * B * B + e];
    red[g][tid & (B - 1)] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (g == 0) {
        c64 t = red[0][tid];
#pragma unroll
        for (int q = 1; q < 8; ++q) t = t + red[q][tid];
        out[e] = t;
    }
}

// M = T^H (sum over row blocks of pm)
__global__ __launch_bounds__(256) void k_msum(const c64 *pm, int nI, const c64 *Tp, c64 *M) {
    __shared__ c64 sS[B][B + 1], sT[B][B + 1];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < B * B; idx += 256) {
        c64 a0 = mk(0.0, 0.0), a1 = mk(0.0, 0.0), a2 = mk(0.0, 0.0), a3 = mk(0.0, 0.0);
        int I = 0;
        for (; I + 3 < nI; I += 4) {
            a0 = a0 + pm[(long)I * B * B + idx];
            a1 = a1 + pm[(long)(I + 1) * B * B + idx];
            a2 = a2 + pm[(long)(I + 2) * B * B + idx];
            a3 = a3 + pm[(long)(I + 3) * B * B + idx];
        }
        for (; I < nI; ++I) a0 = a0 + pm[(long)I * B * B + idx];
        sS[idx % B][idx / B] = (a0 + a1) + (a2 + a3);
        sT[idx % B][idx / B] = Tp[idx];
    }
    __syncthreads();
    for (int idx = tid; idx < B * B; idx += 256) {
        const int qv = idx % B, c = idx / B;
        c64 s = mk(0.0, 0.0);
        for (int q2 = 0; q2 <= qv; ++q2) s = s + cmul(sT[q2][qv], sS[q2][c]);   // conj(T(q2, qv)) S(q2, c)
        M[idx] = s;
    }
}

// W = X - V M / 2 in place on rows >= t0 of the 64-row block (t0 >> 6) + blockIdx.x
template <class T>
__global__ __launch_bounds__(256) void k_wform(cpx<T> *X, const cpx<T> *V, const c64 *M, long ld, int n, int t0) {
    __shared__ c64 sM[B][B + 1];
    __shared__ c64 sV[64][B + 1];
    const int tid = threadIdx.x;
    const int I = (t0 >> 6) + blockIdx.x;
    for (int idx = tid; idx < B * B; idx += 256) sM[idx % B][idx / B] = M[idx];
    {
        c64 rv[64 * B / 256];
#pragma unroll
        for (int u = 0; u < 64 * B / 256; ++u) {
            const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6, row = I * 64 + rr;
            rv[u] = (row >= t0 && row < n) ? ld64(&V[(long)cc * ld + row]) : mk(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 64 * B / 256; ++u) {
            const int idx = tid + 256 * u;
            sV[idx & 63][idx >> 6] = rv[u];
        }
    }
    __syncthreads();
    const int rr = tid & 63, cg = tid >> 6;
    const int row = I * 64 + rr;
    if (row >= n) return;
#pragma unroll
    for (int u = 0; u < B / 4; ++u) {
        const int c = cg + 4 * u;
        c64 w = mk(0.0, 0.0);
        if (row >= t0) {
            c64 s = mk(0.0, 0.0);
            for (int qq = 0; qq < B; ++qq) fma_(s, sV[rr][qq], sM[qq][c]);
            w = ld64(&X[(long)c * ld + row]) - 0.5 * s;
        }
        X[(long)c * ld + row] = cast<T>(w);
    }
}

// A22 -= V W^H + W V^H on lower 64x64 tiles at or below row/column t0.  K = 2 kBW streams
// through double-buffered LDS in chunks of kKC; the MFMAs compute the transposed update, so a
// lane owns 16 consecutive rows of C: the C tile is prefetched into registers in the
// accumulator layout while the K loop runs, and the read-modify-write coalesces.
template <class T>
__global__ __launch_bounds__(256) void k_her2k_band(cpx<T> *A, long ld, int n, int t0, const cpx<T> *V,
                                                    const cpx<T> *W, int Jb, int strip) {
    typedef typename VecT<T>::v4 v4;
    // tiles: the column strip J = strip, I = strip + blockIdx.x (strip >= 0), or the lower triangle of
    // tiles I >= J >= Jb in row order (strip < 0)
    int I, J;
    if (strip >= 0) {
        J = strip;
        I = strip + (int)blockIdx.x;
    } else {
        const long t = blockIdx.x;
        long i = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        while (i * (i + 1) / 2 > t) --i;
        I = (int)i + Jb;
        J = (int)(t - i * (i + 1) / 2) + Jb;
    }
    constexpr int NCH = 2 * B / kKC;
    __shared__ T sAr[2][64][kKC + 1], sAi[2][64][kKC + 1];   // [V W] rows of I: [r][k]
    __shared__ T sBr[2][64][kKC + 1], sBi[2][64][kKC + 1];   // [W V] rows of J: [c][k]
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int c0 = 32 * (wv >> 1), r0 = 32 * (wv & 1);
    // prefetch of this wavefront's 32x32 part of C in the accumulator layout
    cpx<T> cp[2][2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int R = I * 64 + r0 + 16 * q + li, C = J * 64 + c0 + 16 * m + acc_row<T>(lk, e);
                const bool ok = R >= t0 && C >= t0 && R < n && C < n && R >= C;
                cp[m][q][e] = ok ? A[(long)C * ld + R] : mk<T>(0, 0);
            }
    cpx<T> ra[4], rb[4];
    auto load_chunk = [&](int ch) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = tid + 256 * u, rl = idx & 63, k = ch * kKC + (idx >> 6);
            const int R = I * 64 + rl, C = J * 64 + rl;
            cpx<T> a = mk<T>(0, 0), bb = mk<T>(0, 0);
            if (R >= t0 && R < n) a = k < B ? V[(long)k * ld + R] : W[(long)(k - B) * ld + R];
            if (C >= t0 && C < n) bb = k < B ? W[(long)k * ld + C] : V[(long)(k - B) * ld + C];
            ra[u] = a;
            rb[u] = bb;
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = tid + 256 * u, rl = idx & 63, k = idx >> 6;
            sAr[buf][rl][k] = ra[u].re;
            sAi[buf][rl][k] = ra[u].im;
            sBr[buf][rl][k] = rb[u].re;
            sBi[buf][rl][k] = rb[u].im;
        }
    };
    v4 accR[2][2], accI[2][2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            accR[m][q] = v4{0, 0, 0, 0};
            accI[m][q] = v4{0, 0, 0, 0};
        }
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        const int b = ch & 1;
        if (ch + 1 < NCH) load_chunk(ch + 1);
#pragma unroll
        for (int kk = 0; kk < kKC; kk += 4) {
            T ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {   // B operand (k x r): [V W](r, k)
                ar[q] = sAr[b][r0 + 16 * q + li][kk + lk];
                ai[q] = sAi[b][r0 + 16 * q + li][kk + lk];
            }
#pragma unroll
            for (int m = 0; m < 2; ++m) {   // A operand (c x k): [W V](c, k), conjugated below
                br[m] = sBr[b][c0 + 16 * m + li][kk + lk];
                bi[m] = sBi[b][c0 + 16 * m + li][kk + lk];
            }
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    accR[m][q] = mfma(br[m], ar[q], accR[m][q]);
                    accR[m][q] = mfma(bi[m], ai[q], accR[m][q]);
                    accI[m][q] = mfma(br[m], ai[q], accI[m][q]);
                    accI[m][q] = mfma(-bi[m], ar[q], accI[m][q]);
                }
        }
        if (ch + 1 < NCH) store_chunk(b ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int R = I * 64 + r0 + 16 * q + li, C = J * 64 + c0 + 16 * m + acc_row<T>(lk, e);
                if (R < t0 || C < t0 || R >= n || C >= n || R < C) continue;
                cpx<T> v = cp[m][q][e];
                v.re -= accR[m][q][e];
                v.im -= accI[m][q][e];
                if (R == C) v.im = 0;
                A[(long)C * ld + R] = v;
            }
}

// FP32 variant of k_her2k_band on 32x32x2 MFMAs.  A22 -= P Q^H with P = [V W] (rows R of tile I)
// and Q = [W V] (rows C of tile J).  The MFMAs compute the transposed update D(c, r): wavefront w
// owns c0 = 32 (w >> 1), r0 = 32 (w & 1); lane l holds r = r0 + l % 32 and the 16 columns
// c = c0 + acc_row32(l, e).  The C tile is loaded into registers first (unconditionally, from
// clamped addresses); the accumulators sum the update from zero (P stored negated) and C is added
// once at the end, so each rounding of the K sum is relative to the update, not to C.  K = 2 kBW
// passes through one LDS buffer in two halves of kBW (V then W for P, W then V for Q).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_her2k_f32(c32 *A, long ld, int n, int t0, const c32 *V, const c32 *W, int Jb, int strip) {
    typedef float T;
    constexpr int LDK = B + 4;
    // tiles: the column strip J = strip, I = strip + blockIdx.x (strip >= 0), or the lower triangle of
    // tiles I >= J >= Jb in row order (strip < 0)
    int I, J;
    if (strip >= 0) {
        J = strip;
        I = strip + (int)blockIdx.x;
    } else {
        const long t = blockIdx.x;
        long i = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        while (i * (i + 1) / 2 > t) --i;
        I = (int)i + Jb;
        J = (int)(t - i * (i + 1) / 2) + Jb;
    }
    __shared__ __attribute__((aligned(16))) T sPr[64][LDK], sPi[64][LDK], sQr[64][LDK], sQi[64][LDK];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, l32 = lane & 31, h = lane >> 5;
    const int c0 = 32 * (wv >> 1), r0 = 32 * (wv & 1);
    const int R = I * 64 + r0 + l32;
    c32 cv[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int C = J * 64 + c0 + acc_row32(lane, e);
        cv[e] = A[(long)min(C, n - 1) * ld + min(R, n - 1)];
    }
    // conj(q) p in three real products (Gauss): t1 = qr pr, t2 = qi pi, t3 = (qr - qi)(pr + pi);
    // re = t1 + t2, im = t3 - t1 + t2
    f32x16 t1, t2, t3;
#pragma unroll
    for (int e = 0; e < 16; ++e) t1[e] = t2[e] = t3[e] = 0;
    // staging of one half: element u of a thread is row lane, column 8 wv + u (< kBW); written to LDS
    // as two float4 per part (conflict-free: consecutive rows are LDK apart)
    const bool pin = I * 64 >= t0 && I * 64 + 64 <= n, qin = J * 64 >= t0 && J * 64 + 64 <= n;
    c32 rp[B / 4], rq[B / 4];
    auto load_half = [&](int hf) {
        const c32 *Pb = hf == 0 ? V : W, *Qb = hf == 0 ? W : V;
        const int RP = I * 64 + lane, RQ = J * 64 + lane;
        const c32 *p = Pb + (long)(8 * wv) * ld + min(RP, n - 1), *q = Qb + (long)(8 * wv) * ld + min(RQ, n - 1);
#pragma unroll
        for (int u = 0; u < B / 4; ++u) rp[u] = p[(long)u * ld];
#pragma unroll
        for (int u = 0; u < B / 4; ++u) rq[u] = q[(long)u * ld];
        if (!pin) {
            const bool ok = RP >= t0 && RP < n;
#pragma unroll
            for (int u = 0; u < B / 4; ++u) rp[u] = ok ? rp[u] : mk<T>(0, 0);
        }
        if (!qin) {
            const bool ok = RQ >= t0 && RQ < n;
#pragma unroll
            for (int u = 0; u < B / 4; ++u) rq[u] = ok ? rq[u] : mk<T>(0, 0);
        }
    };
    auto store_half = [&]() {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const int k0 = 8 * wv + 4 * g;
            *reinterpret_cast<f32x4 *>(&sPr[lane][k0]) = f32x4{-rp[4 * g].re, -rp[4 * g + 1].re, -rp[4 * g + 2].re, -rp[4 * g + 3].re};
            *reinterpret_cast<f32x4 *>(&sPi[lane][k0]) = f32x4{-rp[4 * g].im, -rp[4 * g + 1].im, -rp[4 * g + 2].im, -rp[4 * g + 3].im};
            *reinterpret_cast<f32x4 *>(&sQr[lane][k0]) = f32x4{rq[4 * g].re, rq[4 * g + 1].re, rq[4 * g + 2].re, rq[4 * g + 3].re};
            *reinterpret_cast<f32x4 *>(&sQi[lane][k0]) = f32x4{rq[4 * g].im, rq[4 * g + 1].im, rq[4 * g + 2].im, rq[4 * g + 3].im};
        }
    };
    // D(c, r) += conj(Q(c, k)) (-P(r, k)); MFMA step j of group g: k = 8 g + 4 h + j
    auto compute = [&]() {
#pragma unroll
        for (int g = 0; g < B / 8; ++g) {
            const int k0 = 8 * g + 4 * h;
            const f32x4 pr = *reinterpret_cast<const f32x4 *>(&sPr[r0 + l32][k0]);
            const f32x4 pi = *reinterpret_cast<const f32x4 *>(&sPi[r0 + l32][k0]);
            const f32x4 qr = *reinterpret_cast<const f32x4 *>(&sQr[c0 + l32][k0]);
            const f32x4 qi = *reinterpret_cast<const f32x4 *>(&sQi[c0 + l32][k0]);
            const f32x4 qs = qr - qi, ps = pr + pi;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                t1 = mfma32(qr[j], pr[j], t1);
                t2 = mfma32(qi[j], pi[j], t2);
                t3 = mfma32(qs[j], ps[j], t3);
            }
        }
    };
    load_half(0);
    store_half();
    load_half(1);
    __syncthreads();
    compute();
    __syncthreads();
    store_half();
    __syncthreads();
    compute();
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int C = J * 64 + c0 + acc_row32(lane, e);
        if (R >= t0 && C >= t0 && R < n && C < n && R >= C)
            A[(long)C * ld + R] = mk<T>(cv[e].re + (t1[e] + t2[e]), R == C ? (T)0 : cv[e].im + (t3[e] - t1[e] + t2[e]));
    }
}

// ------------------------------------------------------------------------------------------
// Application of Q1 = Q_0 Q_1 ... Q_{np-1} (the stage-1 panels, Q_p = I - V_p T_p V_p^H) to a
// block of k <= 64 vectors: X <- Q1 X (adjoint = false) or Q1^H X.  Persistent, cooperative: work-
// group w owns the 64-row blocks b = w, w + G, ... of X; per panel two grid barriers separate
//   (A) per work-group partials of Y = V_p^H X over its rows,
//   (C) Y' = T_p Y (or T_p^H Y) for column l by work-group l,
//   (D) X_rows -= V_rows Y'.
// gb: grid-barrier state (zeroed before the launch).  KU: vectors per thread in (A), k <= 8 KU; the
// products run for all KU (LDS columns >= k hold zeros) so that none sits in a per-lane branch.
// ------------------------------------------------------------------------------------------
template <class T, bool adjoint, int KU>
__global__ __launch_bounds__(256) void k_q1_apply(const cpx<T> *A, long ld, int n, int np, const c64 *Tall, c64 *X,
                                                  long ldx, int k, c64 *part, c64 *Yp, int *sync, int *gb) {
    static_assert(KU >= 1 && KU <= 8, "at most 64 vectors");
    const int G = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
    const int ntb = (n + 63) >> 6;
    __shared__ c64 sV[64][B + 1];
    __shared__ c64 sX[64][65];
    __shared__ c64 sY[B][65];
    __shared__ c64 red[8][B];
    const int c = tid & (B - 1), lg = tid >> 5;   // (reflector c, vectors lg + 8 u)
    int epoch = 0;
    auto grid_barrier = [&]() -> bool { return grid_flags_sync(gb, ++epoch, sync); };
    // all global loads below are unconditional (indices clamped into range, surplus masked after
    // the data arrives), so that a wavefront has every load of a phase in flight at once
    auto load_v = [&](int p0, int t0, int b) {
        c64 rv[64 * B / 256];
#pragma unroll
        for (int u = 0; u < 64 * B / 256; ++u) {
            const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6, r = b * 64 + rr;
            rv[u] = ld64(&A[(long)(p0 + cc) * ld + min(r, n - 1)]);
        }
#pragma unroll
        for (int u = 0; u < 64 * B / 256; ++u) {
            const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6, r = b * 64 + rr;
            sV[rr][cc] = (r >= t0 + cc && r < n) ? rv[u] : mk(0.0, 0.0);
        }
    };
    for (int pp = 0; pp < np; ++pp) {
        const int p = adjoint ? pp : np - 1 - pp;
        const int p0 = p * B, t0 = p0 + B;
        const int bfirst = t0 >> 6;
        const int b0 = bfirst + ((w - bfirst % G) % G + G) % G;   // first own block >= bfirst
        // (A) partials of Y = V^H X over own rows
        c64 acc[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) acc[u] = mk(0.0, 0.0);
        for (int b = b0; b < ntb; b += G) {
            {
                c64 rx[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int idx = tid + 256 * u, rr = idx & 63, l = idx >> 6, r = b * 64 + rr;
                    rx[u] = X[(long)min(l, k - 1) * ldx + min(r, n - 1)];
                }
                load_v(p0, t0, b);
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int idx = tid + 256 * u, r = b * 64 + (idx & 63);
                    sX[idx & 63][idx >> 6] = (r < n && (idx >> 6) < k) ? rx[u] : mk(0.0, 0.0);
                }
            }
            __syncthreads();
            for (int rr = 0; rr < 64; ++rr) {
                const c64 v = conj(sV[rr][c]);
#pragma unroll
                for (int u = 0; u < KU; ++u) fma_(acc[u], v, sX[rr][lg + 8 * u]);
            }
            __syncthreads();
        }
#pragma unroll
        for (int u = 0; u < KU; ++u)
            if (lg + 8 * u < k) st_coh(&part[((long)w * 64 + lg + 8 * u) * B + c], acc[u]);   // [w][l][c]
        if (!grid_barrier()) return;
        // (C) columns l = w, w + G, ...: Y = sum of the partials, Y' = T Y (T^H Y)
        for (int l = w; l < k; l += G) {
            red[lg][c] = sum_coh_wide<16>(part + (long)l * B + c, (long)64 * B, lg, 8, G);
            __syncthreads();
            {   // T of the panel into LDS (sX is free here): sX[row][col] = T(row, col)
                const c64 *Tm = Tall + (long)p * B * B;
                c64 rt[B * B / 256];
#pragma unroll
                for (int u = 0; u < B * B / 256; ++u) rt[u] = Tm[tid + 256 * u];
#pragma unroll
                for (int u = 0; u < B * B / 256; ++u) sX[(tid + 256 * u) % B][(tid + 256 * u) / B] = rt[u];
            }
            if (tid < B) {
                c64 y = mk(0.0, 0.0);
#pragma unroll
                for (int g = 0; g < 8; ++g) y = y + red[g][tid];
                sY[tid][0] = y;
            }
            __syncthreads();
            if (tid < B) {
                c64 y = mk(0.0, 0.0);
                if (!adjoint) {
                    for (int c2 = tid; c2 < B; ++c2) y = y + sX[tid][c2] * sY[c2][0];
                } else {
                    for (int c2 = 0; c2 <= tid; ++c2) y = y + conj(sX[c2][tid]) * sY[c2][0];
                }
                st_coh(&Yp[tid * 64 + l], y);
            }
            __syncthreads();
        }
        if (!grid_barrier()) return;
        // (D) X_rows -= V_rows Y'
        {
            c64 ry[B * 64 / 256];
#pragma unroll
            for (int u = 0; u < B * 64 / 256; ++u) {
                const int idx = tid + 256 * u;
                ry[u] = idx < B * k ? ld_coh(&Yp[(idx % B) * 64 + idx / B]) : mk(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < B * 64 / 256; ++u) {
                const int idx = tid + 256 * u;
                sY[idx % B][idx / B] = idx < B * k ? ry[u] : mk(0.0, 0.0);
            }
        }
        for (int b = b0; b < ntb; b += G) {
            // the X entries of this thread (columns l = tid / 64 + 4 u < 8 KU) are loaded before the
            // arithmetic; the products run for all of them, the stores only for l < k
            constexpr int kXU = 2 * KU;
            c64 xo[kXU];
#pragma unroll
            for (int u = 0; u < kXU; ++u) {
                const int idx = tid + 256 * u, rr = idx & 63, l = idx >> 6, r = b * 64 + rr;
                xo[u] = X[(long)min(l, k - 1) * ldx + min(r, n - 1)];
            }
            load_v(p0, t0, b);
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kXU; ++u) {
                const int idx = tid + 256 * u, rr = idx & 63, l = idx >> 6, r = b * 64 + rr;
                c64 sacc = mk(0.0, 0.0);
#pragma unroll 8
                for (int cc = 0; cc < B; ++cc) fma_(sacc, sV[rr][cc], sY[cc][l]);
                if (l < k && r < n) X[(long)l * ldx + r] = xo[u] - sacc;
            }
            __syncthreads();
        }
        __syncthreads();
    }
}

// the Q1 application variants by KU (1, 2, 4, 5, 8), each forward then adjoint
template <class T>
const std::array<const void *, 10> &q1_kernels() {
    static const std::array<const void *, 10> f = {
        (const void *)k_q1_apply<T, false, 1>, (const void *)k_q1_apply<T, true, 1>,
        (const void *)k_q1_apply<T, false, 2>, (const void *)k_q1_apply<T, true, 2>,
        (const void *)k_q1_apply<T, false, 4>, (const void *)k_q1_apply<T, true, 4>,
        (const void *)k_q1_apply<T, false, 5>, (const void *)k_q1_apply<T, true, 5>,
        (const void *)k_q1_apply<T, false, 8>, (const void *)k_q1_apply<T, true, 8>};
    return f;
}

// ------------------------------------------------------------------------------------------
// Band extraction and unit entries
// ------------------------------------------------------------------------------------------
// AB(d, c) = A(c + d, c) for d <= kBW (zero for the bulge rows d > kBW); diagonal made real.
template <class T>
__global__ __launch_bounds__(256) void k_band_extract(const cpx<T> *A, long ld, int n, c64 *AB) {
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)n * 2 * B) return;
    const int c = (int)(idx / (2 * B)), d = (int)(idx % (2 * B));
    c64 v = mk(0.0, 0.0);
    if (d <= B && c + d < n) v = ld64(&A[(long)c * ld + c + d]);
    if (d == 0) v.im = 0.0;
    AB[idx] = v;
}

template <class T>
__global__ __launch_bounds__(256) void k_set_units(cpx<T> *A, long ld, int n) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < n - B - 1) A[(long)c * ld + c + B] = mk<T>(1, 0);
}

__global__ void k_fill_int_b(int *p, int count, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

__global__ void k_fill_c64(c64 *p, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = mk(0.0, 0.0);
}

}  // namespace

template <class T>
void band_alloc(BandWork<T> &w, int n, cpx<T> *A_external) {
    w.n = n;
    w.ld = ((long)n + 63) / 64 * 64;
    w.ntb = (n + 63) / 64;
    w.A = A_external;
    const long ld = w.ld;
    w.maxS = 32;
    QD_HIP(hipMalloc(&w.V, sizeof(cpx<T>) * ld * B * 2));
    QD_HIP(hipMalloc(&w.gsum, sizeof(c64) * B * B));
    QD_HIP(hipStreamCreateWithFlags(&w.s2, hipStreamNonBlocking));
    QD_HIP(hipEventCreateWithFlags(&w.ev_strip, hipEventDisableTiming));
    QD_HIP(hipEventCreateWithFlags(&w.ev_panel, hipEventDisableTiming));
    QD_HIP(hipMalloc(&w.X, sizeof(cpx<T>) * ld * B));
    QD_HIP(hipMalloc(&w.Xpart, sizeof(cpx<T>) * ld * B * w.maxS));
    QD_HIP(hipMalloc(&w.tau, sizeof(c64) * (n > 0 ? n : 1)));
    w.npan_max = std::max(1, (n + B - 1) / B);
    QD_HIP(hipMalloc(&w.Tall, sizeof(c64) * (long)w.npan_max * B * B));
    QD_HIP(hipMalloc(&w.qpart, sizeof(c64) * 256L * B * 64));
    QD_HIP(hipMalloc(&w.qy, sizeof(c64) * B * 64));
    QD_HIP(hipMalloc(&w.gbar, sizeof(int) * kGridBarrierInts));
    QD_HIP(hipMalloc(&w.part, sizeof(c64) * 4 * ((long)(n + 255) / 256 + 1) * B));
    QD_HIP(hipMalloc(&w.pm, sizeof(c64) * (long)(w.ntb + 1) * B * B));
    QD_HIP(hipMalloc(&w.piv, sizeof(c64) * 4 * B));
    QD_HIP(hipMalloc(&w.gpart, sizeof(c64) * ((long)(n + 255) / 256 + 1) * B * B));
    {   // largest panel grid that is resident at once (cooperative launch of the panel kernel)
        int dev = 0, per_cu = 0;
        hipDeviceProp_t p;
        QD_HIP(hipGetDevice(&dev));
        QD_HIP(hipGetDeviceProperties(&p, dev));
        QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, panel_kernel<T>(), 256, 0));
        w.coop_max = p.cooperativeLaunch ? per_cu * p.multiProcessorCount : 0;
        int per_cu_q = 1 << 30;
        for (const void *f : q1_kernels<T>()) {
            int pc = 0;
            QD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, f, 256, 0));
            per_cu_q = std::min(per_cu_q, pc);
        }
        w.q1_max = p.cooperativeLaunch ? per_cu_q * p.multiProcessorCount : 0;
    }
    QD_HIP(hipMalloc(&w.sc, sizeof(c64) * (4 + B)));
    QD_HIP(hipMalloc(&w.M, sizeof(c64) * B * B));
}

template <class T>
void band_free(BandWork<T> &w) {
    hipFree(w.V); hipFree(w.X); hipFree(w.Xpart); hipFree(w.tau); hipFree(w.Tall); hipFree(w.part); hipFree(w.qpart); hipFree(w.qy); hipFree(w.gbar);
    hipFree(w.pm); hipFree(w.piv); hipFree(w.sc); hipFree(w.M); hipFree(w.gpart); hipFree(w.gsum);
    if (w.ev_strip) hipEventDestroy(w.ev_strip);
    if (w.ev_panel) hipEventDestroy(w.ev_panel);
    if (w.s2) hipStreamDestroy(w.s2);
    w = BandWork<T>();
}

template <class T>
void band_reduce(BandWork<T> &w, c64 *AB, int *sync, hipStream_t s, KernelProfile *prof) {
    const int n = w.n;
    const long ld = w.ld;
    hipLaunchKernelGGL(k_fill_c64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w.tau, (long)n);
    QD_LAUNCH_CHECK();
    w.npan = 0;
    auto vbuf = [&](int p0) { return w.V + (long)((p0 / B) & 1) * ld * B; };   // explicit V, double-buffered
    // panel QR of the panel starting at column p0: reflectors in A, tau, T (Tall), explicit V
    auto panel_qr = [&](int p0, hipStream_t st) {
        const int t0 = p0 + B, m = n - t0, nbp = std::min(B, m - 1);
        c64 *Tp = w.Tall + (long)(p0 / B) * B * B;
        cpx<T> *P = w.A + (long)p0 * ld + t0;
        const int nblk = (m + 255) / 256;
        if (nblk <= w.coop_max && nblk <= 256) {
            c64 *tau_p = w.tau + p0;
            hipLaunchKernelGGL(k_fill_int_b, dim3((kGridBarrierInts + 255) / 256), dim3(256), 0, st, w.gbar, kGridBarrierInts, 0);
            QD_LAUNCH_CHECK();
            void *args[] = {&P, (void *)&ld, (void *)&m, (void *)&nbp, &w.part, &w.piv, &w.gpart, &tau_p, &sync, &w.gbar};
            QD_HIP(hipLaunchCooperativeKernel(panel_kernel<T>(), dim3(nblk), dim3(256), args, 0, st));
            hipLaunchKernelGGL(k_sum_blocks, dim3(B), dim3(256), 0, st, w.gpart, nblk, w.gsum);   // Gram V^H V
            QD_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_panel_tfactor, dim3(1), dim3(256), 0, st, w.gsum, 1, tau_p, nbp, Tp);
            QD_LAUNCH_CHECK();
        } else for (int i = 0; i <= nbp; ++i) {
            hipLaunchKernelGGL(k_pqr_step<T>, dim3(nblk), dim3(256), 0, st, P, ld, m, nbp, i, w.sc, w.part, w.piv);
            QD_LAUNCH_CHECK();
            if (i < nbp) {
                hipLaunchKernelGGL(k_pqr_reduce<T>, dim3(1), dim3(256), 0, st, w.part, nblk, w.piv, i, w.sc, Tp,
                                   w.tau + p0 + i);
                QD_LAUNCH_CHECK();
            }
        }
        hipLaunchKernelGGL(k_pqr_finish<T>, dim3(nblk), dim3(256), 0, st, P, ld, m, nbp, t0, vbuf(p0), w.tau + p0, Tp);
        QD_LAUNCH_CHECK();
    };
    // A22 -= V W^H + W V^H over tiles of column strip `strip` (strip >= 0) or the triangle I >= J >= Jb
    auto her2k = [&](int t0, const cpx<T> *V, int Jb, int strip, long ntiles, hipStream_t st) {
        if (ntiles <= 0) return;
        if constexpr (std::is_same<T, float>::value)
            hipLaunchKernelGGL(k_her2k_f32, dim3((unsigned)ntiles), dim3(256), 0, st, w.A, ld, n, t0, V, w.X, Jb, strip);
        else
            hipLaunchKernelGGL(k_her2k_band<T>, dim3((unsigned)ntiles), dim3(256), 0, st, w.A, ld, n, t0, V, w.X, Jb, strip);
        QD_LAUNCH_CHECK();
    };
    // Look-ahead overlaps the next panel's QR (a latency-bound cooperative kernel) with the trailing
    // update.  Measured on MI355X it removes most of the panel time in FP64; with the FP32 kernels the
    // two do not run concurrently and the split update only adds a launch, so FP32 keeps the serial
    // order.
    constexpr bool kLookAhead = std::is_same<T, double>::value;
    if (n - B >= 2) panel_qr(0, s);
    for (int p0 = 0; n - p0 - B >= 2; p0 += B) {
        const int t0 = p0 + B, m = n - t0;
        w.Tp = w.Tall + (long)(p0 / B) * B * B;   // this panel's T, kept for the back-transformation
        ++w.npan;
        const cpx<T> *V = vbuf(p0);
        const int I0 = t0 >> 6, nI = w.ntb - I0;
        int S = std::max(1, std::min(std::min(nI / 4, w.maxS), (2048 + nI - 1) / nI));
        const int per = (nI + S - 1) / S;
        S = (nI + per - 1) / per;
        const bool timed = prof && prof->on;
        if (timed) prof->begin(s);
        if constexpr (std::is_same<T, float>::value)
            hipLaunchKernelGGL(k_hemm_f32, dim3(nI, S), dim3(256), 0, s, w.A, ld, n, t0, V, w.Xpart, per);
        else
            hipLaunchKernelGGL(k_hemm<T>, dim3(nI, S), dim3(256), 0, s, w.A, ld, n, t0, V, w.Xpart, per);
        QD_LAUNCH_CHECK();
        if (timed) prof->end(s, 8.0 * m * m * B);   // complex m x m times m x B: 8 m^2 B flops
        hipLaunchKernelGGL(k_xfin<T>, dim3(nI), dim3(256), 0, s, w.Xpart, S, ld, n, t0, w.Tp, V, w.X, w.pm);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sum_blocks, dim3(B), dim3(256), 0, s, w.pm, nI, w.pm + (long)nI * B * B);   // V^H X
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_msum, dim3(1), dim3(256), 0, s, w.pm + (long)nI * B * B, 1, w.Tp, w.M);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_wform<T>, dim3(nI), dim3(256), 0, s, w.X, V, w.M, ld, n, t0);
        QD_LAUNCH_CHECK();
        const bool next = n - t0 - B >= 2;
        if (next && kLookAhead) {
            // look-ahead: the tile column that holds the next panel (columns t0 .. t0 + B - 1) is updated
            // first; the next panel's QR then runs on the second stream while the rest of the trailing
            // matrix (tile columns > I0, untouched by that QR) is updated here
            her2k(t0, V, I0, I0, nI, s);
            QD_HIP(hipEventRecord(w.ev_strip, s));
            QD_HIP(hipStreamWaitEvent(w.s2, w.ev_strip, 0));
            panel_qr(t0, w.s2);
            QD_HIP(hipEventRecord(w.ev_panel, w.s2));
            her2k(t0, V, I0 + 1, -1, (long)(nI - 1) * nI / 2, s);
            QD_HIP(hipStreamWaitEvent(s, w.ev_panel, 0));
        } else {
            her2k(t0, V, I0, -1, (long)nI * (nI + 1) / 2, s);
            if (next) panel_qr(t0, s);
        }
    }
    const long tot = (long)n * 2 * B;
    hipLaunchKernelGGL(k_band_extract<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, w.A, ld, n, AB);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_set_units<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w.A, ld, n);
    QD_LAUNCH_CHECK();
}

template <class T>
bool band_apply_q1(BandWork<T> &w, c64 *X, long ldx, int k, bool adjoint, int *sync, hipStream_t s) {
    if (k > 64) throw HipError("too many vectors for the band back-transformation", QD_E_UNSUPPORTED);
    if (w.npan == 0 || k < 1) return true;
    if (w.q1_max <= 0) return false;
    const int G = std::min(std::min(w.q1_max, 256), std::max(1, w.ntb));
    hipLaunchKernelGGL(k_fill_int_b, dim3((kGridBarrierInts + 255) / 256), dim3(256), 0, s, w.gbar, kGridBarrierInts, 0);
    QD_LAUNCH_CHECK();
    const cpx<T> *A = w.A;
    long ld = w.ld;
    int n = w.n, np = w.npan;
    const c64 *Tall = w.Tall;
    void *args[] = {(void *)&A, &ld, &n, &np, (void *)&Tall, &X, &ldx, &k, &w.qpart, &w.qy, &sync, &w.gbar};
    const int v = k <= 8 ? 0 : (k <= 16 ? 1 : (k <= 32 ? 2 : (k <= 40 ? 3 : 4)));   // KU = 1, 2, 4, 5, 8
    QD_HIP(hipLaunchCooperativeKernel(q1_kernels<T>()[2 * v + (adjoint ? 1 : 0)], dim3(G), dim3(256), args, 0, s));
    return true;
}

template void band_alloc<float>(BandWork<float> &, int, cpx<float> *);
template void band_alloc<double>(BandWork<double> &, int, cpx<double> *);
template void band_free<float>(BandWork<float> &);
template void band_free<double>(BandWork<double> &);
template bool band_apply_q1<float>(BandWork<float> &, c64 *, long, int, bool, int *, hipStream_t);
template bool band_apply_q1<double>(BandWork<double> &, c64 *, long, int, bool, int *, hipStream_t);
template void band_reduce<float>(BandWork<float> &, c64 *, int *, hipStream_t, KernelProfile *);
template void band_reduce<double>(BandWork<double> &, c64 *, int *, hipStream_t, KernelProfile *);

}  // namespace qd
