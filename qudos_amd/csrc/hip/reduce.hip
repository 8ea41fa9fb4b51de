// Kernel set (b), part 1: blocked Householder tridiagonalisation of a Hermitian matrix on
// MI355X (LAPACK zhetrd/zlatrd semantics, lower triangle, panels of kNB columns).
//
// Per column c of a panel two kernels run:
//   k_col_update : applies the panel's previous reflectors to column c, forms the norms and the
//                  V^H x, W^H x partials; the last work-group to finish builds the reflector
//                  (beta, tau, scale) and V^H v, W^H v.
//   k_hemv       : y = A22 v over the lower triangle of the trailing matrix (each 64x64 tile
//                  read once, used for both A v and A^H v), per-tile partials in global memory;
//                  the last work-group touching a row block reduces it, applies the panel
//                  corrections and writes w = tau (y - W V^H v - V W^H v); the last row block
//                  forms alpha = -tau/2 w^H v.
// After kNB columns k_her2k applies A22 -= V W^H + W V^H with MFMA (v_mfma_*_16x16x4) tiles
// staged through LDS.  The hemv half streams the trailing matrix once per column and bounds the
// reduction by HBM bandwidth; the her2k half is MFMA work.
#include <algorithm>
#include "qd_internal.hpp"
#include "qd_mfma.hpp"

namespace qd {
namespace {

// Block-wide sum of a complex value over 256 threads (4 wavefronts); result valid in all threads.
__device__ inline c64 block_sum(c64 v, c64 *scratch4) {
    v.re = wave_sum(v.re);
    v.im = wave_sum(v.im);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch4[w] = v;
    __syncthreads();
    c64 s = scratch4[0] + scratch4[1] + scratch4[2] + scratch4[3];
    __syncthreads();
    return s;
}

// ------------------------------------------------------------------------------------------
// Column update + reflector generation
// ------------------------------------------------------------------------------------------
// alpha of the previous panel column from the per-row-block w^H v partials (block-wide sum)
template <class T>
__device__ inline c64 prev_alpha(const ReduceWork<T> &w, int cprev, c64 *red4) {
    const int J0 = (cprev + 1) >> 6;
    c64 v = mk(0.0, 0.0);
    for (int q = J0 + (int)threadIdx.x; q < w.ntb; q += 256) v = v + w.part_wv[q];
    const c64 s = block_sum(v, red4);
    return (-0.5) * (w.tau[cprev] * s);
}

template <class T>
__global__ __launch_bounds__(256) void k_col_update(ReduceWork<T> w, int c, int p0, int j) {
    const int n = w.n;
    const long ld = w.ld;
    __shared__ c64 sVc[kNB], sWc[kNB], red4[4];
    const int tid = threadIdx.x;
    const c64 one = mk(1.0, 0.0);
    const c64 sprev = j > 0 ? w.scal[j - 1] : mk(0.0, 0.0);
    const c64 aprev = j > 0 ? prev_alpha(w, c - 1, red4) : mk(0.0, 0.0);
    if (j > 0 && blockIdx.x == 0 && tid == 0) w.alpha[j - 1] = aprev;
    // row c of the panel's reflectors and of W (W column j-1 fixed by + alpha v)
    if (tid < j) {
        const int q = tid;
        c64 vq = (q == j - 1) ? one : ld64(&w.A[(long)(p0 + q) * ld + c]);
        c64 wq = ld64(&w.W[(long)q * ld + c]);
        if (q == j - 1) wq = wq + aprev;
        sVc[q] = vq;
        sWc[q] = wq;
    }
    __syncthreads();
    const int r = c + blockIdx.x * 256 + tid;
    const bool valid = r < n;
    c64 a = mk(0.0, 0.0);
    if (valid) {
        a = ld64(&w.A[(long)c * ld + r]);
        for (int q = 0; q < j; ++q) {
            c64 vr, wr;
            if (q == j - 1) {
                // previous column: finish its scaling / unit entry and its W fix for this row
                const long ia = (long)(p0 + q) * ld + r;
                vr = (r == c) ? one : sprev * ld64(&w.A[ia]);
                wr = ld64(&w.W[(long)q * ld + r]) + aprev * vr;
                if (r > c) {
                    w.A[ia] = cast<T>(vr);
                    w.W[(long)q * ld + r] = cast<T>(wr);
                }
            } else {
                vr = ld64(&w.A[(long)(p0 + q) * ld + r]);
                wr = ld64(&w.W[(long)q * ld + r]);
            }
            a = a - mulc(vr, sWc[q]) - mulc(wr, sVc[q]);
        }
        if (r == c) {
            a.im = 0.0;
            w.d[c] = a.re;
            if (j > 0) w.A[(long)(p0 + j - 1) * ld + c] = cast<T>(one);
        }
        w.A[(long)c * ld + r] = cast<T>(a);
    }
    // partials over rows r >= c + 2: |x|^2, V^H x, W^H x (wave shuffles, one LDS pass)
    const bool inx = valid && r >= c + 2;
    c64 *part = w.ka_part + (long)blockIdx.x * (2 * kNB + 1);
    __shared__ c64 wred[4][2 * kNB + 1];
    const int wv = tid >> 6, lane = tid & 63;
    {
        double x = inx ? norm2(a) : 0.0;
        x = wave_sum(x);
        if (lane == 0) wred[wv][0] = mk(x, 0.0);
    }
    for (int q = 0; q < j; ++q) {
        c64 pv = mk(0.0, 0.0), pw = mk(0.0, 0.0);
        if (inx) {
            const c64 vr = ld64(&w.A[(long)(p0 + q) * ld + r]);   // already final for this row
            const c64 wr = ld64(&w.W[(long)q * ld + r]);
            pv = cmul(vr, a);
            pw = cmul(wr, a);
        }
        pv.re = wave_sum(pv.re); pv.im = wave_sum(pv.im);
        pw.re = wave_sum(pw.re); pw.im = wave_sum(pw.im);
        if (lane == 0) { wred[wv][1 + q] = pv; wred[wv][1 + kNB + q] = pw; }
    }
    __syncthreads();
    if (tid < 2 * kNB + 1) {
        const bool used = tid == 0 || (tid >= 1 && tid < 1 + j) || (tid >= 1 + kNB && tid < 1 + kNB + j);
        if (used) part[tid] = wred[0][tid] + wred[1][tid] + wred[2][tid] + wred[3][tid];
    }
}

// Householder reflector of column c from the column-update partials (one work-group).
template <class T>
__global__ __launch_bounds__(256) void k_reflector(ReduceWork<T> w, int c, int p0, int j, int nblk) {
    const int n = w.n;
    const long ld = w.ld;
    const int tid = threadIdx.x;
    const c64 one = mk(1.0, 0.0);
    __shared__ c64 sSum[2 * kNB + 1];
    __shared__ c64 s_scale;
    // 3 x 65 threads: value v = tid % 65 summed over every third work-group partial
    constexpr int NV = 2 * kNB + 1;
    __shared__ c64 s3[3][NV];
    if (tid < 3 * NV) {
        const int v = tid % NV, pr = tid / NV;
        c64 acc = mk(0.0, 0.0);
        const bool used = v == 0 || (v >= 1 && v < 1 + j) || (v >= 1 + kNB && v < 1 + kNB + j);
        if (used) {
            c64 a8[8];
            for (int u = 0; u < 8; ++u) a8[u] = mk(0.0, 0.0);
            for (int b0 = pr * 8; b0 < nblk; b0 += 24) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (b0 + u < nblk) a8[u] = a8[u] + w.ka_part[(long)(b0 + u) * NV + v];
            }
            for (int u = 0; u < 8; ++u) acc = acc + a8[u];
        }
        s3[pr][v] = acc;
    }
    __syncthreads();
    if (tid < NV) sSum[tid] = s3[0][tid] + s3[1][tid] + s3[2][tid];
    __syncthreads();
    if (tid == 0) {
        if (c + 1 < n) {
            const c64 alph = ld64(&w.A[(long)c * ld + c + 1]);
            const double xn2 = sSum[0].re;
            c64 tau, sc;
            double beta;
            if (xn2 == 0.0 && alph.im == 0.0) {
                tau = mk(0.0, 0.0);
                beta = alph.re;
                sc = mk(0.0, 0.0);
            } else {
                const double nrm = sqrt(alph.re * alph.re + alph.im * alph.im + xn2);
                beta = alph.re >= 0.0 ? -nrm : nrm;
                tau = mk((beta - alph.re) / beta, -alph.im / beta);
                sc = div_(one, alph - mk(beta, 0.0));
            }
            w.tau[c] = tau;
            w.e[c] = beta;
            w.scal[j] = sc;
            s_scale = sc;
        } else {
            w.tau[c] = mk(0.0, 0.0);
            w.scal[j] = mk(0.0, 0.0);
            s_scale = mk(0.0, 0.0);
        }
    }
    __syncthreads();
    if (tid < j && c + 1 < n) {
        const int q = tid;
        const c64 vq = ld64(&w.A[(long)(p0 + q) * ld + c + 1]);
        const c64 wq = ld64(&w.W[(long)q * ld + c + 1]);
        w.vd[q] = conj(vq) + s_scale * sSum[1 + q];
        w.wd[q] = conj(wq) + s_scale * sSum[1 + kNB + q];
    }
}

// ------------------------------------------------------------------------------------------
// Hermitian matrix-vector product over the trailing matrix + finalisation of w
// ------------------------------------------------------------------------------------------
// Lane layout of the hemv tile loop: each lane reads 16 contiguous bytes of a column (two FP32
// or one FP64 complex rows), so a row group of LR lanes covers one 64-row column segment.
template <class T> struct HemvLayout;
template <> struct HemvLayout<float> { static constexpr int RPT = 2; };
template <> struct HemvLayout<double> { static constexpr int RPT = 1; };

template <class T>
__device__ inline T lane_sum(T v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
__global__ __launch_bounds__(256) void k_hemv(ReduceWork<T> w, int c, int p0, int j) {
    constexpr int RPT = HemvLayout<T>::RPT;   // rows per lane
    constexpr int LR = 64 / RPT;              // lanes per 64-row column segment
    constexpr int NCG = 256 / LR;             // column groups
    constexpr int CPT = 64 / NCG;             // columns per lane
    const int n = w.n;
    const long ld = w.ld;
    const int s0 = c + 1;
    const int J0 = s0 >> 6;
    const int ntb = w.ntb;
    const int tid = threadIdx.x;
    // work-group -> (strip J, segment)
    const int b = blockIdx.x + w.prefix[J0];
    int lo = J0, hi = ntb - 1;
    while (lo < hi) {   // largest J with prefix[J] <= b
        const int mid = (lo + hi + 1) >> 1;
        if (w.prefix[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int J = lo;
    const int seg = b - w.prefix[J];
    const int Ib = J + seg * kSEG, Ie = min(ntb, Ib + kSEG);
    const c64 sc = w.scal[j];
    const cpx<T> *colc = w.A + (long)c * ld;
    auto vval = [&](int R) -> cpx<T> {
        if (R < s0 || R >= n) return mk<T>(0, 0);
        if (R == s0) return mk<T>(1, 0);
        return cast<T>(sc * ld64(&colc[R]));
    };
    const int lr = tid % LR, cgp = tid / LR;
    __shared__ cpx<T> red[NCG][64];
    cpx<T> xJ[CPT], accT[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        xJ[k] = vval(J * 64 + cgp * CPT + k);
        accT[k] = mk<T>(0, 0);
    }
    for (int I = Ib; I < Ie; ++I) {
        const int R0 = I * 64 + lr * RPT;
        cpx<T> xI[RPT], yI[RPT];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            xI[i] = vval(R0 + i);
            yI[i] = mk<T>(0, 0);
        }
        cpx<T> a[CPT][RPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {   // issue all loads of the tile first
            const int C = J * 64 + cgp * CPT + k;
            const cpx<T> *src = &w.A[(long)(C < n ? C : 0) * ld + R0];
#pragma unroll
            for (int i = 0; i < RPT; ++i) a[k][i] = src[i];
        }
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int C = J * 64 + cgp * CPT + k;
            if (C < s0 || C >= n) continue;
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int R = R0 + i;
                if (R < C || R >= n) continue;
                cpx<T> v = a[k][i];
                if (R == C) {
                    v.im = 0;
                    fma_(yI[i], v, xJ[k]);
                } else {
                    fma_(yI[i], v, xJ[k]);
                    fma_(accT[k], conj(v), xI[i]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < RPT; ++i) red[cgp][lr * RPT + i] = yI[i];
        __syncthreads();
        if (tid < 64) {
            cpx<T> sum = mk<T>(0, 0);
#pragma unroll
            for (int q = 0; q < NCG; ++q) sum = sum + red[q][tid];
            w.P[((long)I * (I + 1) / 2 + J) * 64 + tid] = sum;
        }
        __syncthreads();
    }
    // transposed contribution: reduce accT over the LR lanes that share a column group
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        accT[k].re = lane_sum(accT[k].re, LR);
        accT[k].im = lane_sum(accT[k].im, LR);
    }
    if (lr == 0) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) w.PT[((long)J * w.maxseg + seg) * 64 + cgp * CPT + k] = accT[k];
    }
}

// Row block I = J0 + blockIdx.x: sum the hemv partials, apply the panel corrections,
// w = tau (y - V W^H v - W V^H v), and the block's share of w^H v.
template <class T>
__global__ __launch_bounds__(256) void k_finalize(ReduceWork<T> w, int c, int p0, int j) {
    const int n = w.n;
    const long ld = w.ld;
    const int s0 = c + 1;
    const int J0 = s0 >> 6;
    const int ntb = w.ntb;
    const int tid = threadIdx.x;
    const int I = J0 + blockIdx.x;
    const c64 sc = w.scal[j];
    const c64 tau = w.tau[c];
    const int row = tid & 63, part = tid >> 6;
    const int nP = I - J0 + 1, nS = (ntb - I + kSEG - 1) / kSEG, tot = nP + nS;
    c64 acc8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc8[u] = mk(0.0, 0.0);
    for (int q0 = part * 8; q0 < tot; q0 += 32) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u;
            if (q < tot) {
                const cpx<T> *src = q < nP ? &w.P[((long)I * (I + 1) / 2 + J0 + q) * 64]
                                            : &w.PT[((long)I * w.maxseg + (q - nP)) * 64];
                acc8[u] = acc8[u] + ld64(&src[row]);
            }
        }
    }
    c64 acc = mk(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + acc8[u];
    __shared__ c64 s4[4][64];
    s4[part][row] = acc;
    __syncthreads();
    if (tid >= 64) return;
    c64 y = s4[0][row] + s4[1][row] + s4[2][row] + s4[3][row];
    const int R = I * 64 + row;
    c64 wv = mk(0.0, 0.0), dot = mk(0.0, 0.0);
    if (R >= s0 && R < n) {
#pragma unroll 8
        for (int q = 0; q < j; ++q) {
            const c64 vr = ld64(&w.A[(long)(p0 + q) * ld + R]);
            const c64 wr = ld64(&w.W[(long)q * ld + R]);
            y = y - vr * w.wd[q] - wr * w.vd[q];
        }
        wv = tau * y;
        const c64 vR = (R == s0) ? mk(1.0, 0.0) : sc * ld64(&w.A[(long)c * ld + R]);
        dot = cmul(wv, vR);
    }
    if (R < n) w.W[(long)j * ld + R] = cast<T>(wv);
    dot.re = wave_sum(dot.re);
    dot.im = wave_sum(dot.im);
    if (tid == 0) w.part_wv[I] = dot;
}

// Finish the last panel column: scale its reflector, store the unit entry, fix its W column.
template <class T>
__global__ __launch_bounds__(256) void k_panel_end(ReduceWork<T> w, int c, int j) {
    __shared__ c64 red4[4];
    const c64 al = prev_alpha(w, c, red4);
    if (blockIdx.x == 0 && threadIdx.x == 0) w.alpha[j] = al;
    const int r = c + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= w.n) return;
    const long ia = (long)c * w.ld + r;
    const c64 sc = w.scal[j];
    const c64 v = (r == c + 1) ? mk(1.0, 0.0) : sc * ld64(&w.A[ia]);
    w.A[ia] = cast<T>(v);
    const long iw = (long)j * w.ld + r;
    w.W[iw] = cast<T>(ld64(&w.W[iw]) + al * v);
}

// Trailing update A22 -= V W^H + W V^H on the lower 64x64 tiles at or below row/col t0.
template <class T>
__global__ __launch_bounds__(256) void k_her2k(ReduceWork<T> w, int p0, int nbp, int t0) {
    typedef typename VecT<T>::v4 v4;
    const int n = w.n;
    const long ld = w.ld;
    const int Jb = t0 >> 6;
    int I, J;
    {
        const long t = blockIdx.x;
        long i = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        while (i * (i + 1) / 2 > t) --i;
        I = (int)i + Jb;
        J = (int)(t - i * (i + 1) / 2) + Jb;
    }
    const int K = 2 * nbp;
    const int Kp = (K + 3) & ~3;
    constexpr int LDK = 2 * kNB + 1;
    __shared__ T sAr[64][LDK], sAi[64][LDK], sBr[64][LDK], sBi[64][LDK];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < 64 * Kp; idx += 256) {
        const int rl = idx & 63, k = idx >> 6;
        const int R = I * 64 + rl, C = J * 64 + rl;
        cpx<T> a = mk<T>(0, 0), bb = mk<T>(0, 0);
        if (k < K) {
            if (R >= t0 && R < n) a = k < nbp ? w.A[(long)(p0 + k) * ld + R] : w.W[(long)(k - nbp) * ld + R];
            if (C >= t0 && C < n) bb = k < nbp ? w.W[(long)k * ld + C] : w.A[(long)(p0 + k - nbp) * ld + C];
        }
        sAr[rl][k] = a.re;
        sAi[rl][k] = a.im;
        sBr[rl][k] = bb.re;
        sBi[rl][k] = bb.im;
    }
    __syncthreads();
    const int wv = tid >> 6, lane = tid & 63;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    v4 accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) {
            accR[a][b2] = v4{0, 0, 0, 0};
            accI[a][b2] = v4{0, 0, 0, 0};
        }
    for (int kk = 0; kk < Kp; kk += 4) {
        T ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            ar[m] = sAr[wr + m * 16 + li][kk + lk];
            ai[m] = sAi[wr + m * 16 + li][kk + lk];
            br[m] = sBr[wc + m * 16 + li][kk + lk];
            bi[m] = sBi[wc + m * 16 + li][kk + lk];
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                accR[m][q] = mfma(ar[m], br[q], accR[m][q]);
                accR[m][q] = mfma(ai[m], bi[q], accR[m][q]);
                accI[m][q] = mfma(ai[m], br[q], accI[m][q]);
                accI[m][q] = mfma(-ar[m], bi[q], accI[m][q]);
            }
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int R = I * 64 + wr + m * 16 + acc_row<T>(lk, e);
                const int C = J * 64 + wc + q * 16 + li;
                if (R < t0 || C < t0 || R >= n || C >= n || R < C) continue;
                const long ia = (long)C * ld + R;
                cpx<T> v = w.A[ia];
                v.re -= accR[m][q][e];
                v.im -= accI[m][q][e];
                if (R == C) v.im = 0;
                w.A[ia] = v;
            }
}

// C(16x16) = A(16x4) B(4x16) with the operand layout assumed by k_her2k; one wavefront.
template <class T>
__global__ void k_mfma_probe(const T *A, const T *B, T *C) {
    typedef typename VecT<T>::v4 v4;
    const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
    v4 acc = v4{0, 0, 0, 0};
    acc = mfma(A[li * 4 + lk], B[lk * 16 + li], acc);
    for (int e = 0; e < 4; ++e) C[acc_row<T>(lk, e) * 16 + li] = acc[e];
}

// raw dump: lane l provides a[l], b[l]; out[l*4 + e] = accumulator element e of lane l
__global__ void k_mfma_raw64(const double *a, const double *b, double *out) {
    typedef VecT<double>::v4 v4;
    const int lane = threadIdx.x;
    v4 acc = v4{0, 0, 0, 0};
    acc = mfma(a[lane], b[lane], acc);
    for (int e = 0; e < 4; ++e) out[lane * 4 + e] = acc[e];
}

}  // namespace

void mfma_raw64(const double *ha, const double *hb, double *hout) {
    double *da, *db, *dout;
    QD_HIP(hipMalloc(&da, 64 * sizeof(double)));
    QD_HIP(hipMalloc(&db, 64 * sizeof(double)));
    QD_HIP(hipMalloc(&dout, 256 * sizeof(double)));
    QD_HIP(hipMemcpy(da, ha, 64 * sizeof(double), hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(db, hb, 64 * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_raw64, dim3(1), dim3(64), 0, 0, da, db, dout);
    QD_LAUNCH_CHECK();
    QD_HIP(hipMemcpy(hout, dout, 256 * sizeof(double), hipMemcpyDeviceToHost));
    hipFree(da); hipFree(db); hipFree(dout);
}

template <class T>
double mfma_selftest() {
    T hA[64], hB[64], hC[256];
    for (int i = 0; i < 64; ++i) { hA[i] = (T)((i * 7) % 11 - 5); hB[i] = (T)((i * 5) % 13 - 6); }
    T *dA, *dB, *dC;
    QD_HIP(hipMalloc(&dA, sizeof(hA)));
    QD_HIP(hipMalloc(&dB, sizeof(hB)));
    QD_HIP(hipMalloc(&dC, sizeof(hC)));
    QD_HIP(hipMemcpy(dA, hA, sizeof(hA), hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(dB, hB, sizeof(hB), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_probe<T>, dim3(1), dim3(64), 0, 0, dA, dB, dC);
    QD_LAUNCH_CHECK();
    QD_HIP(hipMemcpy(hC, dC, sizeof(hC), hipMemcpyDeviceToHost));
    hipFree(dA); hipFree(dB); hipFree(dC);
    double err = 0;
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) {
            double ref = 0;
            for (int k = 0; k < 4; ++k) ref += (double)hA[i * 4 + k] * (double)hB[k * 16 + j];
            err = std::max(err, std::fabs(ref - (double)hC[i * 16 + j]));
        }
    return err;
}
template double mfma_selftest<float>();
template double mfma_selftest<double>();

template <class T>
void reduce_alloc(ReduceWork<T> &w, int n, cpx<T> *A_external) {
    w.n = n;
    w.ld = ((long)n + 63) / 64 * 64;
    w.ntb = (n + 63) / 64;
    w.maxseg = (w.ntb + kSEG - 1) / kSEG;
    w.A = A_external;
    const long ld = w.ld;
    QD_HIP(hipMalloc(&w.W, sizeof(cpx<T>) * ld * kNB));
    QD_HIP(hipMalloc(&w.d, sizeof(double) * n));
    QD_HIP(hipMalloc(&w.e, sizeof(double) * (n > 1 ? n : 1)));
    QD_HIP(hipMalloc(&w.tau, sizeof(c64) * n));
    QD_HIP(hipMalloc(&w.scal, sizeof(c64) * kNB));
    QD_HIP(hipMalloc(&w.alpha, sizeof(c64) * kNB));
    QD_HIP(hipMalloc(&w.vd, sizeof(c64) * kNB));
    QD_HIP(hipMalloc(&w.wd, sizeof(c64) * kNB));
    const long maxA = (n + 255) / 256 + 1;
    QD_HIP(hipMalloc(&w.ka_part, sizeof(c64) * maxA * (2 * kNB + 1)));
    QD_HIP(hipMalloc(&w.ka_cnt, sizeof(int)));
    QD_HIP(hipMalloc(&w.P, sizeof(cpx<T>) * ((long)w.ntb * (w.ntb + 1) / 2) * 64));
    QD_HIP(hipMalloc(&w.PT, sizeof(cpx<T>) * (long)w.ntb * w.maxseg * 64));
    QD_HIP(hipMalloc(&w.part_wv, sizeof(c64) * w.ntb));
    QD_HIP(hipMalloc(&w.cnt, sizeof(int) * w.ntb));
    QD_HIP(hipMalloc(&w.done, sizeof(int)));
    QD_HIP(hipMalloc(&w.prefix, sizeof(int) * (w.ntb + 1)));
    QD_HIP(hipMemset(w.W, 0, sizeof(cpx<T>) * ld * kNB));
    QD_HIP(hipMemset(w.ka_cnt, 0, sizeof(int)));
    QD_HIP(hipMemset(w.cnt, 0, sizeof(int) * w.ntb));
    QD_HIP(hipMemset(w.done, 0, sizeof(int)));
    std::vector<int> pre(w.ntb + 1, 0);
    for (int J = 0; J < w.ntb; ++J) pre[J + 1] = pre[J] + (w.ntb - J + kSEG - 1) / kSEG;
    QD_HIP(hipMemcpy(w.prefix, pre.data(), sizeof(int) * (w.ntb + 1), hipMemcpyHostToDevice));
}

template <class T>
void reduce_free(ReduceWork<T> &w) {
    hipFree(w.W); hipFree(w.d); hipFree(w.e); hipFree(w.tau); hipFree(w.scal); hipFree(w.alpha);
    hipFree(w.vd); hipFree(w.wd); hipFree(w.ka_part); hipFree(w.ka_cnt); hipFree(w.P); hipFree(w.PT);
    hipFree(w.part_wv); hipFree(w.cnt); hipFree(w.done); hipFree(w.prefix);
    w = ReduceWork<T>();
}

template <class T>
void reduce_run(ReduceWork<T> &w, hipStream_t s, KernelProfile *prof) {
    const int n = w.n;
    std::vector<int> pre(w.ntb + 1, 0);
    for (int J = 0; J < w.ntb; ++J) pre[J + 1] = pre[J] + (w.ntb - J + kSEG - 1) / kSEG;
    for (int p0 = 0; p0 < n; p0 += kNB) {
        const int nbp = std::min(kNB, n - p0);
        for (int j = 0; j < nbp; ++j) {
            const int c = p0 + j;
            const unsigned ga = (unsigned)((n - c + 255) / 256);
            hipLaunchKernelGGL(k_col_update<T>, dim3(ga), dim3(256), 0, s, w, c, p0, j);
            QD_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_reflector<T>, dim3(1), dim3(256), 0, s, w, c, p0, j, (int)ga);
            QD_LAUNCH_CHECK();
            if (c + 1 < n) {
                const int J0 = (c + 1) >> 6;
                const unsigned gb = (unsigned)(pre[w.ntb] - pre[J0]);
                const bool timed = prof && prof->on && (c % kProfStride) == 0;
                if (timed) prof->begin(s);
                hipLaunchKernelGGL(k_hemv<T>, dim3(gb), dim3(256), 0, s, w, c, p0, j);
                QD_LAUNCH_CHECK();
                if (timed) {
                    const double m = n - c - 1;
                    prof->end(s, m * (m + 1) / 2 * sizeof(cpx<T>));
                }
                hipLaunchKernelGGL(k_finalize<T>, dim3(w.ntb - J0), dim3(256), 0, s, w, c, p0, j);
                QD_LAUNCH_CHECK();
            }
        }
        const int cl = p0 + nbp - 1;
        if (cl + 1 < n) {
            const unsigned g = (unsigned)((n - cl - 1 + 255) / 256);
            hipLaunchKernelGGL(k_panel_end<T>, dim3(g), dim3(256), 0, s, w, cl, nbp - 1);
            QD_LAUNCH_CHECK();
        }
        const int t0 = p0 + nbp;
        if (t0 < n) {
            const long tt = w.ntb - (t0 >> 6);
            hipLaunchKernelGGL(k_her2k<T>, dim3((unsigned)(tt * (tt + 1) / 2)), dim3(256), 0, s, w, p0, nbp, t0);
            QD_LAUNCH_CHECK();
        }
    }
}

template void reduce_alloc<float>(ReduceWork<float> &, int, cpx<float> *);
template void reduce_alloc<double>(ReduceWork<double> &, int, cpx<double> *);
template void reduce_free<float>(ReduceWork<float> &);
template void reduce_free<double>(ReduceWork<double> &);
template void reduce_run<float>(ReduceWork<float> &, hipStream_t, KernelProfile *);
template void reduce_run<double>(ReduceWork<double> &, hipStream_t, KernelProfile *);

}  // namespace qd
