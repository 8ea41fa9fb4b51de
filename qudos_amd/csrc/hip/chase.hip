// Kernel set (b), two-stage reduction, stage 2: band (bandwidth kBW) -> real symmetric
// tridiagonal by bulge chasing, and the application of its reflectors Q2 to a block of vectors.
// FP64 in both precision modes (the band is n x 2 kBW, ~50 MB at N_H = 48,552).
//
// Sweep s annihilates column s below the subdiagonal; its step j generates a reflector from
// column c = s (j = 0) or c = s + 1 + (j-1) kBW (the first column of the bulge left by step
// j-1), rows r0 = s + 1 + j kBW .. r0 + kBW - 1, and applies it
//   (i)   from the left to the rest of the bulge block (columns c+1 .. r0-1),
//   (ii)  from both sides to the diagonal block (rows/columns r0 .. r0 + kBW - 1),
//   (iii) from the right to the block below, which creates the next bulge.
// Storage: AB[c * 2 kBW + d] = A(c + d, c), d < 2 kBW (band plus bulge rows).
// k_chase: persistent work-groups claim sweeps in order from a counter; step j of sweep s
// starts once sweep s-1 has finished step j+1.  Step j of sweep s reads columns up to
// c' = s + (j+1) kBW; all of them are final once sweep s-1 has finished step j+1 (its step j+2
// only rewrites A(s + (j+2) kBW, c') and the rows below it in column c', through the reflector
// that it builds during step j+1), so each step writes the first column of its diagonal block
// back as soon as it is final, and the later write-back of that block skips it.
// A work-group keeps the columns of the current diagonal block in LDS from one step to the next
// (no other sweep touches them in between) and writes a block back when the chase moves past it.
// k_q2_apply: Z <- Q2 Z (or Q2^H Z).  Reflectors act on kBW-row windows; for a fixed step j the
// reflectors of kBW consecutive sweeps share a 2 kBW - 1 row window held in LDS.  Step-j
// pipelines run concurrently, each behind its neighbour (j-1 for Q2 Z, j+1 for Q2^H Z) by one
// block of sweeps, which orders every pair of overlapping reflectors as the product requires.
// Waits are bounded: a wait that exceeds its limit raises the abort flag, every waiting
// work-group then exits, and the host reports an error.
#include <algorithm>

#include "qd_internal.hpp"
#include "qd_mfma.hpp"
#include "qd_sync.hpp"

namespace qd {
namespace {

constexpr int B = kBW;
constexpr int LDB = 2 * kBW;
constexpr int kDone = 1 << 29;

// index of the reflector of sweep s, step j (steps j exist for s < n - 1 - j kBW)
__device__ inline long refl_index(int n, int s, int j) {
    return (long)j * (n - 1) - (long)B * j * (j - 1) / 2 + s;
}

__device__ inline c64 shfl_c(c64 v, int src) {
    return mk(__shfl(v.re, src, 64), __shfl(v.im, src, 64));
}

__device__ inline c64 group8_sum(c64 v) {
    v.re = sum8(v.re);
    v.im = sum8(v.im);
    return v;
}

__global__ __launch_bounds__(256) void k_chase(c64 *AB, int n, c64 *V2, c64 *tau2, int *sync, int *prog) {
    __shared__ c64 buf[2][B][LDB];
    __shared__ c64 svb[2][B], sy[B];   // reflector of step j in svb[j & 1]
    __shared__ c64 s_taub[2];
    __shared__ int s_s, s_ok, s_known;
    const int tid = threadIdx.x, lane = tid & 63;
    constexpr int PF = B * LDB / 256;   // prefetched band entries per thread
    while (true) {
        if (tid == 0) s_s = atomicAdd(&sync[0], 1);
        __syncthreads();
        const int s = s_s;
        __syncthreads();
        if (s >= n - 1) return;
        const int ns = (n - 1 - s + B - 1) / B;
        // s_known: progress of sweep s-1 last read by this work-group
        if (tid == 0) s_known = s == 0 ? kDone : -1;
        bool have_cur = false;   // the current step's diagonal block was prefetched into LDS
        bool have_refl = false;  // the current step's reflector was built during the previous step
        for (int j = 0; j < ns; ++j) {
            const int c = j == 0 ? s : s + 1 + (j - 1) * B;
            const int r0 = s + 1 + j * B;
            const int L = min(B, n - r0);
            const int r1 = r0 + L;
            const int L2 = r1 < n ? min(B, n - r1) : 0;
            const int off0 = r0 - c;
            c64(*prev)[LDB] = buf[j & 1];
            c64(*cur)[LDB] = buf[(j + 1) & 1];
            if (!have_cur) {
                if (tid == 0) {
                    int ok = 1;
                    if (s_known < j + 2) {
                        int seen = s_known;
                        ok = wait_at_least(&prog[(long)(s - 1) * kFlagStride], j + 2, sync, &seen, 1);
                        s_known = seen;
                    }
                    s_ok = ok;
                }
                __syncthreads();
                if (!s_ok) return;
                c64 reg[PF];   // all loads of the block in flight together
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    const int idx = tid + 256 * u;
                    reg[u] = idx < L * LDB ? ld_coh(&AB[(long)(r0 + idx / LDB) * LDB + idx % LDB]) : mk(0.0, 0.0);
                }
                const c64 c0 = (j == 0 && tid < LDB) ? ld_coh(&AB[(long)s * LDB + tid]) : mk(0.0, 0.0);
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    const int idx = tid + 256 * u;
                    if (idx < L * LDB) cur[idx / LDB][idx % LDB] = reg[u];
                }
                if (j == 0 && tid < LDB) prev[0][tid] = c0;
                __syncthreads();
            }
            // read the predecessor's progress now; it is needed after the arithmetic of this step
            int pk = 0;
            if (tid == 0 && s_known < kDone)
                pk = __hip_atomic_load(&prog[(long)(s - 1) * kFlagStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // operands of the one-sided updates (i), (iii) and of y = D v (ii), read before the reflector
            // is known: the reflector step writes only column c of prev, which none of them reads
            const int g8 = tid >> 3, sub = tid & 7;
            const bool doI = j > 0 && tid < (B - 1) * 8;
            const int tI = 1 + g8;
            c64 a[B / 8], cc[B / 8], dv[B / 8];
#pragma unroll
            for (int u = 0; u < B / 8; ++u) {
                const int k = sub + 8 * u;
                a[u] = (doI && k < L) ? prev[tI][B + k - tI] : mk(0.0, 0.0);                  // (i): A(r0 + k, c + tI)
                cc[u] = (g8 < L2 && k < L) ? cur[k][B + g8 - k] : mk(0.0, 0.0);              // (iii): C(g8, k)
                dv[u] = (g8 < L && k < L) ? (g8 >= k ? cur[k][g8 - k] : conj(cur[g8][k - g8])) : mk(0.0, 0.0);   // (ii): D(g8, k)
            }
            // reflector of x = A(r0 .. r1-1, c)  (zlarfg): from column c (entries xcol[0 .. Lx-1]) to the
            // unit vector svb[jb], tau s_taub[jb], and the reflector arrays for step jr; the column becomes
            // (beta, 0, ..., 0).  Run by wavefront 0.
            auto reflector = [&](c64 *xcol, int Lx, int jb, int jr) {
                const c64 x = lane < Lx ? xcol[lane] : mk(0.0, 0.0);
                const double xs = sum16(lane >= 1 && lane < Lx ? norm2(x) : 0.0);   // lanes >= 32 hold zeros
                const double xn2 = xs + __shfl_xor(xs, 16, 64);
                const c64 alpha = mk(lane0(x.re), lane0(x.im));
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
                const long ridx = refl_index(n, s, jr);
                if (lane < B) {
                    const c64 v = lane == 0 ? mk(1.0, 0.0) : (lane < Lx ? scl * x : mk(0.0, 0.0));
                    svb[jb][lane] = v;
                    V2[ridx * B + lane] = v;
                }
                if (lane < Lx) xcol[lane] = lane == 0 ? mk(beta, 0.0) : mk(0.0, 0.0);
                if (lane == 0) {
                    s_taub[jb] = tau;
                    tau2[ridx] = tau;
                }
            };
            if (!have_refl) {
                if (tid < 64) reflector(&prev[0][off0], L, j & 1, j);
                __syncthreads();
            }
            const c64 *sv = svb[j & 1];
            const c64 tau = s_taub[j & 1];
            {
                // (i) w = v^H A(:, c + tI) for the columns c+1 .. r0-1 (left application),
                // (iii) u = C v for row g8 of the block below (right application), (ii) y = D v
                c64 wI = mk(0.0, 0.0), uI = mk(0.0, 0.0), yI = mk(0.0, 0.0);
#pragma unroll
                for (int u = 0; u < B / 8; ++u) {
                    const c64 vk = sv[sub + 8 * u];
                    wI = wI + cmul(vk, a[u]);
                    fma_(uI, cc[u], vk);
                    fma_(yI, dv[u], vk);
                }
                wI = group8_sum(wI);
                uI = group8_sum(uI);
                yI = group8_sum(yI);
                if (doI) {
                    const c64 ct = conj(tau) * wI;
#pragma unroll
                    for (int u = 0; u < B / 8; ++u) {
                        const int k = sub + 8 * u;
                        if (k < L) prev[tI][B + k - tI] = a[u] - sv[k] * ct;
                    }
                }
                if (g8 < L2) {
                    const c64 tu = tau * uI;
#pragma unroll
                    for (int u = 0; u < B / 8; ++u) {
                        const int k = sub + 8 * u;
                        if (k < L) cur[k][B + g8 - k] = cc[u] - mulc(tu, sv[k]);
                    }
                }
                if (sub == 0) sy[g8] = yI;
            }
            __syncthreads();
            // the next step's reflector comes from column r0 below the diagonal block, cur[0][B ..], which
            // is final after (iii): wavefront 0 builds it now, off the next step's critical path
            if (j + 1 < ns && tid < 64) reflector(&cur[0][B], L2, (j + 1) & 1, j + 1);
            // the block the chase leaves is final now: its write-back drains during the two-sided update
            if (j == 0) {
                for (int d = tid; d < LDB; d += 256) st_coh(&AB[(long)s * LDB + d], prev[0][d]);
            } else {   // column c itself was written back at the end of the previous step
                for (int idx = LDB + tid; idx < B * LDB; idx += 256)
                    st_coh(&AB[(long)(c + idx / LDB) * LDB + idx % LDB], prev[idx / LDB][idx % LDB]);
            }
            {
                // al = Re(v^H y), summed by every wavefront on its own (no further barrier)
                double pa = lane < L ? cmul(sv[lane], sy[lane]).re : 0.0;   // lanes >= 32: L <= 32
                pa = sum16(pa);
                pa += __shfl_xor(pa, 16, 64);
                const double al = lane0(pa);
                const double t2 = 0.5 * norm2(tau) * al;
                // two-sided update of the diagonal block: entries (i, k) = (tid / 8, tid % 8 + 8 u), lower triangle
                const int i = g8;
                const c64 wi = tau * sy[i] - t2 * sv[i];
#pragma unroll
                for (int u = 0; u < B / 8; ++u) {
                    const int k = sub + 8 * u;
                    const c64 wk = tau * sy[k] - t2 * sv[k];
                    const int d = max(i - k, 0);
                    c64 val = cur[k][d] - mulc(wi, sv[k]) - mulc(sv[i], wk);
                    if (i == k) val.im = 0.0;
                    if (k <= i && i < L) cur[k][d] = val;
                }
            }
            if (tid == 0 && s_known < kDone) s_known = max(s_known, pk);
            __syncthreads();
            // prefetch the next step's diagonal block when the predecessor has released it
            const bool pre = j + 1 < ns && s_known >= j + 3;
            const int Ln = pre ? min(B, n - r1) : 0;
            c64 reg[PF];
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int idx = tid + 256 * u;
                reg[u] = idx < Ln * LDB ? ld_coh(&AB[(long)(r1 + idx / LDB) * LDB + idx % LDB]) : mk(0.0, 0.0);
            }
            // on the last step the current block is final too; otherwise its first column (column r0)
            // is: the successor sweep may read it once this step is published
            if (j == ns - 1) {
                for (int idx = tid; idx < L * LDB; idx += 256)
                    st_coh(&AB[(long)(r0 + idx / LDB) * LDB + idx % LDB], cur[idx / LDB][idx % LDB]);
            } else if (tid < LDB) {
                st_coh(&AB[(long)r0 * LDB + tid], cur[0][tid]);
            }
            wait_vmem();
            __syncthreads();
            if (tid == 0) publish(&prog[(long)s * kFlagStride], j == ns - 1 ? kDone : j + 1);
            if (pre) {
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    const int idx = tid + 256 * u;
                    if (idx < Ln * LDB) prev[idx / LDB][idx % LDB] = reg[u];
                }
                __syncthreads();
            }
            have_cur = pre;
            have_refl = true;
        }
    }
}

__global__ void k_band_tridiag(const c64 *AB, int n, double *d, double *e) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d[i] = AB[(long)i * LDB].re;
    if (i < n - 1) e[i] = AB[(long)i * LDB + 1].re;
}

__global__ void k_fill_int(int *p, int count, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

constexpr int kQG = B;                 // sweeps per block in the Q2 application
constexpr int kQW = kQG + B - 1;       // window rows
constexpr int kQK = 64;                // max vectors
constexpr int kQC = 32;                // vectors (columns of X) per work item

template <bool adjoint>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_q2_apply(const c64 *V2, const c64 *tau2, int n, int nj, c64 *X, long ldx,
                                                  int k, int *sync, int *qprog) {
    // Work item (step j, column group cg): the kQC columns cg kQC .. of X, whose pipeline over the
    // steps is independent of the other groups'.  Lane (l, sub) = (tid / 8, tid % 8) owns column
    // cg kQC + l of the window and its rows r = 8 t + sub, held in registers.  The reflector loop is
    // unrolled over the kQG positions of a block, so every register index is a compile-time
    // constant; only the reflectors live in LDS, each in a row with kQP zeros on both sides: a lane
    // reads entry i = 8 t + sub - b (-7 <= i < B + 8) of reflector b at a compile-time offset from
    // its own base, with no clamp or mask (the stored reflectors are zero past their length).
    constexpr int NR = (kQW + 7) / 8;
    constexpr int kQP = 8, kQR = B + 2 * kQP;
    __shared__ c64 sVb[2][kQG][kQR];   // reflectors of block tt in sVb[tt & 1]
    __shared__ c64 sTb[2][kQG];
    __shared__ int s_q, s_ok;
    const int tid = threadIdx.x, l = tid >> 3, sub = tid & 7;
    const int ng = (k + kQC - 1) / kQC;
    for (int idx = tid; idx < 2 * kQG * 2 * kQP; idx += 256) {   // the padding, never written again
        const int buf = idx / (kQG * 2 * kQP), b = (idx / (2 * kQP)) % kQG, pos = idx % (2 * kQP);
        sVb[buf][b][pos < kQP ? pos : B + pos] = mk(0.0, 0.0);
    }
    while (true) {
        if (tid == 0) s_q = atomicAdd(&sync[0], 1);
        __syncthreads();
        const int q = s_q;
        __syncthreads();
        if (q >= nj * ng) return;
        const int cg = q % ng, jq = q / ng;   // items are claimed step-major
        const int j = adjoint ? nj - 1 - jq : jq;
        const int cl = cg * kQC + l;          // this lane's column of X
        const bool col = cl < k;
        const int smax = n - 2 - j * B;
        const int bmax = smax / kQG;
        // the reflectors of a block (final since the chase ended) are staged into the other LDS buffer at
        // the end of the previous block, their loads overlapping the drain of its stores
        constexpr int PV = kQG * B / 256;
        auto stage = [&](int tt, int buf) __attribute__((always_inline)) {
            const int bk = adjoint ? tt : bmax - tt;
            const int s_lo = bk * kQG, cnt = min((bk + 1) * kQG, smax + 1) - s_lo;
            const c64 *vb = V2 + refl_index(n, s_lo, j) * B;   // the block's reflectors are contiguous
            // unconditional (tid + 256 u < kQG * B always); positions past the block's cnt reflectors
            // get tau = 0, so they act as the identity
#pragma unroll
            for (int u = 0; u < PV; ++u) {
                const int idx = tid + 256 * u;
                sVb[buf][idx / B][kQP + idx % B] = vb[min(idx, cnt * B - 1)];
            }
            const c64 t = tau2[refl_index(n, s_lo + min(tid, cnt - 1), j)];
            if (tid < kQG) sTb[buf][tid] = tid < cnt ? (adjoint ? conj(t) : t) : mk(0.0, 0.0);
        };
        stage(0, 0);
        for (int tt = 0; tt <= bmax; ++tt) {
            const int bk = adjoint ? tt : bmax - tt;
            const int s_lo = bk * kQG, cnt = min((bk + 1) * kQG, smax + 1) - s_lo;
            const int row0 = s_lo + 1 + j * B, nr = min(s_lo + cnt + (j + 1) * B, n) - row0;
            c64(*sV)[kQR] = sVb[tt & 1];
            const c64 *sT = sTb[tt & 1];
            if (tid == 0) {
                int ok = 1;
                if (!adjoint && j > 0) {
                    const int bmd = (n - 2 - (j - 1) * B) / kQG;
                    ok = wait_at_least(&qprog[((long)(j - 1) * ng + cg) * kFlagStride], bmd - bk + 1, sync, nullptr, 1);
                } else if (adjoint && j < nj - 1) {
                    const int bmd = (n - 2 - (j + 1) * B) / kQG;
                    ok = wait_at_least(&qprog[((long)(j + 1) * ng + cg) * kFlagStride], min(bk, bmd) + 1, sync, nullptr, 1);
                }
                s_ok = ok;
            }
            __syncthreads();
            if (!s_ok) return;
            c64 x[NR];
#pragma unroll
            for (int t = 0; t < NR; ++t) {
                const int r = 8 * t + sub;
                x[t] = (col && r < nr) ? ld_coh(&X[(long)cl * ldx + row0 + r]) : mk(0.0, 0.0);
            }
            if (col) {
                constexpr int U = B / 8 + 1;   // register rows t = b/8 .. b/8 + U - 1 cover rows b .. b+B-1
                static_assert((kQG - 1) / 8 + U <= NR, "register rows of every reflector exist");
#pragma unroll
                for (int bb = 0; bb < kQG; ++bb) {
                    const int b = adjoint ? bb : kQG - 1 - bb;   // window offset of reflector s_lo + b
                    // v entries of this lane's rows (zero outside the reflector), all LDS reads
                    // issued before the arithmetic; two partial sums shorten the dependent chain
                    c64 vv[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) vv[u] = sV[b][kQP + 8 * (b / 8 + u) - b + sub];
                    c64 pp[2] = {mk(0.0, 0.0), mk(0.0, 0.0)};
#pragma unroll
                    for (int u = 0; u < U; ++u) pp[u & 1] = pp[u & 1] + cmul(vv[u], x[b / 8 + u]);
                    c64 part = pp[0] + pp[1];
                    part.re = sum8(part.re);
                    part.im = sum8(part.im);
                    const c64 tw = sT[b] * part;
#pragma unroll
                    for (int u = 0; u < U; ++u) x[b / 8 + u] = x[b / 8 + u] - vv[u] * tw;
                }
            }
#pragma unroll
            for (int t = 0; t < NR; ++t) {
                const int r = 8 * t + sub;
                if (col && r < nr) st_coh(&X[(long)cl * ldx + row0 + r], x[t]);
            }
            if (tt < bmax) stage(tt + 1, (tt + 1) & 1);   // includes the wait for this block's stores
            wait_vmem();
            __syncthreads();
            if (tid == 0) publish(&qprog[((long)j * ng + cg) * kFlagStride], tt + 1);
        }
    }
}

__global__ void k_quad_sum_probe(const double *in, double *out) {
    out[threadIdx.x] = quad_sum(in[threadIdx.x]);
    out[64 + threadIdx.x] = sum8(in[threadIdx.x]);
    out[128 + threadIdx.x] = sum16(in[threadIdx.x]);
    out[192 + threadIdx.x] = lane0(in[threadIdx.x]);
    out[256 + threadIdx.x] = wave_sum_dpp(in[threadIdx.x]);
}

int persistent_grid(const ChaseWork &c, int work, int per_cu) { return std::max(1, std::min(work, per_cu * c.cus)); }

}  // namespace

void chase_alloc(ChaseWork &c, int n) {
    c.n = n;
    c.nj = n > 1 ? (n - 1 + B - 1) / B : 0;
    c.nrefl = 0;
    int dev = 0;
    c.cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) c.cus = p.multiProcessorCount;
    }
    for (int j = 0; j < c.nj; ++j) c.nrefl += n - 1 - (long)j * B;
    QD_HIP(hipMalloc(&c.AB, sizeof(c64) * (long)std::max(n, 1) * LDB));
    QD_HIP(hipMalloc(&c.V2, sizeof(c64) * std::max(c.nrefl, 1L) * B));
    QD_HIP(hipMalloc(&c.tau2, sizeof(c64) * std::max(c.nrefl, 1L)));
    QD_HIP(hipMalloc(&c.sync, sizeof(int) * 4));
    QD_HIP(hipMalloc(&c.prog, sizeof(int) * ((long)std::max(n, 1) + 1) * kFlagStride));
    QD_HIP(hipMalloc(&c.d, sizeof(double) * std::max(n, 1)));
    QD_HIP(hipMalloc(&c.e, sizeof(double) * std::max(n, 1)));
    QD_HIP(hipMemset(c.sync, 0, sizeof(int) * 4));
}

void chase_free(ChaseWork &c) {
    hipFree(c.AB); hipFree(c.V2); hipFree(c.tau2); hipFree(c.sync); hipFree(c.prog); hipFree(c.d); hipFree(c.e);
    c = ChaseWork();
}

void chase_clear_error(ChaseWork &c, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.sync, 4, 0);
    QD_LAUNCH_CHECK();
}

void chase_run(ChaseWork &c, hipStream_t s) {
    const int n = c.n;
    if (n > 1) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.sync, 2, 0);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_fill_int, dim3((unsigned)(((long)n * kFlagStride + 255) / 256)), dim3(256), 0, s, c.prog, n * kFlagStride, 0);
        QD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_chase, dim3(persistent_grid(c, n - 1, 2)), dim3(256), 0, s, c.AB, n, c.V2, c.tau2, c.sync, c.prog);
        QD_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_band_tridiag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c.AB, n, c.d, c.e);
    QD_LAUNCH_CHECK();
}

void chase_apply(ChaseWork &c, c64 *X, long ldx, int k, bool adjoint, hipStream_t s) {
    if (k > kQK) throw HipError("too many vectors for the band back-transformation", QD_E_UNSUPPORTED);
    if (c.nj == 0 || k < 1) return;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.sync, 2, 0);
    QD_LAUNCH_CHECK();
    // one progress flag per (step, column group); c.prog holds n + 1 >= nj * groups of them
    const int groups = (k + kQC - 1) / kQC, items = c.nj * groups;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)(((long)items * kFlagStride + 255) / 256)), dim3(256), 0, s, c.prog, items * kFlagStride, 0);
    QD_LAUNCH_CHECK();
    if (adjoint)
        hipLaunchKernelGGL(k_q2_apply<true>, dim3(persistent_grid(c, items, 4)), dim3(256), 0, s, c.V2, c.tau2, c.n, c.nj, X,
                           ldx, k, c.sync, c.prog);
    else
        hipLaunchKernelGGL(k_q2_apply<false>, dim3(persistent_grid(c, items, 4)), dim3(256), 0, s, c.V2, c.tau2, c.n, c.nj, X,
                           ldx, k, c.sync, c.prog);
    QD_LAUNCH_CHECK();
}

void quad_sum_selftest(const double *h_in, double *h_out) {
    double *d = nullptr;
    QD_HIP(hipMalloc(&d, sizeof(double) * 384));
    QD_HIP(hipMemcpy(d, h_in, sizeof(double) * 64, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_quad_sum_probe, dim3(1), dim3(64), 0, 0, d, d + 64);
    QD_LAUNCH_CHECK();
    QD_HIP(hipMemcpy(h_out, d + 64, sizeof(double) * 320, hipMemcpyDeviceToHost));
    (void)hipFree(d);
}

void chase_mark_error(ChaseWork &c, hipStream_t s) {
    // the sticky error flag only: no wait reads it, so every kernel still runs to completion
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.sync + 2, 1, 1);
    QD_LAUNCH_CHECK();
}

bool chase_failed(ChaseWork &c, hipStream_t s) {
    int flag = 0;
    QD_HIP(hipMemcpyAsync(&flag, c.sync + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    return flag != 0;
}

}  // namespace qd
