// Kernel set (b), part 2: FP64 eigenpairs of the real symmetric tridiagonal T = (d, e) for the
// index window il..iu.
//   k_gersh     : Gershgorin interval and pivot floor (dstebz conventions).
//   k_bisect    : one wavefront per eigenvalue; the 64 lanes evaluate Sturm counts at 64 shifts
//                 (multisection), so each round shrinks the bracket 65-fold.
//   k_invit     : inverse iteration for all wanted eigenvalues at once: lane l of one wavefront
//                 owns eigenvalue l, factors T - lambda_l I with partial pivoting (dgttrf
//                 semantics) and solves 4 times; d and e are read once per sweep for all lanes.
//   k_cluster_gs: Gram-Schmidt (twice) inside clusters of (near-)degenerate eigenvalues, i.e.
//                 the Kramers pairs, whose vectors inverse iteration leaves only linearly
//                 independent.
#include "qd_internal.hpp"

namespace qd {
namespace {

__device__ inline double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wmin(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// out[0] = lower bound, out[1] = upper bound, out[2] = pivmin, out[3] = ||T||_1
__global__ __launch_bounds__(256) void k_gersh(const double *d, const double *e, double *e2, int n, double *out) {
    __shared__ double sh[4][4];
    double lo = 1e300, hi = -1e300, emax = 0, nrm = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double er = i < n - 1 ? fabs(e[i]) : 0.0;
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        nrm = fmax(nrm, fabs(d[i]) + el + er);
        if (i < n - 1) {
            e2[i] = e[i] * e[i];
            emax = fmax(emax, e2[i]);
        }
    }
    lo = wmin(lo); hi = wmax(hi); emax = wmax(emax); nrm = wmax(nrm);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[w][0] = lo; sh[w][1] = hi; sh[w][2] = emax; sh[w][3] = nrm; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; ++q) {
            lo = fmin(lo, sh[q][0]); hi = fmax(hi, sh[q][1]); emax = fmax(emax, sh[q][2]); nrm = fmax(nrm, sh[q][3]);
        }
        const double safmin = 2.2250738585072014e-308, ulp = 2.220446049250313e-16;
        const double pivmin = safmin * fmax(1.0, emax);
        const double tnorm = fmax(fabs(lo), fabs(hi));
        const double pad = 2.1 * ulp * tnorm * n + 2.1 * pivmin + 1e-12 * tnorm;
        out[0] = lo - pad;
        out[1] = hi + pad;
        out[2] = pivmin;
        out[3] = nrm;
    }
}

__device__ inline double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Number of eigenvalues of T below x (negative LDL^T pivots); called by all 64 lanes of a
// wavefront, each with its own x.  d and e2 are the same for every lane, so the wavefront loads
// them in chunks of 64 (lane l: element i0 + l, the next chunk in flight while this one is used)
// and the recurrence reads the chunk with v_readlane: one memory round trip per 64 steps instead
// of one per step.  e2 holds max(n - 1, 1) entries.
__device__ inline int sturm(const double *__restrict__ d, const double *__restrict__ e2, int n, double x, double pivmin) {
    const int lane = threadIdx.x & 63, ne = max(n - 2, 0);
    double dn = d[min(lane, n - 1)], en = e2[min(max(lane - 1, 0), ne)];   // e2[i - 1] for i = lane
    double q = 1.0;
    int cnt = 0;
    for (int i0 = 0; i0 < n; i0 += 64) {
        const double dc = dn, ec = en;
        const int i1 = i0 + 64 + lane;
        dn = d[min(i1, n - 1)];
        en = e2[min(i1 - 1, ne)];
#pragma unroll
        for (int j = 0; j < 64; ++j) {
            if (i0 + j >= n) break;
            const double dj = readlane_d(dc, j);
            q = (i0 + j == 0) ? dj - x : dj - x - readlane_d(ec, j) / q;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q <= 0.0;
        }
    }
    return cnt;
}

__global__ __launch_bounds__(64) void k_bisect(const double *d, const double *e2, int n, const double *gb, int il,
                                               double *w, int rounds, int *flags) {
    const int k = il + blockIdx.x;   // 1-based index of the wanted eigenvalue
    const int lane = threadIdx.x;
    double lo = gb[0], hi = gb[1];
    const double pivmin = gb[2];
    const double ulp = 2.220446049250313e-16;
    bool conv = false;
    for (int round = 0; round < rounds; ++round) {
        const double width = hi - lo;
        if (width <= 2.0 * ulp * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin) { conv = true; break; }
        const double x = lo + (lane + 1) * (width / 65.0);
        const int cnt = sturm(d, e2, n, x, pivmin);
        const unsigned long long ge = __ballot(cnt >= k);
        if (ge == 0ull) {
            lo = __shfl(x, 63, 64);
        } else {
            const int first = __ffsll((long long)ge) - 1;
            const double xf = __shfl(x, first, 64);
            const double xp = first > 0 ? __shfl(x, first - 1, 64) : lo;
            lo = first > 0 ? xp : lo;
            hi = xf;
        }
        if (hi - lo >= width) { conv = true; break; }   // no progress (bracket at rounding level)
    }
    if (!conv && hi - lo <= 2.0 * ulp * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin) conv = true;
    if (lane == 0) {
        w[blockIdx.x] = 0.5 * (lo + hi);
        if (!conv) atomicAdd(&flags[0], 1);
    }
}

// Inverse iteration: lane l < k owns eigenvalue w[l].  LU scratch is interleaved [i][field][lane]
// (fields DL, 1/D, DU, DU2, x).  The sequential sweeps over i read their operands in chunks of
// kIC rows issued together, so a chunk waits for one memory latency instead of kIC.
constexpr int kIC = 16;
__global__ __launch_bounds__(64) void k_invit(const double *d, const double *e, int n, const double *w, int k,
                                              const double *gb, double *lu, int *ipiv, double *X, int sweeps) {
    const int l = threadIdx.x;
    const bool act = l < k;
    const double tnorm = gb[3];
    const double eps = 2.220446049250313e-16;
    const double tiny = eps * fmax(tnorm, 1e-300);
    double lam = act ? w[l] : 0.0;
    // separate exactly equal shifts slightly so the factorizations differ (dstein)
    if (act && l > 0 && fabs(w[l] - w[l - 1]) < 10.0 * eps * fabs(w[l])) lam = lam + 10.0 * eps * fmax(fabs(lam), tnorm * 1e-6);
    auto L = [&](long i, int f) -> double & { return lu[(i * 5 + f) * 64 + l]; };
    // --- factorization T - lam I = P L U (dgttrf) ---
    {
        double Dcur = d[0] - lam;
        double DUcur = n > 1 ? e[0] : 0.0;
        for (int i0 = 0; i0 < n - 1; i0 += kIC) {
            double ei[kIC], dn[kIC], en[kIC];
#pragma unroll
            for (int u = 0; u < kIC; ++u) {
                const int i = i0 + u;
                ei[u] = i < n - 1 ? e[i] : 0.0;
                dn[u] = i < n - 1 ? d[i + 1] : 0.0;
                en[u] = i + 1 < n - 1 ? e[i + 1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kIC; ++u) {
                const int i = i0 + u;
                if (i < n - 1) {
                    const double DLi = ei[u];
                    const double Dn = dn[u] - lam;
                    const double DUn = en[u];
                    double Dv = Dcur, DUv = DUcur, DLv, Dnext, DUnext, DU2v = 0.0;
                    int piv = i;
                    if (fabs(Dv) >= fabs(DLi)) {
                        if (fabs(Dv) < tiny) Dv = Dv >= 0 ? tiny : -tiny;
                        const double fact = DLi / Dv;
                        DLv = fact;
                        Dnext = Dn - fact * DUv;
                        DUnext = DUn;
                    } else {
                        const double fact = Dv / DLi;
                        Dv = DLi;
                        DLv = fact;
                        const double temp = DUv;
                        DUv = Dn;
                        Dnext = temp - fact * Dn;
                        DU2v = DUn;
                        DUnext = -fact * DUn;
                        piv = i + 1;
                    }
                    if (act) {
                        L(i, 0) = DLv; L(i, 1) = 1.0 / Dv; L(i, 2) = DUv; L(i, 3) = DU2v;
                        ipiv[(long)i * 64 + l] = piv;
                    }
                    Dcur = Dnext;
                    DUcur = DUnext;
                }
            }
        }
        if (fabs(Dcur) < tiny) Dcur = Dcur >= 0 ? tiny : -tiny;
        if (act) { L(n - 1, 1) = 1.0 / Dcur; L(n - 1, 2) = 0.0; L(n - 1, 3) = 0.0; L(n - 1, 0) = 0.0; }
    }
    // --- start vector: deterministic pseudo-random ---
    double nrm2 = 0.0;
    for (int i = 0; i < n; ++i) {
        unsigned long long h = (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull ^ (unsigned long long)(l + 7) * 0xC2B2AE3D27D4EB4Full;
        h ^= h >> 31; h *= 0xD6E8FEB86659FD93ull; h ^= h >> 32;
        const double v = ((double)(h >> 11) * (1.0 / 9007199254740992.0)) - 0.5;
        if (act) L(i, 4) = v;
        nrm2 += v * v;
    }
    double scale = 1.0 / sqrt(nrm2);
    if (!act) return;
    for (int it = 0; it < sweeps; ++it) {
        // forward: apply P and L^{-1} (b scaled by `scale`)
        double bi = L(0, 4) * scale;
        for (int i0 = 0; i0 < n - 1; i0 += kIC) {
            double bn[kIC], dl[kIC];
            int pv[kIC];
#pragma unroll
            for (int u = 0; u < kIC; ++u) {
                const int i = i0 + u;
                const bool in = i < n - 1;
                bn[u] = in ? L(i + 1, 4) : 0.0;
                dl[u] = in ? L(i, 0) : 0.0;
                pv[u] = in ? ipiv[(long)i * 64 + l] : i;
            }
#pragma unroll
            for (int u = 0; u < kIC; ++u) {
                const int i = i0 + u;
                if (i < n - 1) {
                    const double b1 = bn[u] * scale;
                    double out_i, next;
                    if (pv[u] == i) { out_i = bi; next = b1 - dl[u] * bi; }
                    else { out_i = b1; next = bi - dl[u] * b1; }
                    L(i, 4) = out_i;
                    bi = next;
                }
            }
        }
        L(n - 1, 4) = bi;
        // backward: U x = b
        double x1 = 0.0, x2 = 0.0;   // x[i+1], x[i+2]
        nrm2 = 0.0;
        for (int i0 = n - 1; i0 >= 0; i0 -= kIC) {
            double rd[kIC], du[kIC], du2[kIC], bb[kIC];
#pragma unroll
            for (int u = 0; u < kIC; ++u) {
                const int i = i0 - u;
                const bool in = i >= 0;
                rd[u] = in ? L(i, 1) : 0.0;
                du[u] = in ? L(i, 2) : 0.0;
                du2[u] = in ? L(i, 3) : 0.0;
                bb[u] = in ? L(i, 4) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kIC; ++u) {
                const int i = i0 - u;
                if (i >= 0) {
                    const double v = (bb[u] - du[u] * x1 - du2[u] * x2) * rd[u];
                    L(i, 4) = v;
                    nrm2 += v * v;
                    x2 = x1;
                    x1 = v;
                }
            }
        }
        scale = nrm2 > 0 ? 1.0 / sqrt(nrm2) : 1.0;
    }
    for (int i0 = 0; i0 < n; i0 += kIC) {   // chunks of loads in flight, as in the sweeps
        double xv[kIC];
#pragma unroll
        for (int u = 0; u < kIC; ++u) xv[u] = i0 + u < n ? L(i0 + u, 4) : 0.0;
#pragma unroll
        for (int u = 0; u < kIC; ++u)
            if (i0 + u < n) X[(long)l * n + i0 + u] = xv[u] * scale;
    }
}

__global__ void k_zero_ints(int *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Gram-Schmidt (classical, twice) within clusters |w_u - w_v| <= ctol, in index order.
__global__ __launch_bounds__(256) void k_cluster_gs(double *X, int n, const double *w, int k, const double *gb) {
    __shared__ double red[4];
    const double ctol = 1e-9 * gb[3];
    for (int v = 0; v < k; ++v) {
        int u0 = v;
        while (u0 > 0 && fabs(w[u0 - 1] - w[u0]) <= ctol) --u0;
        double *xv = X + (long)v * n;
        for (int pass = 0; pass < 2; ++pass) {
            for (int u = u0; u < v; ++u) {
                const double *xu = X + (long)u * n;
                double s = 0;
                for (int i = threadIdx.x; i < n; i += 256) s += xu[i] * xv[i];
                s = wsum(s);
                if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
                __syncthreads();
                s = red[0] + red[1] + red[2] + red[3];
                __syncthreads();
                for (int i = threadIdx.x; i < n; i += 256) xv[i] -= s * xu[i];
                __syncthreads();
            }
            double s = 0;
            for (int i = threadIdx.x; i < n; i += 256) s += xv[i] * xv[i];
            s = wsum(s);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
            __syncthreads();
            s = red[0] + red[1] + red[2] + red[3];
            __syncthreads();
            const double inv = 1.0 / sqrt(s);
            for (int i = threadIdx.x; i < n; i += 256) xv[i] *= inv;
            __syncthreads();
        }
    }
}

// Residual of every computed pair: res_l = ||(T - w_l) x_l||_2 / ||T||_1 (x_l has unit norm).  A
// pair above kTriResTol counts as not converged (flags[1]); out[l] keeps the residual.
constexpr double kTriResTol = 1e-11;
__global__ __launch_bounds__(256) void k_tri_residual(const double *d, const double *e, const double *X, int n,
                                                      const double *w, const double *gb, double *out, int *flags) {
    __shared__ double red[4];
    const int l = blockIdx.x;
    const double *x = X + (long)l * n;
    const double lam = w[l];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double r = (d[i] - lam) * x[i];
        if (i > 0) r += e[i - 1] * x[i - 1];
        if (i < n - 1) r += e[i] * x[i + 1];
        s += r * r;
    }
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double res = sqrt(red[0] + red[1] + red[2] + red[3]) / fmax(gb[3], 1e-300);
        out[l] = res;
        if (!(res <= kTriResTol)) atomicAdd(&flags[1], 1);
    }
}

__global__ void k_r2c(const double *Zr, int n, int k, c64 *X, long ld) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * k) return;
    const int col = (int)(idx / n), r = (int)(idx % n);
    X[(long)col * ld + r] = mk(Zr[idx], 0.0);
}

}  // namespace

void tridiag_alloc(TridiagWork &t, int n, int k) {
    t.n = n;
    t.k = k;
    QD_HIP(hipMalloc(&t.e2, sizeof(double) * (n > 1 ? n : 1)));
    QD_HIP(hipMalloc(&t.w, sizeof(double) * (k + 4)));
    QD_HIP(hipMalloc(&t.Z, sizeof(double) * (long)n * k));
    QD_HIP(hipMalloc(&t.lu, sizeof(double) * (long)n * 5 * 64));
    QD_HIP(hipMalloc(&t.ipiv, sizeof(int) * (long)n * 64));
    QD_HIP(hipMalloc(&t.info, sizeof(double) * 4));
    QD_HIP(hipMalloc(&t.flags, sizeof(int) * 2));
    QD_HIP(hipMalloc(&t.res, sizeof(double) * (k > 0 ? k : 1)));
}

void tridiag_free(TridiagWork &t) {
    hipFree(t.e2); hipFree(t.w); hipFree(t.Z); hipFree(t.lu); hipFree(t.ipiv); hipFree(t.info);
    hipFree(t.flags); hipFree(t.res);
    t = TridiagWork();
}

void tridiag_solve(TridiagWork &t, const double *d, const double *e, int il, int iu, hipStream_t s,
                   int bisect_rounds, int invit_sweeps) {
    const int k = iu - il + 1;
    if (k > 64) throw HipError("at most 64 eigenpairs per solve", QD_E_UNSUPPORTED);
    double *gb = (double *)t.info;   // 4 doubles
    // flags are cleared by a kernel, not a memset node, so the sequence stays capturable
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, t.flags, 2);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gersh, dim3(1), dim3(256), 0, s, d, e, t.e2, t.n, gb);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bisect, dim3(k), dim3(64), 0, s, d, t.e2, t.n, gb, il, t.w, bisect_rounds, t.flags);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_invit, dim3(1), dim3(64), 0, s, d, e, t.n, t.w, k, gb, t.lu, t.ipiv, t.Z, invit_sweeps);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cluster_gs, dim3(1), dim3(256), 0, s, t.Z, t.n, t.w, k, gb);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tri_residual, dim3(k), dim3(256), 0, s, d, e, t.Z, t.n, t.w, gb, t.res, t.flags);
    QD_LAUNCH_CHECK();
}

void launch_real_to_complex(const double *Zr, int n, int k, c64 *X, long ld, hipStream_t s) {
    const long tot = (long)n * k;
    hipLaunchKernelGGL(k_r2c, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, Zr, n, k, X, ld);
    QD_LAUNCH_CHECK();
}

}  // namespace qd
