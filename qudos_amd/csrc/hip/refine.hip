// Kernel set (b), part 4: FP64 refinement of the FP32-reduction eigenpairs.
//
// Each sweep (all in FP64 except the reflectors Q, which stay FP32):
//   R   = H X - X Lambda                       (H stored in FP64, GEMM kernel)
//   E_i = Q (T - lambda_i)^{-1} Q^H R_i         (FP32 factorisation as an approximate inverse)
//   E  <- E orthogonalised against X, then orthonormalised (Cholesky QR, twice)
//   Rayleigh-Ritz of H on span[X, E] (Hermitian Jacobi on the 2k x 2k projection), keeping the
//   k Ritz pairs with the largest weight on the previous X.
// This is the clustered-eigenvalue form of Ogita-Aishima refinement: the Kramers pairs are
// treated as 2-dimensional invariant subspaces by the Rayleigh-Ritz step, so no individual
// eigenvector of a degenerate pair needs to converge.
#include "qd_internal.hpp"
#include "qd_refine.hpp"

namespace qd {
namespace {

constexpr int kM = 80;    // max Rayleigh-Ritz dimension (2k)

__device__ inline double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Y(n x m) = H(n x n) B(n x m), split over column segments: work-group (I, seg) sums rows I*64 ..
// +63 over the 16-column chunks of segment seg into Ypart[seg] (k_sum_seg adds the segments).
// The next chunk of H and B is loaded into registers (clamped addresses, values masked when they
// are stored to LDS, so no load sits in a branch) while the current chunk is used.  Wavefront cg
// computes columns cg + 4 t, t < MT, unconditionally (the LDS columns >= m are zero); a per-column
// test would put every product in a branch of its own.
constexpr int kGemmSegMax = 8;
template <int MT>
__global__ __launch_bounds__(256) void k_gemm_h(const c64 *H, long ld, int n, const c64 *B, long ldb, int m, int per,
                                                c64 *Ypart) {
    static_assert(4 * MT <= kM, "columns within the LDS tile");
    __shared__ c64 sH[16][65];
    __shared__ c64 sB[16][kM + 1];
    constexpr int NH = 64 * 16 / 256, NB = (16 * kM + 255) / 256;
    const int tid = threadIdx.x;
    const int row = tid & 63, cg = tid >> 6;
    const int r0 = blockIdx.x * 64, seg = blockIdx.y;
    const int c_begin = seg * per * 16, c_end = min(n, c_begin + per * 16);
    c64 rh[NH], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NH; ++u) {
            const int idx = tid + 256 * u, r = r0 + (idx & 63), c = k0 + (idx >> 6);
            rh[u] = H[(long)min(c, n - 1) * ld + min(r, n - 1)];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int idx = tid + 256 * u, c = k0 + idx % 16, l = idx / 16;
            rb[u] = B[(long)min(l, m - 1) * ldb + min(c, n - 1)];
        }
    };
    auto store = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NH; ++u) {
            const int idx = tid + 256 * u, rr = idx & 63, kk = idx >> 6;
            sH[kk][rr] = (r0 + rr < n && k0 + kk < c_end) ? rh[u] : mk(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int idx = tid + 256 * u, kk = idx % 16, l = idx / 16;   // idx < 16 kM
            sB[kk][l] = (l < m && k0 + kk < c_end) ? rb[u] : mk(0.0, 0.0);
        }
    };
    c64 acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = mk(0.0, 0.0);
    if (c_begin < c_end) load(c_begin);
    for (int k0 = c_begin; k0 < c_end; k0 += 16) {
        store(k0);
        __syncthreads();
        load(k0 + 16);   // the next chunk (clamped past the end): in flight during the products
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            const c64 h = sH[kk][row];
#pragma unroll
            for (int t = 0; t < MT; ++t) fma_(acc[t], h, sB[kk][cg + 4 * t]);
        }
        __syncthreads();
    }
    const int R = r0 + row;
    if (R < n)
#pragma unroll
        for (int t = 0; t < MT; ++t)
            if (cg + 4 * t < m) Ypart[((long)seg * m + cg + 4 * t) * ld + R] = acc[t];
}

// Y(:, l) = sum over the segments of Ypart
__global__ void k_sum_seg(const c64 *Ypart, int S, long ld, int n, int m, c64 *Y, long ldy) {
    const long o = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (o >= (long)n * m) return;
    const int l = (int)(o / n), r = (int)(o % n);
    c64 y = mk(0.0, 0.0);
    for (int sg = 0; sg < S; ++sg) y = y + Ypart[((long)sg * m + l) * ld + r];
    Y[(long)l * ldy + r] = y;
}

// Split-K partials of C = A^H B (A: n x ma, B: n x mb), chunks of 512 rows.
__global__ __launch_bounds__(256) void k_atb(const c64 *A, long lda, int ma, const c64 *B, long ldb, int mb, int n,
                                             c64 *part) {
    __shared__ c64 sA[32][kM + 1];
    __shared__ c64 sB[32][kM + 1];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * 512;
    const int nout = ma * mb;
    c64 acc[25];
    for (int t = 0; t < 25; ++t) acc[t] = mk(0.0, 0.0);
    for (int s = 0; s < 512; s += 32) {
        for (int idx = tid; idx < 32 * ma; idx += 256) {
            const int rr = idx & 31, l = idx >> 5;
            const int r = r0 + s + rr;
            sA[rr][l] = r < n ? A[(long)l * lda + r] : mk(0.0, 0.0);
        }
        for (int idx = tid; idx < 32 * mb; idx += 256) {
            const int rr = idx & 31, l = idx >> 5;
            const int r = r0 + s + rr;
            sB[rr][l] = r < n ? B[(long)l * ldb + r] : mk(0.0, 0.0);
        }
        __syncthreads();
        for (int t = 0; t < 25; ++t) {
            const int o = tid + 256 * t;
            if (o >= nout) break;
            const int i = o % ma, j = o / ma;
            c64 a = acc[t];
            for (int rr = 0; rr < 32; ++rr) fma_(a, conj(sA[rr][i]), sB[rr][j]);
            acc[t] = a;
        }
        __syncthreads();
    }
    for (int t = 0; t < 25; ++t) {
        const int o = tid + 256 * t;
        if (o >= nout) break;
        part[(long)blockIdx.x * kM * kM + o] = acc[t];
    }
}

__global__ void k_sum_parts(const c64 *part, int nchunk, int nout, c64 *C) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    c64 s = mk(0.0, 0.0);
    for (int ch = 0; ch < nchunk; ++ch) s = s + part[(long)ch * kM * kM + o];
    C[o] = s;
}

// Y(n x mb) = beta*Y + alpha*A(n x ma) S(ma x mb), S column-major ma x mb (small, in LDS)
__global__ __launch_bounds__(256) void k_ab(const c64 *A, long lda, int ma, const c64 *S, int mb, int n, double alpha,
                                            double beta, c64 *Y, long ldy) {
    __shared__ c64 sS[kM][kM + 1];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < ma * mb; idx += 256) sS[idx % ma][idx / ma] = S[idx];
    __syncthreads();
    const int r = blockIdx.x * 256 + tid;
    if (r >= n) return;
    for (int j0 = 0; j0 < mb; j0 += 16) {
        c64 acc[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) acc[t] = mk(0.0, 0.0);
        for (int i = 0; i < ma; ++i) {
            const c64 a = A[(long)i * lda + r];
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (j0 + t < mb) fma_(acc[t], a, sS[i][j0 + t]);
        }
        for (int t = 0; t < 16 && j0 + t < mb; ++t) {
            const long iy = (long)(j0 + t) * ldy + r;
            Y[iy] = beta == 0.0 ? alpha * acc[t] : beta * Y[iy] + alpha * acc[t];
        }
    }
}

// R = HX - X diag(lam); also per-column ||R||^2 partials
__global__ void k_resid(const c64 *HX, const c64 *X, long ld, int n, int k, const double *lam, c64 *R) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * k) return;
    const int l = (int)(idx / n), r = (int)(idx % n);
    const long i = (long)l * ld + r;
    R[i] = HX[i] - lam[l] * X[i];
}

// Column 2-norms of an n x k block: one work-group per column.
__global__ __launch_bounds__(256) void k_colnorm(const c64 *X, long ld, int n, double *out) {
    __shared__ double red[4];
    double s = 0;
    for (int r = threadIdx.x; r < n; r += 256) s += norm2(X[(long)blockIdx.x * ld + r]);
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// (T - lam_l I) s_l = rhs_l for complex rhs, lane l (dgttrf / dgtts2 semantics, tiny-pivot guard)
// (T - lam_l I) x_l = s_l for every lane l < k (dgttrf/dgttrs semantics, partial pivoting); the
// sequential sweeps read their operands in chunks of kTC rows issued together.
constexpr int kTC = 16;
__global__ __launch_bounds__(64) void k_trisolve(const double *d, const double *e, int n, const double *lam, int k,
                                                 double tnorm_floor, double *lu, int *ipiv, c64 *S, long ld) {
    const int l = threadIdx.x;
    const bool act = l < k;
    const double eps = 2.220446049250313e-16;
    const double tiny = eps * tnorm_floor;
    const double lm = act ? lam[l] : 0.0;
    auto L = [&](long i, int f) -> double & { return lu[(i * 4 + f) * 64 + l]; };
    double Dcur = d[0] - lm, DUcur = n > 1 ? e[0] : 0.0;
    for (int i0 = 0; i0 < n - 1; i0 += kTC) {
        double ei[kTC], dn[kTC], en[kTC];
#pragma unroll
        for (int u = 0; u < kTC; ++u) {
            const int i = i0 + u;
            ei[u] = i < n - 1 ? e[i] : 0.0;
            dn[u] = i < n - 1 ? d[i + 1] : 0.0;
            en[u] = i + 1 < n - 1 ? e[i + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kTC; ++u) {
            const int i = i0 + u;
            if (i < n - 1) {
                const double DLi = ei[u], Dn = dn[u] - lm, DUn = en[u];
                double Dv = Dcur, DUv = DUcur, DLv, Dnext, DUnext, DU2v = 0.0;
                int piv = i;
                if (fabs(Dv) >= fabs(DLi)) {
                    if (fabs(Dv) < tiny) Dv = Dv >= 0 ? tiny : -tiny;
                    DLv = DLi / Dv;
                    Dnext = Dn - DLv * DUv;
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
                if (act) { L(i, 0) = DLv; L(i, 1) = 1.0 / Dv; L(i, 2) = DUv; L(i, 3) = DU2v; ipiv[(long)i * 64 + l] = piv; }
                Dcur = Dnext;
                DUcur = DUnext;
            }
        }
    }
    if (fabs(Dcur) < tiny) Dcur = Dcur >= 0 ? tiny : -tiny;
    if (act) { L(n - 1, 1) = 1.0 / Dcur; L(n - 1, 2) = 0.0; L(n - 1, 3) = 0.0; }
    if (!act) return;
    c64 *x = S + (long)l * ld;
    c64 bi = x[0];
    for (int i0 = 0; i0 < n - 1; i0 += kTC) {
        c64 bn[kTC];
        double dl[kTC];
        int pv[kTC];
#pragma unroll
        for (int u = 0; u < kTC; ++u) {
            const int i = i0 + u;
            const bool in = i < n - 1;
            bn[u] = in ? x[i + 1] : mk(0.0, 0.0);
            dl[u] = in ? L(i, 0) : 0.0;
            pv[u] = in ? ipiv[(long)i * 64 + l] : i;
        }
#pragma unroll
        for (int u = 0; u < kTC; ++u) {
            const int i = i0 + u;
            if (i < n - 1) {
                if (pv[u] == i) { x[i] = bi; bi = bn[u] - dl[u] * bi; }
                else { x[i] = bn[u]; bi = bi - dl[u] * bn[u]; }
            }
        }
    }
    x[n - 1] = bi;
    c64 x1 = mk(0.0, 0.0), x2 = mk(0.0, 0.0);
    for (int i0 = n - 1; i0 >= 0; i0 -= kTC) {
        double rd[kTC], du[kTC], du2[kTC];
        c64 bb[kTC];
#pragma unroll
        for (int u = 0; u < kTC; ++u) {
            const int i = i0 - u;
            const bool in = i >= 0;
            rd[u] = in ? L(i, 1) : 0.0;
            du[u] = in ? L(i, 2) : 0.0;
            du2[u] = in ? L(i, 3) : 0.0;
            bb[u] = in ? x[i] : mk(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < kTC; ++u) {
            const int i = i0 - u;
            if (i >= 0) {
                const c64 v = rd[u] * (bb[u] - du[u] * x1 - du2[u] * x2);
                x[i] = v;
                x2 = x1;
                x1 = v;
            }
        }
    }
}

// G (m x m Hermitian positive semidefinite) -> Rinv (m x m upper) with G = R^H R; columns whose
// pivot falls below 1e-20 of their diagonal are dropped (zero column of Rinv).
__global__ __launch_bounds__(256) void k_chol_inv(const c64 *G, int m, c64 *Rinv) {
    __shared__ c64 Rs[kM][kM + 1];
    __shared__ int keep[kM];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < m * m; idx += 256) {
        const int a = idx % m, b = idx / m;
        Rs[a][b] = G[idx];
    }
    __syncthreads();
    // right-looking Cholesky, upper factor: Rs(j,j) = sqrt(...), Rs(j, j+1:) = ... / Rs(j,j)
    for (int j = 0; j < m; ++j) {
        if (tid == 0) {
            const double gjj = G[j * m + j].re;
            const double p = Rs[j][j].re;
            keep[j] = (p > 1e-20 * gjj && p > 0.0);
            Rs[j][j] = mk(keep[j] ? sqrt(p) : 1.0, 0.0);
        }
        __syncthreads();
        const double inv = 1.0 / Rs[j][j].re;
        for (int b = j + 1 + tid; b < m; b += 256) Rs[j][b] = keep[j] ? inv * Rs[j][b] : mk(0.0, 0.0);
        __syncthreads();
        for (int idx = tid; idx < m * m; idx += 256) {
            const int a = idx % m, b = idx / m;
            if (a > j && b >= a) Rs[a][b] = Rs[a][b] - cmul(Rs[j][a], Rs[j][b]);
        }
        __syncthreads();
    }
    // invert the upper triangle column by column (thread b owns column b of Rinv, in global memory)
    for (int b = tid; b < m; b += 256) {
        c64 *col = Rinv + (long)b * m;
        for (int a = b + 1; a < m; ++a) col[a] = mk(0.0, 0.0);
        col[b] = mk(keep[b] ? 1.0 / Rs[b][b].re : 0.0, 0.0);
        for (int a = b - 1; a >= 0; --a) {
            c64 s = mk(0.0, 0.0);
            for (int q = a + 1; q <= b; ++q) s = s + Rs[a][q] * col[q];
            col[a] = keep[a] ? (-1.0 / Rs[a][a].re) * s : mk(0.0, 0.0);
        }
    }
}

// Hermitian Jacobi eigensolver for the m x m projection M (m <= kM; an odd m is padded with one
// decoupled zero row and column, which has no weight on the basis), then selection of the k Ritz
// pairs with the largest weight on the first k basis vectors.  Outputs U_sel (m x k, column-major)
// and lam (k ascending).
__global__ __launch_bounds__(256) void k_jacobi_select(const c64 *Min, int m_in, int k, c64 *Usel, double *lam,
                                                      c64 *Ug) {
    const int m = m_in + (m_in & 1);
    __shared__ c64 M[kM][kM + 1];
    auto U = [&](int a, int b) -> c64 & { return Ug[(long)b * kM + a]; };
    __shared__ double cs[kM / 2], sn[kM / 2];
    __shared__ c64 ph[kM / 2];
    __shared__ int pp[kM / 2], qq[kM / 2];
    __shared__ double off_s;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < m * m; idx += 256) {
        const int a = idx % m, b = idx / m;
        // Hermitian part
        c64 v = mk(0.0, 0.0);
        if (a < m_in && b < m_in) v = 0.5 * (Min[b * m_in + a] + conj(Min[a * m_in + b]));
        if (a == b) v.im = 0.0;
        M[a][b] = v;
        U(a, b) = mk(a == b ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    const int half = m / 2;
    // round-robin pairing: after `step` rotations of positions 1 .. m-1, position i >= 1 holds
    // 1 + ((i - 1 - step) mod (m - 1)); a sweep of m - 1 steps returns to the identity
    auto at_pos = [&](int i, int step) { return i == 0 ? 0 : 1 + ((i - 1 - step) % (m - 1) + (m - 1)) % (m - 1); };
    __shared__ double red_s[8];
    for (int sweep = 0; sweep < 30; ++sweep) {
        {   // off-diagonal and total Frobenius mass, reduced over the work-group
            double off = 0, tot = 0;
            for (int idx = tid; idx < m * m; idx += 256) {
                const int a = idx % m, b = idx / m;
                const double x = norm2(M[a][b]);
                tot += x;
                if (a != b) off += x;
            }
            off = wsum(off);
            tot = wsum(tot);
            if ((tid & 63) == 0) { red_s[tid >> 6] = off; red_s[4 + (tid >> 6)] = tot; }
            __syncthreads();
            if (tid == 0) {
                const double o = (red_s[0] + red_s[1]) + (red_s[2] + red_s[3]);
                const double t = (red_s[4] + red_s[5]) + (red_s[6] + red_s[7]);
                off_s = (o <= 1e-30 * t) ? -1.0 : o;
            }
        }
        __syncthreads();
        if (off_s < 0) break;
        for (int step = 0; step < m - 1; ++step) {
            if (tid < half) {
                int p = at_pos(tid, step), q = at_pos(m - 1 - tid, step);
                if (p > q) { const int t = p; p = q; q = t; }
                pp[tid] = p;
                qq[tid] = q;
                const c64 g = M[p][q];
                const double ag = sqrt(norm2(g));
                if (ag < 1e-300) {
                    cs[tid] = 1.0; sn[tid] = 0.0; ph[tid] = mk(1.0, 0.0);
                } else {
                    const double tau = (M[q][q].re - M[p][p].re) / (2.0 * ag);
                    const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    const double c = 1.0 / sqrt(1.0 + t * t);
                    cs[tid] = c;
                    sn[tid] = t * c;
                    ph[tid] = mk(g.re / ag, g.im / ag);   // e^{i phi}
                }
            }
            __syncthreads();
            // rows: M[p,:] <- c M[p,:] - s e^{i phi} M[q,:];  M[q,:] <- s M[p,:] + c e^{i phi} M[q,:]
            for (int idx = tid; idx < half * m; idx += 256) {
                const int pr = idx / m, col = idx % m;
                const int p = pp[pr], q = qq[pr];
                const double c = cs[pr], s = sn[pr];
                const c64 e = ph[pr];
                const c64 mp = M[p][col], mq = M[q][col];
                M[p][col] = c * mp - s * (e * mq);
                M[q][col] = s * mp + c * (e * mq);
            }
            __syncthreads();
            // columns: M[:,p] <- c M[:,p] - s e^{-i phi} M[:,q];  M[:,q] <- s M[:,p] + c e^{-i phi} M[:,q]
            for (int idx = tid; idx < half * m; idx += 256) {
                const int pr = idx / m, row = idx % m;
                const int p = pp[pr], q = qq[pr];
                const double c = cs[pr], s = sn[pr];
                const c64 e = conj(ph[pr]);
                const c64 mp = M[row][p], mq = M[row][q];
                M[row][p] = c * mp - s * (e * mq);
                M[row][q] = s * mp + c * (e * mq);
                const c64 up = U(row, p), uq = U(row, q);
                U(row, p) = c * up - s * (e * uq);
                U(row, q) = s * up + c * (e * uq);
            }
            __syncthreads();
        }
    }
    // selection: weight of Ritz vector j on the first k basis vectors
    __shared__ double wgt[kM], ev[kM];
    __shared__ int sel[kM];
    for (int j = tid; j < m; j += 256) {
        double s = 0;
        for (int a = 0; a < k; ++a) s += norm2(U(a, j));
        wgt[j] = s;
        ev[j] = M[j][j].re;
    }
    __syncthreads();
    if (tid == 0) {
        // pick k largest weights, then sort them by eigenvalue
        int used[kM];
        for (int j = 0; j < m; ++j) used[j] = 0;
        for (int t = 0; t < k; ++t) {
            int best = -1;
            for (int j = 0; j < m; ++j)
                if (!used[j] && (best < 0 || wgt[j] > wgt[best])) best = j;
            used[best] = 1;
            sel[t] = best;
        }
        for (int a = 1; a < k; ++a) {
            const int v = sel[a];
            int b = a - 1;
            while (b >= 0 && ev[sel[b]] > ev[v]) { sel[b + 1] = sel[b]; --b; }
            sel[b + 1] = v;
        }
    }
    __syncthreads();
    for (int idx = tid; idx < m_in * k; idx += 256) {
        const int a = idx % m_in, t = idx / m_in;
        Usel[idx] = U(a, sel[t]);
    }
    for (int t = tid; t < k; t += 256) lam[t] = ev[sel[t]];
}

}  // namespace

void RefineCtx::alloc(int n_, long ld_, int k_) {
    n = n_;
    ld = ld_;
    k = k_;
    const long nb = ld * 2 * k;
    QD_HIP(hipMalloc(&B, sizeof(c64) * nb));
    QD_HIP(hipMalloc(&HB, sizeof(c64) * nb));
    QD_HIP(hipMalloc(&R, sizeof(c64) * ld * k));
    QD_HIP(hipMalloc(&T2, sizeof(c64) * ld * k));
    QD_HIP(hipMalloc(&small, sizeof(c64) * 4 * kM * kM));
    QD_HIP(hipMalloc(&part, sizeof(c64) * (long)((n + 511) / 512 + 1) * kM * kM));
    ypart_cols = kGemmSegMax * std::max(k, 4);
    QD_HIP(hipMalloc(&ypart, sizeof(c64) * ld * ypart_cols));
    QD_HIP(hipMalloc(&lam, sizeof(double) * 2 * kM));
    QD_HIP(hipMalloc(&norms, sizeof(double) * 2 * kM));
    QD_HIP(hipMalloc(&lu, sizeof(double) * (long)n * 4 * 64));
    QD_HIP(hipMalloc(&ipiv, sizeof(int) * (long)n * 64));
    QD_HIP(hipMemset(B, 0, sizeof(c64) * nb));
    QD_HIP(hipMemset(HB, 0, sizeof(c64) * nb));
}

void RefineCtx::release() {
    c64 **cp[] = {&B, &HB, &R, &T2, &small, &part, &ypart};
    for (c64 **p : cp) {
        if (*p) hipFree(*p);
        *p = nullptr;
    }
    double **dp[] = {&lam, &norms, &lu};
    for (double **p : dp) {
        if (*p) hipFree(*p);
        *p = nullptr;
    }
    if (ipiv) hipFree(ipiv);
    ipiv = nullptr;
}

void RefineCtx::atb(const c64 *A, int ma, const c64 *Bm, int mb, c64 *C, hipStream_t s) {
    const int nchunk = (n + 511) / 512;
    hipLaunchKernelGGL(k_atb, dim3(nchunk), dim3(256), 0, s, A, ld, ma, Bm, ld, mb, n, part);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_parts, dim3((ma * mb + 255) / 256), dim3(256), 0, s, part, nchunk, ma * mb, C);
    QD_LAUNCH_CHECK();
}

void RefineCtx::ab(const c64 *A, int ma, const c64 *S, int mb, double alpha, double beta, c64 *Y, hipStream_t s) {
    hipLaunchKernelGGL(k_ab, dim3((n + 255) / 256), dim3(256), 0, s, A, ld, ma, S, mb, n, alpha, beta, Y, ld);
    QD_LAUNCH_CHECK();
}

void RefineCtx::gemm_h(const c64 *H, const c64 *X, int m, c64 *Y, hipStream_t s) {
    if (m < 1 || m > kM || m > ypart_cols) throw HipError("H X: too many columns", QD_E_ARG);
    // enough work-groups to fill the device: segments of the 16-column chunks, at most
    // kGemmSegMax and at most what the partial buffer holds for m columns
    const int nrb = (n + 63) / 64, nck = (n + 15) / 16;
    int S = std::max(1, std::min(std::min(kGemmSegMax, ypart_cols / m), (1536 + nrb - 1) / nrb));
    const int per = (nck + S - 1) / S;
    S = (nck + per - 1) / per;
    const dim3 grid(nrb, S);
    if (m <= 4) hipLaunchKernelGGL(k_gemm_h<1>, grid, dim3(256), 0, s, H, ld, n, X, ld, m, per, ypart);
    else if (m <= 8) hipLaunchKernelGGL(k_gemm_h<2>, grid, dim3(256), 0, s, H, ld, n, X, ld, m, per, ypart);
    else if (m <= 20) hipLaunchKernelGGL(k_gemm_h<5>, grid, dim3(256), 0, s, H, ld, n, X, ld, m, per, ypart);
    else if (m <= 40) hipLaunchKernelGGL(k_gemm_h<10>, grid, dim3(256), 0, s, H, ld, n, X, ld, m, per, ypart);
    else hipLaunchKernelGGL(k_gemm_h<20>, grid, dim3(256), 0, s, H, ld, n, X, ld, m, per, ypart);
    QD_LAUNCH_CHECK();
    const long tot = (long)n * m;
    hipLaunchKernelGGL(k_sum_seg, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, ypart, S, ld, n, m, Y, ld);
    QD_LAUNCH_CHECK();
}

void RefineCtx::residual(const c64 *HX, const c64 *X, const double *lm, c64 *Rout, hipStream_t s) {
    const long tot = (long)n * k;
    hipLaunchKernelGGL(k_resid, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, HX, X, ld, n, k, lm, Rout);
    QD_LAUNCH_CHECK();
}

void RefineCtx::colnorms(const c64 *X, int m, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_colnorm, dim3(m), dim3(256), 0, s, X, ld, n, out);
    QD_LAUNCH_CHECK();
}

void RefineCtx::trisolve(const double *d, const double *e, const double *lm, double tnorm, c64 *S, hipStream_t s) {
    hipLaunchKernelGGL(k_trisolve, dim3(1), dim3(64), 0, s, d, e, n, lm, k, tnorm, lu, ipiv, S, ld);
    QD_LAUNCH_CHECK();
}

void RefineCtx::chol_inv(const c64 *G, int m, c64 *Rinv, hipStream_t s) {
    hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(256), 0, s, G, m, Rinv);
    QD_LAUNCH_CHECK();
}

void RefineCtx::jacobi_select(const c64 *M, int m, c64 *Usel, double *lm, hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi_select, dim3(1), dim3(256), 0, s, M, m, k, Usel, lm, small + 3 * kM * kM);
    QD_LAUNCH_CHECK();
}

}  // namespace qd
