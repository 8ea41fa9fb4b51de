// Secondary outputs of the window states: carrier localisation per shape
// (carrier_localisation.c:23-128 with 8_band_main.c:1427-1545) and the conduction-valence overlaps
// (spatial_overlap.c:23-52 with 8_band_main.c:1373-1415).
//
// Localisation of state n in shape s: L_s(n) = sum_b sum_G conj(psi_b(G)) sum_G' chi_s(G - G') psi_b(G')
// = sum_q chi_s(q) conj(rho_n(q)) with the state's density rho_n(q) = sum_b sum_G conj(psi_b(G))
// psi_b(G + q) on the Delta-G grid (k_level_density of the adjoint), so each shape costs one
// nT-long contraction instead of the reference's per-shape, per-band FFT convolution.
#include "qd_internal.hpp"

namespace qd {

namespace {

// L[n * ns + s] = Re sum_q chi[s][q] conj(rho[n][q]); one work-group per (n, s)
__global__ __launch_bounds__(256) void k_loc_contract(const c64 *__restrict__ chi, const c64 *__restrict__ rho, long nT,
                                                      int ns, double *__restrict__ L) {
    const int n = blockIdx.x / ns, s = blockIdx.x % ns;
    double acc = 0.0;
    for (long q = threadIdx.x; q < nT; q += 256) {
        const c64 a = chi[(long)s * nT + q], b = rho[(long)n * nT + q];
        acc += a.re * b.re + a.im * b.im;   // Re(a conj(b))
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) L[blockIdx.x] = sh[0];
}

// O[i * nb + j] = sum_r conj(A_i(r)) B_j(r); one work-group per (i, j)
__global__ __launch_bounds__(256) void k_col_dot(const c64 *__restrict__ A, const c64 *__restrict__ B, long ld, int n,
                                                 int nb, c64 *__restrict__ O) {
    const int i = blockIdx.x / nb, j = blockIdx.x % nb;
    c64 acc = mk(0.0, 0.0);
    for (int r = threadIdx.x; r < n; r += 256) acc = acc + cmul(A[(long)i * ld + r], B[(long)j * ld + r]);
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
    if (threadIdx.x == 0) O[blockIdx.x] = mk(sr[0], si[0]);
}

}  // namespace

void launch_loc_contract(const c64 *chi, const c64 *rho, long nT, int nstates, int ns, double *L, hipStream_t s) {
    hipLaunchKernelGGL(k_loc_contract, dim3((unsigned)(nstates * ns)), dim3(256), 0, s, chi, rho, nT, ns, L);
    QD_LAUNCH_CHECK();
}

void launch_col_dot(const c64 *A, int na, const c64 *B, int nb, long ld, int n, c64 *O, hipStream_t s) {
    hipLaunchKernelGGL(k_col_dot, dim3((unsigned)(na * nb)), dim3(256), 0, s, A, B, ld, n, nb, O);
    QD_LAUNCH_CHECK();
}

}  // namespace qd
