// Secondary outputs of the window states: carrier localisation per shape
// (carrier_localisation.c:23-128 with 8_band_main.c:1427-1545), the conduction-valence overlaps
// (spatial_overlap.c:23-52 with 8_band_main.c:1373-1415) and the real-space envelope functions
// (envelope_functions.c:23-220, evaluated on the linescan, cross-section and 3-D grids of
// 8_band_main.c:811-1100).
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

// Envelope functions at arbitrary points: F[(s * 8 + b) * npts + p] = scale * sum_G c_b(G)
// exp(i (G + K) . r_p) for the states s of one group of kEnvStates (blockIdx.y) and one point per
// thread.  The phase of each plane wave is computed once per point and used for the group's
// 8 x kEnvStates coefficients, which every lane of the wavefront reads at the same address.
constexpr int kEnvStates = 4;
__global__ __launch_bounds__(256) void k_envelope(const c64 *__restrict__ X, long ld, int count, const double *__restrict__ ux,
                                                  const double *__restrict__ uy, const double *__restrict__ uz, int Nx,
                                                  int Ny, int Nz, const double *__restrict__ px,
                                                  const double *__restrict__ py, const double *__restrict__ pz,
                                                  int npts, double scale, c64 *__restrict__ F) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npts) return;   // no barrier below
    const int s0 = blockIdx.y * kEnvStates;
    const int ns = min(kEnvStates, count - s0);
    const long NG = (long)Nx * Ny * Nz;
    const double x = px[p], y = py[p], z = pz[p];
    c64 acc[kEnvStates][8];
#pragma unroll
    for (int s = 0; s < kEnvStates; ++s)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[s][b] = mk(0.0, 0.0);
    long g = 0;
    for (int ix = 0; ix < Nx; ++ix)
        for (int iy = 0; iy < Ny; ++iy) {
            const double axy = ux[ix] * x + uy[iy] * y;
            for (int iz = 0; iz < Nz; ++iz, ++g) {
                double sn, cs;
                sincos(axy + uz[iz] * z, &sn, &cs);
                const c64 ph = mk(cs, sn);
#pragma unroll
                for (int s = 0; s < kEnvStates; ++s) {
                    if (s >= ns) break;
                    const c64 *col = X + (long)(s0 + s) * ld + g;
#pragma unroll
                    for (int b = 0; b < 8; ++b) fma_(acc[s][b], col[b * NG], ph);
                }
            }
        }
    for (int s = 0; s < ns; ++s)
#pragma unroll
        for (int b = 0; b < 8; ++b) F[((long)(s0 + s) * 8 + b) * npts + p] = scale * acc[s][b];
}

}  // namespace

void launch_envelope(const c64 *X, long ld, int count, const double *const u[3], const int N[3], const double *px,
                     const double *py, const double *pz, int npts, double scale, c64 *F, hipStream_t s) {
    const dim3 grid((unsigned)((npts + 255) / 256), (unsigned)((count + kEnvStates - 1) / kEnvStates));
    hipLaunchKernelGGL(k_envelope, grid, dim3(256), 0, s, X, ld, count, u[0], u[1], u[2], N[0], N[1], N[2], px, py, pz,
                       npts, scale, F);
    QD_LAUNCH_CHECK();
}

void launch_loc_contract(const c64 *chi, const c64 *rho, long nT, int nstates, int ns, double *L, hipStream_t s) {
    hipLaunchKernelGGL(k_loc_contract, dim3((unsigned)(nstates * ns)), dim3(256), 0, s, chi, rho, nT, ns, L);
    QD_LAUNCH_CHECK();
}

void launch_col_dot(const c64 *A, int na, const c64 *B, int nb, long ld, int n, c64 *O, hipStream_t s) {
    hipLaunchKernelGGL(k_col_dot, dim3((unsigned)(na * nb)), dim3(256), 0, s, A, B, ld, n, nb, O);
    QD_LAUNCH_CHECK();
}

}  // namespace qd
