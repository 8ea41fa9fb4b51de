// Adjoint of spectral objectives through the Delta-G tables.
//
// For an objective L of the window eigenpairs, dL/dtheta = Re tr(D dH/dtheta) with a Hermitian
// density D = sum_n g_n psi_n psi_n^H - sum_i (lambda_i psi_i^H + psi_i lambda_i^H): g_n = dL/dE_n
// (Hellmann-Feynman) and lambda_i the Sternheimer solutions carrying the eigenvector terms.
// H(r, c) = sum_t coef_t(r, c) T_t(G_r - G_c), so
//   dL/dtheta = Re sum_t sum_d W_t(d) dT_t(d)/dtheta,
//   W_t(d) = sum_{lower blocks b} w_b sum_{G_r - G_c = d} D(c, r) coef_t(r, c)   (w_b = 2 off the
// band diagonal, 1 on it, where the whole block is summed with the lower formula, exact because
// every table is the transform of a real function).  k_adjoint_bin evaluates W for every table
// and Delta-G bin, one work-group per bin, with the coefficients taken from the Hamiltonian's own
// element formula (h_lower_t with a one-hot table selector), so forward and adjoint cannot
// disagree.  The objective terms of the e-h Coulomb energy use the level densities
// rho(q) = sum_i sum_G conj(psi_i(G)) psi_i(G+q) (k_level_density) and the potential
// application g_i = 1/4 sum_q w(q) conj(rho'(q)) psi_i(G+q) (k_phi_apply).
//
// Reference: the Hamiltonian element formulas of hamiltonian_8_band.c:78-492 (through
// h_lower_t) and the direct Coulomb element of coulomb_integrals.c:229-311 (CME 1 kernel of
// coulomb_matrix_elements.c:392-402).  The adjoint itself has no counterpart in the reference.
#include <algorithm>
#include <vector>

#include "qd_adjoint.hpp"

namespace qd {

namespace {

__device__ inline c64 block_sum_c(c64 v, c64 *sh) {
    for (int o = 32; o > 0; o >>= 1) {
        v.re += __shfl_xor(v.re, o, 64);
        v.im += __shfl_xor(v.im, o, 64);
    }
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wv] = v;
    __syncthreads();
    c64 t = mk(0.0, 0.0);
    for (int i = 0; i < nw; ++i) t = t + sh[i];
    return t;
}

// rho[q] over the (2N-1)^3 grid, q index d = q + N - 1 per axis.
__global__ __launch_bounds__(256) void k_level_density(const c64 *__restrict__ X, long ld, const int *__restrict__ cols,
                                                       int ncols, int Nx, int Ny, int Nz, c64 *__restrict__ rho) {
    __shared__ c64 sh[4];
    const int NTy = 2 * Ny - 1, NTz = 2 * Nz - 1;
    const long d = blockIdx.x;
    const int ox = (int)(d / ((long)NTy * NTz)) - (Nx - 1);
    const int oy = (int)((d / NTz) % NTy) - (Ny - 1);
    const int oz = (int)(d % NTz) - (Nz - 1);
    const int lx = max(0, -ox), hx = min(Nx, Nx - ox);
    const int ly = max(0, -oy), hy = min(Ny, Ny - oy);
    const int lz = max(0, -oz), hz = min(Nz, Nz - oz);
    const long NG = (long)Nx * Ny * Nz;
    const int cy = hy - ly, cz = hz - lz;
    const long cnt = (long)max(0, hx - lx) * max(0, cy) * max(0, cz);
    c64 acc = mk(0.0, 0.0);
    for (long t = threadIdx.x; t < cnt; t += blockDim.x) {
        const int x = lx + (int)(t / ((long)cy * cz)), y = ly + (int)((t / cz) % cy), z = lz + (int)(t % cz);
        const long g = ((long)x * Ny + y) * Nz + z;
        const long gq = ((long)(x + ox) * Ny + (y + oy)) * Nz + (z + oz);
        for (int s = 0; s < ncols; ++s) {
            const c64 *v = X + (long)cols[s] * ld;
#pragma unroll
            for (int b = 0; b < 8; ++b) acc = acc + cmul(v[b * NG + g], v[b * NG + gq]);
        }
    }
    acc = block_sum_c(acc, sh);
    if (threadIdx.x == 0) rho[d] = acc;
}

// G[:, j] = scale * sum_q Phi_j(q) X[:, cols[j]](G + q), band by band; Phi_j = phis[which[j]].
__global__ __launch_bounds__(256) void k_phi_apply(const c64 *__restrict__ X, long ld, const int *__restrict__ cols,
                                                   const int *__restrict__ which, int ncols,
                                                   const c64 *__restrict__ phis, long nT, double scale, int Nx,
                                                   int Ny, int Nz, c64 *__restrict__ G, long ldg) {
    const long NG = (long)Nx * Ny * Nz;
    const int NTy = 2 * Ny - 1, NTz = 2 * Nz - 1;
    const long total = (long)ncols * 8 * NG;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t / (8 * NG));
        const long rem = t - (long)j * 8 * NG;
        const int b = (int)(rem / NG);
        const long g = rem - (long)b * NG;
        const int x = (int)(g / ((long)Ny * Nz)), y = (int)((g / Nz) % Ny), z = (int)(g % Nz);
        const c64 *v = X + (long)cols[j] * ld + b * NG;
        const c64 *phi = phis + (long)which[j] * nT;
        c64 acc = mk(0.0, 0.0);
        for (int x2 = 0; x2 < Nx; ++x2)
            for (int y2 = 0; y2 < Ny; ++y2) {
                const long dxy = ((long)(x2 - x + Nx - 1) * NTy + (y2 - y + Ny - 1)) * NTz + (Nz - 1 - z);
                const long g2 = ((long)x2 * Ny + y2) * Nz;
                for (int z2 = 0; z2 < Nz; ++z2) fma_(acc, phi[dxy + z2], v[g2 + z2]);
            }
        G[(long)j * ldg + b * NG + g] = scale * acc;
    }
}

// W[t][d] for all tables t: one work-group per Delta-G bin d.
__global__ __launch_bounds__(256) void k_adjoint_bin(HGeom geo, const c64 *__restrict__ V, long ldv,
                                                     AdjointTerms terms, AdjointBlocks blocks,
                                                     c64 *__restrict__ W, long nT) {
    __shared__ c64 sh[4];
    const int Nx = geo.N[0], Ny = geo.N[1], Nz = geo.N[2];
    const int NTy = geo.NT[1], NTz = geo.NT[2];
    const long d = blockIdx.x;
    const int ox = (int)(d / ((long)NTy * NTz)) - (Nx - 1);
    const int oy = (int)((d / NTz) % NTy) - (Ny - 1);
    const int oz = (int)(d % NTz) - (Nz - 1);
    const int lx = max(0, -ox), hx = min(Nx, Nx - ox);
    const int ly = max(0, -oy), hy = min(Ny, Ny - oy);
    const int lz = max(0, -oz), hz = min(Nz, Nz - oz);
    const int cy = max(0, hy - ly), cz = max(0, hz - lz);
    const long cnt = (long)max(0, hx - lx) * cy * cz;
    const long NG = geo.NG;
    c64 acc[QD_NTAB];
#pragma unroll
    for (int k = 0; k < QD_NTAB; ++k) acc[k] = mk(0.0, 0.0);
    for (long t = threadIdx.x; t < cnt; t += blockDim.x) {
        const int x = lx + (int)(t / ((long)cy * cz)), y = ly + (int)((t / cz) % cy), z = lz + (int)(t % cz);
        const long gc = ((long)x * Ny + y) * Nz + z;
        const long gr = ((long)(x + ox) * Ny + (y + oy)) * Nz + (z + oz);
        for (int bi = 0; bi < blocks.n; ++bi) {
            const int br = blocks.br[bi], bc = blocks.bc[bi];
            const long r = br * NG + gr, c = bc * NG + gc;
            c64 Dv = mk(0.0, 0.0);
            for (int q = 0; q < terms.n; ++q) {
                const c64 a = V[(long)terms.a[q] * ldv + c];
                const c64 b = V[(long)terms.b[q] * ldv + r];
                Dv = Dv + terms.gamma[q] * mulc(a, b);
            }
            Dv = blocks.w[bi] * Dv;
            unsigned m = blocks.mask[bi];
            while (m) {
                const int k = __ffs(m) - 1;
                m &= m - 1;
                const c64 coef = h_lower_t(
                    geo, [&](int kk, long) { return kk == k ? mk(1.0, 0.0) : mk(0.0, 0.0); }, (int)r, (int)c);
                // unrolled accumulator update (keeps acc[] in registers)
#pragma unroll
                for (int kk = 0; kk < QD_NTAB; ++kk)
                    if (kk == k) acc[kk] = acc[kk] + Dv * coef;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < QD_NTAB; ++k) {
        const c64 v = block_sum_c(acc[k], sh);
        if (threadIdx.x == 0) W[(long)k * nT + d] = v;
    }
}

}  // namespace

AdjointBlocks adjoint_blocks() {
    // probe the element formula with one-hot tables on a small generic geometry: a table
    // enters a block when its coefficient is nonzero
    AdjointBlocks B{};
    double Gx[3] = {-0.31, 0.0, 0.31}, Gy[3] = {-0.27, 0.0, 0.27}, Gz[3] = {-0.23, 0.0, 0.23};
    HGeom g{};
    g.N[0] = g.N[1] = g.N[2] = 2;
    g.NT[0] = g.NT[1] = g.NT[2] = 3;
    g.NG = 8;
    g.K[0] = 0.013; g.K[1] = 0.021; g.K[2] = 0.034;
    g.G[0] = Gx; g.G[1] = Gy; g.G[2] = Gz;
    int n = 0;
    for (int bc = 0; bc < 8; ++bc)
        for (int br = bc; br < 8; ++br) {
            unsigned mask = 0;
            for (int k = 0; k < QD_NTAB; ++k) {
                for (int probe = 0; probe < 3; ++probe) {
                    const int gr = (probe * 3 + 1) % 8, gc = (probe * 5 + 2) % 8;
                    const c64 v = h_lower_t(g, [&](int kk, long) { return kk == k ? mk(1.0, 0.0) : mk(0.0, 0.0); },
                                            br * 8 + gr, bc * 8 + gc);
                    if (v.re != 0.0 || v.im != 0.0) mask |= 1u << k;
                }
            }
            if (!mask) continue;
            B.br[n] = br;
            B.bc[n] = bc;
            B.mask[n] = mask;
            B.w[n] = br == bc ? 1.0 : 2.0;
            ++n;
        }
    B.n = n;
    return B;
}

void launch_level_density(const c64 *X, long ld, const int *cols, int ncols, const int N[3], c64 *rho, hipStream_t s) {
    const long nT = (long)(2 * N[0] - 1) * (2 * N[1] - 1) * (2 * N[2] - 1);
    hipLaunchKernelGGL(k_level_density, dim3((unsigned)nT), dim3(256), 0, s, X, ld, cols, ncols, N[0], N[1], N[2], rho);
    QD_LAUNCH_CHECK();
}

void launch_phi_apply(const c64 *X, long ld, const int *cols, const int *which, int ncols, const c64 *phis, long nT,
                      double scale, const int N[3], c64 *G, long ldg, hipStream_t s) {
    const long total = (long)ncols * 8 * N[0] * N[1] * N[2];
    const unsigned nb = (unsigned)std::min<long>((total + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_phi_apply, dim3(nb), dim3(256), 0, s, X, ld, cols, which, ncols, phis, nT, scale, N[0], N[1],
                       N[2], G, ldg);
    QD_LAUNCH_CHECK();
}

void launch_adjoint_bin(const HGeom &geo, const c64 *V, long ldv, const AdjointTerms &terms,
                        const AdjointBlocks &blocks, c64 *W, long nT, hipStream_t s) {
    hipLaunchKernelGGL(k_adjoint_bin, dim3((unsigned)nT), dim3(256), 0, s, geo, V, ldv, terms, blocks, W, nT);
    QD_LAUNCH_CHECK();
}

}  // namespace qd
