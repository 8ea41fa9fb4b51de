// Adjoint of spectral objectives through the Delta-G tables.
//
// For an objective L of the window eigenpairs, dL/dtheta = Re tr(D dH/dtheta) with a Hermitian
// density D = sum_n g_n psi_n psi_n^H - sum_i (lambda_i psi_i^H + psi_i lambda_i^H): g_n = dL/dE_n
// (Hellmann-Feynman) and lambda_i the Sternheimer solutions carrying the eigenvector terms.
// H(r, c) = sum_t coef_t(r, c) T_t(G_r - G_c), so
//   dL/dtheta = Re sum_t sum_d W_t(d) dT_t(d)/dtheta,
//   W_t(d) = sum_{lower blocks b} w_b sum_{G_r - G_c = d} D(c, r) coef_t(r, c)   (w_b = 2 off the
// band diagonal, 1 on it, where the whole block is summed with the lower formula, exact because
// every table is the transform of a real function).  W is evaluated by FFT correlation of the
// vectors (below), with the coefficients taken from the Hamiltonian's own element formula
// (h_lower_t with a one-hot table selector, bin_coefficients), so forward and adjoint cannot
// disagree.  The objective terms of the e-h Coulomb energy use the level densities
// rho(q) = sum_i sum_G conj(psi_i(G)) psi_i(G+q) (k_level_density) and the potential
// application g_i = 1/4 sum_q w(q) conj(rho'(q)) psi_i(G+q) (k_phi_apply).
//
// Reference: the Hamiltonian element formulas of hamiltonian_8_band.c:78-492 (through
// h_lower_t) and the direct Coulomb element of coulomb_integrals.c:229-311 (CME 1 kernel of
// coulomb_matrix_elements.c:392-402).  The adjoint itself has no counterpart in the reference.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "qd_adjoint.hpp"
#include "qd_coulomb.hpp"

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

// ---- FFT-correlation binning ------------------------------------------------------------
// Every element coefficient is bilinear in u(k_r) and u(k_c), u = (1, kx, ky, kz), k = G + K:
// coef_t(r, c) = sum_ij C_t[i][j] u_i(k_r) u_j(k_c).  Hence
//   W_t(d) = sum_b w_b sum_q gamma_q sum_ij C_t^b[i][j] sum_g A_{bc,j}(g) conj(B_{br,i}(g + d)),
// A_{c,j}(g) = V_a(c, g) u_j(g), B_{r,i}(g) = V_b(r, g) u_i(g): correlations, evaluated by FFT on
// power-of-two grids L >= 2N - 1 per axis (zero padding keeps them linear).  With
// Y = IFFT(conj(FFT A) FFT B), Y(d) = sum_g conj(A(g)) B(g + d), so
//   W_t = conj(IFFT(Q_t)),  Q_t = sum w gamma conj(C) conj(FFT A) FFT B.

// Per-element bodies of the binning kernels, shared with the host emulation used by the CPU tests.
// Zero-padded grids u_j(g) V(band, g) of one vector, grid index band * 4 + j; element i of the
// 32 grids.
__host__ __device__ inline c64 bin_pad_at(long i, const c64 *V, int Nx, int Ny, int Nz, int Ly, int Lz, long grid,
                                          const double *ux, const double *uy, const double *uz) {
    const long NG = (long)Nx * Ny * Nz;
    const int gi = (int)(i / grid);
    const long r = i - (long)gi * grid;
    const int x = (int)(r / ((long)Ly * Lz)), y = (int)((r / Lz) % Ly), z = (int)(r % Lz);
    const int band = gi >> 2, j = gi & 3;
    if (x >= Nx || y >= Ny || z >= Nz) return mk(0.0, 0.0);
    const double u = j == 0 ? 1.0 : (j == 1 ? ux[x] : (j == 2 ? uy[y] : uz[z]));
    return u * V[band * NG + ((long)x * Ny + y) * Nz + z];
}

// Q_t[k] += the contribution of one term at spectral point k: per band block the 16 products
// conj(A_{bc,j}) B_{br,i}, then every table entry of the block.
__host__ __device__ inline void bin_accumulate_at(long k, const c64 *Fa, const c64 *Fb, long grid, double gamma,
                                                  const BinCoefs &cf, c64 *Q) {
    c64 acc[QD_NTAB];
#pragma unroll
    for (int t = 0; t < QD_NTAB; ++t) acc[t] = mk(0.0, 0.0);
    for (int bi = 0; bi < cf.nblocks; ++bi) {
        const int br = cf.br[bi], bc = cf.bc[bi];
        c64 P[4][4];
        c64 A[4], B[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) A[j] = Fa[(long)(bc * 4 + j) * grid + k];
#pragma unroll
        for (int i = 0; i < 4; ++i) B[i] = Fb[(long)(br * 4 + i) * grid + k];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) P[i][j] = cmul(A[j], B[i]);   // conj(A) B
        for (int e = cf.first[bi]; e < cf.first[bi + 1]; ++e) {
            c64 v = mk(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) v = v + cmul(cf.C[e][i][j], P[i][j]);   // conj(C) P
            const int t = cf.t[e];
#pragma unroll
            for (int tt = 0; tt < QD_NTAB; ++tt)
                if (tt == t) acc[tt] = acc[tt] + (cf.w[bi] * gamma) * v;
        }
    }
#pragma unroll
    for (int t = 0; t < QD_NTAB; ++t) Q[(long)t * grid + k] = Q[(long)t * grid + k] + acc[t];
}

// W[t][d] = conj(Q_t[d mod L]) / |L|, d = offset G_r - G_c in [-(N-1), N-1] per axis; element i
// of W (t = i / nT).
__host__ __device__ inline c64 bin_extract_at(long i, const c64 *Q, int Lx, int Ly, int Lz, int Nx, int Ny, int Nz,
                                              long nT) {
    const int NTy = 2 * Ny - 1, NTz = 2 * Nz - 1;
    const long grid = (long)Lx * Ly * Lz;
    const double scale = 1.0 / (double)grid;
    const long t = i / nT, d = i - t * nT;
    int ox = (int)(d / ((long)NTy * NTz)) - (Nx - 1);
    int oy = (int)((d / NTz) % NTy) - (Ny - 1);
    int oz = (int)(d % NTz) - (Nz - 1);
    ox += ox < 0 ? Lx : 0;
    oy += oy < 0 ? Ly : 0;
    oz += oz < 0 ? Lz : 0;
    const c64 v = Q[t * grid + ((long)ox * Ly + oy) * Lz + oz];
    return mk(scale * v.re, -scale * v.im);
}

__global__ void k_bin_pad(const c64 *__restrict__ V, int Nx, int Ny, int Nz, int Lx, int Ly, int Lz,
                          const double *__restrict__ ux, const double *__restrict__ uy, const double *__restrict__ uz,
                          c64 *__restrict__ F) {
    const long grid = (long)Lx * Ly * Lz;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < grid * 32; i += (long)gridDim.x * blockDim.x)
        F[i] = bin_pad_at(i, V, Nx, Ny, Nz, Ly, Lz, grid, ux, uy, uz);
}

__global__ __launch_bounds__(256) void k_bin_accumulate(const c64 *__restrict__ Fa, const c64 *__restrict__ Fb,
                                                        long grid, double gamma, const BinCoefs *__restrict__ cf,
                                                        c64 *__restrict__ Q) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < grid; k += (long)gridDim.x * blockDim.x)
        bin_accumulate_at(k, Fa, Fb, grid, gamma, *cf, Q);
}

__global__ void k_bin_zero(c64 *__restrict__ Q, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        Q[i] = mk(0.0, 0.0);
}

__global__ void k_bin_extract(const c64 *__restrict__ Q, int Lx, int Ly, int Lz, int Nx, int Ny, int Nz,
                              c64 *__restrict__ W, long nT) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nT * QD_NTAB; i += (long)gridDim.x * blockDim.x)
        W[i] = bin_extract_at(i, Q, Lx, Ly, Lz, Nx, Ny, Nz, nT);
}

}  // namespace

BinCoefs bin_coefficients() {
    // geometry whose plane wave m has k = G + K = m per axis (m in {0, 1}), K = 0
    static const double Ga[3] = {0.0, 1.0, 2.0};
    HGeom g{};
    for (int a = 0; a < 3; ++a) { g.N[a] = 2; g.NT[a] = 3; g.K[a] = 0.0; g.G[a] = Ga; }
    g.NG = 8;
    auto pw = [](int mx, int my, int mz) { return (mx * 2 + my) * 2 + mz; };
    const int e[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};   // k = 0, e_x, e_y, e_z
    BinCoefs cf{};
    int ne = 0, nb = 0;
    for (int bc = 0; bc < 8; ++bc)
        for (int br = bc; br < 8; ++br) {
            cf.first[nb] = ne;
            for (int t = 0; t < QD_NTAB; ++t) {
                auto f = [&](int r, int c) {
                    return h_lower_t(g, [&](int kk, long) { return kk == t ? mk(1.0, 0.0) : mk(0.0, 0.0); },
                                     br * 8 + r, bc * 8 + c);
                };
                c64 F[4][4];
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) F[i][j] = f(pw(e[i][0], e[i][1], e[i][2]), pw(e[j][0], e[j][1], e[j][2]));
                c64 C[4][4];
                C[0][0] = F[0][0];
                for (int i = 1; i < 4; ++i) { C[i][0] = F[i][0] - F[0][0]; C[0][i] = F[0][i] - F[0][0]; }
                for (int i = 1; i < 4; ++i)
                    for (int j = 1; j < 4; ++j) C[i][j] = F[i][j] - C[i][0] - C[0][j] - C[0][0];
                bool any = false;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) any = any || C[i][j].re != 0.0 || C[i][j].im != 0.0;
                if (!any) continue;
                // the model must reproduce the formula at every k_r, k_c in {0, 1}^3
                for (int r = 0; r < 8; ++r)
                    for (int c = 0; c < 8; ++c) {
                        const double ur[4] = {1.0, (double)(r >> 2), (double)((r >> 1) & 1), (double)(r & 1)};
                        const double uc[4] = {1.0, (double)(c >> 2), (double)((c >> 1) & 1), (double)(c & 1)};
                        c64 m = mk(0.0, 0.0);
                        for (int i = 0; i < 4; ++i)
                            for (int j = 0; j < 4; ++j) m = m + (ur[i] * uc[j]) * C[i][j];
                        const c64 v = f(r, c);
                        if (std::fabs(m.re - v.re) + std::fabs(m.im - v.im) > 1e-12 * (1.0 + std::fabs(v.re) + std::fabs(v.im)))
                            throw std::runtime_error("element coefficient is not bilinear in (k_r, k_c)");
                    }
                if (ne >= kBinMaxEntries) throw std::runtime_error("too many binning coefficient entries");
                cf.t[ne] = t;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) cf.C[ne][i][j] = C[i][j];
                ++ne;
            }
            if (ne == cf.first[nb]) continue;
            cf.br[nb] = br;
            cf.bc[nb] = bc;
            cf.w[nb] = br == bc ? 1.0 : 2.0;
            ++nb;
        }
    cf.nblocks = nb;
    cf.first[nb] = ne;
    return cf;
}

void launch_adjoint_bin_fft(const HGeom &geo, const double *const u[3], const c64 *V, long ldv, int nv,
                            const AdjointTerms &terms, const BinCoefs *cf_dev, c64 *W, long nT, hipStream_t s) {
    int logL[3], L[3];
    for (int a = 0; a < 3; ++a) {
        logL[a] = std::max(1, ilog2_ceil(2 * geo.N[a] - 1));
        if (logL[a] > 9) throw HipError("adjoint binning grid above 512 points per axis", QD_E_UNSUPPORTED);
        L[a] = 1 << logL[a];
    }
    const long grid = (long)L[0] * L[1] * L[2];
    // which vectors the terms use
    std::vector<int> slot(nv, -1);
    int nused = 0;
    for (int q = 0; q < terms.n; ++q)
        for (int v : {terms.a[q], terms.b[q]})
            if (slot[v] < 0) slot[v] = nused++;
    c64 *buf = nullptr;
    const size_t ntw = (size_t)(L[0] + L[1] + L[2]) / 2;
    const size_t nbuf = (size_t)nused * 32 * grid + (size_t)QD_NTAB * grid + ntw;
    QD_HIP(hipMalloc(&buf, sizeof(c64) * nbuf));
    struct FreeB { c64 *p; ~FreeB() { (void)hipFree(p); } } fb{buf};
    c64 *spec = buf, *Q = buf + (size_t)nused * 32 * grid, *twd = Q + (size_t)QD_NTAB * grid;
    const c64 *tw[3];
    {
        std::vector<c64> h;
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < L[a] / 2; ++k) {
                const double ang = -2.0 * M_PI * (double)k / (double)L[a];
                h.push_back(mk(std::cos(ang), std::sin(ang)));
            }
        QD_HIP(hipMemcpyAsync(twd, h.data(), sizeof(c64) * h.size(), hipMemcpyHostToDevice, s));
        tw[0] = twd;
        tw[1] = twd + L[0] / 2;
        tw[2] = twd + L[0] / 2 + L[1] / 2;
    }
    const unsigned nbp = (unsigned)std::min<long>((32 * grid + 255) / 256, 1 << 16);
    for (int v = 0; v < nv; ++v) {
        if (slot[v] < 0) continue;
        c64 *F = spec + (size_t)slot[v] * 32 * grid;
        hipLaunchKernelGGL(k_bin_pad, dim3(nbp), dim3(256), 0, s, V + (long)v * ldv, geo.N[0], geo.N[1], geo.N[2], L[0],
                           L[1], L[2], u[0], u[1], u[2], F);
        QD_LAUNCH_CHECK();
        fft3_device(F, 32, logL, tw, false, s);
    }
    const unsigned nbq = (unsigned)std::min<long>((QD_NTAB * grid + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_bin_zero, dim3(nbq), dim3(256), 0, s, Q, (long)QD_NTAB * grid);
    QD_LAUNCH_CHECK();
    const unsigned nbk = (unsigned)std::min<long>((grid + 255) / 256, 1 << 14);
    for (int q = 0; q < terms.n; ++q) {
        hipLaunchKernelGGL(k_bin_accumulate, dim3(nbk), dim3(256), 0, s, spec + (size_t)slot[terms.a[q]] * 32 * grid,
                           spec + (size_t)slot[terms.b[q]] * 32 * grid, grid, terms.gamma[q], cf_dev, Q);
        QD_LAUNCH_CHECK();
    }
    fft3_device(Q, QD_NTAB, logL, tw, true, s);
    const unsigned nbe = (unsigned)std::min<long>((QD_NTAB * nT + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_bin_extract, dim3(nbe), dim3(256), 0, s, Q, L[0], L[1], L[2], geo.N[0], geo.N[1], geo.N[2],
                       W, nT);
    QD_LAUNCH_CHECK();
    QD_HIP(hipStreamSynchronize(s));   // buf is freed on return
}

// ---- host emulation of the binning (test support; the product path runs the kernels above) ----
namespace {

// in-place radix-2 transform of one line, the convention of fft3_device: forward
// exp(-2 pi i k n / L), inverse exp(+2 pi i k n / L), unscaled
void host_fft_line(c64 *x, long stride, int L, bool inverse) {
    for (int i = 1, j = 0; i < L; ++i) {
        int bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[(long)i * stride], x[(long)j * stride]);
    }
    for (int len = 2; len <= L; len <<= 1) {
        const double ang = (inverse ? 2.0 : -2.0) * M_PI / len;
        for (int i0 = 0; i0 < L; i0 += len)
            for (int k = 0; k < len / 2; ++k) {
                const c64 w = mk(std::cos(ang * k), std::sin(ang * k));
                c64 &p = x[(long)(i0 + k) * stride], &q = x[(long)(i0 + k + len / 2) * stride];
                const c64 t = w * q;
                q = p - t;
                p = p + t;
            }
    }
}

void host_fft3(c64 *data, long ngrids, const int L[3], bool inverse) {
    const long grid = (long)L[0] * L[1] * L[2];
    for (long gi = 0; gi < ngrids; ++gi) {
        c64 *g = data + gi * grid;
        for (long a = 0; a < (long)L[0] * L[1]; ++a) host_fft_line(g + a * L[2], 1, L[2], inverse);
        for (int x = 0; x < L[0]; ++x)
            for (int z = 0; z < L[2]; ++z) host_fft_line(g + (long)x * L[1] * L[2] + z, L[2], L[1], inverse);
        for (long a = 0; a < (long)L[1] * L[2]; ++a) host_fft_line(g + a, (long)L[1] * L[2], L[0], inverse);
    }
}

}  // namespace

void adjoint_bin_fft_host(const HGeom &geo, const double *const u[3], const c64 *V, long ldv, int nv,
                          const AdjointTerms &terms, const BinCoefs &cf, c64 *W, long nT) {
    int L[3];
    for (int a = 0; a < 3; ++a) L[a] = 1 << std::max(1, ilog2_ceil(2 * geo.N[a] - 1));
    const long grid = (long)L[0] * L[1] * L[2];
    std::vector<c64> spec((size_t)nv * 32 * grid), Q((size_t)QD_NTAB * grid, mk(0.0, 0.0));
    for (int v = 0; v < nv; ++v) {
        c64 *F = spec.data() + (size_t)v * 32 * grid;
        for (long i = 0; i < 32 * grid; ++i)
            F[i] = bin_pad_at(i, V + (long)v * ldv, geo.N[0], geo.N[1], geo.N[2], L[1], L[2], grid, u[0], u[1], u[2]);
        host_fft3(F, 32, L, false);
    }
    for (int q = 0; q < terms.n; ++q)
        for (long k = 0; k < grid; ++k)
            bin_accumulate_at(k, spec.data() + (size_t)terms.a[q] * 32 * grid, spec.data() + (size_t)terms.b[q] * 32 * grid,
                              grid, terms.gamma[q], cf, Q.data());
    host_fft3(Q.data(), QD_NTAB, L, true);
    for (long i = 0; i < nT * QD_NTAB; ++i) W[i] = bin_extract_at(i, Q.data(), L[0], L[1], L[2], geo.N[0], geo.N[1], geo.N[2], nT);
}

// The binning by its definition: W_t(G_r - G_c) = sum over lower band blocks (weight 2 off the
// band diagonal) of gamma_q V_a(c) conj(V_b(r)) coef_t(r, c), the coefficient from the element
// formula with a one-hot table.
void adjoint_bin_direct_host(const HGeom &geo, const c64 *V, long ldv, const AdjointTerms &terms, c64 *W, long nT) {
    const int NG = geo.NG;
    for (long i = 0; i < nT * QD_NTAB; ++i) W[i] = mk(0.0, 0.0);
    for (int bc = 0; bc < 8; ++bc)
        for (int br = bc; br < 8; ++br) {
            const double wb = br == bc ? 1.0 : 2.0;
            for (int gc = 0; gc < NG; ++gc)
                for (int gr = 0; gr < NG; ++gr) {
                    const int r = br * NG + gr, c = bc * NG + gc;
                    const PWIndex R = pw_index(geo, r), C = pw_index(geo, c);
                    const long d = ((long)(R.x - C.x + geo.NT[0] / 2) * geo.NT[1] + (R.y - C.y + geo.NT[1] / 2)) * geo.NT[2] +
                                   (R.z - C.z + geo.NT[2] / 2);
                    c64 Dv = mk(0.0, 0.0);
                    for (int q = 0; q < terms.n; ++q)
                        Dv = Dv + terms.gamma[q] * mulc(V[(long)terms.a[q] * ldv + c], V[(long)terms.b[q] * ldv + r]);
                    for (int t = 0; t < QD_NTAB; ++t) {
                        const c64 coef =
                            h_lower_pw(geo, [&](int kk, long) { return kk == t ? mk(1.0, 0.0) : mk(0.0, 0.0); }, R, C);
                        W[(long)t * nT + d] = W[(long)t * nT + d] + wb * (Dv * coef);
                    }
                }
        }
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

}  // namespace qd
