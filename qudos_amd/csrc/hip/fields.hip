// Coefficient generation on the device (FP64): the strain of the inclusions (continuum-elastic
// Green tensor, strain_fourier_coefficient.c:23-390, summed over shapes as in
// strain_region_reciprocal_space.c:78-150), the piezoelectric polarisation and potential to second
// order (piezo_region_reciprocal_space.c:23-787, piezo_fourier_coefficient.c:24-132), the transfer
// of the fields to the electronic-structure grid in all three cell modes (the same cell; free
// lengths, strain_reciprocal_space.c:23-230; commensurate cells, strain_reciprocal_space_fft.c:23-362
// and piezo_reciprocal_space_fft.c), the convolutions of convolutions.c:24-411 and the QD_NTAB
// Delta-G tables (hamiltonian_8_band_matrix_elements.c:27-969, parameter_matrix_element.c:23-48).
// The shape transforms chi_s are evaluated on the device as well (shapes.hip); only the material
// interpolation comes from the host (materials.cpp).  tables_vjp_device runs the builder's reverse
// pass (tables_vjp of host/fields.cpp) on the same device fields.
//
// Design.  Every truncated linear convolution of the reference is a product of zero-padded
// spectra on a power-of-two grid of L >= 2n - 1 points per axis (the radix-2 transforms of
// fft3_device; the padding keeps the linear convolution exact).  The work is batched per stage:
// all operands of a stage are padded and transformed in one call, one kernel forms every product
// of the stage, one batched inverse transform and one extraction kernel return the central (or,
// for the commensurate transfer, strided) blocks.  Every per-shape convolution of the reference is
// linear in chi_s, so the shapes are summed first with their parameter weights (k_weighted, one
// pass over the shape transforms for all weights), which makes the number of transforms
// independent of the shape count.  The table kernel then forms the 20 tables of one Delta-G point
// per work-item in one pass.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "qd_coulomb.hpp"
#include "qd_host.hpp"
#include "qd_shapes_device.hpp"

namespace qd {

namespace {

constexpr double kPiD = 3.14159265358979323846;
constexpr int kMaxWeights = 16;   // weight vectors per k_weighted pass
constexpr int kMaxAxis = 2048;    // fft3_device: one line of an axis fits the kernel's LDS tile

enum { XX = 0, YY, ZZ, YZ, XZ, XY };

unsigned grid_for(long n, long cap = 1 << 16) {
    return (unsigned)std::max<long>(1, std::min<long>(cap, (n + 255) / 256));
}

// One hipMalloc per buffer, all freed when the builder returns.
struct DevPool {
    std::vector<void *> p;
    template <class T>
    T *take(size_t n) {
        void *q = nullptr;
        QD_HIP(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)));
        p.push_back(q);
        return static_cast<T *>(q);
    }
    template <class T>
    T *upload(const T *h, size_t n, hipStream_t s) {
        T *d = take<T>(n);
        if (n) QD_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
        return d;
    }
    ~DevPool() {
        for (void *q : p) (void)hipFree(q);
    }
};

// out[k][i] = sum_s w[k][s] chi[s][i], s < nsh, k < nw (<= kMaxWeights): every shape transform is
// read once for all weight vectors.
__global__ __launch_bounds__(256) void k_weighted(const c64 *__restrict__ chi, int nsh, long n,
                                                  const double *__restrict__ w, int nw, c64 *__restrict__ out) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        c64 acc[kMaxWeights];
#pragma unroll
        for (int k = 0; k < kMaxWeights; ++k) acc[k] = mk(0.0, 0.0);
        for (int s = 0; s < nsh; ++s) {
            const c64 c = chi[(long)s * n + i];
#pragma unroll
            for (int k = 0; k < kMaxWeights; ++k)
                if (k < nw) acc[k] = acc[k] + w[k * nsh + s] * c;
        }
#pragma unroll
        for (int k = 0; k < kMaxWeights; ++k)
            if (k < nw) out[(long)k * n + i] = acc[k];
    }
}

struct Elastic {
    double C11, C12, C44;
    int iso;
};

// Strain component c of an inclusion with unit misfit at wave vector G (the continuum-elastic
// Green tensor for cubic crystals with the matrix's C_ij; isotropic: C44 = (C11 - C12) / 2).
__device__ double strain_bracket_d(int c, double gx, double gy, double gz, const Elastic &m) {
    const int ci[6] = {0, 1, 2, 1, 0, 0}, cj[6] = {0, 1, 2, 2, 2, 1};
    const int i = ci[c], j = cj[c];
    const double G[3] = {gx, gy, gz};
    const double dij = i == j ? 1.0 : 0.0;
    const double G2 = gx * gx + gy * gy + gz * gz;
    const double an = m.C11 - m.C12 - 2.0 * m.C44;
    if (sqrt(G2) < kEps) {
        if (m.iso) return dij * (1.0 - (3.0 * m.C12 + 2.0 * m.C44) / (3.0 * m.C12 + 6.0 * m.C44));
        const double mean = m.C44 + an / 3.0;
        return dij * (1.0 - (0.5 * (m.C11 + 2.0 * m.C12) / 3.0) * (2.0 / mean) / (1.0 + (m.C12 + m.C44) / mean));
    }
    if (m.iso) return dij - (3.0 * m.C12 + 2.0 * m.C44) * G[i] * G[j] / ((m.C12 + 2.0 * m.C44) * G2);
    double D[3], s = 0.0;
    for (int a = 0; a < 3; ++a) {
        D[a] = m.C44 * G2 + an * G[a] * G[a];
        s += G[a] * G[a] / D[a];
    }
    return dij - (0.5 * (m.C11 + 2.0 * m.C12) * (2.0 * m.C44 * G2 + an * (G[i] * G[i] + G[j] * G[j])) * G[i] * G[j]) /
                     ((1.0 + (m.C12 + m.C44) * s) * D[i] * D[j]);
}

// e[c][i] = bracket_c(G_i) Xeps[i] on the strain grid, the six components XX, YY, ZZ, YZ, XZ, XY
__global__ __launch_bounds__(256) void k_strain(const c64 *__restrict__ Xeps, const double *__restrict__ gx,
                                                const double *__restrict__ gy, const double *__restrict__ gz, int n0,
                                                int n1, int n2, Elastic m, c64 *__restrict__ e) {
    const long n = (long)n0 * n1 * n2;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        const c64 x = Xeps[i];
        for (int c = 0; c < 6; ++c) e[(long)c * n + i] = strain_bracket_d(c, gx[ix], gy[iy], gz[iz], m) * x;
    }
}

// A grid of n0 x n1 x n2 points in an L0 x L1 x L2 grid of zeros: source point (x, y, z) at
// (o + d x, ...) per axis; rev reads the source reversed through its centre, r[j] = a[n - 1 - j]
// (a correlation with a is a convolution with its reversal).
struct PadMap {
    int n[3], o[3], d[3], rev;
};
__global__ __launch_bounds__(256) void k_pad_grid(const c64 *__restrict__ src, PadMap m, int L0, int L1, int L2,
                                                  c64 *__restrict__ dst) {
    const long grid = (long)L0 * L1 * L2;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < grid; t += (long)gridDim.x * 256) {
        const int c[3] = {(int)(t / ((long)L1 * L2)), (int)((t / L2) % L1), (int)(t % L2)};
        int k[3];
        bool in = true;
        for (int a = 0; a < 3; ++a) {
            const int r = c[a] - m.o[a];
            k[a] = r / m.d[a];
            in = in && r >= 0 && r == k[a] * m.d[a] && k[a] < m.n[a];
            if (m.rev) k[a] = m.n[a] - 1 - k[a];
        }
        dst[t] = in ? src[((long)k[0] * m.n[1] + k[1]) * m.n[2] + k[2]] : mk(0.0, 0.0);
    }
}

constexpr int kMaxOut = 16, kMaxTerms = 48;
struct Terms {
    int a[kMaxTerms], b[kMaxTerms];
    int off[kMaxOut + 1];
    int nout;
};

// out[j][t] = sum over the terms k of output j of spec[a_k][t] spec[b_k][t]: the spectra of the
// stage's linear convolutions (and of their sums)
__global__ __launch_bounds__(256) void k_products(const c64 *__restrict__ spec, long grid, Terms tm,
                                                  c64 *__restrict__ out) {
    const int j = blockIdx.y;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < grid; t += (long)gridDim.x * 256) {
        c64 acc = mk(0.0, 0.0);
        for (int k = tm.off[j]; k < tm.off[j + 1]; ++k) fma_(acc, spec[(long)tm.a[k] * grid + t], spec[(long)tm.b[k] * grid + t]);
        out[(long)j * grid + t] = acc;
    }
}

// out[g][k] = scale spec[g][start + step k] per axis, k < nout: the central block (start n / 2,
// step 1) or the strided samples of each of the batch's inverse transforms.
__global__ __launch_bounds__(256) void k_take_block(const c64 *__restrict__ spec, int L0, int L1, int L2, int o0,
                                                    int o1, int o2, int s0, int s1, int s2, int t0, int t1, int t2,
                                                    long count, double scale, c64 *__restrict__ out) {
    const long no = (long)o0 * o1 * o2, grid = (long)L0 * L1 * L2, total = no * count;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
        const long g = t / no, r = t - g * no;
        const int x = (int)(r / ((long)o1 * o2)), y = (int)((r / o2) % o1), z = (int)(r % o2);
        const long src = ((long)(s0 + t0 * x) * L1 + (s1 + t1 * y)) * L2 + (s2 + t2 * z);
        out[t] = scale * spec[g * grid + src];
    }
}

// y[i] = sum_k c_k x_k[i] over up to four terms (no term: zeros); y may be one of the x_k
struct Lin {
    const c64 *x[4];
    double c[4];
    int n;
};
__global__ __launch_bounds__(256) void k_lincomb(c64 *__restrict__ y, long n, Lin l) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        c64 v = mk(0.0, 0.0);
        for (int k = 0; k < l.n; ++k) v = v + l.c[k] * l.x[k][i];
        y[i] = v;
    }
}

// out[s][i] += sum_k c[s][k] G_k[i] (s < nsh, k < nk <= kMaxWeights): gradients of the shape
// transforms from the gradients of the parameter-weighted shape sums
__global__ __launch_bounds__(256) void k_scatter(const c64 *const *__restrict__ G, int nk, const double *__restrict__ c,
                                                 int nsh, long n, c64 *__restrict__ out) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        c64 g[kMaxWeights];
#pragma unroll
        for (int k = 0; k < kMaxWeights; ++k) g[k] = k < nk ? G[k][i] : mk(0.0, 0.0);
        for (int sh = 0; sh < nsh; ++sh) {
            c64 acc = out[(long)sh * n + i];
#pragma unroll
            for (int k = 0; k < kMaxWeights; ++k)
                if (k < nk) acc = acc + c[sh * nk + k] * g[k];
            out[(long)sh * n + i] = acc;
        }
    }
}

// partial[p][blockIdx.x] = Re sum_i A_p[i] B_p[i] over the block's share of i, for pairs p =
// blockIdx.y (the real parts of unconjugated products: dL for complex-linear maps)
constexpr int kDotBlocks = 64;
__global__ __launch_bounds__(256) void k_pair_dots(const c64 *const *__restrict__ A, const c64 *const *__restrict__ B,
                                                   long n, double *__restrict__ partial) {
    const int p = blockIdx.y;
    const c64 *a = A[p], *b = B[p];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        acc += a[i].re * b[i].re - a[i].im * b[i].im;
    __shared__ double sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)p * gridDim.x + blockIdx.x] = sh[0];
}

// Polarisation P_a (first order: 2 e14 e_a + (2 sum_s d e14_s chi_s) (*) e_a, with e_a = e_YZ, e_XZ,
// e_XY; second order: the weighted strain products and their convolutions with the shape sums)
// and the potential phi(G) = -100 i (G . P) / (eps0 1e12 nr0^2 G^2), zero at G = 0.
struct PiezoConsts {
    double e14, B114, B124, B156, nr0;
    int second;
};
__global__ __launch_bounds__(256) void k_piezo_phi(const c64 *__restrict__ e, const c64 *__restrict__ first,
                                                   const c64 *__restrict__ prod, const c64 *__restrict__ conv2,
                                                   const double *__restrict__ gx, const double *__restrict__ gy,
                                                   const double *__restrict__ gz, int n0, int n1, int n2, PiezoConsts m,
                                                   c64 *__restrict__ phi) {
    const long n = (long)n0 * n1 * n2;
    const int comp[3] = {YZ, XZ, XY};
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        const double G[3] = {gx[ix], gy[iy], gz[iz]};
        const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        c64 acc = mk(0.0, 0.0);
        for (int a = 0; a < 3; ++a) {
            c64 P = 2.0 * m.e14 * e[(long)comp[a] * n + i] + first[(long)a * n + i];
            if (m.second) {
                const c64 p0 = prod[(long)(4 * a) * n + i], p12 = prod[(long)(4 * a + 1) * n + i] + prod[(long)(4 * a + 2) * n + i];
                const c64 p3 = prod[(long)(4 * a + 3) * n + i];
                P = P + (2.0 * m.B114) * p0 + (2.0 * m.B124) * p12 + (4.0 * m.B156) * p3;
                P = P + 2.0 * conv2[(long)a * n + i] + 2.0 * conv2[(long)(3 + a) * n + i] + 4.0 * conv2[(long)(6 + a) * n + i];
            }
            const double coef = -100.0 * G[a] / (8.854188 * m.nr0 * m.nr0 * G2);   // times i
            acc = acc + coef * P;
        }
        phi[i] = sqrt(G2) <= kEps ? mk(0.0, 0.0) : mk(-acc.im, acc.re);
    }
}

// partial[(s * 3 + a) * gridDim.x + blockIdx.x] = sum_i G_a(i) Im(g_s[i] chi_s[i]) over the block's
// share of the grid, shape s = blockIdx.y: Re sum_i g_s[i] (-i G_a chi_s[i]), the gradient's
// contraction with d chi_s / d O_a (the centre enters chi_s only through exp(-i G.O)).
__global__ __launch_bounds__(256) void k_centre_dots(const c64 *__restrict__ g, const c64 *__restrict__ chi,
                                                     const double *__restrict__ gx, const double *__restrict__ gy,
                                                     const double *__restrict__ gz, int n0, int n1, int n2,
                                                     double *__restrict__ partial) {
    const long n = (long)n0 * n1 * n2;
    const c64 *gs = g + (long)blockIdx.y * n, *cs = chi + (long)blockIdx.y * n;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        const c64 t = gs[i] * cs[i];
        acc[0] += gx[ix] * t.im;
        acc[1] += gy[iy] * t.im;
        acc[2] += gz[iz] * t.im;
    }
    __shared__ double sh[3][256];
    for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) partial[((long)blockIdx.y * 3 + a) * gridDim.x + blockIdx.x] = sh[a][0];
}

// Reverse of k_piezo_phi's last step: H_a = G_phi i coef_a(G) (zero at G = 0), the gradients of
// the polarisation components.
__global__ __launch_bounds__(256) void k_piezo_adj(const c64 *__restrict__ Gphi, const double *__restrict__ gx,
                                                   const double *__restrict__ gy, const double *__restrict__ gz, int n0,
                                                   int n1, int n2, double nr0, c64 *__restrict__ H) {
    const long n = (long)n0 * n1 * n2;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        const double G[3] = {gx[ix], gy[iy], gz[iz]};
        const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        const c64 g = Gphi[i];
        for (int a = 0; a < 3; ++a) {
            const double coef = sqrt(G2) <= kEps ? 0.0 : -100.0 * G[a] / (8.854188 * nr0 * nr0 * G2);
            H[(long)a * n + i] = mk(-coef * g.im, coef * g.re);
        }
    }
}

// Reverse of k_strain: G_Xeps[i] = sum_c bracket_c(G_i) G_e[c][i]
__global__ __launch_bounds__(256) void k_strain_adj(const c64 *__restrict__ Ge, const double *__restrict__ gx,
                                                    const double *__restrict__ gy, const double *__restrict__ gz, int n0,
                                                    int n1, int n2, Elastic m, c64 *__restrict__ GX) {
    const long n = (long)n0 * n1 * n2;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int ix = (int)(i / ((long)n1 * n2)), iy = (int)((i / n2) % n1), iz = (int)(i % n2);
        c64 acc = mk(0.0, 0.0);
        for (int c = 0; c < 6; ++c) acc = acc + strain_bracket_d(c, gx[ix], gy[iy], gz[iz], m) * Ge[(long)c * n + i];
        GX[i] = acc;
    }
}

// The transform of the uniform field's sawtooth potential along axis a at the Delta-G offsets dd,
// per V/Angstrom of F_a (times i): nonzero only on that axis.
__host__ __device__ inline double sawtooth_coef(int a, const int dd[3], const double Lc[3]) {
    const int o1 = dd[(a + 1) % 3], o2 = dd[(a + 2) % 3];
    if (dd[a] == 0 || o1 != 0 || o2 != 0) return 0.0;
    const double sign = (dd[a] % 2 == 0) ? -1.0 : 1.0;
    return sign / (-(double)dd[a] * 2.0 * kPiD / Lc[a]);
}

// The 20 tables of one Delta-G point.  XE: the 13 shape sums sum_s (p_s - p_host) chi_s of E_cb,
// E_hh, E_lh, E_so, sc, g1, g2, g3, P, ac, av, b, d; eE: the strain on this grid; H: hyd, bia,
// dxy; C: the 13 convolutions hac, hav, bb, rbs, rd, sxz, syz, p6[6]; phi may be null.
enum { W_ECB = 0, W_EHH, W_ELH, W_ESO, W_SC, W_G1, W_G2, W_G3, W_P, W_AC, W_AV, W_B, W_D, W_COUNT };
struct TableConsts {
    double p0[W_COUNT];   // the matrix's parameters (the Delta-G = 0 term of each expansion)
    double b0, d0, ac0, av0, P0;
    double F[3], Lc[3];
    int NT[3], strain;
};
__global__ __launch_bounds__(256) void k_tables(const c64 *__restrict__ XE, const c64 *__restrict__ eE,
                                                const c64 *__restrict__ H, const c64 *__restrict__ C,
                                                const c64 *__restrict__ phi, TableConsts m, c64 *__restrict__ tab) {
    const long n = (long)m.NT[0] * m.NT[1] * m.NT[2];
    const double r3 = sqrt(3.0), rh = sqrt(0.5);
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int dd[3] = {(int)(i / ((long)m.NT[1] * m.NT[2])) - m.NT[0] / 2, (int)((i / m.NT[2]) % m.NT[1]) - m.NT[1] / 2,
                           (int)(i % m.NT[2]) - m.NT[2] / 2};
        const bool centre = dd[0] == 0 && dd[1] == 0 && dd[2] == 0;
        c64 p[W_COUNT];
        for (int k = 0; k < W_COUNT; ++k) {
            p[k] = XE[(long)k * n + i];
            if (centre) p[k].re += m.p0[k];
        }
        // uniform field: the sawtooth potential's transform along one axis (times i)
        double fim = 0.0;
        for (int a = 0; a < 3; ++a) fim += m.F[a] * sawtooth_coef(a, dd, m.Lc);
        c64 Ecb = p[W_ECB], Ehh = p[W_EHH], Elh = p[W_ELH], Eso = p[W_ESO];
        c64 Q = mk(0.0, 0.0), Rb = Q, Rd = Q, Sxz = Q, Syz = Q, P6[6] = {Q, Q, Q, Q, Q, Q};
        if (m.strain) {
            const c64 hyd = H[i], bia = H[n + i], dxy = H[2 * n + i];
            const c64 ph = phi ? phi[i] : mk(0.0, 0.0);
            const c64 hac = C[i], hav = C[n + i], bb = C[2 * n + i], rbs = C[3 * n + i];
            Ecb = Ecb + m.ac0 * hyd + hac + ph;
            Ehh = Ehh + m.av0 * hyd + (0.5 * m.b0) * bia + hav + 0.5 * bb + ph;
            Elh = Elh + m.av0 * hyd - (0.5 * m.b0) * bia + hav - 0.5 * bb + ph;
            Eso = Eso + m.av0 * hyd + hav + ph;
            Q = rh * (m.b0 * bia + bb);
            Rb = (-0.5 * r3) * (m.b0 * dxy + rbs);
            Rd = m.d0 * eE[(long)XY * n + i] + C[4 * n + i];
            Sxz = m.d0 * eE[(long)XZ * n + i] + C[5 * n + i];
            Syz = m.d0 * eE[(long)YZ * n + i] + C[6 * n + i];
            for (int k = 0; k < 6; ++k) P6[k] = m.P0 * eE[(long)k * n + i] + C[(long)(7 + k) * n + i];
        }
        const c64 fld = mk(0.0, fim);
        tab[(long)QD_T_ECB * n + i] = Ecb + fld;
        tab[(long)QD_T_EHH * n + i] = Ehh + fld;
        tab[(long)QD_T_ELH * n + i] = Elh + fld;
        tab[(long)QD_T_ESO * n + i] = Eso + fld;
        tab[(long)QD_T_SC * n + i] = p[W_SC];
        tab[(long)QD_T_G1 * n + i] = p[W_G1];
        tab[(long)QD_T_G2 * n + i] = p[W_G2];
        tab[(long)QD_T_G3 * n + i] = p[W_G3];
        tab[(long)QD_T_P * n + i] = p[W_P];
        tab[(long)QD_T_Q * n + i] = Q;
        tab[(long)QD_T_RB * n + i] = Rb;
        tab[(long)QD_T_RD * n + i] = Rd;
        tab[(long)QD_T_SXZ * n + i] = Sxz;
        tab[(long)QD_T_SYZ * n + i] = Syz;
        for (int k = 0; k < 6; ++k) tab[(long)(QD_T_PXX + k) * n + i] = P6[k];
    }
}

// Truncated linear convolutions of grids of n[3] points on the device.
struct Src {
    const c64 *p;
    PadMap map;
};
using Sum = std::vector<std::pair<int, int>>;   // spectrum index pairs whose products add up

struct ConvPlan {
    int n[3], L[3], logL[3];
    long grid;
    const c64 *tw[3];

    ConvPlan(const int npts[3], DevPool &pool, hipStream_t s) {
        for (int a = 0; a < 3; ++a) {
            n[a] = npts[a];
            logL[a] = std::max(1, ilog2_ceil(2 * n[a] - 1));
            L[a] = 1 << logL[a];
            if (L[a] > kMaxAxis)
                throw StatusError("coefficient grid above " + std::to_string(kMaxAxis / 2) + " points per axis",
                                  QD_E_UNSUPPORTED);
            std::vector<c64> h(L[a] / 2);
            for (int j = 0; j < L[a] / 2; ++j) {
                const double ang = -2.0 * kPiD * (double)j / (double)L[a];
                h[j] = mk(std::cos(ang), std::sin(ang));
            }
            tw[a] = pool.upload(h.data(), h.size(), s);
            QD_HIP(hipStreamSynchronize(s));   // h leaves scope
        }
        grid = (long)L[0] * L[1] * L[2];
    }
    // an n-point grid at the origin (rev: reversed through its centre)
    Src at(const c64 *p, int rev = 0) const { return Src{p, PadMap{{n[0], n[1], n[2]}, {0, 0, 0}, {1, 1, 1}, rev}}; }
    // spectra of the zero-padded grids into spec[k]
    void forward(const std::vector<Src> &srcs, c64 *spec, hipStream_t s) const {
        for (size_t k = 0; k < srcs.size(); ++k) {
            hipLaunchKernelGGL(k_pad_grid, dim3(grid_for(grid)), dim3(256), 0, s, srcs[k].p, srcs[k].map, L[0], L[1], L[2],
                               spec + (long)k * grid);
            QD_LAUNCH_CHECK();
        }
        fft3_device(spec, (long)srcs.size(), logL, tw, false, s);
    }
    // the sums of products of spectrum pairs into prod, inverse transforms, and the blocks
    // start + step k (k < nout) of each into out (nout-point grids, consecutive)
    void combine(const c64 *spec, const std::vector<Sum> &outs, c64 *prod, const int nout[3], const int start[3],
                 const int step[3], c64 *out, hipStream_t s) const {
        const int no = (int)outs.size();
        for (int first = 0; first < no;) {
            Terms tm{};
            int nt = 0, j = 0;
            while (first + j < no && j < kMaxOut && nt + (int)outs[first + j].size() <= kMaxTerms) {
                for (const auto &pr : outs[first + j]) {
                    tm.a[nt] = pr.first;
                    tm.b[nt] = pr.second;
                    ++nt;
                }
                tm.off[++j] = nt;
            }
            if (j == 0) throw std::logic_error("a convolution sum with more terms than one launch takes");
            tm.nout = j;
            hipLaunchKernelGGL(k_products, dim3(grid_for(grid, 4096), j), dim3(256), 0, s, spec, grid, tm,
                               prod + (long)first * grid);
            QD_LAUNCH_CHECK();
            first += j;
        }
        fft3_device(prod, no, logL, tw, true, s);
        const long nb = (long)nout[0] * nout[1] * nout[2];
        hipLaunchKernelGGL(k_take_block, dim3(grid_for(nb * no)), dim3(256), 0, s, prod, L[0], L[1], L[2], nout[0],
                           nout[1], nout[2], start[0], start[1], start[2], step[0], step[1], step[2], (long)no,
                           1.0 / (double)grid, out);
        QD_LAUNCH_CHECK();
    }
    // the central n-point blocks
    void central(const c64 *spec, const std::vector<Sum> &outs, c64 *prod, c64 *out, hipStream_t s) const {
        const int start[3] = {n[0] / 2, n[1] / 2, n[2] / 2}, one[3] = {1, 1, 1};
        combine(spec, outs, prod, n, start, one, out, s);
    }
};

std::vector<Sum> singles(const std::vector<std::pair<int, int>> &pairs) {
    std::vector<Sum> o;
    for (const auto &p : pairs) o.push_back(Sum{p});
    return o;
}

// y = sum_k c_k x_k over up to four terms (none: zeros); y may be one of the x_k
void lin(c64 *y, long n, std::initializer_list<std::pair<double, const c64 *>> terms, hipStream_t s) {
    Lin l{};
    for (const auto &t : terms) {
        if (l.n == 4) throw std::logic_error("lin takes at most four terms");
        l.c[l.n] = t.first;
        l.x[l.n] = t.second;
        ++l.n;
    }
    hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n)), dim3(256), 0, s, y, n, l);
    QD_LAUNCH_CHECK();
}

// sum_s (f(shape s) - f(its host)) chi_s for a list of parameter fields, on one grid
void weighted(const qd_inputs &in, const std::vector<Material> &mp, const c64 *chi, long n,
              const std::vector<std::function<double(const Material &)>> &fs, c64 *out, DevPool &pool, hipStream_t s) {
    const int nsh = in.N_shape - 1, nw = (int)fs.size();   // nsh = 0 writes zeros
    std::vector<double> w((size_t)nw * nsh);
    for (int k = 0; k < nw; ++k)
        for (int sh = 1; sh <= nsh; ++sh) w[(size_t)k * nsh + sh - 1] = fs[k](mp[sh]) - fs[k](mp[in.shapes[sh].embedded]);
    for (int first = 0; first < nw; first += kMaxWeights) {
        const int cnt = std::min(kMaxWeights, nw - first);
        const double *wd = pool.upload(w.data() + (size_t)first * nsh, (size_t)cnt * nsh, s);
        hipLaunchKernelGGL(k_weighted, dim3(grid_for(n)), dim3(256), 0, s, chi, nsh, n, wd, cnt, out + (long)first * n);
        QD_LAUNCH_CHECK();
    }
    QD_HIP(hipStreamSynchronize(s));   // w leaves scope
}

// Re sum_i A_p[i] B_p[i] for the listed pairs of n-point grids (partial sums per work-group, added
// on the host in a fixed order)
std::vector<double> pair_dots(const std::vector<std::pair<const c64 *, const c64 *>> &pairs, long n, DevPool &pool,
                              hipStream_t s) {
    const int np = (int)pairs.size();
    std::vector<double> r(np, 0.0);
    if (np == 0) return r;
    std::vector<const c64 *> A(np), B(np);
    for (int p = 0; p < np; ++p) {
        A[p] = pairs[p].first;
        B[p] = pairs[p].second;
    }
    const c64 *const *Ad = pool.upload(A.data(), A.size(), s);
    const c64 *const *Bd = pool.upload(B.data(), B.size(), s);
    const int nblk = (int)std::min<long>(kDotBlocks, std::max<long>(1, (n + 255) / 256));
    double *part = pool.take<double>((size_t)np * nblk);
    constexpr int kMaxRows = 32768;
    for (int first = 0; first < np; first += kMaxRows) {
        const int cnt = std::min(kMaxRows, np - first);
        hipLaunchKernelGGL(k_pair_dots, dim3(nblk, cnt), dim3(256), 0, s, Ad + first, Bd + first, n,
                           part + (size_t)first * nblk);
        QD_LAUNCH_CHECK();
    }
    std::vector<double> h((size_t)np * nblk);
    QD_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < np; ++p)
        for (int b = 0; b < nblk; ++b) r[p] += h[(size_t)p * nblk + b];
    return r;
}

// Gradient terms G of parameter-weighted shape sums X = sum_s factor (f(s) - f(host of s)) chi_s:
// gchi[s] += factor (f(s) - f(host)) G, and dL/df of shape s and of its host +- factor Re(G . chi_s).
struct ShapeTerm {
    const c64 *G;
    double Material::*f;
    double factor;
};
void shape_terms(const qd_inputs &in, const std::vector<Material> &mp, const std::vector<ShapeTerm> &terms,
                 const c64 *chi, long n, c64 *gchi, std::vector<Material> &gmp, DevPool &pool, hipStream_t s) {
    const int nsh = in.N_shape - 1;
    if (nsh == 0 || terms.empty()) return;
    const int nt = (int)terms.size();
    for (int first = 0; first < nt; first += kMaxWeights) {
        const int cnt = std::min(kMaxWeights, nt - first);
        std::vector<const c64 *> G(cnt);
        std::vector<double> c((size_t)nsh * cnt);
        for (int k = 0; k < cnt; ++k) {
            const ShapeTerm &t = terms[first + k];
            G[k] = t.G;
            for (int sh = 1; sh <= nsh; ++sh) c[(size_t)(sh - 1) * cnt + k] = t.factor * (mp[sh].*t.f - mp[in.shapes[sh].embedded].*t.f);
        }
        const c64 *const *Gd = pool.upload(G.data(), G.size(), s);
        const double *cd = pool.upload(c.data(), c.size(), s);
        hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(256), 0, s, Gd, cnt, cd, nsh, n, gchi);
        QD_LAUNCH_CHECK();
        QD_HIP(hipStreamSynchronize(s));   // G, c leave scope
    }
    std::vector<std::pair<const c64 *, const c64 *>> pairs;
    for (const auto &t : terms)
        for (int sh = 0; sh < nsh; ++sh) pairs.push_back({t.G, chi + (long)sh * n});
    const std::vector<double> R = pair_dots(pairs, n, pool, s);
    for (int k = 0; k < nt; ++k)
        for (int sh = 1; sh <= nsh; ++sh) {
            const double g = terms[k].factor * R[(size_t)k * nsh + sh - 1];
            gmp[sh].*terms[k].f += g;
            gmp[in.shapes[sh].embedded].*terms[k].f -= g;
        }
}

// gc[3 s + a] += Re sum_i g_s[i] (-i G_a chi_s[i]) for the shapes s = 1 .. nsh of one grid
void centre_dots(const c64 *g, const c64 *chi, int nsh, const double *const ax[3], const int n[3],
                 std::vector<double> &gc, DevPool &pool, hipStream_t s) {
    if (nsh == 0) return;
    const long np = (long)n[0] * n[1] * n[2];
    const int nblk = (int)std::min<long>(kDotBlocks, std::max<long>(1, (np + 255) / 256));
    double *part = pool.take<double>((size_t)nsh * 3 * nblk);
    hipLaunchKernelGGL(k_centre_dots, dim3(nblk, nsh), dim3(256), 0, s, g, chi, ax[0], ax[1], ax[2], n[0], n[1], n[2],
                       part);
    QD_LAUNCH_CHECK();
    std::vector<double> h((size_t)nsh * 3 * nblk);
    QD_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    for (int sh = 0; sh < nsh; ++sh)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < nblk; ++b) gc[3 * (sh + 1) + a] += h[((size_t)sh * 3 + a) * nblk + b];
}

ShapeDev device_shape(const qd_inputs &in, int idx, Cell cell) {
    try {
        return shape_device_params(in, idx, cell);
    } catch (const std::runtime_error &e) {   // inputs the host transform rejects as well
        throw StatusError(e.what(), QD_E_INPUT);
    }
}

// chi_s of shapes 1 .. N_shape - 1 on one grid (slot s - 1): the host's grids where it evaluated
// them, the device's for the others (left empty by build_geometry)
c64 *shape_grids(const qd_inputs &in, const std::vector<Grid3> &host, Cell cell, const double *const g[3],
                 const int n[3], const double *gl, DevPool &pool, hipStream_t s) {
    const long np = (long)n[0] * n[1] * n[2];
    const int cnt = in.N_shape - 1;
    c64 *d = pool.take<c64>((size_t)std::max(cnt, 1) * np);
    std::map<int, std::pair<std::vector<ShapeDev>, std::vector<int>>> fam;   // per kernel family
    for (int k = 1; k <= cnt; ++k) {
        if (!host[k].v.empty()) {
            QD_HIP(hipMemcpyAsync(d + (long)(k - 1) * np, host[k].v.data(), sizeof(c64) * np, hipMemcpyHostToDevice, s));
            continue;
        }
        auto &f = fam[shape_family(in.shapes[k].shape)];
        f.first.push_back(device_shape(in, k, cell));
        f.second.push_back(k - 1);
    }
    for (const auto &f : fam) {
        const ShapeDev *sdd = pool.upload(f.second.first.data(), f.second.first.size(), s);
        const int *sl = pool.upload(f.second.second.data(), f.second.second.size(), s);
        launch_shape_grid(sdd, sl, (int)f.second.first.size(), f.first, g, n, gl, d, s);
    }
    QD_HIP(hipStreamSynchronize(s));   // the shape lists leave scope
    return d;
}

// second-order strain products e_a (*) e_b grouped per polarisation component: P_x from products
// 0..3, P_y from 4..7, P_z from 8..11, weights (2 B114, 2 B124, 2 B124, 4 B156)
const int kProdDev[12][2] = {{XX, YZ}, {YY, YZ}, {ZZ, YZ}, {XY, XZ}, {YY, XZ}, {XX, XZ},
                             {ZZ, XZ}, {XY, YZ}, {ZZ, XY}, {XX, XY}, {YY, XY}, {XZ, YZ}};
const int kFirstComp[3] = {YZ, XZ, XY};   // the strain component of each first-order polarisation

// The forward pass's fields on the device, every buffer owned by pool.
struct DevFields {
    DevPool pool;
    const double *gl = nullptr, *gS[3] = {nullptr, nullptr, nullptr}, *gE[3] = {nullptr, nullptr, nullptr};
    const c64 *chiS = nullptr, *chiE = nullptr, *chi0 = nullptr;
    c64 *XS = nullptr;    // strain grid: shape sums of the misfit, 2 e14, B114, B124, B156
    c64 *fS = nullptr;    // strain grid: the six strain components, then phi
    c64 *RA = nullptr;    // strain grid: the first-order convolutions (3), then the second-order products (12)
    c64 *p12 = nullptr;   // strain grid: products 1 + 2 of each polarisation component
    const c64 *eE = nullptr, *phiE = nullptr;   // electronic grid: strain, phi
    c64 *XE = nullptr;    // electronic grid: the W_COUNT shape sums
    c64 *H = nullptr;     // electronic grid: hyd, bia, dxy
    long nSt = 0, nT = 0;
    bool piezo = false, second = false;
};

// Free-length transfer matrices per axis (NT x n_S, box transforms of the strain cell's waves over
// the electronic cell, 1/volume folded into axis 0); transposed for the reverse pass.
std::vector<double> box_matrix(const TableState &st, int a, bool transpose) {
    const int ne = (int)st.gE[a].size(), ns = (int)st.gS[a].size();
    const double inv = 1.0 / (st.Lc[0] * st.Lc[1] * st.Lc[2]);
    std::vector<double> I((size_t)ne * ns);
    for (int g = 0; g < ne; ++g)
        for (int k = 0; k < ns; ++k) {
            const double dG = st.gS[a][k] - st.gE[a][g];
            double v = std::fabs(dG) < kEps ? st.Lc[a] : 2.0 * std::sin(dG * st.Lc[a] / 2.0) / dG;
            if (a == 0) v *= inv;
            if (transpose) I[(size_t)k * ne + g] = v;
            else I[(size_t)g * ns + k] = v;
        }
    return I;
}

void build_fields(const qd_inputs &in, const TableState &st, hipStream_t s, DevFields &f) {
    DevPool &pool = f.pool;
    const std::vector<Material> &mp = st.mp;
    const Material &m0 = mp[0];
    const int *NT = st.NT;
    const long nT = (long)NT[0] * NT[1] * NT[2];
    f.nT = nT;
    const std::vector<double> rules = gauss_legendre_rules();
    f.gl = pool.upload(rules.data(), rules.size(), s);
    for (int a = 0; a < 3; ++a) f.gE[a] = pool.upload(st.gE[a].data(), st.gE[a].size(), s);
    QD_HIP(hipStreamSynchronize(s));   // rules leaves scope
    f.piezo = st.use_strain && in.piezo != 0;
    f.second = f.piezo && (in.piezo == 2 || in.piezo == 3);
    if (st.use_strain) {
        const int *nS = st.nS;
        const long nSt = (long)nS[0] * nS[1] * nS[2];
        f.nSt = nSt;
        for (int a = 0; a < 3; ++a) f.gS[a] = pool.upload(st.gS[a].data(), st.gS[a].size(), s);
        f.chiS = shape_grids(in, st.chiS, Cell::Strain, f.gS, nS, f.gl, pool, s);
        f.XS = pool.take<c64>(5 * nSt);
        const double a0 = m0.a;
        weighted(in, mp, f.chiS, nSt,
                 {[a0](const Material &m) { return -m.a / a0; }, [](const Material &m) { return 2.0 * m.e14; },
                  [](const Material &m) { return m.B114; }, [](const Material &m) { return m.B124; },
                  [](const Material &m) { return m.B156; }},
                 f.XS, pool, s);
        f.fS = pool.take<c64>(7 * nSt);
        const Elastic el{m0.C11, m0.C12, m0.C44, st.iso ? 1 : 0};
        hipLaunchKernelGGL(k_strain, dim3(grid_for(nSt)), dim3(256), 0, s, f.XS, f.gS[0], f.gS[1], f.gS[2], nS[0], nS[1],
                           nS[2], el, f.fS);
        QD_LAUNCH_CHECK();
        const int nf = f.piezo ? 7 : 6;
        if (f.piezo || st.commensurate) {
            const ConvPlan cv(nS, pool, s);
            if (f.piezo) {
                // stage A: 2 sum d e14 chi with the first-order components, the strain products
                c64 *spec = pool.take<c64>((size_t)7 * cv.grid);
                std::vector<Src> src = {cv.at(f.XS + nSt)};
                for (int c = 0; c < 6; ++c) src.push_back(cv.at(f.fS + c * nSt));
                cv.forward(src, spec, s);
                std::vector<std::pair<int, int>> pa;
                for (int a = 0; a < 3; ++a) pa.push_back({0, 1 + kFirstComp[a]});
                if (f.second)
                    for (const auto &q : kProdDev) pa.push_back({1 + q[0], 1 + q[1]});
                c64 *prodA = pool.take<c64>(pa.size() * cv.grid);
                f.RA = pool.take<c64>(pa.size() * nSt);
                cv.central(spec, singles(pa), prodA, f.RA, s);
                c64 *conv2 = nullptr;
                if (f.second) {
                    // stage B: the shape sums of B114, B124, B156 with p0, p1 + p2, p3 per component
                    const c64 *prod = f.RA + 3 * nSt;
                    f.p12 = pool.take<c64>(3 * nSt);
                    for (int a = 0; a < 3; ++a)
                        lin(f.p12 + a * nSt, nSt, {{1.0, prod + (4 * a + 1) * nSt}, {1.0, prod + (4 * a + 2) * nSt}}, s);
                    std::vector<Src> srcB = {cv.at(f.XS + 2 * nSt), cv.at(f.XS + 3 * nSt), cv.at(f.XS + 4 * nSt)};
                    for (int a = 0; a < 3; ++a) srcB.push_back(cv.at(prod + 4 * a * nSt));
                    for (int a = 0; a < 3; ++a) srcB.push_back(cv.at(f.p12 + a * nSt));
                    for (int a = 0; a < 3; ++a) srcB.push_back(cv.at(prod + (4 * a + 3) * nSt));
                    c64 *specB = pool.take<c64>(srcB.size() * cv.grid);
                    cv.forward(srcB, specB, s);
                    std::vector<std::pair<int, int>> pb;
                    for (int a = 0; a < 3; ++a) pb.push_back({3 + a, 0});
                    for (int a = 0; a < 3; ++a) pb.push_back({6 + a, 1});
                    for (int a = 0; a < 3; ++a) pb.push_back({9 + a, 2});
                    c64 *prodB = pool.take<c64>(pb.size() * cv.grid);
                    conv2 = pool.take<c64>(pb.size() * nSt);
                    cv.central(specB, singles(pb), prodB, conv2, s);
                }
                const PiezoConsts pc{m0.e14, m0.B114, m0.B124, m0.B156, m0.nr, f.second ? 1 : 0};
                hipLaunchKernelGGL(k_piezo_phi, dim3(grid_for(nSt)), dim3(256), 0, s, f.fS, f.RA, f.RA + 3 * nSt, conv2,
                                   f.gS[0], f.gS[1], f.gS[2], nS[0], nS[1], nS[2], pc, f.fS + 6 * nSt);
                QD_LAUNCH_CHECK();
            }
            if (st.commensurate) {
                // the electronic cell's transform convolved with every field, sampled every RL
                // strain-grid points from n_S - 1 - RL (N_T / 2)
                c64 *chi0;
                if (!st.chi0S.v.empty()) {
                    chi0 = pool.upload(reinterpret_cast<const c64 *>(st.chi0S.v.data()), st.chi0S.v.size(), s);
                } else {
                    chi0 = pool.take<c64>(nSt);
                    const ShapeDev cell0 = device_shape(in, 0, Cell::Strain);
                    const ShapeDev *sd0 = pool.upload(&cell0, 1, s);
                    const int zero = 0;
                    const int *sl0 = pool.upload(&zero, 1, s);
                    launch_shape_grid(sd0, sl0, 1, shape_family(in.shapes[0].shape), f.gS, nS, f.gl, chi0, s);
                }
                QD_HIP(hipStreamSynchronize(s));   // upload sources
                f.chi0 = chi0;
                std::vector<Src> src;
                for (int k = 0; k < nf; ++k) src.push_back(cv.at(f.fS + k * nSt));
                src.push_back(cv.at(chi0));
                c64 *spec = pool.take<c64>(src.size() * cv.grid);
                cv.forward(src, spec, s);
                std::vector<std::pair<int, int>> pc;
                for (int k = 0; k < nf; ++k) pc.push_back({k, nf});
                c64 *prod = pool.take<c64>(pc.size() * cv.grid);
                c64 *out = pool.take<c64>((size_t)nf * nT);
                int start[3];
                for (int a = 0; a < 3; ++a) start[a] = (nS[a] - 1) - st.RL[a] * (NT[a] / 2);
                cv.combine(spec, singles(pc), prod, NT, start, st.RL, out, s);
                f.eE = out;
                f.phiE = f.piezo ? out + 6 * nT : nullptr;
            }
        }
        if (!st.commensurate) {
            if (st.same) {
                f.eE = f.fS;
                f.phiE = f.piezo ? f.fS + 6 * nSt : nullptr;
            } else {
                // free lengths: the box-transform matrices per axis, z, then y, then x
                const c64 *cur = f.fS;
                int dims[3] = {nS[0], nS[1], nS[2]};
                for (int a = 2; a >= 0; --a) {
                    const std::vector<double> I = box_matrix(st, a, false);
                    const double *M = pool.upload(I.data(), I.size(), s);
                    int od[3] = {dims[0], dims[1], dims[2]};
                    od[a] = NT[a];
                    c64 *nxt = pool.take<c64>((size_t)nf * od[0] * od[1] * od[2]);
                    apply_axis_device(cur, dims, a, NT[a], M, nf, nxt, s);
                    QD_HIP(hipStreamSynchronize(s));   // I leaves scope
                    cur = nxt;
                    dims[a] = NT[a];
                }
                f.eE = cur;
                f.phiE = f.piezo ? cur + 6 * nT : nullptr;
            }
        }
    }
    // the parameter expansions on the electronic grid
    f.chiE = (st.use_strain && st.same) ? f.chiS : shape_grids(in, st.chiE, Cell::Electronic, f.gE, NT, f.gl, pool, s);
    f.XE = pool.take<c64>((size_t)W_COUNT * nT);
    weighted(in, mp, f.chiE, nT,
             {[](const Material &m) { return m.E_cb; }, [](const Material &m) { return m.E_hh; },
              [](const Material &m) { return m.E_lh; }, [](const Material &m) { return m.E_so; },
              [](const Material &m) { return m.sc; }, [](const Material &m) { return m.g1; },
              [](const Material &m) { return m.g2; }, [](const Material &m) { return m.g3; },
              [](const Material &m) { return m.P; }, [](const Material &m) { return m.ac; },
              [](const Material &m) { return m.av; }, [](const Material &m) { return m.b; },
              [](const Material &m) { return m.d; }},
             f.XE, pool, s);
    if (st.use_strain) {
        // hyd = XX + YY + ZZ, bia = XX + YY - 2 ZZ, dxy = XX - YY
        f.H = pool.take<c64>(3 * nT);
        const c64 *e = f.eE;
        lin(f.H, nT, {{1.0, e + XX * nT}, {1.0, e + YY * nT}, {1.0, e + ZZ * nT}}, s);
        lin(f.H + nT, nT, {{1.0, e + XX * nT}, {1.0, e + YY * nT}, {-2.0, e + ZZ * nT}}, s);
        lin(f.H + 2 * nT, nT, {{1.0, e + XX * nT}, {-1.0, e + YY * nT}}, s);
    }
}

}  // namespace

// Fills t exactly as compose_tables(in, build_state(in)) does (host copy, malloc'd like the host
// builder's), with every field and table computed on the device; st comes from build_geometry.
void tables_device(const qd_inputs &in, const TableState &st, hipStream_t s, qd_tables &t) {
    DevFields f;
    build_fields(in, st, s, f);
    DevPool &pool = f.pool;
    const Material &m0 = st.mp[0];
    const int *NT = st.NT;
    const long nT = f.nT;
    c64 *C = nullptr;
    if (st.use_strain) {
        const ConvPlan ce(NT, pool, s);
        // 0 Xac, 1 Xav, 2 Xb, 3 Xd, 4 XP, 5 hyd, 6 bia, 7 dxy, 8 + c strain component c
        const c64 *XE = f.XE, *H = f.H;
        std::vector<Src> src = {ce.at(XE + W_AC * nT), ce.at(XE + W_AV * nT), ce.at(XE + W_B * nT), ce.at(XE + W_D * nT),
                                ce.at(XE + W_P * nT),  ce.at(H),              ce.at(H + nT),         ce.at(H + 2 * nT)};
        for (int c = 0; c < 6; ++c) src.push_back(ce.at(f.eE + c * nT));
        c64 *spec = pool.take<c64>(src.size() * ce.grid);
        ce.forward(src, spec, s);
        std::vector<std::pair<int, int>> pc = {{0, 5}, {1, 5}, {6, 2}, {7, 2}, {8 + XY, 3}, {8 + XZ, 3}, {8 + YZ, 3}};
        for (int k = 0; k < 6; ++k) pc.push_back({8 + k, 4});   // XX, YY, ZZ, YZ, XZ, XY
        c64 *prod = pool.take<c64>(pc.size() * ce.grid);
        C = pool.take<c64>(pc.size() * nT);
        ce.central(spec, singles(pc), prod, C, s);
    }
    TableConsts tc{};
    const double Material::*pf[W_COUNT] = {&Material::E_cb, &Material::E_hh, &Material::E_lh, &Material::E_so,
                                           &Material::sc,   &Material::g1,   &Material::g2,   &Material::g3,
                                           &Material::P,    &Material::ac,   &Material::av,   &Material::b,
                                           &Material::d};
    for (int k = 0; k < W_COUNT; ++k) tc.p0[k] = m0.*pf[k];
    tc.b0 = m0.b; tc.d0 = m0.d; tc.ac0 = m0.ac; tc.av0 = m0.av; tc.P0 = m0.P;
    for (int a = 0; a < 3; ++a) {
        tc.F[a] = in.F[a];
        tc.Lc[a] = st.Lc[a];
        tc.NT[a] = NT[a];
    }
    tc.strain = st.use_strain ? 1 : 0;
    c64 *tab = pool.take<c64>((size_t)QD_NTAB * nT);
    hipLaunchKernelGGL(k_tables, dim3(grid_for(nT)), dim3(256), 0, s, f.XE, f.eE, f.H, C, f.phiE, tc, tab);
    QD_LAUNCH_CHECK();

    // the host copy, laid out as the host builder's
    t.N[0] = in.N[0]; t.N[1] = in.N[1]; t.N[2] = in.N[2];
    for (int a = 0; a < 3; ++a) {
        t.NT[a] = NT[a];
        t.K[a] = in.K[a];
        t.L[a] = st.Lc[a];
        t.G[a] = (double *)std::malloc(sizeof(double) * NT[a]);
        if (!t.G[a]) throw std::bad_alloc();
        std::memcpy(t.G[a], st.gE[a].data(), sizeof(double) * NT[a]);
    }
    t.N_vb = in.N_vb;
    t.N_cb = in.N_cb;
    t.nr0 = m0.nr;
    t.tab = (double *)std::malloc(sizeof(c64) * QD_NTAB * nT);
    if (!t.tab) throw std::bad_alloc();
    QD_HIP(hipMemcpyAsync(t.tab, tab, sizeof(c64) * QD_NTAB * nT, hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
}

// The reverse pass of the table builder on the device (tables_vjp in host/fields.cpp, step by step):
// with dL = Re sum_t sum_i W_t[i] dT_t[i], the gradients of the shape transforms on both grids, of
// the material fields per shape and of the field, after a forward pass on the device that keeps its
// fields.  A correlation c = X (*) Y sends G_X += G_c (*) reversed(Y) (complex-linear maps pass the
// gradient back without conjugation).
void tables_vjp_device(const qd_inputs &in, const TableState &st, const std::complex<double> *Wh, hipStream_t s,
                       int nparams, const int *params, const unsigned *masks, TableGrad &out) {
    DevFields f;
    build_fields(in, st, s, f);
    DevPool &pool = f.pool;
    const std::vector<Material> &mp = st.mp;
    const Material &m0 = mp[0];
    const int ns = in.N_shape, nsh = ns - 1;
    const int *NT = st.NT;
    const long nT = f.nT;
    Material zero;
    std::memset(&zero, 0, sizeof(zero));
    out.gmp.assign(ns, zero);
    out.gchiE.assign(ns, Grid3());
    out.gchiS.assign(ns, Grid3());
    out.gF[0] = out.gF[1] = out.gF[2] = 0.0;

    // the Delta-G = 0 terms of the parameter expansions and the field (a few points: on the host)
    const long centre = ((long)(NT[0] / 2) * NT[1] + NT[1] / 2) * NT[2] + NT[2] / 2;
    const struct { double Material::*f; int t; } pmes[9] = {
        {&Material::E_cb, QD_T_ECB}, {&Material::E_hh, QD_T_EHH}, {&Material::E_lh, QD_T_ELH}, {&Material::E_so, QD_T_ESO},
        {&Material::sc, QD_T_SC},    {&Material::g1, QD_T_G1},    {&Material::g2, QD_T_G2},    {&Material::g3, QD_T_G3},
        {&Material::P, QD_T_P}};
    for (const auto &pm : pmes) out.gmp[0].*pm.f += Wh[(size_t)pm.t * nT + centre].real();
    for (long i = 0; i < nT; ++i) {
        const int dd[3] = {(int)(i / ((long)NT[1] * NT[2])) - NT[0] / 2, (int)((i / NT[2]) % NT[1]) - NT[1] / 2,
                           (int)(i % NT[2]) - NT[2] / 2};
        std::complex<double> wedge = 0.0;
        for (int t = QD_T_ECB; t <= QD_T_ESO; ++t) wedge += Wh[(size_t)t * nT + i];
        for (int a = 0; a < 3; ++a) {
            const double c = sawtooth_coef(a, dd, st.Lc);
            if (c != 0.0) out.gF[a] += -wedge.imag() * c;   // Re(W i c)
        }
    }
    const c64 *W = pool.upload(reinterpret_cast<const c64 *>(Wh), (size_t)QD_NTAB * nT, s);
    auto Wt = [&](int t) { return W + (long)t * nT; };
    c64 *gE = pool.take<c64>((size_t)std::max(nsh, 1) * nT);
    lin(gE, (long)std::max(nsh, 1) * nT, {}, s);
    std::vector<ShapeTerm> eterms;
    for (const auto &pm : pmes) eterms.push_back({Wt(pm.t), pm.f, 1.0});
    std::vector<double> gcentre((size_t)3 * ns, 0.0);   // the centres' gradients, both grids
    const c64 *gSd = nullptr;                           // the strain-grid gradients of the shapes

    if (st.use_strain) {
        const double r3 = std::sqrt(3.0), rh = std::sqrt(0.5);
        const c64 *e = f.eE, *H = f.H;
        // gradients of the combinations: Gac (ac-weighted hydrostatic), Gav, Gb (biaxial), Grb (xx - yy)
        c64 *Gc = pool.take<c64>(4 * nT);
        lin(Gc, nT, {{1.0, Wt(QD_T_ECB)}}, s);
        lin(Gc + nT, nT, {{1.0, Wt(QD_T_EHH)}, {1.0, Wt(QD_T_ELH)}, {1.0, Wt(QD_T_ESO)}}, s);
        lin(Gc + 2 * nT, nT, {{0.5, Wt(QD_T_EHH)}, {-0.5, Wt(QD_T_ELH)}, {rh, Wt(QD_T_Q)}}, s);
        lin(Gc + 3 * nT, nT, {{-0.5 * r3, Wt(QD_T_RB)}}, s);
        const c64 *Gac = Gc, *Gav = Gc + nT, *Gb = Gc + 2 * nT, *Grb = Gc + 3 * nT;
        // strain components on the electronic grid (0..5) and the edge tables' sum for phi (6)
        c64 *GeE = pool.take<c64>(7 * nT);
        for (int c = 0; c < 6; ++c) {
            const int t = c == XY ? QD_T_RD : (c == XZ ? QD_T_SXZ : (c == YZ ? QD_T_SYZ : -1));
            if (t >= 0) lin(GeE + c * nT, nT, {{m0.P, Wt(QD_T_PXX + c)}, {m0.d, Wt(t)}}, s);
            else lin(GeE + c * nT, nT, {{m0.P, Wt(QD_T_PXX + c)}}, s);
        }
        lin(GeE + 6 * nT, nT, {{1.0, Wt(QD_T_ECB)}, {1.0, Wt(QD_T_EHH)}, {1.0, Wt(QD_T_ELH)}, {1.0, Wt(QD_T_ESO)}}, s);
        c64 *Gh = pool.take<c64>(3 * nT);   // Ghyd, Gbia, Gdxy
        lin(Gh, nT, {{m0.ac, Gac}, {m0.av, Gav}}, s);
        lin(Gh + nT, nT, {{m0.b, Gb}}, s);
        lin(Gh + 2 * nT, nT, {{m0.b, Grb}}, s);
        {   // the matrix's parameters
            std::vector<std::pair<const c64 *, const c64 *>> pr = {
                {Gac, H}, {Gav, H}, {Gb, H + nT}, {Grb, H + 2 * nT},
                {Wt(QD_T_RD), e + XY * nT}, {Wt(QD_T_SXZ), e + XZ * nT}, {Wt(QD_T_SYZ), e + YZ * nT}};
            for (int k = 0; k < 6; ++k) pr.push_back({Wt(QD_T_PXX + k), e + k * nT});
            const std::vector<double> r = pair_dots(pr, nT, pool, s);
            out.gmp[0].ac += r[0];
            out.gmp[0].av += r[1];
            out.gmp[0].b += r[2] + r[3];
            out.gmp[0].d += r[4] + r[5] + r[6];
            for (int k = 0; k < 6; ++k) out.gmp[0].P += r[7 + k];
        }
        {   // the shape parts: correlations with the reversed partners
            const ConvPlan ce(NT, pool, s);
            const c64 *XE = f.XE;
            std::vector<Src> src = {ce.at(Gac), ce.at(Gav), ce.at(Gb), ce.at(Grb), ce.at(Wt(QD_T_RD)), ce.at(Wt(QD_T_SXZ)),
                                    ce.at(Wt(QD_T_SYZ))};
            for (int k = 0; k < 6; ++k) src.push_back(ce.at(Wt(QD_T_PXX + k)));                  // 7..12
            for (int k = 0; k < 3; ++k) src.push_back(ce.at(H + k * nT, 1));                      // 13..15
            for (int w : {W_AC, W_AV, W_B, W_D, W_P}) src.push_back(ce.at(XE + w * nT, 1));      // 16..20
            for (int c = 0; c < 6; ++c) src.push_back(ce.at(e + c * nT, 1));                      // 21..26
            c64 *spec = pool.take<c64>(src.size() * ce.grid);
            ce.forward(src, spec, s);
            std::vector<Sum> o = {
                {{0, 13}},                                   // 0 GXac
                {{1, 13}},                                   // 1 GXav
                {{2, 14}, {3, 15}},                          // 2 GXb
                {{4, 21 + XY}, {5, 21 + XZ}, {6, 21 + YZ}},  // 3 GXd
                {},                                          // 4 GXP
                {{0, 16}, {1, 17}},                          // 5 to Ghyd
                {{2, 18}},                                   // 6 to Gbia
                {{3, 18}},                                   // 7 to Gdxy
                {{4, 19}}, {{5, 19}}, {{6, 19}}};            // 8..10 to GeE XY, XZ, YZ
            for (int k = 0; k < 6; ++k) o[4].push_back({7 + k, 21 + k});
            for (int k = 0; k < 6; ++k) o.push_back(Sum{{7 + k, 20}});                             // 11..16 to GeE[k]
            c64 *prod = pool.take<c64>(o.size() * ce.grid);
            c64 *O = pool.take<c64>(o.size() * nT);
            ce.central(spec, o, prod, O, s);
            lin(Gh, nT, {{1.0, Gh}, {1.0, O + 5 * nT}}, s);
            lin(Gh + nT, nT, {{1.0, Gh + nT}, {1.0, O + 6 * nT}}, s);
            lin(Gh + 2 * nT, nT, {{1.0, Gh + 2 * nT}, {1.0, O + 7 * nT}}, s);
            lin(GeE + XY * nT, nT, {{1.0, GeE + XY * nT}, {1.0, O + 8 * nT}}, s);
            lin(GeE + XZ * nT, nT, {{1.0, GeE + XZ * nT}, {1.0, O + 9 * nT}}, s);
            lin(GeE + YZ * nT, nT, {{1.0, GeE + YZ * nT}, {1.0, O + 10 * nT}}, s);
            for (int k = 0; k < 6; ++k) lin(GeE + k * nT, nT, {{1.0, GeE + k * nT}, {1.0, O + (11 + k) * nT}}, s);
            eterms.push_back({O, &Material::ac, 1.0});
            eterms.push_back({O + nT, &Material::av, 1.0});
            eterms.push_back({O + 2 * nT, &Material::b, 1.0});
            eterms.push_back({O + 3 * nT, &Material::d, 1.0});
            eterms.push_back({O + 4 * nT, &Material::P, 1.0});
        }
        // hydrostatic / biaxial / (xx - yy) back to the components
        lin(GeE + XX * nT, nT, {{1.0, GeE + XX * nT}, {1.0, Gh}, {1.0, Gh + nT}, {1.0, Gh + 2 * nT}}, s);
        lin(GeE + YY * nT, nT, {{1.0, GeE + YY * nT}, {1.0, Gh}, {1.0, Gh + nT}, {-1.0, Gh + 2 * nT}}, s);
        lin(GeE + ZZ * nT, nT, {{1.0, GeE + ZZ * nT}, {1.0, Gh}, {-2.0, Gh + nT}}, s);

        // --- the transfer from the strain cell, reversed: GeS (6 components, then phi)
        const int *nS = st.nS;
        const long nSt = f.nSt;
        const int nf = f.piezo ? 7 : 6;
        c64 *GeS = nullptr;
        const ConvPlan cv(nS, pool, s);
        if (st.same) {
            GeS = GeE;
        } else if (st.commensurate) {
            // the electronic-grid gradients placed at the sampled points, correlated with the
            // electronic cell's transform (a convolution with it reversed), the n_S points from n_S - 1
            std::vector<Src> src;
            for (int k = 0; k < nf; ++k) {
                PadMap m{{NT[0], NT[1], NT[2]}, {0, 0, 0}, {st.RL[0], st.RL[1], st.RL[2]}, 0};
                for (int a = 0; a < 3; ++a) m.o[a] = (nS[a] - 1) - st.RL[a] * (NT[a] / 2);
                src.push_back(Src{GeE + k * nT, m});
            }
            src.push_back(cv.at(f.chi0, 1));
            c64 *spec = pool.take<c64>(src.size() * cv.grid);
            cv.forward(src, spec, s);
            std::vector<std::pair<int, int>> pc;
            for (int k = 0; k < nf; ++k) pc.push_back({k, nf});
            c64 *prod = pool.take<c64>(pc.size() * cv.grid);
            GeS = pool.take<c64>((size_t)nf * nSt);
            const int first[3] = {nS[0] - 1, nS[1] - 1, nS[2] - 1}, one[3] = {1, 1, 1};
            cv.combine(spec, singles(pc), prod, nS, first, one, GeS, s);
        } else {
            // free lengths: the transposed box-transform matrices, x, then y, then z
            const c64 *cur = GeE;
            int dims[3] = {NT[0], NT[1], NT[2]};
            c64 *nxt = nullptr;
            for (int a = 0; a < 3; ++a) {
                const std::vector<double> I = box_matrix(st, a, true);
                const double *M = pool.upload(I.data(), I.size(), s);
                int od[3] = {dims[0], dims[1], dims[2]};
                od[a] = nS[a];
                nxt = pool.take<c64>((size_t)nf * od[0] * od[1] * od[2]);
                apply_axis_device(cur, dims, a, nS[a], M, nf, nxt, s);
                QD_HIP(hipStreamSynchronize(s));   // I leaves scope
                cur = nxt;
                dims[a] = nS[a];
            }
            GeS = nxt;
        }
        c64 *gS = gE;
        if (!st.same) {
            gS = pool.take<c64>((size_t)std::max(nsh, 1) * nSt);
            lin(gS, (long)std::max(nsh, 1) * nSt, {}, s);
        }
        gSd = gS;
        std::vector<ShapeTerm> sterms;
        // --- the piezoelectric potential
        if (f.piezo) {
            const c64 *GphiS = GeS + 6 * nSt;
            const c64 *eS = f.fS;
            c64 *Hs = pool.take<c64>(3 * nSt);
            hipLaunchKernelGGL(k_piezo_adj, dim3(grid_for(nSt)), dim3(256), 0, s, GphiS, f.gS[0], f.gS[1], f.gS[2], nS[0],
                               nS[1], nS[2], m0.nr, Hs);
            QD_LAUNCH_CHECK();
            {
                std::vector<std::pair<const c64 *, const c64 *>> pr = {{GphiS, f.fS + 6 * nSt}};
                for (int a = 0; a < 3; ++a) pr.push_back({Hs + a * nSt, eS + kFirstComp[a] * nSt});
                if (f.second)
                    for (int a = 0; a < 3; ++a) {
                        pr.push_back({Hs + a * nSt, f.RA + (3 + 4 * a) * nSt});
                        pr.push_back({Hs + a * nSt, f.p12 + a * nSt});
                        pr.push_back({Hs + a * nSt, f.RA + (3 + 4 * a + 3) * nSt});
                    }
                const std::vector<double> r = pair_dots(pr, nSt, pool, s);
                out.gmp[0].nr += r[0] * (-2.0 / m0.nr);
                for (int a = 0; a < 3; ++a) out.gmp[0].e14 += 2.0 * r[1 + a];
                if (f.second)
                    for (int a = 0; a < 3; ++a) {
                        out.gmp[0].B114 += 2.0 * r[4 + 3 * a];
                        out.gmp[0].B124 += 2.0 * r[5 + 3 * a];
                        out.gmp[0].B156 += 4.0 * r[6 + 3 * a];
                    }
            }
            // first order, matrix: P_a += 2 e14_0 e_first[a]
            for (int a = 0; a < 3; ++a)
                lin(GeS + kFirstComp[a] * nSt, nSt, {{1.0, GeS + kFirstComp[a] * nSt}, {2.0 * m0.e14, Hs + a * nSt}}, s);
            // correlations with the shape sums and the strain (and its products)
            std::vector<Src> src;
            for (int a = 0; a < 3; ++a) src.push_back(cv.at(Hs + a * nSt));          // 0..2
            src.push_back(cv.at(f.XS + nSt, 1));                                      // 3 X_e14
            for (int c = 0; c < 6; ++c) src.push_back(cv.at(eS + c * nSt, 1));        // 4..9
            if (f.second) {
                for (int k = 2; k < 5; ++k) src.push_back(cv.at(f.XS + k * nSt, 1)); // 10..12 X_B114, X_B124, X_B156
                for (int q = 0; q < 12; ++q) src.push_back(cv.at(f.RA + (3 + q) * nSt, 1));   // 13..24
                for (int a = 0; a < 3; ++a) src.push_back(cv.at(f.p12 + a * nSt, 1));         // 25..27
            }
            c64 *spec = pool.take<c64>(src.size() * cv.grid);
            cv.forward(src, spec, s);
            std::vector<Sum> o = {{{0, 4 + kFirstComp[0]}, {1, 4 + kFirstComp[1]}, {2, 4 + kFirstComp[2]}},   // 0 GXe14
                                  {{0, 3}}, {{1, 3}}, {{2, 3}}};                                               // 1..3
            if (f.second) {
                for (int a = 0; a < 3; ++a) o.push_back(Sum{{a, 10}});   // 4..6 g0
                for (int a = 0; a < 3; ++a) o.push_back(Sum{{a, 11}});   // 7..9 g12
                for (int a = 0; a < 3; ++a) o.push_back(Sum{{a, 12}});   // 10..12 g3
                o.push_back(Sum{{0, 13}, {1, 17}, {2, 21}});             // 13 GXB114 / 2
                o.push_back(Sum{{0, 25}, {1, 26}, {2, 27}});             // 14 GXB124 / 2
                o.push_back(Sum{{0, 16}, {1, 20}, {2, 24}});             // 15 GXB156 / 4
            }
            c64 *prod = pool.take<c64>(o.size() * cv.grid);
            c64 *OA = pool.take<c64>(o.size() * nSt);
            cv.central(spec, o, prod, OA, s);
            for (int a = 0; a < 3; ++a)
                lin(GeS + kFirstComp[a] * nSt, nSt, {{1.0, GeS + kFirstComp[a] * nSt}, {1.0, OA + (1 + a) * nSt}}, s);
            sterms.push_back({OA, &Material::e14, 2.0});
            if (f.second) {
                // gradients of the products: the matrix's and the shapes' weights
                c64 *Gp = pool.take<c64>(12 * nSt);
                for (int a = 0; a < 3; ++a) {
                    const c64 *h = Hs + a * nSt;
                    lin(Gp + (4 * a) * nSt, nSt, {{2.0 * m0.B114, h}, {2.0, OA + (4 + a) * nSt}}, s);
                    lin(Gp + (4 * a + 1) * nSt, nSt, {{2.0 * m0.B124, h}, {2.0, OA + (7 + a) * nSt}}, s);
                    lin(Gp + (4 * a + 2) * nSt, nSt, {{2.0 * m0.B124, h}, {2.0, OA + (7 + a) * nSt}}, s);
                    lin(Gp + (4 * a + 3) * nSt, nSt, {{4.0 * m0.B156, h}, {4.0, OA + (10 + a) * nSt}}, s);
                }
                sterms.push_back({OA + 13 * nSt, &Material::B114, 2.0});
                sterms.push_back({OA + 14 * nSt, &Material::B124, 2.0});
                sterms.push_back({OA + 15 * nSt, &Material::B156, 4.0});
                // prod[q] = e_a (*) e_b: each product's gradient to both of its components
                std::vector<Src> srcC;
                for (int q = 0; q < 12; ++q) srcC.push_back(cv.at(Gp + q * nSt));     // 0..11
                for (int c = 0; c < 6; ++c) srcC.push_back(cv.at(eS + c * nSt, 1));   // 12..17
                c64 *specC = pool.take<c64>(srcC.size() * cv.grid);
                cv.forward(srcC, specC, s);
                std::vector<Sum> oc(6);
                for (int q = 0; q < 12; ++q) {
                    oc[kProdDev[q][0]].push_back({q, 12 + kProdDev[q][1]});
                    oc[kProdDev[q][1]].push_back({q, 12 + kProdDev[q][0]});
                }
                c64 *prodC = pool.take<c64>(oc.size() * cv.grid);
                c64 *OC = pool.take<c64>(oc.size() * nSt);
                cv.central(specC, oc, prodC, OC, s);
                for (int c = 0; c < 6; ++c) lin(GeS + c * nSt, nSt, {{1.0, GeS + c * nSt}, {1.0, OC + c * nSt}}, s);
            }
        }
        // --- the strain of the inclusions: e_c = bracket_c sum_s eps0_s chi_s
        c64 *GX = pool.take<c64>(nSt);
        const Elastic el{m0.C11, m0.C12, m0.C44, st.iso ? 1 : 0};
        hipLaunchKernelGGL(k_strain_adj, dim3(grid_for(nSt)), dim3(256), 0, s, GeS, f.gS[0], f.gS[1], f.gS[2], nS[0],
                           nS[1], nS[2], el, GX);
        QD_LAUNCH_CHECK();
        shape_terms(in, mp, sterms, f.chiS, nSt, gS, out.gmp, pool, s);
        if (nsh > 0) {
            // the misfit eps0_s = (a_host - a_s) / a_0: weights -a / a_0, and a_0 itself
            const double a0 = m0.a;
            std::vector<Material> gm(ns, zero);
            shape_terms(in, mp, {{GX, &Material::a, -1.0 / a0}}, f.chiS, nSt, gS, gm, pool, s);
            for (int sh = 0; sh < ns; ++sh) out.gmp[sh].a += gm[sh].a;
            // d eps0_s / d a_0 = -eps0_s / a_0: the dots again, weighted by eps0_s
            std::vector<std::pair<const c64 *, const c64 *>> pr;
            for (int sh = 0; sh < nsh; ++sh) pr.push_back({GX, f.chiS + (long)sh * nSt});
            const std::vector<double> r = pair_dots(pr, nSt, pool, s);
            for (int sh = 1; sh <= nsh; ++sh) {
                const double eps0 = (mp[in.shapes[sh].embedded].a - mp[sh].a) / a0;
                out.gmp[0].a += -r[sh - 1] * eps0 / a0;
            }
        }
        if (!st.same) {
            centre_dots(gS, f.chiS, nsh, f.gS, nS, gcentre, pool, s);
            std::vector<c64> h((size_t)std::max(nsh, 1) * nSt);
            QD_HIP(hipMemcpyAsync(h.data(), gS, sizeof(c64) * h.size(), hipMemcpyDeviceToHost, s));
            QD_HIP(hipStreamSynchronize(s));
            for (int sh = 1; sh <= nsh; ++sh) {
                out.gchiS[sh] = Grid3(nS[0], nS[1], nS[2]);
                std::memcpy(out.gchiS[sh].v.data(), h.data() + (size_t)(sh - 1) * nSt, sizeof(c64) * nSt);
            }
        }
    }
    shape_terms(in, mp, eterms, f.chiE, nT, gE, out.gmp, pool, s);
    centre_dots(gE, f.chiE, nsh, f.gE, NT, gcentre, pool, s);
    out.gcentre = gcentre;
    // the requested shape-length parameters: the gradients contracted with the exact derivatives of
    // the transforms (dual numbers on the device), on both grids when they differ
    out.gparam.assign(nparams, std::nan(""));
    for (int j = 0; j < nparams; ++j) {
        double acc = 0.0;
        bool any = false;
        for (int sh = 1; sh < ns && sh < 32; ++sh) {
            double seed[4];
            if (!(masks[j] & (1u << sh)) || !length_seeds(params[j], in.shapes[sh].shape, seed)) continue;
            any = true;
            const int family = shape_family(in.shapes[sh].shape);
            const int slot = sh - 1;
            auto contract = [&](Cell cell, const double *const ax[3], const int n[3], const c64 *grad) {
                const ShapeDev sd = device_shape(in, sh, cell);
                const ShapeDev *sdd = pool.upload(&sd, 1, s);
                const int *sl = pool.upload(&slot, 1, s);
                const long np = (long)n[0] * n[1] * n[2];
                const int nblk = (int)std::min<long>(kDotBlocks, std::max<long>(1, (np + 255) / 256));
                double *part = pool.take<double>(nblk);
                launch_shape_deriv_dots(sdd, sl, 1, family, ax, n, f.gl, seed, grad, part, nblk, s);
                std::vector<double> h(nblk);
                QD_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nblk, hipMemcpyDeviceToHost, s));
                QD_HIP(hipStreamSynchronize(s));   // sd, slot, h
                for (double v : h) acc += v;
            };
            contract(Cell::Electronic, f.gE, NT, gE);
            if (st.use_strain && !st.same) contract(Cell::Strain, f.gS, st.nS, gSd);
        }
        if (any) out.gparam[j] = acc;
    }
    std::vector<c64> h((size_t)std::max(nsh, 1) * nT);
    QD_HIP(hipMemcpyAsync(h.data(), gE, sizeof(c64) * h.size(), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    for (int sh = 1; sh <= nsh; ++sh) {
        out.gchiE[sh] = Grid3(NT[0], NT[1], NT[2]);
        std::memcpy(out.gchiE[sh].v.data(), h.data() + (size_t)(sh - 1) * nT, sizeof(c64) * nT);
    }
}

}  // namespace qd
