// Coefficient generation on the device (FP64): the strain of the inclusions (continuum-elastic
// Green tensor, strain_fourier_coefficient.c:23-390, summed over shapes as in
// strain_region_reciprocal_space.c:78-150), the piezoelectric polarisation and potential to second
// order (piezo_region_reciprocal_space.c:23-787, piezo_fourier_coefficient.c:24-132), the transfer
// of the fields to the electronic-structure grid in all three cell modes (the same cell; free
// lengths, strain_reciprocal_space.c:23-230; commensurate cells, strain_reciprocal_space_fft.c:23-362
// and piezo_reciprocal_space_fft.c), the convolutions of convolutions.c:24-411 and the QD_NTAB
// Delta-G tables (hamiltonian_8_band_matrix_elements.c:27-969, parameter_matrix_element.c:23-48).
// The shape transforms chi_s with closed forms are evaluated here too (shapes.hip); those that need
// 1-D quadratures of Bessel functions (shape types 3, 4, 11-15) and the material interpolation
// come from the host (shapes.cpp, materials.cpp).
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
#include <string>
#include <vector>

#include "qd_coulomb.hpp"
#include "qd_host.hpp"
#include "qd_shapes_device.hpp"

namespace qd {

namespace {

constexpr double kPiD = 3.14159265358979323846;
constexpr int kMaxWeights = 16;   // weight vectors per k_weighted pass
constexpr int kMaxPairs = 16;     // products per k_products launch
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

// A grid of n0 x n1 x n2 points placed at the origin of an L0 x L1 x L2 grid of zeros.
__global__ __launch_bounds__(256) void k_pad_grid(const c64 *__restrict__ src, int n0, int n1, int n2, int L0, int L1,
                                                  int L2, c64 *__restrict__ dst) {
    const long grid = (long)L0 * L1 * L2;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < grid; t += (long)gridDim.x * 256) {
        const int x = (int)(t / ((long)L1 * L2)), y = (int)((t / L2) % L1), z = (int)(t % L2);
        dst[t] = (x < n0 && y < n1 && z < n2) ? src[((long)x * n1 + y) * n2 + z] : mk(0.0, 0.0);
    }
}

struct Pairs {
    int a[kMaxPairs], b[kMaxPairs];
    int n;
};

// out[p][t] = spec[a_p][t] spec[b_p][t]: the spectra of the stage's linear convolutions
__global__ __launch_bounds__(256) void k_products(const c64 *__restrict__ spec, long grid, Pairs pr,
                                                  c64 *__restrict__ out) {
    const int p = blockIdx.y;
    const c64 *A = spec + (long)pr.a[p] * grid, *B = spec + (long)pr.b[p] * grid;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < grid; t += (long)gridDim.x * 256)
        out[(long)p * grid + t] = A[t] * B[t];
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

// y[i] = a x[i] + b z[i] (b may be 0 with z null)
__global__ __launch_bounds__(256) void k_lincomb(c64 *__restrict__ y, long n, double a, const c64 *__restrict__ x,
                                                 double b, const c64 *__restrict__ z) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        c64 v = a * x[i];
        if (z) v = v + b * z[i];
        y[i] = v;
    }
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
        for (int a = 0; a < 3; ++a) {
            const int o1 = dd[(a + 1) % 3], o2 = dd[(a + 2) % 3];
            if (dd[a] == 0 || o1 != 0 || o2 != 0) continue;
            const double sign = (dd[a] % 2 == 0) ? -1.0 : 1.0;
            fim += m.F[a] * (sign / (-(double)dd[a] * 2.0 * kPiD / m.Lc[a]));
        }
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
    // spectra of the zero-padded grids srcs[k] (n points per axis) into spec[k]
    void forward(const std::vector<const c64 *> &srcs, c64 *spec, hipStream_t s) const {
        for (size_t k = 0; k < srcs.size(); ++k) {
            hipLaunchKernelGGL(k_pad_grid, dim3(grid_for(grid)), dim3(256), 0, s, srcs[k], n[0], n[1], n[2], L[0], L[1],
                               L[2], spec + (long)k * grid);
            QD_LAUNCH_CHECK();
        }
        fft3_device(spec, (long)srcs.size(), logL, tw, false, s);
    }
    // products of the listed spectrum pairs into prod, inverse transforms, and the blocks
    // start + step k (k < nout) of each into out (nout-point grids, consecutive)
    void products(const c64 *spec, const std::vector<std::pair<int, int>> &pairs, c64 *prod, const int nout[3],
                  const int start[3], const int step[3], c64 *out, hipStream_t s) const {
        const int np = (int)pairs.size();
        for (int first = 0; first < np; first += kMaxPairs) {
            Pairs pr;
            pr.n = std::min(kMaxPairs, np - first);
            for (int p = 0; p < pr.n; ++p) {
                pr.a[p] = pairs[first + p].first;
                pr.b[p] = pairs[first + p].second;
            }
            hipLaunchKernelGGL(k_products, dim3(grid_for(grid, 4096), pr.n), dim3(256), 0, s, spec, grid, pr,
                               prod + (long)first * grid);
            QD_LAUNCH_CHECK();
        }
        fft3_device(prod, np, logL, tw, true, s);
        const long no = (long)nout[0] * nout[1] * nout[2];
        hipLaunchKernelGGL(k_take_block, dim3(grid_for(no * np)), dim3(256), 0, s, prod, L[0], L[1], L[2], nout[0],
                           nout[1], nout[2], start[0], start[1], start[2], step[0], step[1], step[2], (long)np,
                           1.0 / (double)grid, out);
        QD_LAUNCH_CHECK();
    }
    // the central n-point blocks of the products
    void central(const c64 *spec, const std::vector<std::pair<int, int>> &pairs, c64 *prod, c64 *out,
                 hipStream_t s) const {
        const int start[3] = {n[0] / 2, n[1] / 2, n[2] / 2}, one[3] = {1, 1, 1};
        products(spec, pairs, prod, n, start, one, out, s);
    }
};

void lincomb(c64 *y, long n, double a, const c64 *x, double b, const c64 *z, hipStream_t s) {
    hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n)), dim3(256), 0, s, y, n, a, x, b, z);
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

ShapeDev device_shape(const qd_inputs &in, int idx, Cell cell) {
    try {
        return shape_device_params(in, idx, cell);
    } catch (const std::runtime_error &e) {   // inputs the host transform rejects as well
        throw StatusError(e.what(), QD_E_INPUT);
    }
}

// chi_s of shapes 1 .. N_shape - 1 on one grid (slot s - 1): the host's grids where it evaluated
// them, the device's for the shapes with closed forms (left empty by build_geometry)
c64 *shape_grids(const qd_inputs &in, const std::vector<Grid3> &host, Cell cell, const double *const g[3],
                 const int n[3], const double *gl, DevPool &pool, hipStream_t s) {
    const long np = (long)n[0] * n[1] * n[2];
    const int cnt = in.N_shape - 1;
    c64 *d = pool.take<c64>((size_t)std::max(cnt, 1) * np);
    std::vector<ShapeDev> sd[2];   // closed forms, quadratures
    std::vector<int> slot[2];
    for (int k = 1; k <= cnt; ++k) {
        if (!host[k].v.empty()) {
            QD_HIP(hipMemcpyAsync(d + (long)(k - 1) * np, host[k].v.data(), sizeof(c64) * np, hipMemcpyHostToDevice, s));
            continue;
        }
        const int fam = shape_needs_quadrature(in.shapes[k].shape) ? 1 : 0;
        sd[fam].push_back(device_shape(in, k, cell));
        slot[fam].push_back(k - 1);
    }
    for (int fam = 0; fam < 2; ++fam) {
        if (sd[fam].empty()) continue;
        const ShapeDev *sdd = pool.upload(sd[fam].data(), sd[fam].size(), s);
        const int *sl = pool.upload(slot[fam].data(), slot[fam].size(), s);
        launch_shape_grid(sdd, sl, (int)sd[fam].size(), fam == 1, g, n, gl, d, s);
    }
    QD_HIP(hipStreamSynchronize(s));   // sd, slot leave scope
    return d;
}

}  // namespace

// Fills t exactly as compose_tables(in, build_state(in)) does (host copy, malloc'd like the host
// builder's), with every field and table computed on the device; st comes from build_geometry.
void tables_device(const qd_inputs &in, const TableState &st, hipStream_t s, qd_tables &t) {
    const std::vector<Material> &mp = st.mp;
    const Material &m0 = mp[0];
    const int *NT = st.NT;
    const long nT = (long)NT[0] * NT[1] * NT[2];
    DevPool pool;
    const std::vector<double> rules = gauss_legendre_rules();
    const double *gl = pool.upload(rules.data(), rules.size(), s);

    const c64 *eE = nullptr, *phiE = nullptr;
    c64 *chiS_d = nullptr;
    if (st.use_strain) {
        const int *nS = st.nS;
        const long nSt = (long)nS[0] * nS[1] * nS[2];
        const double *gS[3];
        for (int a = 0; a < 3; ++a) gS[a] = pool.upload(st.gS[a].data(), st.gS[a].size(), s);
        chiS_d = shape_grids(in, st.chiS, Cell::Strain, gS, nS, gl, pool, s);
        // shape sums on the strain grid: misfit, 2 e14, B114, B124, B156
        c64 *XS = pool.take<c64>(5 * nSt);
        const double a0 = m0.a;
        weighted(in, mp, chiS_d, nSt,
                 {[a0](const Material &m) { return -m.a / a0; }, [](const Material &m) { return 2.0 * m.e14; },
                  [](const Material &m) { return m.B114; }, [](const Material &m) { return m.B124; },
                  [](const Material &m) { return m.B156; }},
                 XS, pool, s);
        // fS: the six strain components, then phi
        c64 *fS = pool.take<c64>(7 * nSt);
        const Elastic el{m0.C11, m0.C12, m0.C44, st.iso ? 1 : 0};
        hipLaunchKernelGGL(k_strain, dim3(grid_for(nSt)), dim3(256), 0, s, XS, gS[0], gS[1], gS[2], nS[0], nS[1], nS[2],
                           el, fS);
        QD_LAUNCH_CHECK();
        const bool piezo = in.piezo != 0, second = in.piezo == 2 || in.piezo == 3;
        const int nf = piezo ? 7 : 6;
        if (piezo || st.commensurate) {
            const ConvPlan cv(nS, pool, s);
            if (piezo) {
                // stage A: spectra of 2 sum d e14 chi and the six strain components
                c64 *spec = pool.take<c64>((size_t)7 * cv.grid);
                cv.forward({XS + nSt, fS, fS + nSt, fS + 2 * nSt, fS + 3 * nSt, fS + 4 * nSt, fS + 5 * nSt}, spec, s);
                std::vector<std::pair<int, int>> pa = {{0, 1 + YZ}, {0, 1 + XZ}, {0, 1 + XY}};
                // second-order products e_a (*) e_b grouped per polarisation component: P_x from
                // products 0..3, P_y from 4..7, P_z from 8..11
                const int prodc[12][2] = {{XX, YZ}, {YY, YZ}, {ZZ, YZ}, {XY, XZ}, {YY, XZ}, {XX, XZ},
                                          {ZZ, XZ}, {XY, YZ}, {ZZ, XY}, {XX, XY}, {YY, XY}, {XZ, YZ}};
                if (second)
                    for (auto &q : prodc) pa.push_back({1 + q[0], 1 + q[1]});
                c64 *prodA = pool.take<c64>(pa.size() * cv.grid);
                c64 *RA = pool.take<c64>(pa.size() * nSt);
                cv.central(spec, pa, prodA, RA, s);
                c64 *conv2 = nullptr;
                if (second) {
                    // stage B: the shape sums of B114, B124, B156 with p0, p1 + p2, p3 per component
                    const c64 *prod = RA + 3 * nSt;
                    c64 *p12 = pool.take<c64>(3 * nSt);
                    for (int a = 0; a < 3; ++a)
                        lincomb(p12 + a * nSt, nSt, 1.0, prod + (4 * a + 1) * nSt, 1.0, prod + (4 * a + 2) * nSt, s);
                    std::vector<const c64 *> src = {XS + 2 * nSt, XS + 3 * nSt, XS + 4 * nSt};
                    for (int a = 0; a < 3; ++a) src.push_back(prod + 4 * a * nSt);
                    for (int a = 0; a < 3; ++a) src.push_back(p12 + a * nSt);
                    for (int a = 0; a < 3; ++a) src.push_back(prod + (4 * a + 3) * nSt);
                    c64 *specB = pool.take<c64>(src.size() * cv.grid);
                    cv.forward(src, specB, s);
                    std::vector<std::pair<int, int>> pb;
                    for (int a = 0; a < 3; ++a) pb.push_back({3 + a, 0});
                    for (int a = 0; a < 3; ++a) pb.push_back({6 + a, 1});
                    for (int a = 0; a < 3; ++a) pb.push_back({9 + a, 2});
                    c64 *prodB = pool.take<c64>(pb.size() * cv.grid);
                    conv2 = pool.take<c64>(pb.size() * nSt);
                    cv.central(specB, pb, prodB, conv2, s);
                }
                const PiezoConsts pc{m0.e14, m0.B114, m0.B124, m0.B156, m0.nr, second ? 1 : 0};
                hipLaunchKernelGGL(k_piezo_phi, dim3(grid_for(nSt)), dim3(256), 0, s, fS, RA, RA + 3 * nSt, conv2, gS[0],
                                   gS[1], gS[2], nS[0], nS[1], nS[2], pc, fS + 6 * nSt);
                QD_LAUNCH_CHECK();
            }
            if (st.commensurate) {
                // the electronic cell's transform convolved with every field, sampled every RL
                // strain-grid points from n_S - 1 - RL (N_T / 2)
                std::vector<const c64 *> src;
                for (int k = 0; k < nf; ++k) src.push_back(fS + k * nSt);
                c64 *chi0;
                if (!st.chi0S.v.empty()) {
                    chi0 = pool.upload(reinterpret_cast<const c64 *>(st.chi0S.v.data()), st.chi0S.v.size(), s);
                } else {
                    chi0 = pool.take<c64>(nSt);
                    const ShapeDev cell0 = device_shape(in, 0, Cell::Strain);
                    const ShapeDev *sd0 = pool.upload(&cell0, 1, s);
                    const int zero = 0;
                    const int *sl0 = pool.upload(&zero, 1, s);
                    launch_shape_grid(sd0, sl0, 1, shape_needs_quadrature(in.shapes[0].shape), gS, nS, gl, chi0, s);
                    QD_HIP(hipStreamSynchronize(s));
                }
                src.push_back(chi0);
                c64 *spec = pool.take<c64>(src.size() * cv.grid);
                cv.forward(src, spec, s);
                std::vector<std::pair<int, int>> pc;
                for (int k = 0; k < nf; ++k) pc.push_back({k, nf});
                c64 *prod = pool.take<c64>(pc.size() * cv.grid);
                c64 *out = pool.take<c64>((size_t)nf * nT);
                int start[3];
                for (int a = 0; a < 3; ++a) start[a] = (nS[a] - 1) - st.RL[a] * (NT[a] / 2);
                cv.products(spec, pc, prod, NT, start, st.RL, out, s);
                QD_HIP(hipStreamSynchronize(s));   // chi0 upload source
                eE = out;
                phiE = piezo ? out + 6 * nT : nullptr;
            }
        }
        if (!st.commensurate) {
            if (st.same) {
                eE = fS;
                phiE = piezo ? fS + 6 * nSt : nullptr;
            } else {
                // free lengths: separable sinc (box-transform) matrices per axis, z, then y, then x
                std::vector<double> I[3];
                const double inv = 1.0 / (st.Lc[0] * st.Lc[1] * st.Lc[2]);
                for (int a = 0; a < 3; ++a) {
                    const int ne = (int)st.gE[a].size(), ns = (int)st.gS[a].size();
                    I[a].resize((size_t)ne * ns);
                    for (int g = 0; g < ne; ++g)
                        for (int k = 0; k < ns; ++k) {
                            const double dG = st.gS[a][k] - st.gE[a][g];
                            double v = std::fabs(dG) < kEps ? st.Lc[a] : 2.0 * std::sin(dG * st.Lc[a] / 2.0) / dG;
                            if (a == 0) v *= inv;
                            I[a][(size_t)g * ns + k] = v;
                        }
                }
                const c64 *cur = fS;
                int dims[3] = {nS[0], nS[1], nS[2]};
                for (int a = 2; a >= 0; --a) {
                    const double *M = pool.upload(I[a].data(), I[a].size(), s);
                    const int m = NT[a];
                    int od[3] = {dims[0], dims[1], dims[2]};
                    od[a] = m;
                    c64 *nxt = pool.take<c64>((size_t)nf * od[0] * od[1] * od[2]);
                    apply_axis_device(cur, dims, a, m, M, nf, nxt, s);
                    cur = nxt;
                    dims[a] = m;
                }
                QD_HIP(hipStreamSynchronize(s));   // I leaves scope
                eE = cur;
                phiE = piezo ? cur + 6 * nT : nullptr;
            }
        }
    }

    // the parameter expansions on the electronic grid
    const c64 *chiE_d = chiS_d;
    if (!(st.use_strain && st.same)) {
        const double *gE[3];
        for (int a = 0; a < 3; ++a) gE[a] = pool.upload(st.gE[a].data(), st.gE[a].size(), s);
        chiE_d = shape_grids(in, st.chiE, Cell::Electronic, gE, NT, gl, pool, s);
    }
    c64 *XE = pool.take<c64>((size_t)W_COUNT * nT);
    weighted(in, mp, chiE_d, nT,
             {[](const Material &m) { return m.E_cb; }, [](const Material &m) { return m.E_hh; },
              [](const Material &m) { return m.E_lh; }, [](const Material &m) { return m.E_so; },
              [](const Material &m) { return m.sc; }, [](const Material &m) { return m.g1; },
              [](const Material &m) { return m.g2; }, [](const Material &m) { return m.g3; },
              [](const Material &m) { return m.P; }, [](const Material &m) { return m.ac; },
              [](const Material &m) { return m.av; }, [](const Material &m) { return m.b; },
              [](const Material &m) { return m.d; }},
             XE, pool, s);
    c64 *H = nullptr, *C = nullptr;
    if (st.use_strain) {
        H = pool.take<c64>(3 * nT);
        // hyd = XX + YY + ZZ, bia = XX + YY - 2 ZZ, dxy = XX - YY
        lincomb(H, nT, 1.0, eE + XX * nT, 1.0, eE + YY * nT, s);
        lincomb(H + nT, nT, 1.0, H, -2.0, eE + ZZ * nT, s);
        lincomb(H, nT, 1.0, H, 1.0, eE + ZZ * nT, s);
        lincomb(H + 2 * nT, nT, 1.0, eE + XX * nT, -1.0, eE + YY * nT, s);
        const ConvPlan ce(NT, pool, s);
        // 0 Xac, 1 Xav, 2 Xb, 3 Xd, 4 XP, 5 hyd, 6 bia, 7 dxy, 8 + c strain component c
        std::vector<const c64 *> src = {XE + W_AC * nT, XE + W_AV * nT, XE + W_B * nT, XE + W_D * nT, XE + W_P * nT,
                                        H, H + nT, H + 2 * nT};
        for (int c = 0; c < 6; ++c) src.push_back(eE + c * nT);
        c64 *spec = pool.take<c64>(src.size() * ce.grid);
        ce.forward(src, spec, s);
        std::vector<std::pair<int, int>> pc = {{0, 5}, {1, 5}, {6, 2}, {7, 2}, {8 + XY, 3}, {8 + XZ, 3}, {8 + YZ, 3}};
        for (int k = 0; k < 6; ++k) pc.push_back({8 + k, 4});   // XX, YY, ZZ, YZ, XZ, XY
        c64 *prod = pool.take<c64>(pc.size() * ce.grid);
        C = pool.take<c64>(pc.size() * nT);
        ce.central(spec, pc, prod, C, s);
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
    hipLaunchKernelGGL(k_tables, dim3(grid_for(nT)), dim3(256), 0, s, XE, eE, H, C, phiE, tc, tab);
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

}  // namespace qd
