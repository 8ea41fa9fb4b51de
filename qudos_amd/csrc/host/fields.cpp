// FP64 reciprocal-space fields on the strain and electronic-structure grids and the Delta-G
// tables of the Hamiltonian kernel.
//   strain:       strain_fourier_coefficient.c:23-390 (continuum-elastic Green tensor, matrix C_ij),
//                 summed over shapes as in strain_region_reciprocal_space.c:78-150
//   piezo:        piezo_region_reciprocal_space.c:23-787 + piezo_fourier_coefficient.c:24-132
//   downsampling: strain_reciprocal_space.c:23-230 / piezo_reciprocal_space.c (free lengths),
//                 strain_reciprocal_space_fft.c:23-362 / piezo_reciprocal_space_fft.c (commensurate)
//   convolutions: convolutions.c:24-411 (central block of zero-padded products)
//   tables:       the per-element sums of hamiltonian_8_band_matrix_elements.c:27-969 and
//                 parameter_matrix_element.c:23-48, pre-combined per Delta-G
// The convolutions use exact separable DFTs of the zero-padded grids (sizes 2N_T-1 are small
// odd numbers, so a direct transform per axis is cheap and needs no FFT library).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "qd_host.hpp"

namespace qd {
namespace {
constexpr double kPi = 3.14159265358979323846;
}

Convolver::Convolver(const int n[3]) {
    for (int a = 0; a < 3; ++a) {
        n_[a] = n[a];
        m_[a] = 2 * n[a] - 1;
        tw_[a].resize(m_[a]);
        for (int t = 0; t < m_[a]; ++t) {
            const double ang = -2.0 * kPi * t / m_[a];
            tw_[a][t] = cplx(std::cos(ang), std::sin(ang));
        }
    }
}

void Convolver::forward(const Grid3 &a, std::vector<cplx> &spec) const {
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2], mx = m_[0], my = m_[1], mz = m_[2];
    if (nx > mx || ny > my || nz > mz) throw std::logic_error("Convolver::forward: grid larger than the padding");
    std::vector<cplx> s1((size_t)nx * ny * mz), s2((size_t)nx * my * mz);
    // z
    parallel_for_cost((long)nx * ny, (double)mz * nz, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r)
            for (int k = 0; k < mz; ++k) {
                cplx acc = 0;
                const cplx *src = &a.v[(size_t)r * nz];
                for (int j = 0; j < nz; ++j) acc += src[j] * tw_[2][(size_t)j * k % mz];
                s1[(size_t)r * mz + k] = acc;
            }
    });
    // y
    parallel_for_cost((long)nx * my, (double)ny * mz, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r) {
            const long x = r / my, k = r % my;
            cplx *dst = &s2[((size_t)x * my + k) * mz];
            for (int z = 0; z < mz; ++z) dst[z] = 0;
            for (int j = 0; j < ny; ++j) {
                const cplx w = tw_[1][(size_t)j * k % my];
                const cplx *src = &s1[((size_t)x * ny + j) * mz];
                for (int z = 0; z < mz; ++z) dst[z] += src[z] * w;
            }
        }
    });
    // x
    spec.assign((size_t)mx * my * mz, cplx(0, 0));
    parallel_for_cost((long)mx, (double)nx * my * mz, [&](long lo, long hi) {
        for (long k = lo; k < hi; ++k) {
            cplx *dst = &spec[(size_t)k * my * mz];
            for (int j = 0; j < nx; ++j) {
                const cplx w = tw_[0][(size_t)j * k % mx];
                const cplx *src = &s2[(size_t)j * my * mz];
                for (size_t q = 0; q < (size_t)my * mz; ++q) dst[q] += src[q] * w;
            }
        }
    });
}

void Convolver::inverse_central(const std::vector<cplx> &spec, Grid3 &out) const {
    const int start[3] = {n_[0] / 2, n_[1] / 2, n_[2] / 2}, step[3] = {1, 1, 1};
    inverse_at(spec, n_, start, step, out);
}

void Convolver::inverse_at(const std::vector<cplx> &spec, const int nout[3], const int start[3], const int step[3],
                           Grid3 &out) const {
    const int nx = nout[0], ny = nout[1], nz = nout[2], mx = m_[0], my = m_[1], mz = m_[2];
    for (int a = 0; a < 3; ++a)
        if (nout[a] < 1 || start[a] < 0 || start[a] + step[a] * (nout[a] - 1) >= m_[a] || step[a] < 1)
            throw std::logic_error("Convolver::inverse_at: output points outside the padded grid");
    const int cx = start[0], cy = start[1], cz = start[2], sx = step[0], sy = step[1], sz = step[2];
    std::vector<cplx> s1((size_t)nx * my * mz), s2((size_t)nx * ny * mz);
    // x: keep outputs cx + sx k, k < nx (conjugate twiddles = inverse transform)
    parallel_for_cost((long)nx, (double)mx * my * mz, [&](long lo, long hi) {
        for (long k = lo; k < hi; ++k) {
            cplx *dst = &s1[(size_t)k * my * mz];
            for (size_t q = 0; q < (size_t)my * mz; ++q) dst[q] = 0;
            for (int j = 0; j < mx; ++j) {
                const cplx w = std::conj(tw_[0][(size_t)j * (cx + sx * k) % mx]);
                const cplx *src = &spec[(size_t)j * my * mz];
                for (size_t q = 0; q < (size_t)my * mz; ++q) dst[q] += src[q] * w;
            }
        }
    });
    parallel_for_cost((long)nx * ny, (double)my * mz, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r) {
            const long x = r / ny, k = r % ny;
            cplx *dst = &s2[((size_t)x * ny + k) * mz];
            for (int z = 0; z < mz; ++z) dst[z] = 0;
            for (int j = 0; j < my; ++j) {
                const cplx w = std::conj(tw_[1][(size_t)j * (cy + sy * k) % my]);
                const cplx *src = &s1[((size_t)x * my + j) * mz];
                for (int z = 0; z < mz; ++z) dst[z] += src[z] * w;
            }
        }
    });
    out = Grid3(nx, ny, nz);
    const double scale = 1.0 / ((double)mx * my * mz);
    parallel_for_cost((long)nx * ny, (double)nz * mz, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r)
            for (int k = 0; k < nz; ++k) {
                cplx acc = 0;
                const cplx *src = &s2[(size_t)r * mz];
                for (int j = 0; j < mz; ++j) acc += src[j] * std::conj(tw_[2][(size_t)j * (cz + sz * k) % mz]);
                out.v[(size_t)r * nz + k] = acc * scale;
            }
    });
}

Grid3 Convolver::convolve(const Grid3 &a, const Grid3 &b) const {
    std::vector<cplx> A, B;
    forward(a, A);
    forward(b, B);
    for (size_t i = 0; i < A.size(); ++i) A[i] *= B[i];
    Grid3 out;
    inverse_central(A, out);
    return out;
}

std::vector<double> reciprocal_axis(int N, double L) {
    const int NT = 2 * N - 1;
    std::vector<double> g(NT);
    for (int i = 0; i < NT; ++i) g[i] = (2.0 * kPi / L) * (0.5 * (1.0 - NT) + i);
    return g;
}

namespace {

enum { XX = 0, YY, ZZ, YZ, XZ, XY };
const int kComp[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};

// Bracket of the continuum-elastic strain of an inclusion with unit misfit (matrix C_ij).
double strain_bracket(int c, const double G[3], const Material &m0, bool iso) {
    const int i = kComp[c][0], j = kComp[c][1];
    const double dij = (i == j) ? 1.0 : 0.0;
    const double C11 = m0.C11, C12 = m0.C12, C44 = m0.C44;
    const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    if (std::sqrt(G2) < kEps) {
        if (iso) return dij * (1.0 - (3.0 * C12 + 2.0 * C44) / (3.0 * C12 + 6.0 * C44));
        const double an = C11 - C12 - 2.0 * C44;
        const double s0 = 1.0 / (C44 + an / 3.0);
        return dij * (1.0 - (0.5 * (C11 + 2.0 * C12) / 3.0) * (2.0 / (C44 + an / 3.0)) / (1.0 + (C12 + C44) * s0));
    }
    if (iso) return dij - (3.0 * C12 + 2.0 * C44) * G[i] * G[j] / ((C12 + 2.0 * C44) * G2);
    const double an = C11 - C12 - 2.0 * C44;
    double D[3], s = 0;
    for (int a = 0; a < 3; ++a) {
        D[a] = C44 * G2 + an * G[a] * G[a];
        s += G[a] * G[a] / D[a];
    }
    return dij - (0.5 * (C11 + 2.0 * C12) * (2.0 * C44 * G2 + an * (G[i] * G[i] + G[j] * G[j])) * G[i] * G[j]) /
                     ((1.0 + (C12 + C44) * s) * D[i] * D[j]);
}

struct StrainField { Grid3 e[6]; };

// Separable sinc downsampling from the strain cell to the electronic-structure Delta-G grid.
Grid3 downsample(const Grid3 &src, const std::vector<double> gS[3], const std::vector<double> gE[3],
                 const double L[3], bool same) {
    if (same) return src;
    const int nE[3] = {(int)gE[0].size(), (int)gE[1].size(), (int)gE[2].size()};
    const int nS[3] = {src.n[0], src.n[1], src.n[2]};
    std::vector<double> I[3];
    for (int a = 0; a < 3; ++a) {
        I[a].resize((size_t)nE[a] * nS[a]);
        for (int g = 0; g < nE[a]; ++g)
            for (int s = 0; s < nS[a]; ++s) {
                const double dG = gS[a][s] - gE[a][g];
                I[a][(size_t)g * nS[a] + s] = std::fabs(dG) < kEps ? L[a] : 2.0 * std::sin(dG * L[a] / 2.0) / dG;
            }
    }
    const double inv = 1.0 / (L[0] * L[1] * L[2]);
    // contract z, then y, then x
    std::vector<cplx> t1((size_t)nS[0] * nS[1] * nE[2]), t2((size_t)nS[0] * nE[1] * nE[2]);
    for (int x = 0; x < nS[0]; ++x)
        for (int y = 0; y < nS[1]; ++y)
            for (int k = 0; k < nE[2]; ++k) {
                cplx acc = 0;
                for (int z = 0; z < nS[2]; ++z) acc += src(x, y, z) * I[2][(size_t)k * nS[2] + z];
                t1[((size_t)x * nS[1] + y) * nE[2] + k] = acc;
            }
    for (int x = 0; x < nS[0]; ++x)
        for (int k = 0; k < nE[1]; ++k)
            for (int z = 0; z < nE[2]; ++z) {
                cplx acc = 0;
                for (int y = 0; y < nS[1]; ++y) acc += t1[((size_t)x * nS[1] + y) * nE[2] + z] * I[1][(size_t)k * nS[1] + y];
                t2[((size_t)x * nE[1] + k) * nE[2] + z] = acc;
            }
    Grid3 out(nE[0], nE[1], nE[2]);
    for (int k = 0; k < nE[0]; ++k)
        for (int y = 0; y < nE[1]; ++y)
            for (int z = 0; z < nE[2]; ++z) {
                cplx acc = 0;
                for (int x = 0; x < nS[0]; ++x) acc += t2[((size_t)x * nE[1] + y) * nE[2] + z] * I[0][(size_t)k * nS[0] + x];
                out(k, y, z) = acc * inv;
            }
    return out;
}

// Transpose of downsample: the strain-cell grid whose downsampling pairs with g (the separable
// sinc matrices are real, so no conjugation).
Grid3 downsample_adjoint(const Grid3 &g, const std::vector<double> gS[3], const std::vector<double> gE[3],
                         const double L[3], bool same) {
    if (same) return g;
    const int nE[3] = {(int)gE[0].size(), (int)gE[1].size(), (int)gE[2].size()};
    const int nS[3] = {(int)gS[0].size(), (int)gS[1].size(), (int)gS[2].size()};
    std::vector<double> I[3];
    for (int a = 0; a < 3; ++a) {
        I[a].resize((size_t)nE[a] * nS[a]);
        for (int k = 0; k < nE[a]; ++k)
            for (int s = 0; s < nS[a]; ++s) {
                const double dG = gS[a][s] - gE[a][k];
                I[a][(size_t)k * nS[a] + s] = std::fabs(dG) < kEps ? L[a] : 2.0 * std::sin(dG * L[a] / 2.0) / dG;
            }
    }
    const double inv = 1.0 / (L[0] * L[1] * L[2]);
    std::vector<cplx> t1((size_t)nS[0] * nE[1] * nE[2]), t2((size_t)nS[0] * nS[1] * nE[2]);
    for (int x = 0; x < nS[0]; ++x)
        for (int y = 0; y < nE[1]; ++y)
            for (int z = 0; z < nE[2]; ++z) {
                cplx acc = 0;
                for (int k = 0; k < nE[0]; ++k) acc += g(k, y, z) * I[0][(size_t)k * nS[0] + x];
                t1[((size_t)x * nE[1] + y) * nE[2] + z] = acc * inv;
            }
    for (int x = 0; x < nS[0]; ++x)
        for (int y = 0; y < nS[1]; ++y)
            for (int z = 0; z < nE[2]; ++z) {
                cplx acc = 0;
                for (int k = 0; k < nE[1]; ++k) acc += t1[((size_t)x * nE[1] + k) * nE[2] + z] * I[1][(size_t)k * nS[1] + y];
                t2[((size_t)x * nS[1] + y) * nE[2] + z] = acc;
            }
    Grid3 out(nS[0], nS[1], nS[2]);
    for (int x = 0; x < nS[0]; ++x)
        for (int y = 0; y < nS[1]; ++y)
            for (int z = 0; z < nS[2]; ++z) {
                cplx acc = 0;
                for (int k = 0; k < nE[2]; ++k) acc += t2[((size_t)x * nS[1] + y) * nE[2] + k] * I[2][(size_t)k * nS[2] + z];
                out(x, y, z) = acc;
            }
    return out;
}

// a reversed through the centre on every axis: r[j] = a[n - 1 - j] (a correlation with a is a
// truncated convolution with r)
Grid3 reversed(const Grid3 &a) {
    Grid3 r(a.n[0], a.n[1], a.n[2]);
    for (int x = 0; x < a.n[0]; ++x)
        for (int y = 0; y < a.n[1]; ++y)
            for (int z = 0; z < a.n[2]; ++z) r(a.n[0] - 1 - x, a.n[1] - 1 - y, a.n[2] - 1 - z) = a(x, y, z);
    return r;
}

// Downsampling from a commensurate strain cell (lengths RL x the electronic cell) as in
// strain_reciprocal_space_fft.c:23-362 and piezo_reciprocal_space_fft.c: the full linear
// convolution of the field with the electronic cell's transform chi0 on the strain grid (zero
// padding to 2 n_S - 1 points per axis), sampled every RL points from n_S - 1 - RL (N_T / 2):
// strain-grid offset RL e is electronic-grid offset e.  apply() is that map; adjoint() its
// transpose (G placed at the sampled points, correlated with chi0 = convolved with chi0 reversed;
// the circular transform of size 2 n_S - 1 is exact for the n_S points kept).
class CommensurateDown {
  public:
    explicit CommensurateDown(const TableState &st) : conv_(st.nS) {
        for (int a = 0; a < 3; ++a) {
            nS_[a] = st.nS[a];
            nE_[a] = st.NT[a];
            step_[a] = st.RL[a];
            start_[a] = (st.nS[a] - 1) - st.RL[a] * (st.NT[a] / 2);
        }
        conv_.forward(st.chi0S, chi0_);
        conv_.forward(reversed(st.chi0S), rev_);
    }
    Grid3 apply(const Grid3 &src) const {
        std::vector<cplx> s;
        conv_.forward(src, s);
        for (size_t i = 0; i < s.size(); ++i) s[i] *= chi0_[i];
        Grid3 out;
        conv_.inverse_at(s, nE_, start_, step_, out);
        return out;
    }
    Grid3 adjoint(const Grid3 &g) const {
        const int *m = conv_.padded();
        Grid3 z(m[0], m[1], m[2]);
        for (int x = 0; x < nE_[0]; ++x)
            for (int y = 0; y < nE_[1]; ++y)
                for (int k = 0; k < nE_[2]; ++k)
                    z(start_[0] + step_[0] * x, start_[1] + step_[1] * y, start_[2] + step_[2] * k) = g(x, y, k);
        std::vector<cplx> s;
        conv_.forward(z, s);
        for (size_t i = 0; i < s.size(); ++i) s[i] *= rev_[i];
        const int first[3] = {nS_[0] - 1, nS_[1] - 1, nS_[2] - 1}, one[3] = {1, 1, 1};
        Grid3 out;
        conv_.inverse_at(s, nS_, first, one, out);
        return out;
    }

  private:
    Convolver conv_;
    std::vector<cplx> chi0_, rev_;
    int nS_[3], nE_[3], start_[3], step_[3];
};

void axpy(Grid3 &y, cplx a, const Grid3 &x) {
    for (size_t i = 0; i < y.v.size(); ++i) y.v[i] += a * x.v[i];
}

// Re sum_i a[i] b[i]
double re_dot(const Grid3 &a, const Grid3 &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.v.size(); ++i) s += a.v[i].real() * b.v[i].real() - a.v[i].imag() * b.v[i].imag();
    return s;
}

// Sums of truncated convolutions: add(a, b) queues a (*) b, result() returns the sum.  Operands
// used several times are transformed once by the caller (spec) and passed as spectra.
class ConvSum {
  public:
    using Spec = std::vector<cplx>;
    explicit ConvSum(const Convolver &c) : c_(c) {}
    Spec spec(const Grid3 &a) const {
        Spec s;
        c_.forward(a, s);
        return s;
    }
    void add(const Spec &A, const Spec &B) {
        if (acc_.empty()) acc_.assign(A.size(), cplx(0, 0));
        for (size_t i = 0; i < A.size(); ++i) acc_[i] += A[i] * B[i];
    }
    void add(const Grid3 &a, const Grid3 &b) { add(spec(a), spec(b)); }
    void add(const Grid3 &a, const Spec &B) { add(spec(a), B); }
    Grid3 result() {
        Grid3 out;
        if (acc_.empty()) return out;
        c_.inverse_central(acc_, out);
        acc_.clear();
        return out;
    }

  private:
    const Convolver &c_;
    std::vector<cplx> acc_;
};

// second-order strain products e_a (*) e_b, grouped per polarisation component: P_x from
// products 0..3, P_y from 4..7, P_z from 8..11, weights (2 B114, 2 B124, 2 B124, 4 B156)
const int kProd[12][2] = {{XX, YZ}, {YY, YZ}, {ZZ, YZ}, {XY, XZ}, {YY, XZ}, {XX, XZ},
                          {ZZ, XZ}, {XY, YZ}, {ZZ, XY}, {XX, XY}, {YY, XY}, {XZ, YZ}};

}  // namespace

// ------------------------------------------------------------------------------------------
// Forward pass.  Every per-shape convolution of the reference (chi_s (*) strain, chi_s (*) strain
// products, each weighted by a parameter difference of shape s and its host) is linear in chi_s,
// so the shapes are summed first: X_w = sum_s w_s chi_s, then one convolution per weight.  The
// number of convolutions is independent of the shape count; the values equal the per-shape sums
// to rounding.
// ------------------------------------------------------------------------------------------
namespace {

// sum_s (f(shape s) - f(its host)) chi_s
template <class F>
Grid3 weighted_chi(const qd_inputs &in, const std::vector<Material> &mp, const std::vector<Grid3> &chi, const F &f,
                   const int n[3]) {
    Grid3 X(n[0], n[1], n[2]);
    for (int s = 1; s < in.N_shape; ++s) axpy(X, f(mp[s]) - f(mp[in.shapes[s].embedded]), chi[s]);
    return X;
}

double piezo_coef(double Ga, double G2, double nr0) {   // -100 i G_a / (eps0 1e12 nr0^2 G^2): imaginary part
    return -100.0 * Ga / (8.854188 * nr0 * nr0 * G2);
}

}  // namespace

void build_geometry(const qd_inputs &in, TableState &st, const std::function<bool(int)> &on_device) {
    const int ns = in.N_shape;
    st.mp = interpolate_materials(in);
    std::vector<Material> &mp = st.mp;
    st.use_strain = in.strain != 0;
    st.iso = in.isotropic != 0;
    if (st.use_strain && st.iso)
        for (auto &m : mp) m.C44 = 0.5 * (m.C11 - m.C12);
    const qd_shape &c0 = in.shapes[0];
    st.Lc[0] = c0.L1; st.Lc[1] = c0.L2; st.Lc[2] = c0.L3;
    for (int a = 0; a < 3; ++a) {
        st.gE[a] = reciprocal_axis(in.N[a], st.Lc[a]);
        st.NT[a] = 2 * in.N[a] - 1;
    }
    st.chiE.assign(ns, Grid3());
    st.chiS.assign(ns, Grid3());
    st.same = false;
    st.commensurate = false;
    if (st.use_strain) {
        // 8_band_main.c:158-200: a commensurate cell of relative lengths 1 and the same plane waves
        // is the electronic cell itself
        int strain_type = in.strain_type;
        if (strain_type == 1 && in.RL[0] == 1 && in.RL[1] == 1 && in.RL[2] == 1 && in.N_S[0] == in.N[0] &&
            in.N_S[1] == in.N[1] && in.N_S[2] == in.N[2])
            strain_type = 0;
        st.commensurate = strain_type == 1;
        st.same = !st.commensurate;
        for (int a = 0; a < 3; ++a) {
            st.gS[a] = reciprocal_axis(in.N_S[a], in.SL[a]);
            st.nS[a] = (int)st.gS[a].size();
            st.RL[a] = st.commensurate ? in.RL[a] : 1;
            st.same = st.same && in.N_S[a] == in.N[a] && in.SL[a] == st.Lc[a];
            // the strided extraction reaches RL (N_T / 2) strain-grid steps from the centre of the
            // (2 n_S - 1)-point convolution: M_S >= M RL / 2, rounded up
            if (st.commensurate && (in.RL[a] < 1 || (long)in.RL[a] * (in.N[a] - 1) > 2L * (in.N_S[a] - 1)))
                throw std::runtime_error(std::string("M_S_G") + "xyz"[a] +
                                         " is too small for the commensurate strain supercell: M_S_G" + "xyz"[a] +
                                         " must be >= M_G" + "xyz"[a] + " x (relative length) / 2, rounded up");
        }
        for (int s = 1; s < ns; ++s) {
            if (on_device && on_device(in.shapes[s].shape)) continue;
            st.chiS[s] = Grid3(st.nS[0], st.nS[1], st.nS[2]);
            shape_grid(in, s, st.gS[0], st.gS[1], st.gS[2], Cell::Strain, st.chiS[s].v);
        }
        if (st.commensurate && !(on_device && on_device(in.shapes[0].shape))) {
            // the electronic cell (shape 0) on the strain grid, normalised by its own volume
            st.chi0S = Grid3(st.nS[0], st.nS[1], st.nS[2]);
            shape_grid(in, 0, st.gS[0], st.gS[1], st.gS[2], Cell::Strain, st.chi0S.v);
        }
    }
    for (int s = 1; s < ns; ++s) {
        if (st.use_strain && st.same) { st.chiE[s] = st.chiS[s]; continue; }
        if (on_device && on_device(in.shapes[s].shape)) continue;
        st.chiE[s] = Grid3(st.NT[0], st.NT[1], st.NT[2]);
        shape_grid(in, s, st.gE[0], st.gE[1], st.gE[2], Cell::Electronic, st.chiE[s].v);
    }
    st.have_phi = false;
    st.second = false;
}

void build_state(const qd_inputs &in, TableState &st) {
    build_geometry(in, st);
    if (!st.use_strain) return;
    const int ns = in.N_shape;
    const std::vector<Material> &mp = st.mp;
    const int *nS = st.nS;
    const long nSt = (long)nS[0] * nS[1] * nS[2];
    // strain of the inclusions: e_c = bracket_c(G) sum_s eps0_s chi_s
    Grid3 Xeps(nS[0], nS[1], nS[2]);
    for (int s = 1; s < ns; ++s) axpy(Xeps, (mp[in.shapes[s].embedded].a - mp[s].a) / mp[0].a, st.chiS[s]);
    for (int c = 0; c < 6; ++c) {
        st.br[c].assign(nSt, 0.0);
        st.eS[c] = Grid3(nS[0], nS[1], nS[2]);
    }
    parallel_for(nSt, [&](long lo, long hi) {
        for (long i = lo; i < hi; ++i) {
            const long ix = i / ((long)nS[1] * nS[2]), iy = (i / nS[2]) % nS[1], iz = i % nS[2];
            const double G[3] = {st.gS[0][ix], st.gS[1][iy], st.gS[2][iz]};
            for (int c = 0; c < 6; ++c) {
                st.br[c][i] = strain_bracket(c, G, mp[0], st.iso);
                st.eS[c].v[i] = st.br[c][i] * Xeps.v[i];
            }
        }
    });
    Grid3 phiS;
    if (in.piezo) {
        const Convolver conv(nS);
        ConvSum cs(conv);
        st.have_phi = true;
        st.second = in.piezo == 2 || in.piezo == 3;
        const Material &m0 = mp[0];
        // polarisation P_a on the strain grid
        Grid3 P[3];
        const int first[3] = {YZ, XZ, XY};
        const ConvSum::Spec sXe14 = cs.spec(weighted_chi(in, mp, st.chiS, [](const Material &m) { return 2.0 * m.e14; }, nS));
        ConvSum::Spec seS[6];
        for (int c = 0; c < 6; ++c) seS[c] = cs.spec(st.eS[c]);
        for (int a = 0; a < 3; ++a) {
            P[a] = st.eS[first[a]];
            for (auto &v : P[a].v) v *= 2.0 * m0.e14;
            cs.add(sXe14, seS[first[a]]);
            axpy(P[a], 1.0, cs.result());
        }
        if (st.second) {
            for (int q = 0; q < 12; ++q) {
                cs.add(seS[kProd[q][0]], seS[kProd[q][1]]);
                st.prod[q] = cs.result();
            }
            const ConvSum::Spec sB114 = cs.spec(weighted_chi(in, mp, st.chiS, [](const Material &m) { return m.B114; }, nS));
            const ConvSum::Spec sB124 = cs.spec(weighted_chi(in, mp, st.chiS, [](const Material &m) { return m.B124; }, nS));
            const ConvSum::Spec sB156 = cs.spec(weighted_chi(in, mp, st.chiS, [](const Material &m) { return m.B156; }, nS));
            for (int a = 0; a < 3; ++a) {
                const Grid3 &p0 = st.prod[4 * a], &p1 = st.prod[4 * a + 1], &p2 = st.prod[4 * a + 2], &p3 = st.prod[4 * a + 3];
                Grid3 p12 = p1;
                axpy(p12, 1.0, p2);
                axpy(P[a], 2.0 * m0.B114, p0);
                axpy(P[a], 2.0 * m0.B124, p12);
                axpy(P[a], 4.0 * m0.B156, p3);
                cs.add(p0, sB114);
                axpy(P[a], 2.0, cs.result());
                cs.add(p12, sB124);
                axpy(P[a], 2.0, cs.result());
                cs.add(p3, sB156);
                axpy(P[a], 4.0, cs.result());
            }
        }
        // phi = -100 i (G . P) / (eps0 1e12 nr0^2 G^2), zero at G = 0
        phiS = Grid3(nS[0], nS[1], nS[2]);
        const double nr0 = m0.nr;
        parallel_for(nSt, [&](long lo, long hi) {
            for (long i = lo; i < hi; ++i) {
                const long ix = i / ((long)nS[1] * nS[2]), iy = (i / nS[2]) % nS[1], iz = i % nS[2];
                const double G[3] = {st.gS[0][ix], st.gS[1][iy], st.gS[2][iz]};
                const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
                if (std::sqrt(G2) <= kEps) continue;
                cplx acc = 0;
                for (int a = 0; a < 3; ++a) acc += piezo_coef(G[a], G2, nr0) * P[a].v[i];
                phiS.v[i] = cplx(0.0, 1.0) * acc;
            }
        });
        st.phiS = phiS;
    }
    if (st.commensurate) {
        const CommensurateDown down(st);
        for (int c = 0; c < 6; ++c) st.eE[c] = down.apply(st.eS[c]);
        if (st.have_phi) st.phiE = down.apply(phiS);
        return;
    }
    for (int c = 0; c < 6; ++c) st.eE[c] = downsample(st.eS[c], st.gS, st.gE, st.Lc, st.same);
    if (st.have_phi) st.phiE = downsample(phiS, st.gS, st.gE, st.Lc, st.same);
}

namespace {
// the per-element field transform at Delta-G offsets (dx, dy, dz), per V/Angstrom of F_a
double field_coef(int a, const int dd[3], const double Lc[3]) {
    const int o1 = dd[(a + 1) % 3], o2 = dd[(a + 2) % 3];
    if (dd[a] == 0 || o1 != 0 || o2 != 0) return 0.0;
    const double sign = (dd[a] % 2 == 0) ? -1.0 : 1.0;
    return sign / (-(double)dd[a] * 2.0 * kPi / Lc[a]);   // times i
}
}  // namespace

// Composes the QD_NTAB tables from the state (the per-element sums of
// hamiltonian_8_band_matrix_elements.c:27-969 and parameter_matrix_element.c:23-48).
void compose_tables(const qd_inputs &in, const TableState &st, qd_tables &t) {
    const int ns = in.N_shape;
    const std::vector<Material> &mp = st.mp;
    const int *NT = st.NT;
    const size_t nT = (size_t)NT[0] * NT[1] * NT[2];
    t.N[0] = in.N[0]; t.N[1] = in.N[1]; t.N[2] = in.N[2];
    for (int a = 0; a < 3; ++a) {
        t.NT[a] = NT[a];
        t.K[a] = in.K[a];
        t.L[a] = st.Lc[a];
        t.G[a] = (double *)std::malloc(sizeof(double) * NT[a]);
        std::memcpy(t.G[a], st.gE[a].data(), sizeof(double) * NT[a]);
    }
    t.N_vb = in.N_vb;
    t.N_cb = in.N_cb;
    t.nr0 = mp[0].nr;
    t.tab = (double *)std::calloc((size_t)QD_NTAB * nT * 2, sizeof(double));
    if (!t.tab) throw std::runtime_error("table allocation failed");
    // shape parts of the strain terms: one convolution per weight
    Grid3 hac, hav, bb, rbs, rd, sxz, syz, p6[6], hyd, bia, dxy;
    if (st.use_strain) {
        const Convolver conv(NT);
        ConvSum cs(conv);
        hyd = st.eE[XX]; axpy(hyd, 1.0, st.eE[YY]); axpy(hyd, 1.0, st.eE[ZZ]);
        bia = st.eE[XX]; axpy(bia, 1.0, st.eE[YY]); axpy(bia, -2.0, st.eE[ZZ]);
        dxy = st.eE[XX]; axpy(dxy, -1.0, st.eE[YY]);
        const Grid3 Xac = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.ac; }, NT);
        const Grid3 Xav = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.av; }, NT);
        const Grid3 Xb = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.b; }, NT);
        const Grid3 Xd = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.d; }, NT);
        const Grid3 XP = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.P; }, NT);
        const ConvSum::Spec shyd = cs.spec(hyd), sXb = cs.spec(Xb), sXd = cs.spec(Xd), sXP = cs.spec(XP);
        cs.add(Xac, shyd); hac = cs.result();
        cs.add(Xav, shyd); hav = cs.result();
        cs.add(bia, sXb); bb = cs.result();
        cs.add(dxy, sXb); rbs = cs.result();
        cs.add(st.eE[XY], sXd); rd = cs.result();
        cs.add(st.eE[XZ], sXd); sxz = cs.result();
        cs.add(st.eE[YZ], sXd); syz = cs.result();
        const int order[6] = {XX, YY, ZZ, YZ, XZ, XY};
        for (int k = 0; k < 6; ++k) { cs.add(st.eE[order[k]], sXP); p6[k] = cs.result(); }
    }
    auto put = [&](int k, size_t i, cplx v) {
        t.tab[2 * ((size_t)k * nT + i)] = v.real();
        t.tab[2 * ((size_t)k * nT + i) + 1] = v.imag();
    };
    const int cx = NT[0] / 2, cy = NT[1] / 2, cz = NT[2] / 2;
    const double r3 = std::sqrt(3.0), rh = std::sqrt(0.5);
    parallel_for((long)nT, [&](long lo, long hi) {
        for (long i = lo; i < hi; ++i) {
            const int ix = (int)(i / ((long)NT[1] * NT[2])), iy = (int)((i / NT[2]) % NT[1]), iz = (int)(i % NT[2]);
            const int dd[3] = {ix - cx, iy - cy, iz - cz};
            const bool centre = dd[0] == 0 && dd[1] == 0 && dd[2] == 0;
            // parameter expansion: p0 delta + sum_s (p_s - p_emb) chi_s
            auto pme = [&](double Material::*f) {
                cplx v = centre ? cplx(mp[0].*f, 0.0) : cplx(0, 0);
                for (int s = 1; s < ns; ++s) v += (mp[s].*f - mp[in.shapes[s].embedded].*f) * st.chiE[s].v[i];
                return v;
            };
            cplx Ecb = pme(&Material::E_cb), Ehh = pme(&Material::E_hh), Elh = pme(&Material::E_lh),
                 Eso = pme(&Material::E_so);
            // uniform field: sawtooth potential transform along one axis
            double fim = 0.0;
            for (int a = 0; a < 3; ++a) fim += in.F[a] * field_coef(a, dd, st.Lc);
            const cplx fld(0.0, fim);
            cplx Q = 0, Rb = 0, Rd = 0, Sxz = 0, Syz = 0, P6[6] = {0, 0, 0, 0, 0, 0};
            if (st.use_strain) {
                const Material &m0 = mp[0];
                const cplx ph = st.have_phi ? st.phiE.v[i] : cplx(0, 0);
                Ecb += m0.ac * hyd.v[i] + hac.v[i] + ph;
                Ehh += m0.av * hyd.v[i] + 0.5 * m0.b * bia.v[i] + hav.v[i] + 0.5 * bb.v[i] + ph;
                Elh += m0.av * hyd.v[i] - 0.5 * m0.b * bia.v[i] + hav.v[i] - 0.5 * bb.v[i] + ph;
                Eso += m0.av * hyd.v[i] + hav.v[i] + ph;
                Q = rh * (m0.b * bia.v[i] + bb.v[i]);
                Rb = -0.5 * r3 * (m0.b * dxy.v[i] + rbs.v[i]);
                Rd = m0.d * st.eE[XY].v[i] + rd.v[i];
                Sxz = m0.d * st.eE[XZ].v[i] + sxz.v[i];
                Syz = m0.d * st.eE[YZ].v[i] + syz.v[i];
                const int order[6] = {XX, YY, ZZ, YZ, XZ, XY};
                for (int k = 0; k < 6; ++k) P6[k] = m0.P * st.eE[order[k]].v[i] + p6[k].v[i];
            }
            put(QD_T_ECB, i, Ecb + fld);
            put(QD_T_EHH, i, Ehh + fld);
            put(QD_T_ELH, i, Elh + fld);
            put(QD_T_ESO, i, Eso + fld);
            put(QD_T_SC, i, pme(&Material::sc));
            put(QD_T_G1, i, pme(&Material::g1));
            put(QD_T_G2, i, pme(&Material::g2));
            put(QD_T_G3, i, pme(&Material::g3));
            put(QD_T_P, i, pme(&Material::P));
            put(QD_T_Q, i, Q);
            put(QD_T_RB, i, Rb);
            put(QD_T_RD, i, Rd);
            put(QD_T_SXZ, i, Sxz);
            put(QD_T_SYZ, i, Syz);
            for (int k = 0; k < 6; ++k) put(QD_T_PXX + k, i, P6[k]);
        }
    });
}

// Builds the QD_NTAB Delta-G tables.  Throws std::runtime_error on unsupported options.
void build_tables(const qd_inputs &in, qd_tables &t) {
    TableState st;
    build_state(in, st);
    compose_tables(in, st, t);
}

// ------------------------------------------------------------------------------------------
// Reverse pass.  With dL = Re sum_t sum_i W_t[i] dT_t[i], every intermediate X gets G_X such that
// dL = Re sum G_X dX (complex-linear maps pass G back without conjugation: for c = a (*) b,
// G_a = G_c (*) reversed(b)); real scalars p get dL/dp.
// ------------------------------------------------------------------------------------------
namespace {

// material fields that the tables use, as member pointers
double Material::*const kMatFields[] = {&Material::a,   &Material::e14, &Material::B114, &Material::B124,
                                        &Material::B156, &Material::nr, &Material::sc,   &Material::g1,
                                        &Material::g2,  &Material::g3,  &Material::P,    &Material::ac,
                                        &Material::av,  &Material::b,   &Material::d,    &Material::E_cb,
                                        &Material::E_hh, &Material::E_lh, &Material::E_so};

void add_field(Material &g, double Material::*f, double v) { g.*f += v; }

}  // namespace

void tables_vjp(const qd_inputs &in, const TableState &st, const cplx *W, TableGrad &out) {
    const int ns = in.N_shape;
    const std::vector<Material> &mp = st.mp;
    const int *NT = st.NT;
    const size_t nT = (size_t)NT[0] * NT[1] * NT[2];
    Material zero{};
    std::memset(&zero, 0, sizeof(zero));
    out.gmp.assign(ns, zero);
    out.gchiE.assign(ns, Grid3());
    out.gchiS.assign(ns, Grid3());
    for (int s = 1; s < ns; ++s) out.gchiE[s] = Grid3(NT[0], NT[1], NT[2]);
    out.gF[0] = out.gF[1] = out.gF[2] = 0.0;
    auto Wt = [&](int t) {
        Grid3 g(NT[0], NT[1], NT[2]);
        for (size_t i = 0; i < nT; ++i) g.v[i] = W[(size_t)t * nT + i];
        return g;
    };
    Grid3 Wg[QD_NTAB];
    for (int t = 0; t < QD_NTAB; ++t) Wg[t] = Wt(t);
    const int cx = NT[0] / 2, cy = NT[1] / 2, cz = NT[2] / 2;
    const size_t centre = ((size_t)cx * NT[1] + cy) * NT[2] + cz;
    // --- parameter expansions: T_f += p0 delta + sum_s (p_s - p_emb) chi_s
    const struct { double Material::*f; int t; } pmes[9] = {
        {&Material::E_cb, QD_T_ECB}, {&Material::E_hh, QD_T_EHH}, {&Material::E_lh, QD_T_ELH}, {&Material::E_so, QD_T_ESO},
        {&Material::sc, QD_T_SC},    {&Material::g1, QD_T_G1},    {&Material::g2, QD_T_G2},    {&Material::g3, QD_T_G3},
        {&Material::P, QD_T_P}};
    for (const auto &pm : pmes) {
        const Grid3 &Wf = Wg[pm.t];
        add_field(out.gmp[0], pm.f, Wf.v[centre].real());
        for (int s = 1; s < ns; ++s) {
            const int e = in.shapes[s].embedded;
            axpy(out.gchiE[s], mp[s].*pm.f - mp[e].*pm.f, Wf);
            const double g = re_dot(Wf, st.chiE[s]);
            add_field(out.gmp[s], pm.f, g);
            add_field(out.gmp[e], pm.f, -g);
        }
    }
    // --- field (linear in F)
    Grid3 Wedge = Wg[QD_T_ECB];
    axpy(Wedge, 1.0, Wg[QD_T_EHH]);
    axpy(Wedge, 1.0, Wg[QD_T_ELH]);
    axpy(Wedge, 1.0, Wg[QD_T_ESO]);
    for (size_t i = 0; i < nT; ++i) {
        const int ix = (int)(i / ((size_t)NT[1] * NT[2])), iy = (int)((i / NT[2]) % NT[1]), iz = (int)(i % NT[2]);
        const int dd[3] = {ix - cx, iy - cy, iz - cz};
        for (int a = 0; a < 3; ++a) {
            const double c = field_coef(a, dd, st.Lc);
            if (c != 0.0) out.gF[a] += -Wedge.v[i].imag() * c;   // Re(W i c)
        }
    }
    if (!st.use_strain) return;
    const Material &m0 = mp[0];
    const double r3 = std::sqrt(3.0), rh = std::sqrt(0.5);
    // --- strain terms on the electronic grid
    Grid3 hyd = st.eE[XX], bia = st.eE[XX], dxy = st.eE[XX];
    axpy(hyd, 1.0, st.eE[YY]); axpy(hyd, 1.0, st.eE[ZZ]);
    axpy(bia, 1.0, st.eE[YY]); axpy(bia, -2.0, st.eE[ZZ]);
    axpy(dxy, -1.0, st.eE[YY]);
    // output gradients of the combinations
    Grid3 Gac = Wg[QD_T_ECB];                                      // multiplies ac-weighted hydrostatic
    Grid3 Gav = Wg[QD_T_EHH]; axpy(Gav, 1.0, Wg[QD_T_ELH]); axpy(Gav, 1.0, Wg[QD_T_ESO]);
    Grid3 Gb = Wg[QD_T_EHH]; axpy(Gb, -1.0, Wg[QD_T_ELH]);
    for (auto &v : Gb.v) v *= 0.5;
    axpy(Gb, rh, Wg[QD_T_Q]);                                      // multiplies b-weighted biaxial
    Grid3 Grb = Wg[QD_T_RB];
    for (auto &v : Grb.v) v *= -0.5 * r3;                          // multiplies b-weighted (xx - yy)
    const int order[6] = {XX, YY, ZZ, YZ, XZ, XY};
    // matrix parts
    Grid3 GeE[6];
    for (int c = 0; c < 6; ++c) GeE[c] = Grid3(NT[0], NT[1], NT[2]);
    Grid3 Ghyd = Gac, Gbia = Gb, Gdxy = Grb;
    for (auto &v : Ghyd.v) v *= m0.ac;
    axpy(Ghyd, m0.av, Gav);
    for (auto &v : Gbia.v) v *= m0.b;
    for (auto &v : Gdxy.v) v *= m0.b;
    out.gmp[0].ac += re_dot(Gac, hyd);
    out.gmp[0].av += re_dot(Gav, hyd);
    out.gmp[0].b += re_dot(Gb, bia) + re_dot(Grb, dxy);
    out.gmp[0].d += re_dot(Wg[QD_T_RD], st.eE[XY]) + re_dot(Wg[QD_T_SXZ], st.eE[XZ]) + re_dot(Wg[QD_T_SYZ], st.eE[YZ]);
    axpy(GeE[XY], m0.d, Wg[QD_T_RD]);
    axpy(GeE[XZ], m0.d, Wg[QD_T_SXZ]);
    axpy(GeE[YZ], m0.d, Wg[QD_T_SYZ]);
    for (int k = 0; k < 6; ++k) {
        axpy(GeE[order[k]], m0.P, Wg[QD_T_PXX + k]);
        out.gmp[0].P += re_dot(Wg[QD_T_PXX + k], st.eE[order[k]]);
    }
    // shape parts: c = X_w (*) Y  ->  G_X += G_c (*) rev(Y), G_Y += G_c (*) rev(X)
    {
        const Convolver conv(NT);
        ConvSum cs(conv);
        const Grid3 Xac = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.ac; }, NT);
        const Grid3 Xav = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.av; }, NT);
        const Grid3 Xb = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.b; }, NT);
        const Grid3 Xd = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.d; }, NT);
        const Grid3 XP = weighted_chi(in, mp, st.chiE, [](const Material &m) { return m.P; }, NT);
        const ConvSum::Spec rhyd = cs.spec(reversed(hyd)), rbia = cs.spec(reversed(bia)), rdxy = cs.spec(reversed(dxy));
        const ConvSum::Spec rXac = cs.spec(reversed(Xac)), rXav = cs.spec(reversed(Xav)), rXb = cs.spec(reversed(Xb));
        const ConvSum::Spec rXd = cs.spec(reversed(Xd)), rXP = cs.spec(reversed(XP));
        ConvSum::Spec reE[6], sW[QD_NTAB];
        for (int c = 0; c < 6; ++c) reE[c] = cs.spec(reversed(st.eE[c]));
        for (int t = QD_T_RD; t <= QD_T_PXY; ++t) sW[t] = cs.spec(Wg[t]);
        const ConvSum::Spec sGac = cs.spec(Gac), sGav = cs.spec(Gav), sGb = cs.spec(Gb), sGrb = cs.spec(Grb);
        cs.add(sGac, rhyd); const Grid3 GXac = cs.result();
        cs.add(sGav, rhyd); const Grid3 GXav = cs.result();
        cs.add(sGb, rbia); cs.add(sGrb, rdxy); const Grid3 GXb = cs.result();
        cs.add(sW[QD_T_RD], reE[XY]); cs.add(sW[QD_T_SXZ], reE[XZ]); cs.add(sW[QD_T_SYZ], reE[YZ]);
        const Grid3 GXd = cs.result();
        for (int k = 0; k < 6; ++k) cs.add(sW[QD_T_PXX + k], reE[order[k]]);
        const Grid3 GXP = cs.result();
        cs.add(sGac, rXac); cs.add(sGav, rXav); axpy(Ghyd, 1.0, cs.result());
        cs.add(sGb, rXb); axpy(Gbia, 1.0, cs.result());
        cs.add(sGrb, rXb); axpy(Gdxy, 1.0, cs.result());
        cs.add(sW[QD_T_RD], rXd); axpy(GeE[XY], 1.0, cs.result());
        cs.add(sW[QD_T_SXZ], rXd); axpy(GeE[XZ], 1.0, cs.result());
        cs.add(sW[QD_T_SYZ], rXd); axpy(GeE[YZ], 1.0, cs.result());
        for (int k = 0; k < 6; ++k) { cs.add(sW[QD_T_PXX + k], rXP); axpy(GeE[order[k]], 1.0, cs.result()); }
        // back to the shapes
        const struct { double Material::*f; const Grid3 *G; } wts[5] = {
            {&Material::ac, &GXac}, {&Material::av, &GXav}, {&Material::b, &GXb}, {&Material::d, &GXd}, {&Material::P, &GXP}};
        for (const auto &w : wts)
            for (int s = 1; s < ns; ++s) {
                const int e = in.shapes[s].embedded;
                axpy(out.gchiE[s], mp[s].*w.f - mp[e].*w.f, *w.G);
                const double g = re_dot(*w.G, st.chiE[s]);
                add_field(out.gmp[s], w.f, g);
                add_field(out.gmp[e], w.f, -g);
            }
    }
    // hydrostatic / biaxial / (xx - yy) combinations back to the components
    axpy(GeE[XX], 1.0, Ghyd); axpy(GeE[YY], 1.0, Ghyd); axpy(GeE[ZZ], 1.0, Ghyd);
    axpy(GeE[XX], 1.0, Gbia); axpy(GeE[YY], 1.0, Gbia); axpy(GeE[ZZ], -2.0, Gbia);
    axpy(GeE[XX], 1.0, Gdxy); axpy(GeE[YY], -1.0, Gdxy);
    // --- downsampling (strain cell -> electronic cell)
    const int *nS = st.nS;
    Grid3 GeS[6];
    std::unique_ptr<CommensurateDown> down;
    if (st.commensurate) down.reset(new CommensurateDown(st));
    auto down_adjoint = [&](const Grid3 &g) {
        return down ? down->adjoint(g) : downsample_adjoint(g, st.gS, st.gE, st.Lc, st.same);
    };
    for (int c = 0; c < 6; ++c) GeS[c] = down_adjoint(GeE[c]);
    for (int s = 1; s < ns; ++s) out.gchiS[s] = Grid3(nS[0], nS[1], nS[2]);
    const long nSt = (long)nS[0] * nS[1] * nS[2];
    // --- piezo potential
    if (st.have_phi) {
        const Grid3 GphiS = down_adjoint(Wedge);
        const double nr0 = m0.nr;
        out.gmp[0].nr += re_dot(GphiS, st.phiS) * (-2.0 / nr0);
        // G_{P_a} = G_phi i coef_a
        Grid3 H[3];
        for (int a = 0; a < 3; ++a) H[a] = Grid3(nS[0], nS[1], nS[2]);
        for (long i = 0; i < nSt; ++i) {
            const long ix = i / ((long)nS[1] * nS[2]), iy = (i / nS[2]) % nS[1], iz = i % nS[2];
            const double G[3] = {st.gS[0][ix], st.gS[1][iy], st.gS[2][iz]};
            const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
            if (std::sqrt(G2) <= kEps) continue;
            for (int a = 0; a < 3; ++a) H[a].v[i] = GphiS.v[i] * cplx(0.0, piezo_coef(G[a], G2, nr0));
        }
        const Convolver conv(nS);
        ConvSum cs(conv);
        const int first[3] = {YZ, XZ, XY};
        // first order, matrix: P_a += 2 e14_0 e_first[a]
        for (int a = 0; a < 3; ++a) {
            axpy(GeS[first[a]], 2.0 * m0.e14, H[a]);
            out.gmp[0].e14 += 2.0 * re_dot(H[a], st.eS[first[a]]);
        }
        // first order, shapes: P_a += X_e14 (*) e_first[a], X_e14 = sum_s 2 (e14_s - e14_emb) chi_s
        const ConvSum::Spec rXe14 =
            cs.spec(reversed(weighted_chi(in, mp, st.chiS, [](const Material &m) { return 2.0 * m.e14; }, nS)));
        ConvSum::Spec reS[6], sH[3];
        for (int c = 0; c < 6; ++c) reS[c] = cs.spec(reversed(st.eS[c]));
        for (int a = 0; a < 3; ++a) sH[a] = cs.spec(H[a]);
        for (int a = 0; a < 3; ++a) cs.add(sH[a], reS[first[a]]);
        const Grid3 GXe14 = cs.result();
        for (int a = 0; a < 3; ++a) { cs.add(sH[a], rXe14); axpy(GeS[first[a]], 1.0, cs.result()); }
        for (int s = 1; s < ns; ++s) {
            const int e = in.shapes[s].embedded;
            axpy(out.gchiS[s], 2.0 * (mp[s].e14 - mp[e].e14), GXe14);
            const double g = 2.0 * re_dot(GXe14, st.chiS[s]);
            out.gmp[s].e14 += g;
            out.gmp[e].e14 -= g;
        }
        if (st.second) {
            Grid3 Gprod[12];
            for (int q = 0; q < 12; ++q) Gprod[q] = Grid3(nS[0], nS[1], nS[2]);
            const ConvSum::Spec rB114 = cs.spec(reversed(weighted_chi(in, mp, st.chiS, [](const Material &m) { return m.B114; }, nS)));
            const ConvSum::Spec rB124 = cs.spec(reversed(weighted_chi(in, mp, st.chiS, [](const Material &m) { return m.B124; }, nS)));
            const ConvSum::Spec rB156 = cs.spec(reversed(weighted_chi(in, mp, st.chiS, [](const Material &m) { return m.B156; }, nS)));
            Grid3 p12[3];
            ConvSum::Spec rp[12], rp12[3];
            for (int q = 0; q < 12; ++q) rp[q] = cs.spec(reversed(st.prod[q]));
            for (int a = 0; a < 3; ++a) {
                p12[a] = st.prod[4 * a + 1];
                axpy(p12[a], 1.0, st.prod[4 * a + 2]);
                rp12[a] = cs.spec(reversed(p12[a]));
            }
            for (int a = 0; a < 3; ++a) {
                // matrix: P_a += 2 B114_0 p0 + 2 B124_0 (p1 + p2) + 4 B156_0 p3
                axpy(Gprod[4 * a], 2.0 * m0.B114, H[a]);
                axpy(Gprod[4 * a + 1], 2.0 * m0.B124, H[a]);
                axpy(Gprod[4 * a + 2], 2.0 * m0.B124, H[a]);
                axpy(Gprod[4 * a + 3], 4.0 * m0.B156, H[a]);
                out.gmp[0].B114 += 2.0 * re_dot(H[a], st.prod[4 * a]);
                out.gmp[0].B124 += 2.0 * re_dot(H[a], p12[a]);
                out.gmp[0].B156 += 4.0 * re_dot(H[a], st.prod[4 * a + 3]);
                // shapes: P_a += 2 XB114 (*) p0 + 2 XB124 (*) (p1 + p2) + 4 XB156 (*) p3
                cs.add(sH[a], rB114); const Grid3 g0 = cs.result();
                cs.add(sH[a], rB124); const Grid3 g12 = cs.result();
                cs.add(sH[a], rB156); const Grid3 g3 = cs.result();
                axpy(Gprod[4 * a], 2.0, g0);
                axpy(Gprod[4 * a + 1], 2.0, g12);
                axpy(Gprod[4 * a + 2], 2.0, g12);
                axpy(Gprod[4 * a + 3], 4.0, g3);
            }
            for (int a = 0; a < 3; ++a) cs.add(sH[a], rp[4 * a]);
            Grid3 GXB114 = cs.result();
            for (auto &v : GXB114.v) v *= 2.0;
            for (int a = 0; a < 3; ++a) cs.add(sH[a], rp12[a]);
            Grid3 GXB124 = cs.result();
            for (auto &v : GXB124.v) v *= 2.0;
            for (int a = 0; a < 3; ++a) cs.add(sH[a], rp[4 * a + 3]);
            Grid3 GXB156 = cs.result();
            for (auto &v : GXB156.v) v *= 4.0;
            const struct { double Material::*f; const Grid3 *G; } wts[3] = {
                {&Material::B114, &GXB114}, {&Material::B124, &GXB124}, {&Material::B156, &GXB156}};
            for (const auto &w : wts)
                for (int s = 1; s < ns; ++s) {
                    const int e = in.shapes[s].embedded;
                    axpy(out.gchiS[s], mp[s].*w.f - mp[e].*w.f, *w.G);
                    const double g = re_dot(*w.G, st.chiS[s]);
                    add_field(out.gmp[s], w.f, g);
                    add_field(out.gmp[e], w.f, -g);
                }
            // products: prod[q] = e_a (*) e_b
            ConvSum::Spec sGp[12];
            for (int q = 0; q < 12; ++q) sGp[q] = cs.spec(Gprod[q]);
            for (int c = 0; c < 6; ++c) {
                bool any = false;
                for (int q = 0; q < 12; ++q) {
                    if (kProd[q][0] == c) { cs.add(sGp[q], reS[kProd[q][1]]); any = true; }
                    if (kProd[q][1] == c) { cs.add(sGp[q], reS[kProd[q][0]]); any = true; }
                }
                if (any) axpy(GeS[c], 1.0, cs.result());
            }
        }
    }
    // --- strain of the inclusions: e_c = bracket_c sum_s eps0_s chi_s
    Grid3 GXeps(nS[0], nS[1], nS[2]);
    for (int c = 0; c < 6; ++c)
        for (long i = 0; i < nSt; ++i) GXeps.v[i] += st.br[c][i] * GeS[c].v[i];
    for (int s = 1; s < ns; ++s) {
        const int e = in.shapes[s].embedded;
        const double eps0 = (mp[e].a - mp[s].a) / mp[0].a;
        axpy(out.gchiS[s], eps0, GXeps);
        const double g = re_dot(GXeps, st.chiS[s]);   // dL / d eps0_s
        out.gmp[s].a += -g / mp[0].a;
        out.gmp[e].a += g / mp[0].a;
        out.gmp[0].a += -g * eps0 / mp[0].a;
    }
    if (st.same)
        for (int s = 1; s < ns; ++s) {
            axpy(out.gchiE[s], 1.0, out.gchiS[s]);
            out.gchiS[s] = Grid3();
        }
}

}  // namespace qd
