// FP64 reciprocal-space fields on the strain and electronic-structure grids and the Delta-G
// tables of the Hamiltonian kernel.
//   strain:       strain_fourier_coefficient.c:23-390 (continuum-elastic Green tensor, matrix C_ij),
//                 summed over shapes as in strain_region_reciprocal_space.c:78-150
//   piezo:        piezo_region_reciprocal_space.c:23-787 + piezo_fourier_coefficient.c:24-132
//   downsampling: strain_reciprocal_space.c:23-230 / piezo_reciprocal_space.c (free lengths)
//   convolutions: convolutions.c:24-411 (central block of zero-padded products)
//   tables:       the per-element sums of hamiltonian_8_band_matrix_elements.c:27-969 and
//                 parameter_matrix_element.c:23-48, pre-combined per Delta-G
// The convolutions use exact separable DFTs of the zero-padded grids (sizes 2N_T-1 are small
// odd numbers, so a direct transform per axis is cheap and needs no FFT library).
#include <cmath>
#include <cstring>
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
    const int nx = n_[0], ny = n_[1], nz = n_[2], mx = m_[0], my = m_[1], mz = m_[2];
    std::vector<cplx> s1((size_t)nx * ny * mz), s2((size_t)nx * my * mz);
    // z
    parallel_for((long)nx * ny, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r)
            for (int k = 0; k < mz; ++k) {
                cplx acc = 0;
                const cplx *src = &a.v[(size_t)r * nz];
                for (int j = 0; j < nz; ++j) acc += src[j] * tw_[2][(size_t)j * k % mz];
                s1[(size_t)r * mz + k] = acc;
            }
    });
    // y
    parallel_for((long)nx * my, [&](long lo, long hi) {
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
    parallel_for((long)mx, [&](long lo, long hi) {
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
    const int nx = n_[0], ny = n_[1], nz = n_[2], mx = m_[0], my = m_[1], mz = m_[2];
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    std::vector<cplx> s1((size_t)nx * my * mz), s2((size_t)nx * ny * mz);
    // x: keep outputs k + cx, k < nx (conjugate twiddles = inverse transform)
    parallel_for((long)nx, [&](long lo, long hi) {
        for (long k = lo; k < hi; ++k) {
            cplx *dst = &s1[(size_t)k * my * mz];
            for (size_t q = 0; q < (size_t)my * mz; ++q) dst[q] = 0;
            for (int j = 0; j < mx; ++j) {
                const cplx w = std::conj(tw_[0][(size_t)j * (k + cx) % mx]);
                const cplx *src = &spec[(size_t)j * my * mz];
                for (size_t q = 0; q < (size_t)my * mz; ++q) dst[q] += src[q] * w;
            }
        }
    });
    parallel_for((long)nx * ny, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r) {
            const long x = r / ny, k = r % ny;
            cplx *dst = &s2[((size_t)x * ny + k) * mz];
            for (int z = 0; z < mz; ++z) dst[z] = 0;
            for (int j = 0; j < my; ++j) {
                const cplx w = std::conj(tw_[1][(size_t)j * (k + cy) % my]);
                const cplx *src = &s1[((size_t)x * my + j) * mz];
                for (int z = 0; z < mz; ++z) dst[z] += src[z] * w;
            }
        }
    });
    out = Grid3(nx, ny, nz);
    const double scale = 1.0 / ((double)mx * my * mz);
    parallel_for((long)nx * ny, [&](long lo, long hi) {
        for (long r = lo; r < hi; ++r)
            for (int k = 0; k < nz; ++k) {
                cplx acc = 0;
                const cplx *src = &s2[(size_t)r * mz];
                for (int j = 0; j < mz; ++j) acc += src[j] * std::conj(tw_[2][(size_t)j * (k + cz) % mz]);
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

}  // namespace

// Builds the QD_NTAB Delta-G tables.  Throws std::runtime_error on unsupported options.
void build_tables(const qd_inputs &in, qd_tables &t) {
    const int ns = in.N_shape;
    std::vector<Material> mp = interpolate_materials(in);
    const bool use_strain = in.strain != 0;
    const bool iso = in.isotropic != 0;
    if (use_strain && iso)
        for (auto &m : mp) m.C44 = 0.5 * (m.C11 - m.C12);
    const qd_shape &c0 = in.shapes[0];
    const double Lc[3] = {c0.L1, c0.L2, c0.L3};
    std::vector<double> gE[3];
    int NT[3];
    for (int a = 0; a < 3; ++a) {
        gE[a] = reciprocal_axis(in.N[a], Lc[a]);
        NT[a] = 2 * in.N[a] - 1;
    }
    const size_t nT = (size_t)NT[0] * NT[1] * NT[2];

    // electronic-structure-cell chi per shape (copied from the strain cell when identical)
    std::vector<Grid3> chiE(ns);
    std::vector<Grid3> chiS(ns);
    std::vector<double> gS[3];
    bool same = false;
    if (use_strain) {
        int strain_type = in.strain_type;
        if (strain_type == 1 && in.RL[0] == 1 && in.RL[1] == 1 && in.RL[2] == 1 && in.N_S[0] == in.N[0] &&
            in.N_S[1] == in.N[1] && in.N_S[2] == in.N[2])
            strain_type = 0;
        if (strain_type != 0)
            throw std::runtime_error("commensurate strain supercell (Strain_supercell_type 1 with relative "
                                     "lengths > 1) is not supported; use free lengths");
        same = true;
        for (int a = 0; a < 3; ++a) {
            gS[a] = reciprocal_axis(in.N_S[a], in.SL[a]);
            same = same && in.N_S[a] == in.N[a] && in.SL[a] == Lc[a];
        }
        for (int s = 1; s < ns; ++s) {
            chiS[s] = Grid3((int)gS[0].size(), (int)gS[1].size(), (int)gS[2].size());
            shape_grid(in, s, gS[0], gS[1], gS[2], Cell::Strain, chiS[s].v);
        }
    }
    for (int s = 1; s < ns; ++s) {
        if (use_strain && same) { chiE[s] = chiS[s]; continue; }
        chiE[s] = Grid3(NT[0], NT[1], NT[2]);
        shape_grid(in, s, gE[0], gE[1], gE[2], Cell::Electronic, chiE[s].v);
    }

    Grid3 eE[6], phiE;
    bool have_phi = false;
    if (use_strain) {
        const int nS[3] = {(int)gS[0].size(), (int)gS[1].size(), (int)gS[2].size()};
        const long nSt = (long)nS[0] * nS[1] * nS[2];
        StrainField eS;
        for (int c = 0; c < 6; ++c) eS.e[c] = Grid3(nS[0], nS[1], nS[2]);
        parallel_for(nSt, [&](long lo, long hi) {
            for (long i = lo; i < hi; ++i) {
                const long ix = i / ((long)nS[1] * nS[2]), iy = (i / nS[2]) % nS[1], iz = i % nS[2];
                const double G[3] = {gS[0][ix], gS[1][iy], gS[2][iz]};
                for (int c = 0; c < 6; ++c) {
                    const double br = strain_bracket(c, G, mp[0], iso);
                    cplx acc = 0;
                    for (int s = 1; s < ns; ++s) {
                        const double eps0 = (mp[in.shapes[s].embedded].a - mp[s].a) / mp[0].a;
                        acc += eps0 * chiS[s].v[i] * br;
                    }
                    eS.e[c].v[i] = acc;
                }
            }
        });
        Grid3 phiS;
        if (in.piezo) {
            const Convolver conv(nS);
            phiS = Grid3(nS[0], nS[1], nS[2]);
            const double nr0 = mp[0].nr;
            auto add_phi = [&](const cplx *Px, const cplx *Py, const cplx *Pz, double sx, double sy, double sz) {
                // phi += -100 i (G . P)/(eps0[F/m] 1e12 nr0^2 G^2), zero at G = 0
                parallel_for(nSt, [&](long lo, long hi) {
                    for (long i = lo; i < hi; ++i) {
                        const long ix = i / ((long)nS[1] * nS[2]), iy = (i / nS[2]) % nS[1], iz = i % nS[2];
                        const double Gx = gS[0][ix], Gy = gS[1][iy], Gz = gS[2][iz];
                        const double G2 = Gx * Gx + Gy * Gy + Gz * Gz;
                        if (std::sqrt(G2) <= kEps) continue;
                        const cplx dot = Gx * sx * Px[i] + Gy * sy * Py[i] + Gz * sz * Pz[i];
                        phiS.v[i] += cplx(0.0, -100.0) * dot / (8.854188 * nr0 * nr0 * G2);
                    }
                });
            };
            // first order, matrix
            const double e0 = 2.0 * mp[0].e14;
            add_phi(eS.e[YZ].v.data(), eS.e[XZ].v.data(), eS.e[XY].v.data(), e0, e0, e0);
            // second-order strain products (central block of e_a (*) e_b)
            const bool second = in.piezo == 2 || in.piezo == 3;
            // product list per polarisation component: {B114 term, B124 a, B124 b, B156 term}
            const int prod[12][2] = {{XX, YZ}, {YY, YZ}, {ZZ, YZ}, {XY, XZ}, {YY, XZ}, {XX, XZ},
                                     {ZZ, XZ}, {XY, YZ}, {ZZ, XY}, {XX, XY}, {YY, XY}, {XZ, YZ}};
            std::vector<std::vector<cplx>> specE(6), specP(12);
            Grid3 prodG[12];
            if (second) {
                for (int c = 0; c < 6; ++c) conv.forward(eS.e[c], specE[c]);
                for (int q = 0; q < 12; ++q) {
                    std::vector<cplx> sp = specE[prod[q][0]];
                    const auto &B = specE[prod[q][1]];
                    for (size_t i = 0; i < sp.size(); ++i) sp[i] *= B[i];
                    conv.inverse_central(sp, prodG[q]);
                    conv.forward(prodG[q], specP[q]);
                }
            }
            auto add_second = [&](const Grid3 *P, double B114, double B124, double B156) {
                Grid3 px(nS[0], nS[1], nS[2]), py = px, pz = px;
                for (long i = 0; i < nSt; ++i) {
                    px.v[i] = 2 * B114 * P[0].v[i] + 2 * B124 * P[1].v[i] + 2 * B124 * P[2].v[i] + 4 * B156 * P[3].v[i];
                    py.v[i] = 2 * B114 * P[4].v[i] + 2 * B124 * P[5].v[i] + 2 * B124 * P[6].v[i] + 4 * B156 * P[7].v[i];
                    pz.v[i] = 2 * B114 * P[8].v[i] + 2 * B124 * P[9].v[i] + 2 * B124 * P[10].v[i] + 4 * B156 * P[11].v[i];
                }
                add_phi(px.v.data(), py.v.data(), pz.v.data(), 1.0, 1.0, 1.0);
            };
            if (second) add_second(prodG, mp[0].B114, mp[0].B124, mp[0].B156);
            for (int s = 1; s < ns; ++s) {
                const Material &ms = mp[s], &me = mp[in.shapes[s].embedded];
                std::vector<cplx> specChi;
                conv.forward(chiS[s], specChi);
                auto conv_with = [&](const std::vector<cplx> &spec) {
                    std::vector<cplx> sp = specChi;
                    for (size_t i = 0; i < sp.size(); ++i) sp[i] *= spec[i];
                    Grid3 r;
                    conv.inverse_central(sp, r);
                    return r;
                };
                if (specE[0].empty())
                    for (int c = 0; c < 6; ++c) conv.forward(eS.e[c], specE[c]);
                const Grid3 cyz = conv_with(specE[YZ]), cxz = conv_with(specE[XZ]), cxy = conv_with(specE[XY]);
                const double de = 2.0 * (ms.e14 - me.e14);
                add_phi(cyz.v.data(), cxz.v.data(), cxy.v.data(), de, de, de);
                if (second) {
                    Grid3 cp[12];
                    for (int q = 0; q < 12; ++q) cp[q] = conv_with(specP[q]);
                    add_second(cp, ms.B114 - me.B114, ms.B124 - me.B124, ms.B156 - me.B156);
                }
            }
        }
        for (int c = 0; c < 6; ++c) eE[c] = downsample(eS.e[c], gS, gE, Lc, same);
        if (in.piezo) { phiE = downsample(phiS, gS, gE, Lc, same); have_phi = true; }
    }

    // ---- compose the Delta-G tables ----
    t.N[0] = in.N[0]; t.N[1] = in.N[1]; t.N[2] = in.N[2];
    for (int a = 0; a < 3; ++a) {
        t.NT[a] = NT[a];
        t.K[a] = in.K[a];
        t.L[a] = Lc[a];
        t.G[a] = (double *)std::malloc(sizeof(double) * NT[a]);
        std::memcpy(t.G[a], gE[a].data(), sizeof(double) * NT[a]);
    }
    t.N_vb = in.N_vb;
    t.N_cb = in.N_cb;
    t.nr0 = mp[0].nr;
    t.tab = (double *)std::calloc((size_t)QD_NTAB * nT * 2, sizeof(double));
    if (!t.tab) throw std::runtime_error("table allocation failed");
    auto put = [&](int k, size_t i, cplx v) {
        t.tab[2 * ((size_t)k * nT + i)] = v.real();
        t.tab[2 * ((size_t)k * nT + i) + 1] = v.imag();
    };
    // per-shape convolutions chi_s (*) e_ij on the electronic grid
    std::vector<Grid3> cv[6];
    if (use_strain) {
        const Convolver conv(NT);
        std::vector<cplx> specE[6];
        for (int c = 0; c < 6; ++c) conv.forward(eE[c], specE[c]);
        for (int c = 0; c < 6; ++c) cv[c].resize(ns);
        for (int s = 1; s < ns; ++s) {
            std::vector<cplx> specChi;
            conv.forward(chiE[s], specChi);
            for (int c = 0; c < 6; ++c) {
                std::vector<cplx> sp = specChi;
                for (size_t i = 0; i < sp.size(); ++i) sp[i] *= specE[c][i];
                conv.inverse_central(sp, cv[c][s]);
            }
        }
    }
    const int cx = NT[0] / 2, cy = NT[1] / 2, cz = NT[2] / 2;
    const double r2 = std::sqrt(2.0), r3 = std::sqrt(3.0), rh = std::sqrt(0.5);
    (void)r2;
    parallel_for((long)nT, [&](long lo, long hi) {
        for (long i = lo; i < hi; ++i) {
            const int ix = (int)(i / ((long)NT[1] * NT[2])), iy = (int)((i / NT[2]) % NT[1]), iz = (int)(i % NT[2]);
            const int dx = ix - cx, dy = iy - cy, dz = iz - cz;
            const bool centre = dx == 0 && dy == 0 && dz == 0;
            // parameter expansion: p0 delta + sum_s (p_s - p_emb) chi_s
            auto pme = [&](double Material::*f) {
                cplx v = centre ? cplx(mp[0].*f, 0.0) : cplx(0, 0);
                for (int s = 1; s < ns; ++s) v += (mp[s].*f - mp[in.shapes[s].embedded].*f) * chiE[s].v[i];
                return v;
            };
            auto shape_sum = [&](double Material::*f, int c) {
                cplx v = 0;
                for (int s = 1; s < ns; ++s) v += (mp[s].*f - mp[in.shapes[s].embedded].*f) * cv[c][s].v[i];
                return v;
            };
            cplx Ecb = pme(&Material::E_cb), Ehh = pme(&Material::E_hh), Elh = pme(&Material::E_lh),
                 Eso = pme(&Material::E_so);
            // uniform field: sawtooth potential transform along one axis
            cplx fld = 0;
            const double Fv[3] = {in.F[0], in.F[1], in.F[2]};
            const int dd[3] = {dx, dy, dz};
            for (int a = 0; a < 3; ++a) {
                const int o1 = dd[(a + 1) % 3], o2 = dd[(a + 2) % 3];
                if (dd[a] != 0 && o1 == 0 && o2 == 0 && Fv[a] != 0.0) {
                    const double sign = (dd[a] % 2 == 0) ? -1.0 : 1.0;
                    fld += cplx(0.0, Fv[a] * sign) / (-(double)dd[a] * 2.0 * kPi / Lc[a]);
                }
            }
            cplx Q = 0, Rb = 0, Rd = 0, Sxz = 0, Syz = 0, P6[6] = {0, 0, 0, 0, 0, 0};
            if (use_strain) {
                const cplx hyd = eE[XX].v[i] + eE[YY].v[i] + eE[ZZ].v[i];
                const cplx bia = eE[XX].v[i] + eE[YY].v[i] - 2.0 * eE[ZZ].v[i];
                cplx chh = 0, cbia = 0;
                // shape-weighted hydrostatic / biaxial parts
                const cplx ph = have_phi ? phiE.v[i] : cplx(0, 0);
                cplx hac = 0, hav = 0, bb = 0;
                for (int s = 1; s < ns; ++s) {
                    const Material &ms = mp[s], &me = mp[in.shapes[s].embedded];
                    chh = cv[XX][s].v[i] + cv[YY][s].v[i] + cv[ZZ][s].v[i];
                    cbia = cv[XX][s].v[i] + cv[YY][s].v[i] - 2.0 * cv[ZZ][s].v[i];
                    hac += (ms.ac - me.ac) * chh;
                    hav += (ms.av - me.av) * chh;
                    bb += (ms.b - me.b) * cbia;
                }
                const Material &m0 = mp[0];
                Ecb += m0.ac * hyd + hac + ph;
                Ehh += m0.av * hyd + 0.5 * m0.b * bia + hav + 0.5 * bb + ph;
                Elh += m0.av * hyd - 0.5 * m0.b * bia + hav - 0.5 * bb + ph;
                Eso += m0.av * hyd + hav + ph;
                Q = rh * (m0.b * bia + bb);
                Rb = -0.5 * r3 * (m0.b * (eE[XX].v[i] - eE[YY].v[i]) + shape_sum(&Material::b, XX) -
                                  shape_sum(&Material::b, YY));
                Rd = m0.d * eE[XY].v[i] + shape_sum(&Material::d, XY);
                Sxz = m0.d * eE[XZ].v[i] + shape_sum(&Material::d, XZ);
                Syz = m0.d * eE[YZ].v[i] + shape_sum(&Material::d, YZ);
                const int order[6] = {XX, YY, ZZ, YZ, XZ, XY};
                for (int k = 0; k < 6; ++k) P6[k] = m0.P * eE[order[k]].v[i] + shape_sum(&Material::P, order[k]);
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
            put(QD_T_PXX, i, P6[0]);
            put(QD_T_PYY, i, P6[1]);
            put(QD_T_PZZ, i, P6[2]);
            put(QD_T_PYZ, i, P6[3]);
            put(QD_T_PXZ, i, P6[4]);
            put(QD_T_PXY, i, P6[5]);
        }
    });
}

}  // namespace qd
