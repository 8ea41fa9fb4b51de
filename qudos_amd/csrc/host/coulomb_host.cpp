// FP64 host parts of the Coulomb matrix elements (see qd_coulomb_host.hpp).
//
// Reference (code/source/8_band): Madelung term coulomb_matrix_elements.c:250-334 (its Ewald
// sums run on the device, coulomb.hip); free-length upsampling B_coulomb_supercell.c:23-100;
// contractions and corrections coulomb_integrals.c:23-356; output
// write_coulomb_matrix_elements.c:23-395.
#include <cmath>
#include <cstdio>
#include <memory>

#include "qd_coulomb_host.hpp"
#include "qd_host.hpp"

namespace qd {

namespace {
constexpr double kESQEPS = 180.9513;   // e^2/eps0 [eV Angstrom], 8_band.h:52
constexpr double kEPS = 1.0e-08;       // 8_band.h:59
constexpr double kPi = 3.14159265358979323846;

}  // namespace

void coulomb_setup(const qd_inputs &in, const qd_tables &t, CoulombSetup &cs) {
    if (in.cme < 1 || in.cme > 4) throw StatusError("CME level must be 1..4", QD_E_ARG);
    const double L[3] = {t.L[0], t.L[1], t.L[2]};
    double CL[3] = {in.CL[0], in.CL[1], in.CL[2]};
    cs.cme = in.cme;
    for (int a = 0; a < 3; ++a) {
        cs.N[a] = t.N[a];
        cs.NT[a] = 2 * t.N[a] - 1;
        cs.NC[a] = in.N_C[a];
        cs.NTC[a] = 2 * in.N_C[a] - 1;
        if (CL[a] <= 0) throw StatusError("Coulomb supercell lengths missing", QD_E_INPUT);
    }
    if (in.coulomb_type != 0 && in.coulomb_type != 1) throw StatusError("Coulomb_supercell_type must be 0 or 1", QD_E_INPUT);
    const bool commensurate = in.coulomb_type == 1;
    if (commensurate) {
        const qd_shape c0 = in.shapes ? in.shapes[0] : qd_shape{};
        if (c0.Ox != 0.0 || c0.Oy != 0.0 || c0.Oz != 0.0 || c0.alpha != 0.0 || c0.beta != 0.0 || c0.gamma != 0.0)
            throw StatusError("commensurate Coulomb supercell with a displaced or rotated electronic cell is not "
                              "supported", QD_E_UNSUPPORTED);
        for (int a = 0; a < 3; ++a)
            if (in.RCL[a] < 1) throw StatusError("Coulomb length (relative) must be >= 1", QD_E_INPUT);
    }
    cs.omega_c = CL[0] * CL[1] * CL[2];
    // Upsampling matrices, one per axis.  Free lengths (B_coulomb_supercell.c:37-89): the box
    // transform I0(g) = 2 sin(g L/2)/g between Coulomb and electronic-structure wave vectors, zero
    // beyond the Coulomb basis.  Commensurate lengths CL = RCL L (B_coulomb_supercell_fft.c:23-100,
    // characteristic_function_coulomb_supercell.c): B placed at Coulomb offsets RCL e, linearly
    // convolved with the electronic cell's transform on the Coulomb grid (1/Omega_C prod I0, times
    // the softening Gaussian when enabled) and scaled by Omega_C/Omega_es; that transform is a
    // product over axes, so the convolution is the same separable map with entries
    // I0(G_C[d]) at d = Q - RCL e, |d| <= N_C - 1 (the transform's own grid).  An axis whose
    // plane-wave count and length both equal the electronic-structure ones is copied.
    const double sig2 = in.softening ? in.sigma_softening * in.sigma_softening : 0.0;
    for (int a = 0; a < 3; ++a) {
        cs.copy[a] = (cs.NC[a] == cs.N[a] && L[a] == CL[a]) && (!commensurate || in.RCL[a] == 1);
        cs.up[a].clear();
        if (cs.copy[a]) continue;
        const std::vector<double> G = reciprocal_axis(cs.N[a], L[a]);
        const std::vector<double> GC = reciprocal_axis(cs.NC[a], CL[a]);
        cs.up[a].assign((size_t)cs.NTC[a] * cs.NT[a], 0.0);
        const int cC = cs.NTC[a] / 2, cE = cs.NT[a] / 2;
        for (int i = 0; i < cs.NTC[a]; ++i)
            for (int j = 0; j < cs.NT[a]; ++j) {
                double v;
                if (commensurate) {
                    const int d = (i - cC) - in.RCL[a] * (j - cE);
                    if (d < -cC || d > cC) continue;
                    const double g = GC[cC + d];
                    v = std::fabs(g) < kEPS ? L[a] : 2.0 * std::sin(g * L[a] / 2.0) / g;
                    if (sig2 > 0.0) v *= std::exp(-0.5 * sig2 * g * g);
                } else {
                    const double g = GC[i] - G[j];
                    if (std::fabs(g) < kEPS) v = L[a];
                    else if (std::fabs(g) > GC[cs.NTC[a] - 1] + kEPS) v = 0.0;
                    else v = 2.0 * std::sin(g * L[a] / 2.0) / g;
                }
                cs.up[a][(size_t)i * cs.NT[a] + j] = v;
            }
    }
    // fold 1/Omega_es and the copied axes' factor L into the first non-copied matrix
    double scale = 1.0 / (L[0] * L[1] * L[2]);
    for (int a = 0; a < 3; ++a)
        if (cs.copy[a]) scale *= L[a];
    if (!cs.identity()) {
        for (int a = 0; a < 3; ++a)
            if (!cs.copy[a]) {
                for (double &v : cs.up[a]) v *= scale;
                break;
            }
    }
    // Coulomb-grid axes and the Ewald parameters of the Madelung constant; the Poisson kernel, the
    // dipole / quadrupole vectors and both Ewald sums are evaluated on the device (coulomb.hip)
    for (int a = 0; a < 3; ++a) {
        cs.gq[a] = reciprocal_axis(cs.NC[a], CL[a]);
        cs.CL[a] = CL[a];
    }
    cs.nr0 = t.nr0;
    cs.pre = kESQEPS / (4.0 * kPi * t.nr0 * t.nr0);
    cs.eta = std::sqrt(kPi / (CL[0] * CL[1]));
    cs.a_mad = 0.0;
    // real-space part: rank r > 0 of the reference sums the flat index slice of rank r-1, rank 0
    // none (:287-305), over (mx, my, mz) with my < NyT/2, mz < NzT/2: the device sums flat indices
    // below mad_hi
    const int ranks = in.madelung_ranks > 0 ? in.madelung_ranks : 1;
    const long nq = cs.nq();
    cs.mad_hy = cs.NTC[1] / 2;
    cs.mad_hz = cs.NTC[2] / 2;
    cs.mad_hi = 0;
    for (int p = 0; p < ranks - 1; ++p) cs.mad_hi += nq / ranks + ((p < nq % ranks) ? 1 : 0);
    if (cs.mad_hy == 0 || cs.mad_hz == 0) cs.mad_hi = 0;
}

double madelung_constant(const CoulombSetup &cs, double reciprocal_sum, double real_sum) {
    return -2.0 * cs.eta / std::sqrt(kPi) - kPi / (cs.eta * cs.eta * cs.omega_c) + reciprocal_sum + real_sum;
}

void coulomb_assemble(const CoulombSetup &cs, const zc *C, int P, int R, int nh, int ne, zc *hh, zc *ee, zc *ed,
                      zc *ex) {
    const int S = nh + ne;
    const int cme = cs.cme;
    const double om = cs.omega_c;
    const double fdip = cs.pre * (4.0 * kPi / (3.0 * om * om * om));
    const double fquad = -cs.pre * (2.0 * kPi / (3.0 * om * om));
    auto Cat = [&](long p, long r) { return C[p * R + r]; };
    // V_ijkl for i <= j (coulomb_integrals.c:75-86): zdotu with the reversed, kernel-weighted
    // B of pair (k, l) when l >= k, zdotc with that of pair (l, k) otherwise
    auto V = [&](int i, int j, int k, int l) {
        const long pij = pair_index(i, j);
        return l >= k ? Cat(pij, pair_index(k, l)) : Cat(pij, P + pair_index(l, k));
    };
    // dipole moments d[ax][l][k] (coulomb_matrix_elements.c:517-537)
    std::vector<zc> D;
    if (cme >= 3) {
        D.assign((size_t)3 * S * S, zc(0, 0));
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < S; ++k)
                for (int l = 0; l < S; ++l)
                    D[((size_t)ax * S + l) * S + k] =
                        l >= k ? Cat(pair_index(k, l), 2L * P + ax) : std::conj(Cat(pair_index(l, k), 2L * P + 3 + ax));
    }
    auto d = [&](int ax, int a, int b) { return D[((size_t)ax * S + a) * S + b]; };
    auto Q = [&](long p) { return Cat(p, 2L * P + 6); };
    auto corr = [&](int i, int j, int k, int l) {
        zc v = 0;
        if (cme >= 2 && i == j && k == l) v += -cs.pre * cs.a_mad;
        if (cme >= 3) v += fdip * (d(0, j, i) * d(0, l, k) + d(1, j, i) * d(1, l, k) + d(2, j, i) * d(2, l, k));
        if (cme >= 4) {
            if (i == j && l >= k) v += fquad * Q(pair_index(k, l));
            else if (i == j && l < k) v += fquad * std::conj(Q(pair_index(l, k)));
            if (k == l) v += fquad * Q(pair_index(i, j));
        }
        return v;
    };
    auto lstart = [](int i, int j, int k) { return j == i ? k : (k == i ? j : (k < j ? i + 1 : i)); };
    const size_t nh4 = (size_t)nh * nh * nh * nh, ne4 = (size_t)ne * ne * ne * ne, neh = (size_t)nh * nh * ne * ne;
    std::fill(hh, hh + nh4, zc(0, 0));
    std::fill(ee, ee + ne4, zc(0, 0));
    std::fill(ed, ed + neh, zc(0, 0));
    std::fill(ex, ex + neh, zc(0, 0));
    // hh (:47-143)
    for (int i = 0; i < nh; ++i)
        for (int j = i; j < nh; ++j)
            for (int k = i; k < nh; ++k)
                for (int l = lstart(i, j, k); l < nh; ++l)
                    hh[(((size_t)i * nh + j) * nh + k) * nh + l] = V(i, j, k, l) + corr(i, j, k, l);
    // ee (:145-227)
    const int o = nh;
    for (int i = 0; i < ne; ++i)
        for (int j = i; j < ne; ++j)
            for (int k = i; k < ne; ++k)
                for (int l = lstart(i, j, k); l < ne; ++l)
                    ee[(((size_t)i * ne + j) * ne + k) * ne + l] =
                        V(o + i, o + j, o + k, o + l) + corr(o + i, o + j, o + k, o + l);
    // eh direct (:229-311): i, j holes, k, l electrons
    for (int i = 0; i < nh; ++i)
        for (int j = i; j < nh; ++j)
            for (int k = 0; k < ne; ++k)
                for (int l = (j == i ? k : 0); l < ne; ++l)
                    ed[(((size_t)i * nh + j) * ne + k) * ne + l] = V(i, j, o + k, o + l) + corr(i, j, o + k, o + l);
    // eh exchange (:313-349): zdotc of pair (l, o+k) against pair (i, o+j), dipole term only
    for (int i = 0; i < nh; ++i)
        for (int j = 0; j < ne; ++j)
            for (int k = j; k < ne; ++k)
                for (int l = (k == j ? i : 0); l < nh; ++l) {
                    zc v = Cat(pair_index(i, o + j), P + pair_index(l, o + k));
                    if (cme >= 3)
                        v += fdip * (d(0, o + j, i) * d(0, l, o + k) + d(1, o + j, i) * d(1, l, o + k) +
                                     d(2, o + j, i) * d(2, l, o + k));
                    ex[(((size_t)i * ne + j) * ne + k) * nh + l] = v;
                }
}

namespace {
struct FileCloser {
    void operator()(FILE *f) const { if (f) std::fclose(f); }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

FilePtr open_out(const std::string &dir, const char *name) {
    const std::string path = (dir.empty() ? std::string(".") : dir) + "/" + name;
    FilePtr f(std::fopen(path.c_str(), "w"));
    if (!f) throw StatusError("cannot write " + path, QD_E_INPUT);
    return f;
}

void put(FILE *f, zc v, bool conj) {
    std::fprintf(f, "%17.10e\t%17.10e\t%17.10e\t", v.real(), conj ? -v.imag() : v.imag(), std::abs(v));
}

bool unique_same(int i, int j, int k, int l) {
    return j >= i && k >= i && l >= (j == i ? k : (k == i ? j : (k < j ? i + 1 : i)));
}

// hh / ee files: Patra Eq. 2.86 order i k l j, with the symmetry look-ups of :60-117
void write_same(FILE *f, const zc *T, int n, bool descending) {
    auto at = [&](int i, int j, int k, int l) { return T[(((size_t)i * n + j) * n + k) * n + l]; };
    auto idx = [&](int c) { return descending ? n - 1 - c : c; };
    auto lab = [&](int x) { return descending ? n - x : x + 1; };
    for (int ci = 0; ci < n; ++ci)
        for (int ck = 0; ck < n; ++ck) {
            for (int cl = 0; cl < n; ++cl) {
                for (int cj = 0; cj < n; ++cj) {
                    const int i = idx(ci), k = idx(ck), l = idx(cl), j = idx(cj);
                    std::fprintf(f, "%d\t%d\t%d\t%d\t", lab(i), lab(k), lab(l), lab(j));
                    if (unique_same(i, j, k, l)) put(f, at(i, j, k, l), false);
                    else if (unique_same(k, l, i, j)) put(f, at(k, l, i, j), false);
                    else if (unique_same(j, i, l, k)) put(f, at(j, i, l, k), true);
                    else if (unique_same(l, k, j, i)) put(f, at(l, k, j, i), true);
                    std::fprintf(f, "\n");
                }
                std::fprintf(f, "\n");
            }
        }
}
}  // namespace

void coulomb_write(const std::string &dir, int nh, int ne, const zc *hh, const zc *ee, const zc *ed, const zc *ex) {
    {
        FilePtr f = open_out(dir, "coulomb_matrix_elements_hh.dat");
        write_same(f.get(), hh, nh, true);
    }
    {
        FilePtr f = open_out(dir, "coulomb_matrix_elements_ee.dat");
        write_same(f.get(), ee, ne, false);
    }
    {   // :229-300
        FilePtr f = open_out(dir, "coulomb_matrix_elements_eh_direct.dat");
        auto at = [&](int i, int j, int k, int l) { return ed[(((size_t)i * nh + j) * ne + k) * ne + l]; };
        for (int k = 0; k < ne; ++k)
            for (int i = nh - 1; i >= 0; --i)
                for (int j = nh - 1; j >= 0; --j) {
                    for (int l = 0; l < ne; ++l) {
                        std::fprintf(f.get(), "%d\t%d\t%d\t%d\t", k + 1, nh - i, nh - j, l + 1);
                        if (j >= i && l >= (j == i ? k : 0)) put(f.get(), at(i, j, k, l), false);
                        else if (i >= j && k >= (i == j ? l : 0)) put(f.get(), at(j, i, l, k), true);
                        std::fprintf(f.get(), "\n");
                    }
                    std::fprintf(f.get(), "\n");
                }
    }
    {   // :302-395
        FilePtr f = open_out(dir, "coulomb_matrix_elements_eh_exchange.dat");
        auto at = [&](int i, int j, int k, int l) { return ex[(((size_t)i * ne + j) * ne + k) * nh + l]; };
        for (int k = 0; k < ne; ++k)
            for (int i = nh - 1; i >= 0; --i)
                for (int j = 0; j < ne; ++j) {
                    for (int l = nh - 1; l >= 0; --l) {
                        std::fprintf(f.get(), "%d\t%d\t%d\t%d\t", k + 1, nh - i, j + 1, nh - l);
                        if (k >= j && l >= (k == j ? i : 0)) put(f.get(), at(i, j, k, l), false);
                        else if (j >= k && i >= (j == k ? l : 0)) put(f.get(), at(l, k, j, i), true);
                        std::fprintf(f.get(), "\n");
                    }
                    std::fprintf(f.get(), "\n");
                }
    }
}

}  // namespace qd

using namespace qd;

extern "C" int qd_write_coulomb(const char *dir, int nh, int ne, const double *hh, const double *ee,
                                const double *eh_direct, const double *eh_exchange) {
    if (!hh || !ee || !eh_direct || !eh_exchange || nh < 1 || ne < 1) {
        set_error("bad argument");
        return QD_E_ARG;
    }
    try {
        coulomb_write(dir ? dir : ".", nh, ne, reinterpret_cast<const zc *>(hh), reinterpret_cast<const zc *>(ee),
                      reinterpret_cast<const zc *>(eh_direct), reinterpret_cast<const zc *>(eh_exchange));
        return QD_OK;
    } catch (const StatusError &e) {
        set_error(e.what());
        return e.code;
    }
}

// Test support, host only: the upsampling of P pair grids B (electronic-structure Delta-G grid)
// to the Coulomb grid with the matrices of coulomb_setup, axis by axis in the order and form of the
// device's k_apply_axis (copied axes skipped).
extern "C" int qd_selftest_coulomb_upsample_host(const qd_inputs *in, const qd_tables *t, int P, const double *B,
                                                 double *out) {
    if (!in || !t || !B || !out || P < 1) {
        set_error("bad argument");
        return QD_E_ARG;
    }
    try {
        qd_inputs in1 = *in;
        in1.cme = 1;   // the multipole vectors and the Madelung sum do not enter the upsampling
        CoulombSetup cs;
        coulomb_setup(in1, *t, cs);
        int dims[3] = {cs.NT[0], cs.NT[1], cs.NT[2]};
        std::vector<zc> cur(reinterpret_cast<const zc *>(B), reinterpret_cast<const zc *>(B) + (size_t)P * cs.nes());
        for (int k = 0; k < 3; ++k) {
            if (cs.copy[k]) continue;
            int od[3] = {dims[0], dims[1], dims[2]};
            od[k] = cs.NTC[k];
            const long osz = (long)od[0] * od[1] * od[2], isz = (long)dims[0] * dims[1] * dims[2];
            const long istride = k == 0 ? (long)dims[1] * dims[2] : (k == 1 ? dims[2] : 1);
            std::vector<zc> nxt((size_t)P * osz);
            for (long b = 0; b < P; ++b)
                for (int x = 0; x < od[0]; ++x)
                    for (int y = 0; y < od[1]; ++y)
                        for (int z = 0; z < od[2]; ++z) {
                            const int i = k == 0 ? x : (k == 1 ? y : z);
                            long b0 = b * isz;
                            if (k == 0) b0 += (long)y * dims[2] + z;
                            else if (k == 1) b0 += (long)x * dims[1] * dims[2] + z;
                            else b0 += ((long)x * dims[1] + y) * dims[2];
                            zc acc = 0;
                            for (int j = 0; j < dims[k]; ++j) acc += cs.up[k][(size_t)i * dims[k] + j] * cur[b0 + j * istride];
                            nxt[b * osz + ((long)x * od[1] + y) * od[2] + z] = acc;
                        }
            cur.swap(nxt);
            dims[k] = cs.NTC[k];
        }
        if ((long)cur.size() != (long)P * cs.nq()) throw StatusError("copied axis with a different grid", QD_E_INPUT);
        std::copy(cur.begin(), cur.end(), reinterpret_cast<zc *>(out));
        return QD_OK;
    } catch (const StatusError &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_error(e.what());
        return QD_E_INPUT;
    }
}
