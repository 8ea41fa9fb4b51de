// Alloy material parameters for each shape (restates material_parameters.c:23-1542 and
// interpolate_quaternary.c:23-160).  The numeric values are published physical data
// (Vurgaftman, Meyer, Ram-Mohan, JAP 89, 5815 (2001); Caro, Schulz, O'Reilly, PRB 91, 075203
// (2015)), tabulated in QUDOS at code/source/parameters/material_parameters.h:137-355.
#include <cmath>
#include <stdexcept>

#include "qd_host.hpp"

namespace qd {
namespace {

// One binary compound.  Energies eV, lengths Angstrom, elastic constants GPa, piezo C/m^2.
struct Binary {
    const char *name;
    double a, C11, C12, C44, e14, e14_calc, B114, B124, B156;
    double ac, av, b, d;
    double Ev_avg, Eg0, Dso, mc, Ep, gL1, gL2, gL3, nr;
    double varshni_alpha_meV, varshni_beta;
};

// clang-format off
const Binary kBinaries[15] = {
 {"AlN",  4.3800,304.00,160.00,193.00, 0.548, 0.548,-6.810,-5.040,-4.150, -6.000,3.400,-1.900,-10.00, -2.640,4.9000,0.0190,0.2500,27.100,1.9200,0.4700,0.8500,2.1600, 0.5930,600.00},
 {"AlP",  5.4672,133.00, 63.00, 61.50, 0.014, 0.014,-2.020,-2.760,-1.430, -5.700,3.000,-1.500,-4.600, -0.940,3.6300,0.0700,0.2200,17.700,3.3500,0.7100,1.2300,2.7500, 0.5771,372.00},
 {"AlAs", 5.6611,125.00, 53.40, 54.20,-0.055,-0.055,-1.610,-2.590,-1.320, -5.640,2.470,-2.300,-3.400, -0.530,3.0990,0.2800,0.1500,21.100,3.7600,0.8200,1.4200,3.0000, 0.8850,530.00},
 {"AlSb", 6.1355, 87.69, 43.41, 40.76,-0.094,-0.094,-0.760,-1.990,-0.820, -4.500,1.400,-1.350,-4.300,  0.390,2.3860,0.6760,0.1400,18.700,5.1800,1.1900,1.9700,3.1900, 0.4200,140.00},
 {"AlBi", 6.3417, 72.20, 36.20, 33.40,-0.094,-0.094,-0.760,-1.990,-0.820, -4.500,1.400,-1.350,-4.300,  1.100,0.042 ,2.1240,99.999,99.999,99.999,99.999,99.999,3.1900, 0.4200,140.00},
 {"GaN",  4.5000,293.00,159.00,155.00, 0.366, 0.366,-5.380,-6.730,-3.180, -2.200,5.200,-2.200,-3.400, -1.840,3.2990,0.0170,0.1500,25.000,2.6700,0.7500,1.1000,2.4000, 0.5930,600.00},
 {"GaP",  5.4505,140.50, 62.03, 70.33,-0.121,-0.121,-1.230,-3.270,-1.380, -8.200,1.700,-1.600,-4.600, -0.470,2.8860,0.0800,0.1300,31.400,4.0500,0.4900,2.9300,2.9000, 0.0000,  0.00},
 {"GaAs", 5.6503,118.8 , 53.8 , 59.4 ,-0.160,-0.230,-0.439,-3.765,-0.492, -8.013,0.220,-1.824,-5.062,  0.000,1.5180,0.3400,0.0667,25.7  ,7.1   ,2.0200,2.9100,3.6304, 0.5405,204.00},
 {"GaSb", 6.0959, 88.42, 40.26, 43.22,-0.216,-0.216,-0.310,-2.770,-0.700, -7.500,0.800,-2.000,-4.700,  0.770,0.8120,0.7600,0.0390,27.000,13.400,4.7000,6.0000,3.7900, 0.4170,140.00},
 {"GaBi", 6.3240, 73.00, 32.70, 36.30,-0.216,-0.216,-0.310,-2.770,-0.700, -7.500,0.800,-2.000,-4.700,  1.100,-1.450,2.1500,99.999,99.999,99.999,99.999,99.999,3.7900, 0.4170,140.00},
 {"InN",  4.9800,187.00,125.00, 86.00, 0.593, 0.593,-5.960,-6.320,-2.000, -1.850,1.500,-1.200,-9.300, -1.580,1.9400,0.0060,0.1200,25.000,3.7200,1.2600,1.6300,2.5900, 0.2450,624.00},
 {"InP",  5.8697,101.10, 56.10, 45.60, 0.016, 0.016,-1.540,-3.620,-1.020, -6.000,0.600,-2.000,-5.000, -0.140,1.4240,0.1080,0.0795,20.700,5.0800,1.6000,2.1000,3.1000, 0.3630,162.00},
 {"InAs", 6.0553, 83.3 , 45.3 , 39.6 ,-0.045,-0.115,-0.531,-4.076,-0.120, -5.080,1.000,-1.800,-3.600,  0.173,0.4050,0.3800,0.02226,21.846,19.67,8.4000,9.3000,3.8210, 0.2760, 93.00},
 {"InSb", 6.4794, 68.47, 37.35, 31.11,-0.161,-0.161,-0.620,-4.040,-0.160, -6.940,0.360,-2.000,-4.700,  0.800,0.2350,0.8100,0.0135,23.300,34.800,15.000,16.500,3.9500, 0.4200,170.00},
 {"InBi", 6.6860, 58.50, 30.40, 27.50,-0.161,-0.161,-0.620,-4.040,-0.160, -6.940,0.360,-2.000,-4.700,  1.100,-1.630,2.1500,99.999,99.999,99.999,99.999,99.999,3.9500, 0.4200,170.00},
};
// clang-format on

// Ternary bowing of the gap / spin-orbit splitting, indexed as the 18 ternaries below.
const double kBowEg[18] = {0.000, 0.528, 0.566, 0.650, 0.477, 0.415, -0.48, 0.700, 0.430,
                           0.220, 2.700, 0.800, 0.190, 2.700, 1.430, 0.100, 1.900, 0.670};
const double kBowDso[18] = {0.000, 0.000, 0.300, 0.000, 0.150, 0.100, -0.19, 0.150, 0.250,
                            0.000, 0.000, 0.150, 0.000, 0.000, 0.600, 0.160, 0.750, 1.200};
// The two end-point binaries of each ternary: AlGaP AlGaAs AlGaSb GaInP GaInAs GaInSb AlInP
// AlInAs AlInSb AlPAs AlPSb AlAsSb GaPAs GaPSb GaAsSb InPAs InPSb InAsSb.
const int kTernEnds[18][2] = {{1, 6}, {2, 7}, {3, 8}, {6, 11}, {7, 12}, {8, 13}, {1, 11}, {2, 12},
                              {3, 13}, {1, 2}, {1, 3}, {2, 3}, {6, 7}, {6, 8}, {7, 8}, {11, 12},
                              {11, 13}, {12, 13}};

// Alloy system: form 0 = A(x)B(1-x)C(y)D(1-y), 1 = A(x)B(y)C(1-x-y)D, 2 = AB(x)C(y)D(1-x-y).
struct System {
    int form;
    int bin[4];   // form 0: AC AD BC BD; forms 1/2: three binaries
    int tern[4];  // form 0: ABC ABD ACD BCD; forms 1/2: three ternaries
};
const System kSystems[15] = {
    {0, {1, 2, 6, 7}, {0, 1, 9, 12}},     {0, {1, 3, 6, 8}, {0, 2, 10, 13}},
    {0, {2, 3, 7, 8}, {1, 2, 11, 14}},    {0, {6, 7, 11, 12}, {3, 4, 12, 15}},
    {0, {6, 8, 11, 13}, {3, 5, 13, 16}},  {0, {7, 8, 12, 13}, {4, 5, 14, 17}},
    {0, {1, 2, 11, 12}, {6, 7, 9, 15}},   {0, {1, 3, 11, 13}, {6, 8, 10, 16}},
    {0, {2, 3, 12, 13}, {7, 8, 11, 17}},  {1, {1, 6, 11, -1}, {0, 6, 3, -1}},
    {1, {2, 7, 12, -1}, {1, 2, 4, -1}},   {1, {3, 8, 13, -1}, {2, 8, 5, -1}},
    {2, {1, 2, 3, -1}, {9, 10, 11, -1}},  {2, {6, 7, 8, -1}, {12, 13, 14, -1}},
    {2, {11, 12, 13, -1}, {15, 16, 17, -1}},
};

// Composition weights of the alloy in `sys` at (x, y).
double mix_linear(const System &sys, double x, double y, const double *P) {
    if (sys.form == 0)
        return x * y * P[0] + x * (1 - y) * P[1] + (1 - x) * y * P[2] + (1 - x) * (1 - y) * P[3];
    return x * P[0] + y * P[1] + (1 - x - y) * P[2];
}

// Interpolation with ternary bowings B[0..3] and quaternary bowing Bq.
double mix_bowed(const System &sys, double x, double y, const double *P, const double *B, double Bq) {
    if (sys.form == 0) {
        // exact binaries (the general formula divides by x(1-x)+y(1-y))
        if (std::fabs(x - 1) < kEps && std::fabs(y - 1) < kEps) return P[0];
        if (std::fabs(x - 1) < kEps && std::fabs(y) < kEps) return P[1];
        if (std::fabs(x) < kEps && std::fabs(y - 1) < kEps) return P[2];
        if (std::fabs(x) < kEps && std::fabs(y) < kEps) return P[3];
        const double ux = x * (1 - x), uy = y * (1 - y);
        const double tABC = x * P[0] + (1 - x) * P[2] - ux * B[0];
        const double tABD = x * P[1] + (1 - x) * P[3] - ux * B[1];
        const double tACD = y * P[0] + (1 - y) * P[1] - uy * B[2];
        const double tBCD = y * P[2] + (1 - y) * P[3] - uy * B[3];
        double r = ux * (y * tABC + (1 - y) * tABD) + uy * (x * tACD + (1 - x) * tBCD);
        r /= ux + uy;
        return r - ux * uy * Bq;
    }
    const double z = 1 - x - y;
    return x * P[0] + y * P[1] + z * P[2] - x * y * B[0] - x * z * B[1] - y * z * B[2] - x * y * z * Bq;
}

// d/dx of mix_linear.
double dmix_linear(const System &sys, double y, const double *P) {
    if (sys.form == 0) return y * P[0] + (1 - y) * P[1] - y * P[2] - (1 - y) * P[3];
    return P[0] - P[2];
}

// d/dx of mix_bowed, the derivative of the smooth interpolant (at the exact binaries, where the
// value is the tabulated compound, its one-sided limit).
double dmix_bowed(const System &sys, double x, double y, const double *P, const double *B, double Bq) {
    if (sys.form == 0) {
        const double dux = 1 - 2 * x;
        if (std::fabs(y) < kEps) return P[1] - P[3] - dux * B[1];          // r = tABD
        if (std::fabs(y - 1) < kEps) return P[0] - P[2] - dux * B[0];      // r = tABC
        const double ux = x * (1 - x), uy = y * (1 - y);
        const double tABC = x * P[0] + (1 - x) * P[2] - ux * B[0];
        const double tABD = x * P[1] + (1 - x) * P[3] - ux * B[1];
        const double tACD = y * P[0] + (1 - y) * P[1] - uy * B[2];
        const double tBCD = y * P[2] + (1 - y) * P[3] - uy * B[3];
        const double dtABC = P[0] - P[2] - dux * B[0], dtABD = P[1] - P[3] - dux * B[1];
        const double N = ux * (y * tABC + (1 - y) * tABD) + uy * (x * tACD + (1 - x) * tBCD);
        const double dN = dux * (y * tABC + (1 - y) * tABD) + ux * (y * dtABC + (1 - y) * dtABD) + uy * (tACD - tBCD);
        const double D = ux + uy, dD = dux;
        return (dN * D - N * dD) / (D * D) - dux * uy * Bq;
    }
    const double z = 1 - x - y;
    return P[0] - P[2] - y * B[0] - (z - x) * B[1] + y * B[2] - (y * z - x * y) * Bq;
}

}  // namespace

Material material_dx(const qd_inputs &in, int s) {
    if (s < 1 || s >= in.N_shape) throw std::runtime_error("composition derivative needs a shape index >= 1");
    const std::vector<Material> mp = interpolate_materials(in);
    double sc[15], g1[15], g2[15], g3[15], Eg[15];
    for (int i = 0; i < 15; ++i) {
        const Binary &B = kBinaries[i];
        sc[i] = 1.0 / B.mc - (B.Ep / 3.0) * (2.0 / B.Eg0 + 1.0 / (B.Eg0 + B.Dso));
        g1[i] = B.gL1 - B.Ep / (3.0 * B.Eg0 + B.Dso);
        g2[i] = B.gL2 - B.Ep / (6.0 * B.Eg0 + 2.0 * B.Dso);
        g3[i] = B.gL3 - B.Ep / (6.0 * B.Eg0 + 2.0 * B.Dso);
        Eg[i] = B.Eg0;
        if (in.T > 0.001) {
            if (i != 6) Eg[i] -= B.varshni_alpha_meV * 0.001 * in.T * in.T / (in.T + B.varshni_beta);
            else Eg[i] += 0.1081 * (1.0 - std::cosh(164.0 / in.T) / std::sinh(164.0 / in.T));
        }
    }
    const qd_shape &sh = in.shapes[s];
    const System &sys = kSystems[sh.material];
    const int nb = sys.form == 0 ? 4 : 3;
    auto lin = [&](auto getter) {
        double P[4] = {0, 0, 0, 0};
        for (int k = 0; k < nb; ++k) P[k] = getter(sys.bin[k]);
        return dmix_linear(sys, sh.y, P);
    };
    Material d{};
    d.a = lin([](int i) { return kBinaries[i].a; });
    d.C11 = lin([](int i) { return kBinaries[i].C11; });
    d.C12 = lin([](int i) { return kBinaries[i].C12; });
    d.C44 = lin([](int i) { return kBinaries[i].C44; });
    d.e14 = in.piezo == 1 ? lin([](int i) { return kBinaries[i].e14; }) : lin([](int i) { return kBinaries[i].e14_calc; });
    d.B114 = lin([](int i) { return kBinaries[i].B114; });
    d.B124 = lin([](int i) { return kBinaries[i].B124; });
    d.B156 = lin([](int i) { return kBinaries[i].B156; });
    d.ac = lin([](int i) { return kBinaries[i].ac; });
    d.av = lin([](int i) { return kBinaries[i].av; });
    d.b = lin([](int i) { return kBinaries[i].b; });
    d.d = lin([](int i) { return kBinaries[i].d; });
    d.nr = lin([](int i) { return kBinaries[i].nr; });
    d.sc = lin([&](int i) { return sc[i]; });
    d.Ep = lin([](int i) { return kBinaries[i].Ep; });
    d.P = 0.25 * kHbsqm * d.Ep / mp[s].P;                  // P = sqrt(hbar^2/m0 Ep / 2)
    d.g1 = lin([&](int i) { return g1[i]; });
    d.g2 = lin([&](int i) { return g2[i]; });
    d.g3 = lin([&](int i) { return g3[i]; });
    double bowEv[18];
    for (int t = 0; t < 18; ++t) {
        const Binary &A = kBinaries[kTernEnds[t][0]], &B = kBinaries[kTernEnds[t][1]];
        bowEv[t] = (3.0 / mp[0].a) * (A.av - B.av) * (A.a - B.a);
    }
    double Pev[4] = {0}, Peg[4] = {0}, Pso[4] = {0}, Bev[4] = {0}, Beg[4] = {0}, Bso[4] = {0};
    for (int k = 0; k < nb; ++k) {
        Pev[k] = kBinaries[sys.bin[k]].Ev_avg;
        Peg[k] = Eg[sys.bin[k]];
        Pso[k] = kBinaries[sys.bin[k]].Dso;
        Bev[k] = bowEv[sys.tern[k]];
        Beg[k] = kBowEg[sys.tern[k]];
        Bso[k] = kBowDso[sys.tern[k]];
    }
    d.Ev_avg = dmix_bowed(sys, sh.x, sh.y, Pev, Bev, 0.0);
    d.Eg = dmix_bowed(sys, sh.x, sh.y, Peg, Beg, 0.0);
    d.Dso = dmix_bowed(sys, sh.x, sh.y, Pso, Bso, 0.0);
    d.E_hh = d.Ev_avg + d.Dso / 3.0;     // the energy zero belongs to the matrix (shape 0)
    d.E_lh = d.E_hh;
    d.E_so = d.E_hh - d.Dso;
    d.E_cb = d.E_hh + d.Eg;
    return d;
}

std::vector<Material> interpolate_materials(const qd_inputs &in) {
    const int ns = in.N_shape;
    // Per-binary derived quantities: renormalised inverse mass / Luttinger parameters from the
    // T = 0 gaps, then the Varshni-shifted gaps.
    double sc[15], g1[15], g2[15], g3[15], Eg[15];
    for (int i = 0; i < 15; ++i) {
        const Binary &B = kBinaries[i];
        sc[i] = 1.0 / B.mc - (B.Ep / 3.0) * (2.0 / B.Eg0 + 1.0 / (B.Eg0 + B.Dso));
        g1[i] = B.gL1 - B.Ep / (3.0 * B.Eg0 + B.Dso);
        g2[i] = B.gL2 - B.Ep / (6.0 * B.Eg0 + 2.0 * B.Dso);
        g3[i] = B.gL3 - B.Ep / (6.0 * B.Eg0 + 2.0 * B.Dso);
        Eg[i] = B.Eg0;
        if (in.T > 0.001) {
            if (i != 6) {
                const double aT = B.varshni_alpha_meV * 0.001;
                Eg[i] -= aT * in.T * in.T / (in.T + B.varshni_beta);
            } else {  // GaP: the tabulated non-Varshni form
                Eg[i] += 0.1081 * (1.0 - std::cosh(164.0 / in.T) / std::sinh(164.0 / in.T));
            }
        }
    }
    std::vector<Material> out(ns);
    for (int s = 0; s < ns; ++s) {
        const qd_shape &sh = in.shapes[s];
        if (sh.material < 0 || sh.material > 14)
            throw std::runtime_error("Requested material system not recognised");
        const System &sys = kSystems[sh.material];
        const int nb = sys.form == 0 ? 4 : 3;
        auto lin = [&](auto getter) {
            double P[4] = {0, 0, 0, 0};
            for (int k = 0; k < nb; ++k) P[k] = getter(sys.bin[k]);
            return mix_linear(sys, sh.x, sh.y, P);
        };
        Material &m = out[s];
        m.a = lin([](int i) { return kBinaries[i].a; });
        m.C11 = lin([](int i) { return kBinaries[i].C11; });
        m.C12 = lin([](int i) { return kBinaries[i].C12; });
        m.C44 = lin([](int i) { return kBinaries[i].C44; });
        if (in.piezo == 1)
            m.e14 = lin([](int i) { return kBinaries[i].e14; });
        else
            m.e14 = lin([](int i) { return kBinaries[i].e14_calc; });
        m.B114 = lin([](int i) { return kBinaries[i].B114; });
        m.B124 = lin([](int i) { return kBinaries[i].B124; });
        m.B156 = lin([](int i) { return kBinaries[i].B156; });
        m.ac = lin([](int i) { return kBinaries[i].ac; });
        m.av = lin([](int i) { return kBinaries[i].av; });
        m.b = lin([](int i) { return kBinaries[i].b; });
        m.d = lin([](int i) { return kBinaries[i].d; });
        m.nr = lin([](int i) { return kBinaries[i].nr; });
        m.sc = lin([&](int i) { return sc[i]; });
        m.Ep = lin([](int i) { return kBinaries[i].Ep; });
        m.P = std::sqrt(0.5 * kHbsqm * m.Ep);
        m.g1 = lin([&](int i) { return g1[i]; });
        m.g2 = lin([&](int i) { return g2[i]; });
        m.g3 = lin([&](int i) { return g3[i]; });
    }
    // Valence-band-average bowing, referenced to the matrix lattice constant.
    double bowEv[18];
    for (int t = 0; t < 18; ++t) {
        const Binary &A = kBinaries[kTernEnds[t][0]], &B = kBinaries[kTernEnds[t][1]];
        bowEv[t] = (3.0 / out[0].a) * (A.av - B.av) * (A.a - B.a);
    }
    for (int s = 0; s < ns; ++s) {
        const qd_shape &sh = in.shapes[s];
        const System &sys = kSystems[sh.material];
        const int nb = sys.form == 0 ? 4 : 3;
        double Pev[4] = {0}, Peg[4] = {0}, Pso[4] = {0}, Bev[4] = {0}, Beg[4] = {0}, Bso[4] = {0};
        for (int k = 0; k < nb; ++k) {
            Pev[k] = kBinaries[sys.bin[k]].Ev_avg;
            Peg[k] = Eg[sys.bin[k]];
            Pso[k] = kBinaries[sys.bin[k]].Dso;
            Bev[k] = bowEv[sys.tern[k]];
            Beg[k] = kBowEg[sys.tern[k]];
            Bso[k] = kBowDso[sys.tern[k]];
        }
        Material &m = out[s];
        m.Ev_avg = mix_bowed(sys, sh.x, sh.y, Pev, Bev, 0.0);
        m.Eg = mix_bowed(sys, sh.x, sh.y, Peg, Beg, 0.0);   // quaternary bowings tabulate as 0
        m.Dso = mix_bowed(sys, sh.x, sh.y, Pso, Bso, 0.0);
    }
    // Band edges with the zero of energy at the unstrained matrix valence-band edge.
    const double ref = out[0].Ev_avg + out[0].Dso / 3.0;
    for (auto &m : out) {
        m.E_hh = m.Ev_avg + m.Dso / 3.0 - ref;
        m.E_lh = m.E_hh;
        m.E_so = m.E_hh - m.Dso;
        m.E_cb = m.E_hh + m.Eg;
    }
    return out;
}

}  // namespace qd
