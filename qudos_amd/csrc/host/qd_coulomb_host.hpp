// Host-side FP64 parts of the Coulomb matrix elements: the Coulomb-cell geometry (axes, Ewald
// parameters, box-transform upsampling matrices), the assembly of the four V_ijkl arrays from the
// device contraction matrix, and the text writer.
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qudos_gpu.h"

namespace qd {

using zc = std::complex<double>;

// An error that carries a QD_E_* status.
struct StatusError : std::runtime_error {
    int code;
    StatusError(const std::string &m, int c) : std::runtime_error(m), code(c) {}
};

struct CoulombSetup {
    int N[3] = {0, 0, 0}, NT[3] = {0, 0, 0};     // es-cell plane waves and 2N-1
    int NC[3] = {0, 0, 0}, NTC[3] = {0, 0, 0};   // Coulomb-cell plane waves and 2N_C-1
    bool copy[3] = {true, true, true};           // axis copied unchanged (no upsampling)
    std::vector<double> up[3];                   // NTC x NT box transform (row-major), scaled
    std::vector<double> gq[3];                   // Coulomb-grid wave vectors per axis (NTC[a])
    double CL[3] = {0, 0, 0};                    // Coulomb-cell lengths
    double a_mad = 0, pre = 0, omega_c = 0, nr0 = 1, eta = 0;
    long mad_hi = 0, mad_hy = 0, mad_hz = 0;     // real-space Ewald sum: flat indices < mad_hi
    int cme = 1;
    long nq() const { return (long)NTC[0] * NTC[1] * NTC[2]; }
    long nes() const { return (long)NT[0] * NT[1] * NT[2]; }
    bool identity() const { return copy[0] && copy[1] && copy[2]; }
};

// Packed pair index for a <= b (B_convolutions.c:427).
inline long pair_index(int a, int b) { return (long)b * (b + 1) / 2 + a; }

void coulomb_setup(const qd_inputs &in, const qd_tables &t, CoulombSetup &cs);

// Madelung constant (coulomb_matrix_elements.c:250-334) from the device's reciprocal- and
// real-space Ewald sums.
double madelung_constant(const CoulombSetup &cs, double reciprocal_sum, double real_sum);

// C: P x R row-major contraction matrix, R = 2P (+7 with dipole/quadrupole vectors):
// C[p][r] = sum_q B_p(q) Y_r(q), Y = {B(-q) w, conj(B(-q)) w, Ix, Iy, Iz, conj(I*), J}.
void coulomb_assemble(const CoulombSetup &cs, const zc *C, int P, int R, int nh, int ne, zc *hh, zc *ee,
                      zc *ed, zc *ex);

void coulomb_write(const std::string &dir, int nh, int ne, const zc *hh, const zc *ee, const zc *ed,
                   const zc *ex);

}  // namespace qd
