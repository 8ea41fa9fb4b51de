// Host-side (CPU, FP64) pieces of the QUDOS single-particle solve: input parsing, material
// interpolation, shape Fourier transforms, strain / piezoelectric fields and the Delta-G tables
// consumed by the MI355X Hamiltonian kernel.  Everything here is O(grid) or O(grid * shapes);
// the O(N_H^2) and O(N_H^3) work runs on the GPU.
#pragma once

#include <complex>
#include <string>
#include <vector>
#include <functional>

#include "qudos_gpu.h"

namespace qd {

using cplx = std::complex<double>;

constexpr double kHbsqm = 7.61996;   // hbar^2 / m0 [eV A^2]      (8_band.h:51)
constexpr double kEsqeps = 180.9513; // e^2 / eps0  [eV A]        (8_band.h:52)
constexpr double kEps = 1.0e-8;      // equality threshold used by the branch tests (8_band.h:59)

void set_error(const std::string &msg);

// Per-shape material parameters (struct material_params, 8_band.h:200).
struct Material {
    double a, C11, C12, C44, e14, B114, B124, B156, Eg, Dso, Ep, P, sc, g1, g2, g3;
    double ac, av, b, d, Ev_avg, E_so, E_lh, E_hh, E_cb, nr;
};

// Throws std::runtime_error on an unknown material system.
std::vector<Material> interpolate_materials(const qd_inputs &in);
// d Material(shape s) / d x_s for a shape s >= 1 (nothing else depends on x_s): linear and bowed
// interpolants differentiated in closed form, the renormalised binaries being constants.
Material material_dx(const qd_inputs &in, int s);

// Reciprocal-lattice vectors G_a[i] = (2 pi / L)(i - (NT-1)/2), i < NT = 2N-1 (basis_set.c:23-90).
std::vector<double> reciprocal_axis(int N, double L);

// Which supercell normalises a shape transform (characteristic_function_fourier.c:72-91).
enum class Cell { Electronic = 0, Strain = 1, Coulomb = 3 };

// chi_s(G) = (1/Omega) int_s exp(-i G.r) d^3r for shape `idx` of `in`.
cplx shape_transform(const qd_inputs &in, int idx, double Gx, double Gy, double Gz, Cell cell);

// d chi_s(G) / d theta for a geometry parameter of shape `idx` (QD_DP_L1, L2, L3, L12, OX, OY, OZ),
// exact (forward-mode dual numbers through the same formulas and quadrature nodes).
cplx shape_transform_deriv(const qd_inputs &in, int idx, double Gx, double Gy, double Gz, Cell cell, int param);
void shape_grid_deriv(const qd_inputs &in, int idx, const std::vector<double> &gx, const std::vector<double> &gy,
                      const std::vector<double> &gz, Cell cell, int param, std::vector<cplx> &out);

// The quadrature of the shape transforms' 1-D integrals: the 40- and 64-point Gauss-Legendre
// rules on [-1, 1] whose difference controls the interval halving, as x40, w40, x64, w64.
std::vector<double> gauss_legendre_rules();
constexpr int kGLLo = 40, kGLHi = 64;

// Fill chi_s on a 3-D grid spanned by the three axes (row-major x, y, z).
void shape_grid(const qd_inputs &in, int idx, const std::vector<double> &gx,
                const std::vector<double> &gy, const std::vector<double> &gz, Cell cell,
                std::vector<cplx> &out);

// Simple thread pool loop: body(i) for i in [0, n), split across hardware threads.
void parallel_for(long n, const std::function<void(long, long)> &body);
// n items of about `cost` arithmetic operations each
void parallel_for_cost(long n, double cost, const std::function<void(long, long)> &body);

// 3-D complex array on an odd grid with row-major (x, y, z) layout.
struct Grid3 {
    int n[3] = {0, 0, 0};
    std::vector<cplx> v;
    Grid3() = default;
    Grid3(int a, int b, int c) : v((size_t)a * b * c) { n[0] = a; n[1] = b; n[2] = c; }
    size_t size() const { return v.size(); }
    cplx &operator()(int i, int j, int k) { return v[((size_t)i * n[1] + j) * n[2] + k]; }
    const cplx &operator()(int i, int j, int k) const { return v[((size_t)i * n[1] + j) * n[2] + k]; }
};

// Truncated linear convolution of two grids of equal odd shape: out[k] = sum_{i+j=k+c} a[i] b[j]
// with c the centre index, i.e. the central block of the zero-padded (2n-1)^3 product.
class Convolver {
  public:
    explicit Convolver(const int n[3]);
    // spectrum of a zero-padded grid (size (2n-1)^3; a may have up to 2n-1 points per axis)
    void forward(const Grid3 &a, std::vector<cplx> &spec) const;
    // central n^3 block of the inverse transform of spec
    void inverse_central(const std::vector<cplx> &spec, Grid3 &out) const;
    // points start[a] + step[a] k, k < nout[a], of the inverse transform of spec
    void inverse_at(const std::vector<cplx> &spec, const int nout[3], const int start[3], const int step[3],
                    Grid3 &out) const;
    const int *padded() const { return m_; }
    Grid3 convolve(const Grid3 &a, const Grid3 &b) const;
    size_t spec_size() const { return (size_t)m_[0] * m_[1] * m_[2]; }

  private:
    int n_[3], m_[3];
    std::vector<cplx> tw_[3];   // exp(-2 pi i k / m) tables
};

}  // namespace qd

namespace qd {

// Intermediates of the table builder that its reverse pass needs.
struct TableState {
    std::vector<Material> mp;
    bool use_strain = false, iso = false, same = false, have_phi = false, second = false;
    bool commensurate = false;            // strain cell = RL x the electronic cell (FFT downsampling)
    int NT[3] = {0, 0, 0}, nS[3] = {0, 0, 0}, RL[3] = {1, 1, 1};
    double Lc[3] = {0, 0, 0};
    std::vector<double> gE[3], gS[3];
    std::vector<Grid3> chiE, chiS;        // per shape (index 0 unused); chiE == chiS when the cells agree
    std::vector<double> br[6];            // strain brackets on the strain grid
    Grid3 eS[6], eE[6];                   // strain on the strain / electronic grids
    Grid3 prod[12];                       // second-order strain products (piezo 2, 3)
    Grid3 phiS, phiE;                     // piezoelectric potential
    Grid3 chi0S;                          // the electronic cell on the strain grid (commensurate)
};

// Gradients of L = Re sum W T with respect to the builder's inputs.
struct TableGrad {
    std::vector<Grid3> gchiE, gchiS;      // dL/dchi per shape (gchiS empty when the cells agree)
    std::vector<Material> gmp;            // dL/d(material field) per shape
    double gF[3] = {0, 0, 0};             // dL/dF (per V/Angstrom)
    // dL/dO_x, dL/dO_y, dL/dO_z per shape (3 per shape, index 0 unused), when the pass computed them
    // from the centre's phase (d chi_s / d O_a = -i G_a chi_s); empty otherwise
    std::vector<double> gcentre;
    // dL/dtheta_j of the requested shape-length parameters when the pass contracted them (NaN for
    // the others); empty otherwise
    std::vector<double> gparam;
};

// Materials, cell flags, reciprocal axes and the shape transforms chi_s of every grid the fields
// need (the first part of build_state, which the device builder shares).  Shapes whose type
// on_device accepts are left out (empty grids): the device builder evaluates those itself.
void build_geometry(const qd_inputs &in, TableState &st, const std::function<bool(int)> &on_device = nullptr);
void build_state(const qd_inputs &in, TableState &st);
void compose_tables(const qd_inputs &in, const TableState &st, qd_tables &t);
void tables_vjp(const qd_inputs &in, const TableState &st, const cplx *W, TableGrad &out);
// The argument checks of qd_tables_vjp (QD_OK or the status, with the message set).
int vjp_check_params(const qd_inputs *in, const double *W, int nparams, const int *params, const unsigned *masks,
                     const double *grad);
// grad[j] = dL/dtheta_j from the builder's gradients: the field directly, compositions through the
// materials' derivatives, shape geometry through the exact d chi_s / d theta (dual numbers).
void vjp_contract(const qd_inputs &in, const TableState &st, const TableGrad &g, int nparams, const int *params,
                  const unsigned *masks, double *grad);

}  // namespace qd
