// Adjoint kernels (adjoint.hip): Delta-G binning of Re tr(D dH) and the e-h Coulomb terms.
#pragma once

#include "qd_internal.hpp"

namespace qd {

constexpr int kAdjMaxTerms = 64;
constexpr int kAdjMaxBlocks = 36;

// D = sum_q gamma[q] V[:, a[q]] V[:, b[q]]^H
struct AdjointTerms {
    int n;
    int a[kAdjMaxTerms], b[kAdjMaxTerms];
    double gamma[kAdjMaxTerms];
};



// Bilinear coefficients of the element formula per lower band block and table:
// coef_t(r, c) = sum_ij C[i][j] u_i(k_r) u_j(k_c), u = (1, kx, ky, kz), probed from h_lower_t.
constexpr int kBinMaxEntries = 128;
struct BinCoefs {
    int nblocks;
    int br[kAdjMaxBlocks], bc[kAdjMaxBlocks], first[kAdjMaxBlocks + 1];
    double w[kAdjMaxBlocks];
    int t[kBinMaxEntries];
    c64 C[kBinMaxEntries][4][4];
};
BinCoefs bin_coefficients();
// W (QD_NTAB x nT) by FFT correlation; u[a]: N[a] values G + K of the plane waves along axis a
// (device); cf_dev: bin_coefficients() in device memory; V: nv columns (ldv-strided).
void launch_adjoint_bin_fft(const HGeom &geo, const double *const u[3], const c64 *V, long ldv, int nv,
                            const AdjointTerms &terms, const BinCoefs *cf_dev, c64 *W, long nT, hipStream_t s);
// Host emulation of launch_adjoint_bin_fft (the kernels' per-element bodies, a host FFT) and the
// binning by its definition, for the CPU tests; host pointers throughout.
void adjoint_bin_fft_host(const HGeom &geo, const double *const u[3], const c64 *V, long ldv, int nv,
                          const AdjointTerms &terms, const BinCoefs &cf, c64 *W, long nT);
void adjoint_bin_direct_host(const HGeom &geo, const c64 *V, long ldv, const AdjointTerms &terms, c64 *W, long nT);
void launch_level_density(const c64 *X, long ld, const int *cols, int ncols, const int N[3], c64 *rho, hipStream_t s);
void launch_phi_apply(const c64 *X, long ld, const int *cols, const int *which, int ncols, const c64 *phis, long nT,
                      double scale, const int N[3], c64 *G, long ldg, hipStream_t s);

}  // namespace qd
