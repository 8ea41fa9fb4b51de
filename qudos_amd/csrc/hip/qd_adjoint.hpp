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

// Lower band blocks (br >= bc) of H with the tables they reference and their weight.
struct AdjointBlocks {
    int n;
    int br[kAdjMaxBlocks], bc[kAdjMaxBlocks];
    unsigned mask[kAdjMaxBlocks];
    double w[kAdjMaxBlocks];
};

AdjointBlocks adjoint_blocks();
void launch_level_density(const c64 *X, long ld, const int *cols, int ncols, const int N[3], c64 *rho, hipStream_t s);
void launch_phi_apply(const c64 *X, long ld, const int *cols, const int *which, int ncols, const c64 *phis, long nT,
                      double scale, const int N[3], c64 *G, long ldg, hipStream_t s);
void launch_adjoint_bin(const HGeom &geo, const c64 *V, long ldv, const AdjointTerms &terms,
                        const AdjointBlocks &blocks, c64 *W, long nT, hipStream_t s);

}  // namespace qd
