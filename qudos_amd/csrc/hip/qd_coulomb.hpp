// Device entry of the Coulomb matrix elements (coulomb.hip).
#pragma once

#include <vector>

#include "qd_coulomb_host.hpp"
#include "qd_internal.hpp"

namespace qd {

// X: ldx x S complex FP64 coefficient vectors on the device (holes first).  Returns the P x R
// contraction matrix (row-major) described in qd_coulomb_host.hpp, the FFT grid and stage times.
void coulomb_device(const c64 *X, long ldx, int S, const CoulombSetup &cs, hipStream_t s, std::vector<zc> &C,
                    int &P, int &R, int fft_len[3], double *t_pairs, double *t_contract);

}  // namespace qd
