// Device entry of the Coulomb matrix elements (coulomb.hip).
#pragma once

#include <vector>

#include "qd_coulomb_host.hpp"
#include "qd_internal.hpp"

namespace qd {

// smallest l with 2^l >= n
int ilog2_ceil(int n);
// In-place radix-2 FFTs of ngrids contiguous row-major grids of 2^logL[0] x 2^logL[1] x 2^logL[2]
// (each axis at most 2^11 long: one line fits the kernel's LDS tile), unnormalised; tw[a]:
// exp(-2 pi i k / L_a), k < L_a / 2, on the device.  inverse = true uses the conjugate twiddles.
void fft3_device(c64 *data, long ngrids, const int logL[3], const c64 *const tw[3], bool inverse, hipStream_t s);
// out[..., i, ...] = sum_j M[i][j] in[..., j, ...] along `axis` of `batch` consecutive grids of
// dims[0] x dims[1] x dims[2]; M: m x dims[axis] row-major (device); out has m points on that axis.
void apply_axis_device(const c64 *in, const int dims[3], int axis, int m, const double *M, long batch, c64 *out,
                       hipStream_t s);

// Device scratch (bytes) of coulomb_device for S states: FFT grids of every state and band, pair
// spectra and densities, the upsampled densities, contraction operands and split-K partials.
size_t coulomb_scratch_bytes(int S, const CoulombSetup &cs);

// X: ldx x S complex FP64 coefficient vectors on the device (holes first); scratch: device memory
// of at least coulomb_scratch_bytes(S, cs) bytes (every piece is written before it is read, so it
// needs no clearing); ev: three events for the stage times.  Returns the P x R contraction matrix
// (row-major) described in qd_coulomb_host.hpp, the FFT grid, the stage times and the Madelung
// constant (0 below CME 2), whose Ewald sums run on the device with the Coulomb-cell vectors.
void coulomb_device(const c64 *X, long ldx, int S, const CoulombSetup &cs, unsigned char *scratch,
                    size_t scratch_bytes, const hipEvent_t ev[3], hipStream_t s, std::vector<zc> &C, int &P, int &R,
                    int fft_len[3], double *t_pairs, double *t_contract, double *a_mad);

}  // namespace qd
