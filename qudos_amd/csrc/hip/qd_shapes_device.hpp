// Shape transforms on the device (shapes.hip) for the shapes with closed-form transforms.
#pragma once

#include "qd_device.hpp"
#include "qd_host.hpp"

namespace qd {

// One shape on one cell, resolved on the host: centre, lengths (Angstrom), the Euler rotation of
// the wave vector, 1 / cell volume, 0.5 sigma^2 of the softening (0 without), the side count of
// the regular pyramids and prisms, and the box lengths of the cuboid, quantum well and wire with
// the length each of them is (lbmap: index into L, -1 for a cell length).
struct ShapeDev {
    int kind, nside;
    int lbmap[3];
    double O[3], L[4], Lb[3];
    double rot[9];
    double inv_vol, soft;
};

// true for the shape types with a device form (all of 0-18)
bool shape_on_device(int kind);
// the shapes whose transforms need 1-D integrals (3, 4, 11-15): their own kernel instances
bool shape_needs_quadrature(int kind);
// the kernel instance a shape type runs in: -1 for the closed forms, else its own type
int shape_family(int kind);
// Throws std::runtime_error on the inputs the host transform rejects (side counts, directions).
ShapeDev shape_device_params(const qd_inputs &in, int idx, Cell cell);
// out[slot[j] * n0 n1 n2 + i] = chi of sd[j] at point i of g[0] x g[1] x g[2]; every sd[j] of one
// family (shape_family); gl: the uploaded gauss_legendre_rules() (device arrays)
void launch_shape_grid(const ShapeDev *sd, const int *slot, int count, int family, const double *const g[3],
                       const int n[3], const double *gl, c64 *out, hipStream_t s);
// The derivative seeds of the lengths L1..L4 for a QD_DP_L* parameter of a shape of type kind
// (false for the other parameters)
bool length_seeds(int param, int kind, double seed[4]);
// out[i] = chi of sd at G[3 i .. 3 i + 2], or with seed (host array of 4) its exact derivative along
// the seeded lengths (forward-mode dual numbers through the same formulas and quadrature nodes)
void launch_shape_points(const ShapeDev &sd, const double *G, int n, const double *gl, const double *seed, c64 *out,
                         hipStream_t s);
// partial[j * nblk + b] (b < nblk): Re sum_i grad[slot[j]][i] d chi_j[i] / d theta over work-group b's
// share of the grid g[0] x g[1] x g[2], for the shapes sd[j] of one family; seed: host array of 4
void launch_shape_deriv_dots(const ShapeDev *sd, const int *slot, int count, int family, const double *const g[3],
                             const int n[3], const double *gl, const double seed[4], const c64 *grad, double *partial,
                             int nblk, hipStream_t s);

}  // namespace qd
