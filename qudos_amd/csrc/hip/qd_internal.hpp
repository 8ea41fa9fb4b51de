// Internal declarations shared by the HIP translation units of libqdgpu.
#pragma once

#include <vector>

#include "qd_device.hpp"

namespace qd {

constexpr int kNB = 32;    // Householder panel width (columns reduced per trailing update)
constexpr int kSEG = 8;    // 64x64 tiles per hemv work-group along a column strip

// Delta-G tables resident on the device.
struct DeviceTables {
    HGeom geom{};          // G pointers are device pointers
    c64 *tab = nullptr;    // QD_NTAB x nT
    long nT = 0;
    double *G[3] = {nullptr, nullptr, nullptr};
    int NH = 0;
};

// Device workspace of the tridiagonal reduction of an n x n Hermitian matrix (lower triangle
// of A, column-major, leading dimension ld = n rounded up to 64).
template <class T>
struct ReduceWork {
    int n = 0, ntb = 0, maxseg = 0;
    long ld = 0;
    cpx<T> *A = nullptr;      // matrix, overwritten by the reflectors (unit entries stored)
    cpx<T> *W = nullptr;      // ld x kNB panel of the blocked update
    double *d = nullptr;      // diagonal of T
    double *e = nullptr;      // off-diagonal of T (beta of each reflector)
    c64 *tau = nullptr;       // reflector scalars
    c64 *scal = nullptr;      // per panel column: 1/(alpha - beta)
    c64 *alpha = nullptr;     // per panel column: -tau/2 (w^H v)
    c64 *vd = nullptr, *wd = nullptr;   // V^H v and W^H v of the current column
    c64 *ka_part = nullptr;   // per work-group partials of the column update
    int *ka_cnt = nullptr;
    cpx<T> *P = nullptr;      // hemv partials, triangular (I, J) x 64
    cpx<T> *PT = nullptr;     // transposed partials, (J, seg) x 64
    c64 *part_wv = nullptr;   // per row block w^H v partials
    int *cnt = nullptr;       // per row block arrival counters
    int *done = nullptr;      // finished row blocks
    int *prefix = nullptr;    // work-group prefix over column strips
};

// Optional sampled timing of the hemv kernel (every kProfStride-th column).
constexpr int kProfStride = 16;
struct HemvProfile {
    bool on = false;
    std::vector<hipEvent_t> ev;      // pairs
    std::vector<double> bytes;       // per timed launch
    size_t used = 0;
};

template <class T> void reduce_alloc(ReduceWork<T> &w, int n, cpx<T> *A_external);
template <class T> double mfma_selftest();
void mfma_raw64(const double *a, const double *b, double *out);
template <class T> void reduce_free(ReduceWork<T> &w);
// Householder tridiagonalisation Q^H A Q = T (LAPACK zhetrd, uplo 'L', semantics).
template <class T> void reduce_run(ReduceWork<T> &w, hipStream_t s, HemvProfile *prof = nullptr);

template <class T> void launch_build_lower(const DeviceTables &t, int n, long ld, cpx<T> *H, hipStream_t s);
void launch_build_full(const DeviceTables &t, int n, long ld, c64 *H, hipStream_t s);

// ---- tridiagonal eigenproblem (FP64) ----
struct TridiagWork {
    int n = 0, k = 0;
    double *e2 = nullptr;       // e^2
    double *w = nullptr;        // k eigenvalues
    double *Z = nullptr;        // n x k real eigenvectors of T (column-major)
    double *lu = nullptr;       // per vector LU scratch: 5 n doubles
    int *ipiv = nullptr;        // per vector pivots: n ints
    int *info = nullptr;
};
void tridiag_alloc(TridiagWork &t, int n, int k);
void tridiag_free(TridiagWork &t);
// bisection for indices il..iu (1-based) then inverse iteration; d, e device arrays
void tridiag_solve(TridiagWork &t, const double *d, const double *e, int il, int iu, hipStream_t s);

// ---- back-transformation with the stored reflectors ----
template <class T>
struct BackWork {
    int n = 0, k = 0, npan = 0;
    c64 *Tm = nullptr;          // per panel kNB x kNB triangular factors (column-major)
    c64 *part = nullptr;        // partial sums
    c64 *Y = nullptr;           // kNB x k
    int *cnt = nullptr;
};
template <class T> void back_alloc(BackWork<T> &b, int n, int k);
template <class T> void back_free(BackWork<T> &b);
// Build the per-panel triangular factors (needs the reduction's A and tau).
template <class T> void back_factors(BackWork<T> &b, const ReduceWork<T> &r, hipStream_t s);
// X <- Q X (adjoint = false) or X <- Q^H X (adjoint = true); X is n x k complex FP64, ld = r.ld
template <class T> void back_apply(BackWork<T> &b, const ReduceWork<T> &r, c64 *X, int k, bool adjoint, hipStream_t s);

// real n x k (column-major, ld n) -> complex n x k with ld
void launch_real_to_complex(const double *Zr, int n, int k, c64 *X, long ld, hipStream_t s);

// ---- refinement ----
struct RefineStats {
    int iters = 0;
    double max_res = 0, max_orth = 0, max_dlambda = 0;
};

}  // namespace qd
