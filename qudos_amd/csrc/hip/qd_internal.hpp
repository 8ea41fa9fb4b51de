// Internal declarations shared by the HIP translation units of libqdgpu.
#pragma once

#include <vector>

#include "qd_device.hpp"

namespace qd {

constexpr int kNB = 32;    // Householder panel width (columns reduced per trailing update)
constexpr int kBW = 32;    // band width of the two-stage reduction (stage-1 panel width)
constexpr int kSEG = 8;    // 64x64 tiles per hemv work-group along a column strip

// Delta-G tables resident on the device.
struct DeviceTables {
    HGeom geom{};          // G pointers are device pointers
    c64 *tab = nullptr;    // QD_NTAB x nT
    long nT = 0;
    double *G[3] = {nullptr, nullptr, nullptr};
    double *u[3] = {nullptr, nullptr, nullptr};   // G + K of the N[a] plane waves along each axis
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

// Optional timing of the reduction's dominant kernel with HIP events on its stream: the one-stage
// hemv (every kProfStride-th column, work = bytes read) or the two-stage hemm (every panel, work =
// floating-point operations).
constexpr int kProfStride = 16;
struct KernelProfile {
    bool on = false;
    std::vector<hipEvent_t> ev;      // pairs
    std::vector<double> work;        // per timed launch
    size_t used = 0;
    void begin(hipStream_t s) {
        if (used + 2 > ev.size()) {
            for (int q = 0; q < 256; ++q) {
                hipEvent_t e;
                QD_HIP(hipEventCreate(&e));
                ev.push_back(e);
            }
        }
        QD_HIP(hipEventRecord(ev[used], s));
    }
    void end(hipStream_t s, double units) {
        QD_HIP(hipEventRecord(ev[used + 1], s));
        work.push_back(units);
        used += 2;
    }
};

template <class T> void reduce_alloc(ReduceWork<T> &w, int n, cpx<T> *A_external);
template <class T> double mfma_selftest();
void mfma_raw64(const double *a, const double *b, double *out);
template <class T> void reduce_free(ReduceWork<T> &w);
// Householder tridiagonalisation Q^H A Q = T (LAPACK zhetrd, uplo 'L', semantics).
template <class T> void reduce_run(ReduceWork<T> &w, hipStream_t s, KernelProfile *prof = nullptr);

// ---- two-stage reduction, stage 1: dense -> band (bandwidth kBW) ----
template <class T>
struct BandWork {
    int n = 0, ntb = 0, maxS = 0;
    long ld = 0;
    cpx<T> *A = nullptr;      // matrix; ends holding the band and, below it, the panel reflectors
    cpx<T> *V = nullptr;      // 2 x ld x kBW explicit reflectors (panel p in buffer p % 2)
    cpx<T> *X = nullptr;      // ld x kBW: A22 V T, then W
    cpx<T> *Xpart = nullptr;  // maxS x ld x kBW column-segment partials of A22 V
    c64 *tau = nullptr;       // n reflector scalars (0 where a column has no reflector)
    c64 *Tp = nullptr;        // kBW x kBW triangular factor of the current panel
    c64 *part = nullptr;      // per work-group partial sums of the panel QR
    c64 *pm = nullptr;        // per row block partials of V^H X
    c64 *piv = nullptr;       // pivot row of the panel
    c64 *sc = nullptr;        // reflector scalars and v^H a_k of the last panel column
    c64 *M = nullptr;         // kBW x kBW
    c64 *gpart = nullptr;     // per work-group Gram partials V^H V of the panel
    c64 *Tall = nullptr;      // npan_max x kBW x kBW: T factor of every panel (Tp points into it)
    c64 *qpart = nullptr;     // 256 x kBW x 64 partials of the Q1 application
    c64 *qy = nullptr;        // kBW x 64 (T Y of the Q1 application)
    int npan = 0, npan_max = 0;
    int coop_max = 0;         // largest co-resident grid of the persistent panel kernel
    int q1_max = 0;           // largest co-resident grid of the persistent Q1 application
    int *gbar = nullptr;      // grid-barrier state of the persistent kernels
    c64 *gsum = nullptr;      // kBW x kBW Gram matrix of the panel being factorised
    hipStream_t s2 = nullptr;                          // look-ahead panel factorisation
    hipEvent_t ev_strip = nullptr, ev_panel = nullptr; // next panel's columns updated / next panel factorised
};
template <class T> void band_alloc(BandWork<T> &w, int n, cpx<T> *A_external);
template <class T> void band_free(BandWork<T> &w);
// Reduce A to band form, copy the band (FP64) into AB (n x 2 kBW) and store the unit entries of
// the reflectors in A.  sync: the ChaseWork flags (abort / error / barrier counter, zeroed by
// chase_clear_error before the reduction).
template <class T> void band_reduce(BandWork<T> &w, c64 *AB, int *sync, hipStream_t s, KernelProfile *prof = nullptr);
// X <- Q1 X (or Q1^H X) with the stage-1 panels and their stored T factors (persistent kernel);
// false if the device cannot co-schedule the kernel's grid (the caller then uses back_apply).
template <class T> bool band_apply_q1(BandWork<T> &w, c64 *X, long ldx, int k, bool adjoint, int *sync, hipStream_t s);

// ---- two-stage reduction, stage 2: band -> tridiagonal (FP64) ----
struct ChaseWork {
    int n = 0, nj = 0, cus = 256;
    long nrefl = 0;
    c64 *AB = nullptr;        // n x 2 kBW band storage
    c64 *V2 = nullptr;        // nrefl x kBW reflectors (unit first entry stored)
    c64 *tau2 = nullptr;      // nrefl reflector scalars
    int *sync = nullptr;      // [0] work claim counter, [1] abort, [2] sticky error, [3] grid-barrier count
    int *prog = nullptr;      // per sweep (chase) or per step (application) progress
    double *d = nullptr, *e = nullptr;
};
void chase_alloc(ChaseWork &c, int n);
void chase_free(ChaseWork &c);
void chase_clear_error(ChaseWork &c, hipStream_t s);
// band AB -> d, e and the reflectors
void chase_run(ChaseWork &c, hipStream_t s);
// X <- Q2 X (adjoint = false) or Q2^H X; X is n x k complex FP64 with leading dimension ldx
void chase_apply(ChaseWork &c, c64 *X, long ldx, int k, bool adjoint, hipStream_t s);
// true if a bounded wait of the chase or of an application gave up (synchronises the stream)
bool chase_failed(ChaseWork &c, hipStream_t s);
// sets the sticky error flag that chase_failed reads, as a wait that gave up would (test support)
void chase_mark_error(ChaseWork &c, hipStream_t s);
// 64 doubles -> 4-, 8-, 16-lane group sums, lane-0 broadcast and 64-lane sum (DPP self-test)
void quad_sum_selftest(const double *in, double *out);

// Householder reflectors stored column by column in a matrix: reflector of column c has its
// unit entry at row c + off and acts on rows c + off .. n-1 (off = 1 for the one-stage
// reduction, kBW for stage 1 of the two-stage reduction).
template <class T>
struct Reflectors {
    const cpx<T> *A = nullptr;
    long ld = 0;
    int n = 0, off = 1;
    const c64 *tau = nullptr;
};

template <class T> void launch_build_lower(const DeviceTables &t, int n, long ld, cpx<T> *H, hipStream_t s);
void launch_build_full(const DeviceTables &t, int n, long ld, c64 *H, hipStream_t s);
// momentum matrix elements (momentum.hip)
size_t momentum_scratch_bytes(int n, int nc);
void momentum_device(const DeviceTables &t, int n, const c64 *V, int nv, const c64 *C, int nc, long ld,
                     unsigned char *scratch, c64 *P, hipStream_t s);
void momentum_host(const HGeom &g, const c64 *tab, long nT, int n, const c64 *V, int nv, const c64 *C, int nc,
                   long ld, c64 *P);
void momentum_order(const c64 *P, int nv, int nc, double *p);
// secondary outputs (outputs.hip)
void launch_loc_contract(const c64 *chi, const c64 *rho, long nT, int nstates, int ns, double *L, hipStream_t s);
void launch_col_dot(const c64 *A, int na, const c64 *B, int nb, long ld, int n, c64 *O, hipStream_t s);
// envelope functions of `count` columns of X at npts points (scale = 1/sqrt(cell volume))
void launch_envelope(const c64 *X, long ld, int count, const double *const u[3], const int N[3], const double *px,
                     const double *py, const double *pz, int npts, double scale, c64 *F, hipStream_t s);

// ---- tridiagonal eigenproblem (FP64) ----
struct TridiagWork {
    int n = 0, k = 0;
    double *e2 = nullptr;       // e^2
    double *w = nullptr;        // k eigenvalues
    double *Z = nullptr;        // n x k real eigenvectors of T (column-major)
    double *lu = nullptr;       // per vector LU scratch: 5 n doubles
    int *ipiv = nullptr;        // per vector pivots: n ints
    int *info = nullptr;        // 4 doubles: Gershgorin bounds, pivot floor, ||T||_1
    int *flags = nullptr;       // [0] eigenvalues whose bisection hit the round limit,
                                // [1] pairs whose residual ||(T - w) x|| / ||T||_1 exceeds 1e-11
    double *res = nullptr;      // k residuals ||(T - w) x|| / ||T||_1
};
void tridiag_alloc(TridiagWork &t, int n, int k);
void tridiag_free(TridiagWork &t);
// bisection for indices il..iu (1-based) then inverse iteration; d, e device arrays.  Limits:
// multisection rounds (40 reach rounding level from any Gershgorin interval) and inverse
// iteration sweeps (4); non-convergence is counted in t.flags.
void tridiag_solve(TridiagWork &t, const double *d, const double *e, int il, int iu, hipStream_t s,
                   int bisect_rounds = 40, int invit_sweeps = 4);

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
// Build the per-block triangular factors of the stored reflectors.
template <class T> void back_factors(BackWork<T> &b, const Reflectors<T> &r, hipStream_t s);
// X <- Q X (adjoint = false) or X <- Q^H X (adjoint = true); X is n x k complex FP64, ld = r.ld
template <class T> void back_apply(BackWork<T> &b, const Reflectors<T> &r, c64 *X, int k, bool adjoint, hipStream_t s);
template <class T> Reflectors<T> reflectors_of(const ReduceWork<T> &r) { return {r.A, r.ld, r.n, 1, r.tau}; }
template <class T> Reflectors<T> reflectors_of(const BandWork<T> &w) { return {w.A, w.ld, w.n, kBW, w.tau}; }

// real n x k (column-major, ld n) -> complex n x k with ld
void launch_real_to_complex(const double *Zr, int n, int k, c64 *X, long ld, hipStream_t s);

// ---- refinement ----
struct RefineStats {
    int iters = 0;
    double max_res = 0, max_orth = 0, max_dlambda = 0;
};

}  // namespace qd
