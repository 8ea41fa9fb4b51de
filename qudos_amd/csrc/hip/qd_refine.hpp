// Device buffers and small dense building blocks of the FP64 eigenpair refinement.
#pragma once

#include "qd_internal.hpp"

namespace qd {

struct RefineCtx {
    int n = 0, k = 0;
    long ld = 0;
    c64 *B = nullptr;      // ld x 2k basis [X | E]
    c64 *HB = nullptr;     // H B
    c64 *R = nullptr;      // ld x k residual / correction work
    c64 *T2 = nullptr;     // ld x k second work block
    c64 *small = nullptr;  // 4 small (<= 80x80) matrices
    c64 *part = nullptr;   // split-K partials
    c64 *ypart = nullptr;  // column-segment partials of H X: ld x ypart_cols
    int ypart_cols = 0;
    double *lam = nullptr; // current Ritz values (k) + scratch
    double *norms = nullptr;
    double *lu = nullptr;
    int *ipiv = nullptr;

    void alloc(int n, long ld, int k);
    void release();
    // C (ma x mb, column-major) = A^H B, A/B are ld-strided n x ma / n x mb
    void atb(const c64 *A, int ma, const c64 *B, int mb, c64 *C, hipStream_t s);
    // Y = beta Y + alpha A S
    void ab(const c64 *A, int ma, const c64 *S, int mb, double alpha, double beta, c64 *Y, hipStream_t s);
    // Y = H X (H full FP64, ld-strided)
    void gemm_h(const c64 *H, const c64 *X, int m, c64 *Y, hipStream_t s);
    void residual(const c64 *HX, const c64 *X, const double *lam, c64 *R, hipStream_t s);
    void colnorms(const c64 *X, int m, double *out, hipStream_t s);
    void trisolve(const double *d, const double *e, const double *lam, double tnorm, c64 *S, hipStream_t s);
    void chol_inv(const c64 *G, int m, c64 *Rinv, hipStream_t s);
    void jacobi_select(const c64 *M, int m, c64 *Usel, double *lam, hipStream_t s);
};

}  // namespace qd
