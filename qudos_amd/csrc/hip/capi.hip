// C ABI of the device path: handles, table upload, Hamiltonian assembly and the windowed
// eigensolve (FP64, or FP32 reduction + FP64 refinement).  Replaces eigenstates()
// (eigenstates.c:23-117) with diagonalise_hamiltonian() (diagonalise_hamiltonian.c:24-295,
// zheevr RANGE='I' semantics) and collect_eigenstates() (collect_eigenstates.c:23-75).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <functional>
#include <map>
#include <vector>

#include "qd_adjoint.hpp"
#include "qd_coulomb.hpp"
#include "qd_internal.hpp"
#include "qd_refine.hpp"
#include "qd_host.hpp"
#include "qd_shapes_device.hpp"

namespace qd {
void set_error(const std::string &msg);
}

using namespace qd;

struct qd_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceTables tab;
    bool has_tables = false;
    // current problem size
    int n = 0;
    long ld = 0;
    // matrix storage
    cpx<float> *A32 = nullptr;
    cpx<double> *A64 = nullptr;   // lower (FP64 path) -- reduced in place
    c64 *Hfull = nullptr;         // full FP64 H (refinement)
    size_t a32_elems = 0, a64_elems = 0, hf_elems = 0;
    ReduceWork<float> r32;
    ReduceWork<double> r64;
    BandWork<float> w32;          // two-stage reduction, stage 1
    BandWork<double> w64;
    ChaseWork ch;                 // two-stage reduction, stage 2
    bool two_stage = true;        // false: one-stage Householder tridiagonalisation
    BackWork<float> b32;
    BackWork<double> b64;
    TridiagWork tw;
    RefineCtx rc;
    int work_n = -1, work_k = -1, work_prec = -1;
    // results
    c64 *X = nullptr;             // ld x k eigenvectors
    double *wres = nullptr;       // k eigenvalues (device)
    int k = 0;
    hipEvent_t ev[8];
    hipEvent_t evc[3];              // Coulomb stage times
    // adjoint work (allocated on first use)
    RefineCtx adj;
    c64 *adj_H = nullptr;           // full FP64 H for the FP64 path's Sternheimer residuals
    size_t adj_h_elems = 0;
    BinCoefs *bin_cf = nullptr;     // element coefficients of the FFT binning (device)
    KernelProfile prof;
    double prof_sum_s = 0, prof_sum_b = 0;
    long prof_n = 0;
    // iteration limits (qd_debug_set_limits) and the refinement stop (qd_set_refine_tolerance)
    int lim_refine = 6, lim_stern = 12, lim_invit = 4, lim_bisect = 40;
    double refine_rtol = 1e-10;   // max_i ||H x_i - lambda_i x_i|| <= refine_rtol ||T||_1
};

namespace {

template <class F>
int guarded(F &&f) {
    (void)hipGetLastError();   // start every entry point from a clean runtime error state
    try {
        return f();
    } catch (const HipError &e) {
        set_error(e.what());
        return e.code;
    } catch (const StatusError &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return QD_E_ALLOC;
    } catch (const std::exception &e) {
        set_error(e.what());
        return QD_E_DEVICE;
    }
}

template <class P>
void ensure(P *&p, size_t &have, size_t need) {
    if (have >= need && p) return;
    if (p) hipFree(p);
    p = nullptr;
    QD_HIP(hipMalloc(&p, sizeof(P) * need));
    have = need;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    return ms * 1e-3f;
}

void free_work(qd_handle *h) {
    h->adj.release();
    if (h->adj_H) hipFree(h->adj_H);
    h->adj_H = nullptr;
    h->adj_h_elems = 0;
    if (h->work_n < 0) return;
    reduce_free(h->r32);
    reduce_free(h->r64);
    band_free(h->w32);
    band_free(h->w64);
    chase_free(h->ch);
    back_free(h->b32);
    back_free(h->b64);
    tridiag_free(h->tw);
    h->rc.release();
    if (h->X) hipFree(h->X);
    if (h->wres) hipFree(h->wres);
    h->X = nullptr;
    h->wres = nullptr;
    h->work_n = h->work_k = h->work_prec = -1;
}

void ensure_work(qd_handle *h, int n, int k, int prec) {
    if (h->work_n == n && h->work_k == k && h->work_prec == prec) return;
    free_work(h);
    h->n = n;
    h->ld = ((long)n + 63) / 64 * 64;
    const long ld = h->ld;
    if (prec == QD_FP64) {
        ensure(h->A64, h->a64_elems, (size_t)ld * ld);
        QD_HIP(hipMemset(h->A64, 0, sizeof(c64) * ld * ld));
        if (h->two_stage) band_alloc(h->w64, n, h->A64); else reduce_alloc(h->r64, n, h->A64);
        back_alloc(h->b64, n, k);
    } else {
        ensure(h->A32, h->a32_elems, (size_t)ld * ld);
        QD_HIP(hipMemset(h->A32, 0, sizeof(cpx<float>) * ld * ld));
        ensure(h->Hfull, h->hf_elems, (size_t)ld * n);
        if (h->two_stage) band_alloc(h->w32, n, h->A32); else reduce_alloc(h->r32, n, h->A32);
        back_alloc(h->b32, n, k);
        h->rc.alloc(n, ld, k);
    }
    if (h->two_stage) chase_alloc(h->ch, n);
    tridiag_alloc(h->tw, n, k);
    QD_HIP(hipMalloc(&h->X, sizeof(c64) * ld * k));
    QD_HIP(hipMalloc(&h->wres, sizeof(double) * k));
    h->work_n = n;
    h->work_k = k;
    h->work_prec = prec;
}

__global__ void k_copy_cols(const c64 *src, long lds, c64 *dst, long ldd, int n, int k) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * k) return;
    const int l = (int)(idx / n), r = (int)(idx % n);
    dst[(long)l * ldd + r] = src[(long)l * lds + r];
}

__global__ void k_axpby_cols(const c64 *src, double a, c64 *dst, double b, long ld, int n, int k) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * k) return;
    const long i = (idx / n) * ld + idx % n;
    dst[i] = a * src[i] + b * dst[i];
}

// dst = a src + b dst on n x k ld-strided blocks
void axpby_cols(const c64 *src, double a, c64 *dst, double b, long ld, int n, int k, hipStream_t s) {
    const long tot = (long)n * k;
    hipLaunchKernelGGL(k_axpby_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, a, dst, b, ld, n, k);
    QD_LAUNCH_CHECK();
}

void copy_cols(const c64 *src, c64 *dst, long ld, int n, int k, hipStream_t s) {
    const long tot = (long)n * k;
    hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, ld, dst, ld, n, k);
    QD_LAUNCH_CHECK();
}

// Tridiagonal T = Q^H H Q of the last reduction (device arrays).
const double *tri_d(const qd_handle *h) {
    if (h->two_stage) return h->ch.d;
    return h->work_prec == QD_FP64 ? h->r64.d : h->r32.d;
}
const double *tri_e(const qd_handle *h) {
    if (h->two_stage) return h->ch.e;
    return h->work_prec == QD_FP64 ? h->r64.e : h->r32.e;
}

// X <- Q X (adjoint = false) or Q^H X with the factorisation of the last reduction;
// two-stage: Q = Q1 Q2 (stage-1 panels, then the bulge-chasing reflectors).
void apply_q(qd_handle *h, c64 *X, int k, bool adjoint, hipStream_t s) {
    const bool f64 = h->work_prec == QD_FP64;
    if (h->two_stage) {
        if (!adjoint) chase_apply(h->ch, X, h->ld, k, false, s);
        const bool done = f64 ? band_apply_q1(h->w64, X, h->ld, k, adjoint, h->ch.sync, s)
                              : band_apply_q1(h->w32, X, h->ld, k, adjoint, h->ch.sync, s);
        if (!done) {
            if (f64) back_apply(h->b64, reflectors_of(h->w64), X, k, adjoint, s);
            else back_apply(h->b32, reflectors_of(h->w32), X, k, adjoint, s);
        }
        if (adjoint) chase_apply(h->ch, X, h->ld, k, true, s);
    } else {
        if (f64) back_apply(h->b64, reflectors_of(h->r64), X, k, adjoint, s);
        else back_apply(h->b32, reflectors_of(h->r32), X, k, adjoint, s);
    }
}

// FP64 refinement of (X, w) for the FP32 path; fills the statistics and returns true when every
// residual ||H x - lambda x|| is at most h->refine_rtol ||T||_1 at the end.
bool refine(qd_handle *h, int k, qd_stats *st) {
    RefineCtx &rc = h->rc;
    hipStream_t s = h->stream;
    const int n = h->n;
    c64 *Xb = rc.B, *Eb = rc.B + h->ld * k;
    c64 *HXb = rc.HB, *HEb = rc.HB + h->ld * k;
    c64 *S1 = rc.small, *S2 = rc.small + 80 * 80, *S3 = rc.small + 2 * 80 * 80;
    double tnorm = 0;
    {
        double gb[4];
        QD_HIP(hipMemcpyAsync(gb, h->tw.info, sizeof(gb), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        tnorm = gb[3];
    }
    copy_cols(h->X, Xb, h->ld, n, k, s);
    QD_HIP(hipMemcpyAsync(rc.lam, h->tw.w, sizeof(double) * k, hipMemcpyDeviceToDevice, s));
    std::vector<double> lam_old(k), lam_new(k);
    QD_HIP(hipMemcpyAsync(lam_old.data(), h->tw.w, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    int it = 0;
    double dl = 1e300;
    // converged when the Ritz values moved by < 1e-10 eV in the last sweep and every residual
    // ||H x - lambda x|| is below refine_rtol ||T||_1 (eigenvectors good to ~ that / the gap)
    const int max_it = h->lim_refine;
    const double rtol = h->refine_rtol * (tnorm > 0 ? tnorm : 1.0);
    std::vector<double> rn(k);
    // orthonormalise the back-transformed X (Cholesky QR, twice) and form HX
    for (int pass = 0; pass < 2; ++pass) {
        rc.atb(Xb, k, Xb, k, S1, s);
        rc.chol_inv(S1, k, S2, s);
        rc.ab(Xb, k, S2, k, 1.0, 0.0, rc.R, s);
        copy_cols(rc.R, Xb, h->ld, n, k, s);
    }
    rc.gemm_h(h->Hfull, Xb, k, HXb, s);
    // Rayleigh-Ritz on X alone: the shifts of the first correction become Rayleigh quotients of H.
    // The eigenvalues of T itself would make T - lambda singular to working precision, and the
    // huge component of the correction along the current vector would leave cancellation noise
    // after the projection against X.
    rc.atb(Xb, k, HXb, k, S1, s);
    rc.jacobi_select(S1, k, S2, rc.lam, s);
    rc.ab(Xb, k, S2, k, 1.0, 0.0, rc.R, s);
    rc.ab(HXb, k, S2, k, 1.0, 0.0, rc.T2, s);
    copy_cols(rc.R, Xb, h->ld, n, k, s);
    copy_cols(rc.T2, HXb, h->ld, n, k, s);
    QD_HIP(hipMemcpyAsync(lam_old.data(), rc.lam, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    // a sweep that does not halve the largest residual has reached the floor of the FP64
    // residual (measured ~2e-12 ||T||_1 after two sweeps): stop there, converged or not
    double mr_prev = 1e300;
    for (; it < max_it; ++it) {
        rc.residual(HXb, Xb, rc.lam, rc.R, s);
        if (dl < 1e-10) {
            rc.colnorms(rc.R, k, rc.norms, s);
            QD_HIP(hipMemcpyAsync(rn.data(), rc.norms, sizeof(double) * k, hipMemcpyDeviceToHost, s));
            QD_HIP(hipStreamSynchronize(s));
            double mr = 0;
            for (int i = 0; i < k; ++i) mr = std::max(mr, rn[i]);
            if (mr < rtol || mr > 0.5 * mr_prev) break;
            mr_prev = mr;
        }
        // E = Q (T - lambda)^{-1} Q^H R
        apply_q(h, rc.R, k, true, s);
        rc.trisolve(tri_d(h), tri_e(h), rc.lam, tnorm, rc.R, s);
        apply_q(h, rc.R, k, false, s);
        copy_cols(rc.R, Eb, h->ld, n, k, s);
        for (int pass = 0; pass < 2; ++pass) {   // E -= X (X^H E)
            rc.atb(Xb, k, Eb, k, S1, s);
            rc.ab(Xb, k, S1, k, -1.0, 1.0, Eb, s);
        }
        for (int pass = 0; pass < 2; ++pass) {   // E <- E R^{-1}
            rc.atb(Eb, k, Eb, k, S1, s);
            rc.chol_inv(S1, k, S2, s);
            rc.ab(Eb, k, S2, k, 1.0, 0.0, rc.R, s);
            copy_cols(rc.R, Eb, h->ld, n, k, s);
        }
        rc.gemm_h(h->Hfull, Eb, k, HEb, s);
        // Rayleigh-Ritz on [X E]
        rc.atb(rc.B, 2 * k, rc.HB, 2 * k, S3, s);
        rc.jacobi_select(S3, 2 * k, S1, rc.lam, s);
        rc.ab(rc.B, 2 * k, S1, k, 1.0, 0.0, rc.R, s);
        rc.ab(rc.HB, 2 * k, S1, k, 1.0, 0.0, rc.T2, s);
        copy_cols(rc.R, Xb, h->ld, n, k, s);
        copy_cols(rc.T2, HXb, h->ld, n, k, s);
        QD_HIP(hipMemcpyAsync(lam_new.data(), rc.lam, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        dl = 0;
        for (int i = 0; i < k; ++i) dl = std::max(dl, std::fabs(lam_new[i] - lam_old[i]));
        lam_old = lam_new;
    }
    // final residual statistics
    rc.residual(HXb, Xb, rc.lam, rc.R, s);
    rc.colnorms(rc.R, k, rc.norms, s);
    rc.atb(Xb, k, Xb, k, S1, s);
    std::vector<c64> G((size_t)k * k);
    QD_HIP(hipMemcpyAsync(rn.data(), rc.norms, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    QD_HIP(hipMemcpyAsync(G.data(), S1, sizeof(c64) * k * k, hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    double mr = 0, mo = 0;
    for (int i = 0; i < k; ++i) mr = std::max(mr, rn[i]);
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            const c64 g = G[(size_t)b * k + a];
            mo = std::max(mo, std::sqrt((g.re - (a == b)) * (g.re - (a == b)) + g.im * g.im));
        }
    copy_cols(Xb, h->X, h->ld, n, k, s);
    QD_HIP(hipMemcpyAsync(h->wres, rc.lam, sizeof(double) * k, hipMemcpyDeviceToDevice, s));
    if (st) {
        st->refine_iters = it;
        st->max_residual = mr / (tnorm > 0 ? tnorm : 1.0);
        st->max_orth = mo;
    }
    return mr <= rtol;
}

int solve_impl(qd_handle *h, int n, int il, int iu, int prec, qd_stats *st, const double *A_host) {
    const int k = iu - il + 1;
    if (il < 1 || iu > n || k < 1) { set_error("eigen window outside 1..N_H"); return QD_E_ARG; }
    if (prec == QD_FP32_REFINE && k > 40) { set_error("FP32+refine supports at most 40 eigenpairs"); return QD_E_UNSUPPORTED; }
    if (k > 64) { set_error("at most 64 eigenpairs per solve"); return QD_E_UNSUPPORTED; }
    if (prec != QD_FP64 && prec != QD_FP32_REFINE) { set_error("unknown precision"); return QD_E_ARG; }
    ensure_work(h, n, k, prec);
    hipStream_t s = h->stream;
    const long ld = h->ld;
    QD_HIP(hipEventRecord(h->ev[0], s));
    if (A_host) {
        // caller matrix (lower, column-major, ld n): widen / pad on the host, upload
        if (prec == QD_FP64) {
            std::vector<c64> tmp((size_t)ld * n, mk(0.0, 0.0));
            for (long c = 0; c < n; ++c)
                for (long r = c; r < n; ++r) tmp[c * ld + r] = mk(A_host[2 * (c * n + r)], A_host[2 * (c * n + r) + 1]);
            QD_HIP(hipMemcpyAsync(h->A64, tmp.data(), sizeof(c64) * ld * n, hipMemcpyHostToDevice, s));
            QD_HIP(hipStreamSynchronize(s));
        } else {
            std::vector<cpx<float>> t32((size_t)ld * n, mk(0.0f, 0.0f));
            std::vector<c64> full((size_t)ld * n, mk(0.0, 0.0));
            for (long c = 0; c < n; ++c)
                for (long r = c; r < n; ++r) {
                    const c64 v = mk(A_host[2 * (c * n + r)], A_host[2 * (c * n + r) + 1]);
                    t32[c * ld + r] = cast<float>(v);
                    full[c * ld + r] = v;
                    full[r * ld + c] = conj(v);
                }
            QD_HIP(hipMemcpyAsync(h->A32, t32.data(), sizeof(cpx<float>) * ld * n, hipMemcpyHostToDevice, s));
            QD_HIP(hipMemcpyAsync(h->Hfull, full.data(), sizeof(c64) * ld * n, hipMemcpyHostToDevice, s));
            QD_HIP(hipStreamSynchronize(s));
        }
    } else {
        if (prec == QD_FP64) {
            launch_build_lower<double>(h->tab, n, ld, h->A64, s);
        } else {
            launch_build_lower<float>(h->tab, n, ld, h->A32, s);
            launch_build_full(h->tab, n, ld, h->Hfull, s);
        }
    }
    QD_HIP(hipEventRecord(h->ev[1], s));
    if (h->two_stage) {
        chase_clear_error(h->ch, s);
        if (prec == QD_FP64) band_reduce(h->w64, h->ch.AB, h->ch.sync, s, &h->prof);
        else band_reduce(h->w32, h->ch.AB, h->ch.sync, s, &h->prof);
        QD_HIP(hipEventRecord(h->ev[6], s));
        chase_run(h->ch, s);
    } else {
        if (prec == QD_FP64) reduce_run(h->r64, s, &h->prof); else reduce_run(h->r32, s, &h->prof);
        QD_HIP(hipEventRecord(h->ev[6], s));
    }
    QD_HIP(hipEventRecord(h->ev[2], s));
    tridiag_solve(h->tw, tri_d(h), tri_e(h), il, iu, s, h->lim_bisect, h->lim_invit);
    QD_HIP(hipEventRecord(h->ev[3], s));
    launch_real_to_complex(h->tw.Z, n, k, h->X, ld, s);
    // T factors of the stored reflectors: the two-stage path keeps those of its panels (the
    // compact-WY blocks of back_apply are only needed when the persistent kernel cannot run)
    const bool q1_persistent = h->two_stage && (prec == QD_FP64 ? h->w64.q1_max > 0 : h->w32.q1_max > 0);
    if (!q1_persistent) {
        if (prec == QD_FP64) {
            if (h->two_stage) back_factors(h->b64, reflectors_of(h->w64), s);
            else back_factors(h->b64, reflectors_of(h->r64), s);
        } else {
            if (h->two_stage) back_factors(h->b32, reflectors_of(h->w32), s);
            else back_factors(h->b32, reflectors_of(h->r32), s);
        }
    }
    apply_q(h, h->X, k, false, s);
    QD_HIP(hipMemcpyAsync(h->wres, h->tw.w, sizeof(double) * k, hipMemcpyDeviceToDevice, s));
    QD_HIP(hipEventRecord(h->ev[4], s));
    qd_stats local{};
    bool refined = true;
    if (prec == QD_FP32_REFINE) refined = refine(h, k, &local);
    QD_HIP(hipEventRecord(h->ev[5], s));
    QD_HIP(hipEventSynchronize(h->ev[5]));
    if (h->two_stage && chase_failed(h->ch, s))
        throw HipError("band reduction: a bounded wait between work-groups gave up", QD_E_DEVICE);
    int tflags[2] = {0, 0};
    std::vector<double> tres(k);
    QD_HIP(hipMemcpy(tflags, h->tw.flags, sizeof(tflags), hipMemcpyDeviceToHost));
    QD_HIP(hipMemcpy(tres.data(), h->tw.res, sizeof(double) * k, hipMemcpyDeviceToHost));
    local.tri_residual = *std::max_element(tres.begin(), tres.end());
    h->k = k;
    if (h->prof.on && h->prof.used) {   // fold this solve's kernel timings into the totals
        for (size_t i = 0; i + 1 < h->prof.used; i += 2) {
            h->prof_sum_s += elapsed(h->prof.ev[i], h->prof.ev[i + 1]);
            h->prof_sum_b += h->prof.work[i / 2];
            ++h->prof_n;
        }
        h->prof.used = 0;
        h->prof.work.clear();
    }
    if (st) {
        *st = local;
        st->t_hamiltonian = elapsed(h->ev[0], h->ev[1]);
        st->t_reduce = elapsed(h->ev[1], h->ev[2]);
        st->t_band = elapsed(h->ev[1], h->ev[6]);
        st->t_tridiag = elapsed(h->ev[2], h->ev[3]);
        st->t_backtransform = elapsed(h->ev[3], h->ev[4]);
        st->t_refine = elapsed(h->ev[4], h->ev[5]);
        st->t_total = elapsed(h->ev[0], h->ev[5]);
    }
    // the results stay fetchable; the status says they missed a tolerance
    if (tflags[0] || tflags[1]) {
        set_error(std::to_string(tflags[0]) + " eigenvalue(s) of T did not converge in " +
                  std::to_string(h->lim_bisect) + " bisection rounds, " + std::to_string(tflags[1]) +
                  " inverse-iteration vector(s) above the residual tolerance after " + std::to_string(h->lim_invit) +
                  " sweeps");
        return QD_E_NOCONV;
    }
    if (!refined) {
        set_error("FP64 refinement: residual above " + std::to_string(h->refine_rtol) + " ||T||_1 after " +
                  std::to_string(h->lim_refine) + " sweeps");
        return QD_E_NOCONV;
    }
    return QD_OK;
}

}  // namespace

extern "C" int qd_create(int device, qd_handle **out) {
    if (!out) { set_error("null handle pointer"); return QD_E_ARG; }
    return guarded([&] {
        int ndev = 0;
        QD_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) { set_error("no such HIP device"); return QD_E_DEVICE; }
        QD_HIP(hipSetDevice(device));
        qd_handle *h = new qd_handle();
        h->device = device;
        QD_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        for (auto &e : h->ev) QD_HIP(hipEventCreate(&e));
        for (auto &e : h->evc) QD_HIP(hipEventCreate(&e));
        *out = h;
        return QD_OK;
    });
}

extern "C" int qd_destroy(qd_handle *h) {
    if (!h) return QD_OK;
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        hipStreamSynchronize(h->stream);
        free_work(h);
        if (h->A32) hipFree(h->A32);
        if (h->A64) hipFree(h->A64);
        if (h->Hfull) hipFree(h->Hfull);
        if (h->has_tables) {
            hipFree(h->tab.tab);
            for (int a = 0; a < 3; ++a) { hipFree(h->tab.G[a]); hipFree(h->tab.u[a]); }
        }
        if (h->bin_cf) hipFree(h->bin_cf);
        for (auto &e : h->ev) hipEventDestroy(e);
        for (auto &e : h->prof.ev) hipEventDestroy(e);
        for (auto &e : h->evc) hipEventDestroy(e);
        hipStreamDestroy(h->stream);
        delete h;
        return QD_OK;
    });
}

extern "C" int qd_upload_tables(qd_handle *h, const qd_tables *t) {
    if (!h || !t || !t->tab) { set_error("null argument"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        if (h->has_tables) {
            hipFree(h->tab.tab);
            for (int a = 0; a < 3; ++a) { hipFree(h->tab.G[a]); hipFree(h->tab.u[a]); }
            h->has_tables = false;
        }
        DeviceTables &d = h->tab;
        d.nT = (long)t->NT[0] * t->NT[1] * t->NT[2];
        QD_HIP(hipMalloc(&d.tab, sizeof(c64) * QD_NTAB * d.nT));
        QD_HIP(hipMemcpy(d.tab, t->tab, sizeof(c64) * QD_NTAB * d.nT, hipMemcpyHostToDevice));
        for (int a = 0; a < 3; ++a) {
            QD_HIP(hipMalloc(&d.G[a], sizeof(double) * t->NT[a]));
            QD_HIP(hipMemcpy(d.G[a], t->G[a], sizeof(double) * t->NT[a], hipMemcpyHostToDevice));
            std::vector<double> u(t->N[a]);
            for (int m = 0; m < t->N[a]; ++m) u[m] = t->G[a][t->NT[a] / 2 - t->N[a] / 2 + m] + t->K[a];
            QD_HIP(hipMalloc(&d.u[a], sizeof(double) * t->N[a]));
            QD_HIP(hipMemcpy(d.u[a], u.data(), sizeof(double) * t->N[a], hipMemcpyHostToDevice));
            d.geom.N[a] = t->N[a];
            d.geom.NT[a] = t->NT[a];
            d.geom.K[a] = t->K[a];
            d.geom.G[a] = d.G[a];
        }
        d.geom.NG = t->N[0] * t->N[1] * t->N[2];
        d.NH = 8 * d.geom.NG;
        h->has_tables = true;
        return QD_OK;
    });
}

namespace qd {
// hip/fields.hip
void tables_device(const qd_inputs &in, const TableState &st, hipStream_t s, qd_tables &t);
void tables_vjp_device(const qd_inputs &in, const TableState &st, const std::complex<double> *W, hipStream_t s,
                       int nparams, const int *params, const unsigned *masks, TableGrad &out);
}

extern "C" int qd_tables_vjp_device(qd_handle *h, const qd_inputs *in, const double *W, int nparams, const int *params,
                                    const unsigned *masks, double *grad) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    const int rc = vjp_check_params(in, W, nparams, params, masks, grad);
    if (rc != QD_OK) return rc;
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        TableState st;
        try {
            build_geometry(*in, st, shape_on_device);   // host: materials, axes
        } catch (const std::runtime_error &e) {
            const std::string m = e.what();
            throw StatusError(m, m.find("not supported") != std::string::npos ? QD_E_UNSUPPORTED : QD_E_INPUT);
        }
        TableGrad g;
        tables_vjp_device(*in, st, reinterpret_cast<const std::complex<double> *>(W), h->stream, nparams, params, masks, g);
        try {
            vjp_contract(*in, st, g, nparams, params, masks, grad);   // host: alloy and shape derivatives
        } catch (const std::runtime_error &e) {
            const std::string m = e.what();
            throw StatusError(m, m.find("not supported") != std::string::npos ? QD_E_UNSUPPORTED : QD_E_INPUT);
        }
        return QD_OK;
    });
}

extern "C" int qd_tables_build_device(qd_handle *h, const qd_inputs *in, qd_tables *out) {
    if (!h || !in || !out || !in->shapes || in->N_shape < 1) { set_error("null or empty inputs"); return QD_E_ARG; }
    for (int a = 0; a < 3; ++a)
        if (in->N[a] < 1 || (in->N[a] % 2) == 0) { set_error("plane-wave counts must be odd and positive"); return QD_E_ARG; }
    std::memset(out, 0, sizeof(*out));
    const int rc = guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        TableState st;
        try {
            build_geometry(*in, st, shape_on_device);   // host: materials, axes, the quadrature shapes
        } catch (const std::runtime_error &e) {
            const std::string m = e.what();
            throw StatusError(m, m.find("not supported") != std::string::npos ? QD_E_UNSUPPORTED : QD_E_INPUT);
        }
        tables_device(*in, st, h->stream, *out);
        const size_t n = (size_t)QD_NTAB * out->NT[0] * out->NT[1] * out->NT[2] * 2;
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(out->tab[i])) throw StatusError("non-finite value in coefficient tables", QD_E_NAN);
        return QD_OK;
    });
    const int up = rc == QD_OK ? qd_upload_tables(h, out) : rc;
    if (up != QD_OK) qd_tables_free(out);
    return up;
}

extern "C" int qd_selftest_shape_device(qd_handle *h, const qd_inputs *in, int idx, int cell, int param, int n,
                                        const double *G, double *out) {
    if (!h || !in || !in->shapes || !G || !out || n < 0 || idx < 0 || idx >= in->N_shape) { set_error("bad argument"); return QD_E_ARG; }
    if (cell != 0 && cell != 1 && cell != 3) { set_error("cell must be 0, 1 or 3"); return QD_E_ARG; }
    double seed[4];
    const bool centre = param == QD_DP_OX || param == QD_DP_OY || param == QD_DP_OZ;
    if (param >= 0 && !centre && !length_seeds(param, in->shapes[idx].shape, seed)) {
        set_error("not a shape-geometry parameter");
        return QD_E_ARG;
    }
    if (!shape_on_device(in->shapes[idx].shape)) {
        set_error("shape type " + std::to_string(in->shapes[idx].shape) + " is not supported");
        return QD_E_UNSUPPORTED;
    }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        ShapeDev sd;
        try {
            sd = shape_device_params(*in, idx, cell == 0 ? Cell::Electronic : (cell == 1 ? Cell::Strain : Cell::Coulomb));
        } catch (const std::runtime_error &e) {
            throw StatusError(e.what(), QD_E_INPUT);
        }
        double *Gd = nullptr, *gl = nullptr;
        c64 *od = nullptr;
        struct Free {
            double **a, **c;
            c64 **b;
            ~Free() { (void)hipFree(*a); (void)hipFree(*b); (void)hipFree(*c); }
        } fr{&Gd, &gl, &od};
        const std::vector<double> rules = gauss_legendre_rules();
        QD_HIP(hipMalloc(&Gd, sizeof(double) * 3 * std::max(n, 1)));
        QD_HIP(hipMalloc(&od, sizeof(c64) * std::max(n, 1)));
        QD_HIP(hipMalloc(&gl, sizeof(double) * rules.size()));
        QD_HIP(hipMemcpyAsync(Gd, G, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
        QD_HIP(hipMemcpyAsync(gl, rules.data(), sizeof(double) * rules.size(), hipMemcpyHostToDevice, h->stream));
        launch_shape_points(sd, Gd, n, gl, (param >= 0 && !centre) ? seed : nullptr, od, h->stream);
        QD_HIP(hipMemcpyAsync(out, od, sizeof(c64) * n, hipMemcpyDeviceToHost, h->stream));
        QD_HIP(hipStreamSynchronize(h->stream));
        if (centre)   // the centre enters only through exp(-i G.O): d chi / d O_a = -i G_a chi
            for (int i = 0; i < n; ++i) {
                const double ga = G[3 * i + (param - QD_DP_OX)], re = out[2 * i], im = out[2 * i + 1];
                out[2 * i] = ga * im;
                out[2 * i + 1] = -ga * re;
            }
        return QD_OK;
    });
}

extern "C" int qd_hamiltonian(qd_handle *h, int precision, double *H_host) {
    if (!h || !H_host) { set_error("null argument"); return QD_E_ARG; }
    if (!h->has_tables) { set_error("no tables uploaded"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        const int n = h->tab.NH;
        const long ld = ((long)n + 63) / 64 * 64;
        if (precision == QD_FP64) {
            c64 *buf = nullptr;
            QD_HIP(hipMalloc(&buf, sizeof(c64) * ld * n));
            QD_HIP(hipMemsetAsync(buf, 0, sizeof(c64) * ld * n, h->stream));
            launch_build_lower<double>(h->tab, n, ld, buf, h->stream);
            std::vector<c64> tmp((size_t)ld * n);
            QD_HIP(hipMemcpyAsync(tmp.data(), buf, sizeof(c64) * ld * n, hipMemcpyDeviceToHost, h->stream));
            QD_HIP(hipStreamSynchronize(h->stream));
            hipFree(buf);
            for (long c = 0; c < n; ++c)
                for (long r = c; r < n; ++r) {
                    H_host[2 * (c * n + r)] = tmp[c * ld + r].re;
                    H_host[2 * (c * n + r) + 1] = tmp[c * ld + r].im;
                }
        } else if (precision == QD_FP32_REFINE) {
            cpx<float> *buf = nullptr;
            QD_HIP(hipMalloc(&buf, sizeof(cpx<float>) * ld * n));
            QD_HIP(hipMemsetAsync(buf, 0, sizeof(cpx<float>) * ld * n, h->stream));
            launch_build_lower<float>(h->tab, n, ld, buf, h->stream);
            std::vector<cpx<float>> tmp((size_t)ld * n);
            QD_HIP(hipMemcpyAsync(tmp.data(), buf, sizeof(cpx<float>) * ld * n, hipMemcpyDeviceToHost, h->stream));
            QD_HIP(hipStreamSynchronize(h->stream));
            hipFree(buf);
            for (long c = 0; c < n; ++c)
                for (long r = c; r < n; ++r) {
                    H_host[2 * (c * n + r)] = tmp[c * ld + r].re;
                    H_host[2 * (c * n + r) + 1] = tmp[c * ld + r].im;
                }
        } else {
            set_error("unknown precision");
            return QD_E_ARG;
        }
        return QD_OK;
    });
}

extern "C" int qd_solve_device(qd_handle *h, int il, int iu, int precision, qd_stats *st) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    if (!h->has_tables) { set_error("no tables uploaded"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        return solve_impl(h, h->tab.NH, il, iu, precision, st, nullptr);
    });
}

extern "C" int qd_fetch_results(qd_handle *h, double *w, double *Z) {
    if (!h || h->k == 0) { set_error("no results"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        const int n = h->n, k = h->k;
        if (w) QD_HIP(hipMemcpyAsync(w, h->wres, sizeof(double) * k, hipMemcpyDeviceToHost, h->stream));
        if (Z) {
            std::vector<c64> tmp((size_t)h->ld * k);
            QD_HIP(hipMemcpyAsync(tmp.data(), h->X, sizeof(c64) * h->ld * k, hipMemcpyDeviceToHost, h->stream));
            QD_HIP(hipStreamSynchronize(h->stream));
            for (long l = 0; l < k; ++l)
                for (long r = 0; r < n; ++r) {
                    Z[2 * (l * n + r)] = tmp[l * h->ld + r].re;
                    Z[2 * (l * n + r) + 1] = tmp[l * h->ld + r].im;
                }
        }
        QD_HIP(hipStreamSynchronize(h->stream));
        return QD_OK;
    });
}

extern "C" int qd_solve(qd_handle *h, int il, int iu, int precision, double *w, double *Z, qd_stats *st) {
    const int rc = qd_solve_device(h, il, iu, precision, st);
    if (rc != QD_OK && rc != QD_E_NOCONV) return rc;
    const std::string msg = rc == QD_E_NOCONV ? std::string(qd_last_error()) : std::string();
    const int rf = qd_fetch_results(h, w, Z);
    if (rf != QD_OK) return rf;
    if (rc == QD_E_NOCONV) set_error(msg);
    return rc;
}

extern "C" int qd_solve_matrix(qd_handle *h, int n, const double *A, int il, int iu, int precision, double *w,
                               double *Z, qd_stats *st) {
    if (!h || !A || n < 2) { set_error("bad argument"); return QD_E_ARG; }
    const int rc = guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        return solve_impl(h, n, il, iu, precision, st, A);
    });
    if (rc != QD_OK && rc != QD_E_NOCONV) return rc;
    const std::string msg = rc == QD_E_NOCONV ? std::string(qd_last_error()) : std::string();
    const int rf = qd_fetch_results(h, w, Z);
    if (rf != QD_OK) return rf;
    if (rc == QD_E_NOCONV) set_error(msg);
    return rc;
}

extern "C" int qd_set_refine_tolerance(qd_handle *h, double rel_residual) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    if (!(rel_residual > 0.0) || !(rel_residual < 1.0)) { set_error("tolerance must lie in (0, 1)"); return QD_E_ARG; }
    h->refine_rtol = rel_residual;
    return QD_OK;
}

extern "C" int qd_debug_set_limits(qd_handle *h, int refine_sweeps, int sternheimer_iters, int invit_sweeps,
                                   int bisect_rounds) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    h->lim_refine = refine_sweeps < 0 ? 6 : refine_sweeps;
    h->lim_stern = sternheimer_iters < 0 ? 12 : sternheimer_iters;
    h->lim_invit = invit_sweeps < 0 ? 4 : invit_sweeps;
    h->lim_bisect = bisect_rounds < 0 ? 40 : bisect_rounds;
    return QD_OK;
}

extern "C" int qd_profile_enable(qd_handle *h, int enable) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    h->prof.on = enable != 0;
    h->prof.used = 0;
    h->prof.work.clear();
    h->prof_sum_s = h->prof_sum_b = 0;
    h->prof_n = 0;
    return QD_OK;
}

extern "C" int qd_set_reduction(qd_handle *h, int method) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    if (method != QD_REDUCE_TWO_STAGE && method != QD_REDUCE_ONE_STAGE) { set_error("unknown reduction"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        const bool two = method == QD_REDUCE_TWO_STAGE;
        if (two != h->two_stage) {
            QD_HIP(hipStreamSynchronize(h->stream));
            free_work(h);
            h->two_stage = two;
        }
        return QD_OK;
    });
}

extern "C" int qd_profile_read(qd_handle *h, int *launches, double *mean_seconds, double *mean_work) {
    if (!h) { set_error("null handle"); return QD_E_ARG; }
    if (launches) *launches = (int)h->prof_n;
    if (mean_seconds) *mean_seconds = h->prof_n ? h->prof_sum_s / h->prof_n : 0.0;
    if (mean_work) *mean_work = h->prof_n ? h->prof_sum_b / h->prof_n : 0.0;
    return QD_OK;
}

extern "C" int qd_debug_tridiagonal(qd_handle *h, double *d, double *e) {
    if (!h || h->work_n < 0 || !d || !e) { set_error("no reduction available"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipMemcpy(d, tri_d(h), sizeof(double) * h->n, hipMemcpyDeviceToHost));
        if (h->n > 1) QD_HIP(hipMemcpy(e, tri_e(h), sizeof(double) * (h->n - 1), hipMemcpyDeviceToHost));
        return QD_OK;
    });
}

extern "C" int qd_debug_apply_q(qd_handle *h, int k, int adjoint, double *X) {
    if (!h || h->work_n < 0 || !X || k < 1 || k > 64) { set_error("bad argument or no reduction available"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        hipStream_t s = h->stream;
        const int n = h->n;
        const long ld = h->ld;
        c64 *dX = nullptr;
        QD_HIP(hipMalloc(&dX, sizeof(c64) * ld * k));
        struct FreeX { c64 *p; ~FreeX() { (void)hipFree(p); } } fx{dX};
        QD_HIP(hipMemcpy2DAsync(dX, sizeof(c64) * ld, X, sizeof(c64) * n, sizeof(c64) * n, k, hipMemcpyHostToDevice, s));
        apply_q(h, dX, k, adjoint != 0, s);
        QD_HIP(hipMemcpy2DAsync(X, sizeof(c64) * n, dX, sizeof(c64) * ld, sizeof(c64) * n, k, hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        if (h->two_stage && chase_failed(h->ch, s)) throw HipError("band back-transformation: a bounded wait gave up", QD_E_DEVICE);
        return QD_OK;
    });
}

extern "C" int qd_debug_mark_wait_failure(qd_handle *h) {
    if (!h || h->work_n < 0 || !h->two_stage) { set_error("no two-stage reduction on this handle"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        chase_mark_error(h->ch, h->stream);
        QD_HIP(hipStreamSynchronize(h->stream));
        return QD_OK;
    });
}

extern "C" int qd_selftest_quad_sum(const double *in, double *out) {
    if (!in || !out) { set_error("null argument"); return QD_E_ARG; }
    return guarded([&] {
        quad_sum_selftest(in, out);
        return QD_OK;
    });
}

extern "C" int qd_selftest_adjoint_bin_host(const qd_tables *t, int nv, const double *V, int nterms, const int *a,
                                            const int *b, const double *gamma, double *W_fft, double *W_direct) {
    if (!t || !V || !a || !b || !gamma || !W_fft || !W_direct || nv < 1) { set_error("bad argument"); return QD_E_ARG; }
    if (nterms < 1 || nterms > kAdjMaxTerms) { set_error("nterms out of range"); return QD_E_ARG; }
    for (int q = 0; q < nterms; ++q)
        if (a[q] < 0 || a[q] >= nv || b[q] < 0 || b[q] >= nv) { set_error("term index out of range"); return QD_E_ARG; }
    try {
        HGeom g{};
        std::vector<double> u[3];
        const double *up[3];
        for (int k = 0; k < 3; ++k) {
            g.N[k] = t->N[k];
            g.NT[k] = t->NT[k];
            g.K[k] = t->K[k];
            g.G[k] = t->G[k];
            u[k].resize(t->N[k]);
            for (int m = 0; m < t->N[k]; ++m) u[k][m] = t->G[k][t->NT[k] / 2 - t->N[k] / 2 + m] + t->K[k];
            up[k] = u[k].data();
        }
        g.NG = t->N[0] * t->N[1] * t->N[2];
        const long nT = (long)t->NT[0] * t->NT[1] * t->NT[2];
        AdjointTerms terms{};
        terms.n = nterms;
        for (int q = 0; q < nterms; ++q) { terms.a[q] = a[q]; terms.b[q] = b[q]; terms.gamma[q] = gamma[q]; }
        const BinCoefs cf = bin_coefficients();
        const c64 *Vc = reinterpret_cast<const c64 *>(V);
        adjoint_bin_fft_host(g, up, Vc, 8L * g.NG, nv, terms, cf, reinterpret_cast<c64 *>(W_fft), nT);
        adjoint_bin_direct_host(g, Vc, 8L * g.NG, terms, reinterpret_cast<c64 *>(W_direct), nT);
    } catch (const std::exception &e) {
        set_error(e.what());
        return QD_E_DEVICE;
    }
    return QD_OK;
}

extern "C" int qd_selftest_momentum_host(const qd_tables *t, int nv, const double *V, int nc, const double *C,
                                         double *p) {
    if (!t || !V || !C || !p || nv < 1 || nc < 1) { set_error("bad argument"); return QD_E_ARG; }
    try {
        HGeom g{};
        for (int k = 0; k < 3; ++k) { g.N[k] = t->N[k]; g.NT[k] = t->NT[k]; g.K[k] = t->K[k]; g.G[k] = t->G[k]; }
        g.NG = t->N[0] * t->N[1] * t->N[2];
        const long nT = (long)t->NT[0] * t->NT[1] * t->NT[2];
        const int n = 8 * g.NG;
        std::vector<c64> P((size_t)3 * nc * nv);
        momentum_host(g, reinterpret_cast<const c64 *>(t->tab), nT, n, reinterpret_cast<const c64 *>(V), nv,
                      reinterpret_cast<const c64 *>(C), nc, n, P.data());
        momentum_order(P.data(), nv, nc, p);
    } catch (const std::exception &e) {
        set_error(e.what());
        return QD_E_DEVICE;
    }
    return QD_OK;
}

namespace qd {
void parallel_for(long n, const std::function<void(long, long)> &body);   // host/shapes.cpp
}

extern "C" int qd_selftest_hamiltonian_host(const qd_tables *t, double *H) {
    if (!t || !H || !t->tab) { set_error("bad argument"); return QD_E_ARG; }
    try {
        HGeom g{};
        for (int k = 0; k < 3; ++k) { g.N[k] = t->N[k]; g.NT[k] = t->NT[k]; g.K[k] = t->K[k]; g.G[k] = t->G[k]; }
        g.NG = t->N[0] * t->N[1] * t->N[2];
        const long nT = (long)t->NT[0] * t->NT[1] * t->NT[2];
        const int n = 8 * g.NG;
        const c64 *tab = reinterpret_cast<const c64 *>(t->tab);
        c64 *Hc = reinterpret_cast<c64 *>(H);
        qd::parallel_for(n, [&](long lo, long hi) {
            for (long c = lo; c < hi; ++c)
                for (int r = 0; r < n; ++r) Hc[c * n + r] = h_full(g, tab, nT, r, (int)c);
        });
    } catch (const std::exception &e) {
        set_error(e.what());
        return QD_E_DEVICE;
    }
    return QD_OK;
}

extern "C" int qd_selftest_mfma(double *out) {
    if (!out) { set_error("null output"); return QD_E_ARG; }
    return guarded([&] {
        out[0] = mfma_selftest<float>();
        out[1] = mfma_selftest<double>();
        // raw f64 layout probes: (one-hot a at lane 5, b = lane+1) and (a = lane+1, one-hot b at lane 5)
        double a[64], b[64];
        for (int l = 0; l < 64; ++l) { a[l] = (l == 5); b[l] = l + 1; }
        mfma_raw64(a, b, out + 2);
        for (int l = 0; l < 64; ++l) { a[l] = l + 1; b[l] = (l == 5); }
        mfma_raw64(a, b, out + 2 + 256);
        return QD_OK;
    });
}

extern "C" int qd_coulomb(qd_handle *h, const qd_inputs *in, const qd_tables *t, int nh, int ne, const double *Z,
                          double *hh, double *ee, double *eh_direct, double *eh_exchange, qd_coulomb_stats *st) {
    if (!h || !in || !t || !hh || !ee || !eh_direct || !eh_exchange) { set_error("null argument"); return QD_E_ARG; }
    if (nh < 1 || ne < 1) { set_error("need at least one hole and one electron state"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        auto t0 = std::chrono::steady_clock::now();
        CoulombSetup cs;
        coulomb_setup(*in, *t, cs);
        const int S = nh + ne;
        const long NH = 8L * t->N[0] * t->N[1] * t->N[2];
        hipStream_t s = h->stream;
        const c64 *X = nullptr;
        long ldx = 0;
        c64 *Xtmp = nullptr;
        struct Free {
            c64 **p;
            ~Free() { if (*p) (void)hipFree(*p); }
        } fr{&Xtmp};
        if (Z) {
            QD_HIP(hipMalloc(&Xtmp, sizeof(c64) * NH * S));
            QD_HIP(hipMemcpyAsync(Xtmp, Z, sizeof(c64) * NH * S, hipMemcpyHostToDevice, s));
            X = Xtmp;
            ldx = NH;
        } else {
            // device-resident window of the last solve: holes are the nh highest of the first
            // 2 N_vb states, electrons the ne lowest of the rest (window order, ascending)
            const int nvb = 2 * in->N_vb;
            if (h->k == 0 || h->n != NH) throw StatusError("no solve of this problem on the handle", QD_E_ARG);
            if (nh > nvb || nvb + ne > h->k)
                throw StatusError("requested states exceed the solved window", QD_E_ARG);
            QD_HIP(hipMalloc(&Xtmp, sizeof(c64) * h->ld * S));
            QD_HIP(hipMemcpyAsync(Xtmp, h->X + (long)(nvb - nh) * h->ld, sizeof(c64) * h->ld * nh,
                                  hipMemcpyDeviceToDevice, s));
            QD_HIP(hipMemcpyAsync(Xtmp + (long)nh * h->ld, h->X + (long)nvb * h->ld, sizeof(c64) * h->ld * ne,
                                  hipMemcpyDeviceToDevice, s));
            X = Xtmp;
            ldx = h->ld;
        }
        std::vector<zc> C;
        int P = 0, R = 0, fl[3];
        double tp = 0, tc = 0;
        const size_t nbytes = coulomb_scratch_bytes(S, cs);
        unsigned char *arena = nullptr;
        QD_HIP(hipMalloc(&arena, nbytes));
        struct FreeA { unsigned char *p; ~FreeA() { (void)hipFree(p); } } fa{arena};
        coulomb_device(X, ldx, S, cs, arena, nbytes, h->evc, s, C, P, R, fl, &tp, &tc, &cs.a_mad);
        coulomb_assemble(cs, C.data(), P, R, nh, ne, reinterpret_cast<zc *>(hh), reinterpret_cast<zc *>(ee),
                         reinterpret_cast<zc *>(eh_direct), reinterpret_cast<zc *>(eh_exchange));
        if (st) {
            st->t_pairs = tp;
            st->t_contract = tc;
            st->t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            st->n_pairs = P;
            for (int a = 0; a < 3; ++a) st->fft[a] = fl[a];
            st->nq = (long)cs.nq();
        }
        return QD_OK;
    });
}

extern "C" int qd_momentum(qd_handle *h, const qd_inputs *in, int nv, int nc, double *p, qd_momentum_stats *st) {
    if (!h || !in || !p) { set_error("null argument"); return QD_E_ARG; }
    if (nv < 1 || nc < 1) { set_error("need at least one valence and one conduction state"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        auto t0 = std::chrono::steady_clock::now();
        if (!h->has_tables) throw StatusError("no tables on the handle", QD_E_ARG);
        const int nvb = 2 * in->N_vb;
        if (h->k == 0 || h->n != h->tab.NH) throw StatusError("no solve of this problem on the handle", QD_E_ARG);
        if (nv > nvb || nvb + nc > h->k) throw StatusError("requested states exceed the solved window", QD_E_ARG);
        hipStream_t s = h->stream;
        const size_t sb = momentum_scratch_bytes(h->n, nc);
        unsigned char *scratch = nullptr;
        c64 *P = nullptr;
        QD_HIP(hipMalloc(&scratch, sb));
        struct FreeA { unsigned char *p; ~FreeA() { (void)hipFree(p); } } fa{scratch};
        QD_HIP(hipMalloc(&P, sizeof(c64) * 3 * nc * nv));
        struct FreeP { c64 *p; ~FreeP() { (void)hipFree(p); } } fp{P};
        // valence states: the nv highest of the first 2 N_vb window columns (ascending there)
        momentum_device(h->tab, h->n, h->X + (long)(nvb - nv) * h->ld, nv, h->X + (long)nvb * h->ld, nc, h->ld,
                        scratch, P, s);
        std::vector<c64> Ph((size_t)3 * nc * nv);
        QD_HIP(hipMemcpyAsync(Ph.data(), P, sizeof(c64) * Ph.size(), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        momentum_order(Ph.data(), nv, nc, p);
        if (st) st->t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return QD_OK;
    });
}

namespace qd {
}

extern "C" int qd_localisation(qd_handle *h, const qd_inputs *in, double *loc_vb, double *loc_cb, double *seconds) {
    if (!h || !in || !in->shapes || !loc_vb || !loc_cb) { set_error("null argument"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        auto t0 = std::chrono::steady_clock::now();
        const int nvb = 2 * in->N_vb, ncb = 2 * in->N_cb, ns = in->N_shape;
        if (!h->has_tables || h->k == 0 || h->n != h->tab.NH) throw StatusError("no solve of this problem on the handle", QD_E_ARG);
        for (int a = 0; a < 3; ++a)
            if (in->N[a] != h->tab.geom.N[a]) throw StatusError("inputs do not match the resident tables", QD_E_ARG);
        if (nvb + ncb > h->k) throw StatusError("the solved window holds fewer states than N_VB + N_CB", QD_E_ARG);
        const long nT = (long)h->tab.nT;
        const int nst = nvb + ncb;
        hipStream_t s = h->stream;
        std::vector<void *> bufs;
        struct FreeAll {
            std::vector<void *> &b;
            ~FreeAll() {
                for (void *p : b) (void)hipFree(p);
            }
        } fa{bufs};
        auto dev = [&](size_t bytes) {
            void *p = nullptr;
            QD_HIP(hipMalloc(&p, std::max<size_t>(bytes, 1)));
            bufs.push_back(p);
            return p;
        };
        // chi_s of every shape on the electronic grid (characteristic_functions_reciprocal_space.c),
        // evaluated on the device; slot 0 (the matrix) stays zero
        std::map<int, std::pair<std::vector<ShapeDev>, std::vector<int>>> fam;   // per kernel family
        for (int sh = 1; sh < ns; ++sh) {
            auto &f = fam[shape_family(in->shapes[sh].shape)];
            try {
                f.first.push_back(shape_device_params(*in, sh, Cell::Electronic));
            } catch (const std::runtime_error &e) {
                throw StatusError(e.what(), QD_E_INPUT);
            }
            f.second.push_back(sh);
        }
        const double Lc[3] = {in->shapes[0].L1, in->shapes[0].L2, in->shapes[0].L3};
        const double *gax[3];
        std::vector<double> axes[3];
        for (int a = 0; a < 3; ++a) {
            axes[a] = reciprocal_axis(in->N[a], Lc[a]);
            double *d = static_cast<double *>(dev(sizeof(double) * axes[a].size()));
            QD_HIP(hipMemcpyAsync(d, axes[a].data(), sizeof(double) * axes[a].size(), hipMemcpyHostToDevice, s));
            gax[a] = d;
        }
        const int NT[3] = {(int)axes[0].size(), (int)axes[1].size(), (int)axes[2].size()};
        if ((long)NT[0] * NT[1] * NT[2] != nT) throw StatusError("inputs do not match the resident tables", QD_E_ARG);
        const std::vector<double> rules = gauss_legendre_rules();
        double *gl = static_cast<double *>(dev(sizeof(double) * rules.size()));
        QD_HIP(hipMemcpyAsync(gl, rules.data(), sizeof(double) * rules.size(), hipMemcpyHostToDevice, s));
        c64 *chid = static_cast<c64 *>(dev(sizeof(c64) * ns * nT));
        QD_HIP(hipMemsetAsync(chid, 0, sizeof(c64) * nT, s));
        for (const auto &f : fam) {
            const auto &sd = f.second.first;
            const auto &slot = f.second.second;
            ShapeDev *sdd = static_cast<ShapeDev *>(dev(sizeof(ShapeDev) * sd.size()));
            int *sl = static_cast<int *>(dev(sizeof(int) * slot.size()));
            QD_HIP(hipMemcpyAsync(sdd, sd.data(), sizeof(ShapeDev) * sd.size(), hipMemcpyHostToDevice, s));
            QD_HIP(hipMemcpyAsync(sl, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice, s));
            launch_shape_grid(sdd, sl, (int)sd.size(), f.first, gax, NT, gl, chid, s);
        }
        c64 *rho = static_cast<c64 *>(dev(sizeof(c64) * nT * nst));
        double *Ld = static_cast<double *>(dev(sizeof(double) * ns * nst));
        int *cols = static_cast<int *>(dev(sizeof(int) * nst));
        std::vector<int> hc(nst);
        for (int i = 0; i < nst; ++i) hc[i] = i;
        QD_HIP(hipMemcpyAsync(cols, hc.data(), sizeof(int) * nst, hipMemcpyHostToDevice, s));
        for (int i = 0; i < nst; ++i) launch_level_density(h->X, h->ld, cols + i, 1, in->N, rho + (long)i * nT, s);
        launch_loc_contract(chid, rho, nT, nst, ns, Ld, s);
        std::vector<double> L((size_t)ns * nst);
        QD_HIP(hipMemcpyAsync(L.data(), Ld, sizeof(double) * L.size(), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        // rows: valence states highest first, conduction states lowest first; column 0 is the rest
        // of the cell, 1 - the sum over the shapes embedded in the matrix (8_band_main.c:1531-1540)
        for (int r = 0; r < nst; ++r) {
            const int n = r < nvb ? nvb - 1 - r : r;
            double *row = r < nvb ? loc_vb + (size_t)r * ns : loc_cb + (size_t)(r - nvb) * ns;
            double sum = 0.0;
            for (int sh = 1; sh < ns; ++sh) {
                row[sh] = L[(size_t)n * ns + sh];
                if (in->shapes[sh].embedded == 0) sum += row[sh];
            }
            row[0] = 1.0 - sum;
        }
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return QD_OK;
    });
}

extern "C" int qd_overlaps(qd_handle *h, const qd_inputs *in, double *ov) {
    if (!h || !in || !ov) { set_error("null argument"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        const int nvb = 2 * in->N_vb, ncb = 2 * in->N_cb;
        if (!h->has_tables || h->k == 0 || h->n != h->tab.NH) throw StatusError("no solve of this problem on the handle", QD_E_ARG);
        if (nvb + ncb > h->k) throw StatusError("the solved window holds fewer states than N_VB + N_CB", QD_E_ARG);
        hipStream_t s = h->stream;
        c64 *O = nullptr;
        QD_HIP(hipMalloc(&O, sizeof(c64) * ncb * nvb));
        struct FreeO { c64 *p; ~FreeO() { (void)hipFree(p); } } fo{O};
        launch_col_dot(h->X + (long)nvb * h->ld, ncb, h->X, nvb, h->ld, h->n, O, s);
        std::vector<c64> Oh((size_t)ncb * nvb);
        QD_HIP(hipMemcpyAsync(Oh.data(), O, sizeof(c64) * Oh.size(), hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        // ov[c][v] = <cb_c|vb_v>, v = 0 the highest valence state (spatial_overlap.c, 8_band_main.c:1396)
        for (int c = 0; c < ncb; ++c)
            for (int v = 0; v < nvb; ++v) {
                const c64 x = Oh[(size_t)c * nvb + (nvb - 1 - v)];
                ov[2 * ((size_t)c * nvb + v)] = x.re;
                ov[2 * ((size_t)c * nvb + v) + 1] = x.im;
            }
        return QD_OK;
    });
}

extern "C" int qd_envelopes(qd_handle *h, const qd_inputs *in, int first, int count, int npts, const double *x,
                            const double *y, const double *z, double *F) {
    if (!h || !in || !in->shapes || !x || !y || !z || !F) { set_error("null argument"); return QD_E_ARG; }
    if (count < 1 || npts < 1 || first < 0) { set_error("need at least one state and one point"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        if (!h->has_tables || h->k == 0 || h->n != h->tab.NH) throw StatusError("no solve of this problem on the handle", QD_E_ARG);
        for (int a = 0; a < 3; ++a)
            if (in->N[a] != h->tab.geom.N[a]) throw StatusError("inputs do not match the resident tables", QD_E_ARG);
        if (first + count > h->k) throw StatusError("requested states exceed the solved window", QD_E_ARG);
        const double vol = in->shapes[0].L1 * in->shapes[0].L2 * in->shapes[0].L3;
        if (!(vol > 0.0)) throw StatusError("supercell lengths missing", QD_E_INPUT);
        hipStream_t s = h->stream;
        double *pts = nullptr;
        c64 *Fd = nullptr;
        struct Free {
            double **p;
            c64 **f;
            ~Free() { (void)hipFree(*p); (void)hipFree(*f); }
        } fr{&pts, &Fd};
        QD_HIP(hipMalloc(&pts, sizeof(double) * 3 * npts));
        QD_HIP(hipMalloc(&Fd, sizeof(c64) * 8 * (size_t)count * npts));
        QD_HIP(hipMemcpyAsync(pts, x, sizeof(double) * npts, hipMemcpyHostToDevice, s));
        QD_HIP(hipMemcpyAsync(pts + npts, y, sizeof(double) * npts, hipMemcpyHostToDevice, s));
        QD_HIP(hipMemcpyAsync(pts + 2 * (long)npts, z, sizeof(double) * npts, hipMemcpyHostToDevice, s));
        const double *u[3] = {h->tab.u[0], h->tab.u[1], h->tab.u[2]};
        launch_envelope(h->X + (long)first * h->ld, h->ld, count, u, in->N, pts, pts + npts, pts + 2 * (long)npts, npts,
                        1.0 / std::sqrt(vol), Fd, s);
        QD_HIP(hipMemcpyAsync(F, Fd, sizeof(c64) * 8 * (size_t)count * npts, hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        return QD_OK;
    });
}

namespace {

// W (QD_NTAB x nT) for D = sum_q gamma_q V[:, a_q] V[:, b_q]^H with V ldv-strided on the device.
void adjoint_bin_device(qd_handle *h, const c64 *V, long ldv, int nv, const std::vector<int> &a,
                        const std::vector<int> &b, const std::vector<double> &gam, std::vector<c64> &W) {
    if ((int)a.size() > kAdjMaxTerms) throw StatusError("too many adjoint terms", QD_E_ARG);
    AdjointTerms terms{};
    terms.n = (int)a.size();
    for (int q = 0; q < terms.n; ++q) { terms.a[q] = a[q]; terms.b[q] = b[q]; terms.gamma[q] = gam[q]; }
    const HGeom &geo = h->tab.geom;
    const long nT = (long)geo.NT[0] * geo.NT[1] * geo.NT[2];
    c64 *Wd = nullptr;
    QD_HIP(hipMalloc(&Wd, sizeof(c64) * QD_NTAB * nT));
    struct Free { c64 *p; ~Free() { (void)hipFree(p); } } fr{Wd};
    if (!h->bin_cf) {
        static const BinCoefs cf = bin_coefficients();
        QD_HIP(hipMalloc(&h->bin_cf, sizeof(BinCoefs)));
        QD_HIP(hipMemcpy(h->bin_cf, &cf, sizeof(BinCoefs), hipMemcpyHostToDevice));
    }
    const double *u[3] = {h->tab.u[0], h->tab.u[1], h->tab.u[2]};
    launch_adjoint_bin_fft(geo, u, V, ldv, nv, terms, h->bin_cf, Wd, nT, h->stream);
    W.resize((size_t)QD_NTAB * nT);
    QD_HIP(hipMemcpyAsync(W.data(), Wd, sizeof(c64) * QD_NTAB * nT, hipMemcpyDeviceToHost, h->stream));
    QD_HIP(hipStreamSynchronize(h->stream));
}

// P x: remove the components of the ncol columns of Xc along the two columns of L (per column
// block of two), twice (classical Gram-Schmidt).
// Xc (ncol ld-strided columns) minus its components along the nl orthonormal columns L, twice.
void project_out(RefineCtx &rc, const c64 *L, int nl, c64 *Xc, int ncol, c64 *S, hipStream_t s) {
    for (int pass = 0; pass < 2; ++pass) {
        rc.atb(L, nl, Xc, ncol, S, s);
        rc.ab(L, nl, S, ncol, -1.0, 1.0, Xc, s);
    }
}

constexpr int kAdjMaxZ = 32;          // eigenvector gradients per qd_adjoint_z call
constexpr double kLevelTol = 1e-8;    // eV: window states closer than this form one degenerate level

// A degenerate level of the window: columns first .. first + count - 1 of the resident eigenvectors.
struct Level {
    int first, count;
};

// Sternheimer solutions lambda_j = (H - E_j)^+ P_j b_j of m right-hand sides (the columns of Gv,
// overwritten by the solutions), P_j the projector onto the complement of level lev[j] of the
// window: X += P Q (T - tau)^-1 Q^H R with the reduction's factorisation (tau = E_j + eta off the
// level), R = b - P (H - E) X with the FP64 H, until max_j ||R_j|| / ||b_j|| < 1e-12 or
// h->lim_stern iterations.  Returns that residual (1e300 when no iteration ran).
double sternheimer(qd_handle *h, c64 *Gv, int m, const std::vector<double> &El, const std::vector<Level> &lev,
                   int *iters, hipStream_t s) {
    const int n = h->n;
    const long ld = h->ld;
    RefineCtx &rc = h->adj;
    if (rc.n != n || rc.ld != ld || rc.k != m || !rc.B) { rc.release(); rc.alloc(n, ld, m); }
    const bool fp64 = h->work_prec == QD_FP64;
    const c64 *Hf = h->Hfull;
    if (fp64) {
        ensure(h->adj_H, h->adj_h_elems, (size_t)ld * n);
        launch_build_full(h->tab, n, ld, h->adj_H, s);
        Hf = h->adj_H;
    }
    double gb[4];
    QD_HIP(hipMemcpyAsync(gb, h->tw.info, sizeof(gb), hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    const double tnorm = gb[3];
    const double eta = fp64 ? 1e-6 : 1e-4;   // eV: shift off the level in the tridiagonal solve
    std::vector<double> tau(m);
    for (int j = 0; j < m; ++j) tau[j] = El[j] + eta;
    double *dEl = rc.lam, *dtau = rc.lam + m;
    QD_HIP(hipMemcpyAsync(dEl, El.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    QD_HIP(hipMemcpyAsync(dtau, tau.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    c64 *Bv = rc.B, *X = rc.B + (long)m * ld, *R = rc.R, *HX = rc.HB, *S = rc.small;
    // every column against its own level, runs of columns with the same level together
    auto project = [&](c64 *Y) {
        for (int j = 0; j < m;) {
            int e = j + 1;
            while (e < m && lev[e].first == lev[j].first && lev[e].count == lev[j].count) ++e;
            project_out(rc, h->X + (long)lev[j].first * ld, lev[j].count, Y + (long)j * ld, e - j, S, s);
            j = e;
        }
    };
    // b = P g
    copy_cols(Gv, Bv, ld, n, m, s);
    project(Bv);
    rc.colnorms(Bv, m, rc.norms, s);
    std::vector<double> bn(m), rn(m);
    QD_HIP(hipMemcpyAsync(bn.data(), rc.norms, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    QD_HIP(hipMemsetAsync(X, 0, sizeof(c64) * ld * m, s));
    copy_cols(Bv, R, ld, n, m, s);
    double res_max = 1e300;   // no iteration: not converged
    *iters = 0;
    for (int it = 0; it < h->lim_stern; ++it) {
        // X += P Q (T - tau)^{-1} Q^H R
        copy_cols(R, rc.T2, ld, n, m, s);
        apply_q(h, rc.T2, m, true, s);
        rc.trisolve(tri_d(h), tri_e(h), dtau, tnorm, rc.T2, s);
        apply_q(h, rc.T2, m, false, s);
        project(rc.T2);
        axpby_cols(rc.T2, 1.0, X, 1.0, ld, n, m, s);     // X += correction
        // R = b - P (H - E) X
        rc.gemm_h(Hf, X, m, HX, s);
        rc.residual(HX, X, dEl, R, s);
        project(R);
        axpby_cols(Bv, 1.0, R, -1.0, ld, n, m, s);      // R = b - R
        rc.colnorms(R, m, rc.norms, s);
        QD_HIP(hipMemcpyAsync(rn.data(), rc.norms, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        res_max = 0.0;
        for (int j = 0; j < m; ++j) res_max = std::max(res_max, bn[j] > 0 ? rn[j] / bn[j] : 0.0);
        *iters = it + 1;
        if (res_max < 1e-12) break;
    }
    if (h->two_stage && chase_failed(h->ch, s))
        throw HipError("adjoint: a bounded wait of the back-transformation gave up", QD_E_DEVICE);
    if (!(res_max < 1e-12))
        throw StatusError("Sternheimer solve: relative residual " + std::to_string(res_max) + " after " +
                              std::to_string(*iters) + " iterations (tolerance 1e-12)",
                          QD_E_NOCONV);
    copy_cols(X, Gv, ld, n, m, s);
    return res_max;
}

}  // namespace

extern "C" int qd_adjoint(qd_handle *h, const qd_inputs *in, const qd_tables *t, const double *gE, int hole_level,
                          int electron_level, double wJ, double *W, qd_adjoint_stats *st) {
    if (!h || !in || !t || !gE || !W) { set_error("null argument"); return QD_E_ARG; }
    if (h->k == 0 || !h->has_tables) { set_error("no solve on this handle"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        hipStream_t s = h->stream;
        const int n = h->n, k = h->k;
        const long ld = h->ld;
        const HGeom &geo = h->tab.geom;
        const long nT = (long)geo.NT[0] * geo.NT[1] * geo.NT[2];
        if (n != h->tab.NH) throw StatusError("last solve was not of the uploaded tables", QD_E_ARG);
        auto t0 = std::chrono::steady_clock::now();
        auto secs = [](std::chrono::steady_clock::time_point a) {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        };
        std::vector<double> w(k);
        QD_HIP(hipMemcpyAsync(w.data(), h->wres, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        // vectors of D: the window eigenvectors with nonzero weight, then the Sternheimer solutions
        std::vector<int> eig;
        for (int i = 0; i < k; ++i)
            if (gE[i] != 0.0) eig.push_back(i);
        const bool coul = wJ != 0.0;
        const int nv = (int)eig.size() + (coul ? 8 : 0);
        if (nv == 0) throw StatusError("objective has no terms", QD_E_ARG);
        c64 *V = nullptr;
        QD_HIP(hipMalloc(&V, sizeof(c64) * ld * std::max(nv, 1)));
        struct FreeV { c64 *p; ~FreeV() { (void)hipFree(p); } } fv{V};
        QD_HIP(hipMemsetAsync(V, 0, sizeof(c64) * ld * nv, s));
        std::vector<int> ta, tb;
        std::vector<double> tg;
        for (size_t q = 0; q < eig.size(); ++q) {
            copy_cols(h->X + (long)eig[q] * ld, V + (long)q * ld, ld, n, 1, s);
            ta.push_back((int)q); tb.push_back((int)q); tg.push_back(gE[eig[q]]);
        }
        double J = 0.0, t_dens = 0.0, t_st = 0.0, res_max = 0.0;
        int st_iters = 0;
        if (coul) {
            if (hole_level < 0 || hole_level + 1 >= k || electron_level < 0 || electron_level + 1 >= k)
                throw StatusError("level index outside the solved window", QD_E_ARG);
            for (int a = 0; a < 3; ++a)
                if (in->N_C[a] != t->N[a] || std::fabs(in->CL[a] - t->L[a]) > 1e-9 * t->L[a])
                    throw StatusError("the Coulomb term of the adjoint needs the Coulomb cell equal to the "
                                      "electronic-structure cell", QD_E_UNSUPPORTED);
            auto td = std::chrono::steady_clock::now();
            const int N[3] = {geo.N[0], geo.N[1], geo.N[2]};
            // vectors: psi_h1a psi_h1b psi_e1a psi_e1b at base, lambda at base + 4
            const int base = (int)eig.size();
            const int lv[4] = {hole_level, hole_level + 1, electron_level, electron_level + 1};
            for (int j = 0; j < 4; ++j) copy_cols(h->X + (long)lv[j] * ld, V + (long)(base + j) * ld, ld, n, 1, s);
            int *dcols = nullptr;
            QD_HIP(hipMalloc(&dcols, sizeof(int) * 8));
            struct FreeI { int *p; ~FreeI() { (void)hipFree(p); } } fi{dcols};
            const int hcols[8] = {base, base + 1, base + 2, base + 3, 0, 0, 1, 1};   // cols, then which[]
            QD_HIP(hipMemcpyAsync(dcols, hcols, sizeof(hcols), hipMemcpyHostToDevice, s));
            c64 *rho = nullptr;
            QD_HIP(hipMalloc(&rho, sizeof(c64) * 2 * nT));
            struct FreeR { c64 *p; ~FreeR() { (void)hipFree(p); } } frr{rho};
            launch_level_density(V, ld, dcols, 2, N, rho, s);
            launch_level_density(V, ld, dcols + 2, 2, N, rho + nT, s);
            std::vector<c64> rh(2 * nT);
            QD_HIP(hipMemcpyAsync(rh.data(), rho, sizeof(c64) * 2 * nT, hipMemcpyDeviceToHost, s));
            QD_HIP(hipStreamSynchronize(s));
            // Poisson kernel on the Delta-G grid (identical Coulomb cell), J and the potentials
            std::vector<c64> phi(2 * nT);
            const double om = t->L[0] * t->L[1] * t->L[2];
            const double nr0 = t->nr0;
            for (long q = 0; q < nT; ++q) {
                const int ix = (int)(q / ((long)geo.NT[1] * geo.NT[2])), iy = (int)((q / geo.NT[2]) % geo.NT[1]),
                          iz = (int)(q % geo.NT[2]);
                const double gx = t->G[0][ix], gy = t->G[1][iy], gz = t->G[2][iz];
                const double g2 = gx * gx + gy * gy + gz * gz;
                const bool cen = ix == geo.NT[0] / 2 && iy == geo.NT[1] / 2 && iz == geo.NT[2] / 2;
                const double wq = cen ? 0.0 : 180.9513 / (om * g2 * nr0 * nr0);
                const c64 a = rh[q], b = rh[nT + q];
                J += 0.25 * wq * (a.re * b.re + a.im * b.im);
                phi[q] = wq * conj(b);          // acts on the hole level: w conj(rho_e)
                phi[nT + q] = wq * conj(a);     // acts on the electron level: w conj(rho_h)
            }
            QD_HIP(hipMemcpyAsync(rho, phi.data(), sizeof(c64) * 2 * nT, hipMemcpyHostToDevice, s));
            // g_i = wJ/4 Phi psi_i into the lambda slots
            c64 *Gv = V + (long)(base + 4) * ld;
            launch_phi_apply(V, ld, dcols, dcols + 4, 4, rho, nT, 0.25 * wJ, N, Gv, ld, s);
            QD_HIP(hipStreamSynchronize(s));
            t_dens = secs(td);
            // ---- Sternheimer: lambda_i = (H - E_i)^+ P_i g_i, in place of g_i ----
            auto tsn = std::chrono::steady_clock::now();
            const std::vector<double> El = {w[lv[0]], w[lv[1]], w[lv[2]], w[lv[3]]};
            const std::vector<Level> lev = {{hole_level, 2}, {hole_level, 2}, {electron_level, 2}, {electron_level, 2}};
            res_max = sternheimer(h, Gv, 4, El, lev, &st_iters, s);
            t_st = secs(tsn);
            for (int j = 0; j < 4; ++j) {
                ta.push_back(base + 4 + j); tb.push_back(base + j); tg.push_back(-1.0);
                ta.push_back(base + j); tb.push_back(base + 4 + j); tg.push_back(-1.0);
            }
        }
        auto tb0 = std::chrono::steady_clock::now();
        std::vector<c64> Wh;
        adjoint_bin_device(h, V, ld, nv, ta, tb, tg, Wh);
        const double t_bin = secs(tb0);
        std::memcpy(W, Wh.data(), sizeof(c64) * Wh.size());
        if (st) {
            st->t_density = t_dens;
            st->t_sternheimer = t_st;
            st->t_bin = t_bin;
            st->t_total = secs(t0);
            st->sternheimer_iters = st_iters;
            st->sternheimer_residual = res_max;
            st->J = J;
        }
        return QD_OK;
    });
}

extern "C" int qd_adjoint_z(qd_handle *h, const double *gE, int ncols, const int *cols, const double *gZ, double *W,
                            qd_adjoint_stats *st) {
    if (!h || !W || ncols < 0 || (ncols > 0 && (!cols || !gZ))) { set_error("bad argument"); return QD_E_ARG; }
    if (h->k == 0 || !h->has_tables) { set_error("no solve on this handle"); return QD_E_ARG; }
    if (ncols > kAdjMaxZ) { set_error("at most 32 eigenvector gradients per call"); return QD_E_UNSUPPORTED; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        hipStream_t s = h->stream;
        const int n = h->n, k = h->k;
        const long ld = h->ld;
        if (n != h->tab.NH) throw StatusError("last solve was not of the uploaded tables", QD_E_ARG);
        auto t0 = std::chrono::steady_clock::now();
        std::vector<double> w(k);
        QD_HIP(hipMemcpyAsync(w.data(), h->wres, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        QD_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < ncols; ++j)
            if (cols[j] < 0 || cols[j] >= k) throw StatusError("column outside the solved window", QD_E_ARG);
        // vectors of D: the eigenvectors with energy weight, then psi_j and lambda_j of every column
        std::vector<int> eig;
        for (int i = 0; gE && i < k; ++i)
            if (gE[i] != 0.0) eig.push_back(i);
        const int base = (int)eig.size(), nv = base + 2 * ncols;
        if (nv == 0) throw StatusError("objective has no terms", QD_E_ARG);
        c64 *V = nullptr;
        QD_HIP(hipMalloc(&V, sizeof(c64) * ld * nv));
        struct FreeV { c64 *p; ~FreeV() { (void)hipFree(p); } } fv{V};
        QD_HIP(hipMemsetAsync(V, 0, sizeof(c64) * ld * nv, s));
        std::vector<int> ta, tb;
        std::vector<double> tg;
        for (int q = 0; q < base; ++q) {
            copy_cols(h->X + (long)eig[q] * ld, V + (long)q * ld, ld, n, 1, s);
            ta.push_back(q); tb.push_back(q); tg.push_back(gE[eig[q]]);
        }
        double res = 0.0;
        int iters = 0;
        auto tsn = std::chrono::steady_clock::now();
        if (ncols > 0) {
            c64 *Gv = V + (long)(base + ncols) * ld;
            QD_HIP(hipMemcpy2DAsync(Gv, sizeof(c64) * ld, gZ, sizeof(c64) * n, sizeof(c64) * n, ncols,
                                    hipMemcpyHostToDevice, s));
            // the degenerate level of each column: the window states within kLevelTol of its energy
            // (contiguous, the window being sorted)
            std::vector<double> El(ncols);
            std::vector<Level> lev(ncols);
            for (int j = 0; j < ncols; ++j) {
                copy_cols(h->X + (long)cols[j] * ld, V + (long)(base + j) * ld, ld, n, 1, s);
                const double e = w[cols[j]];
                int lo = cols[j], hi = cols[j];
                while (lo > 0 && std::fabs(w[lo - 1] - e) <= kLevelTol) --lo;
                while (hi + 1 < k && std::fabs(w[hi + 1] - e) <= kLevelTol) ++hi;
                if (hi - lo + 1 > 4) throw StatusError("more than four degenerate states in one level", QD_E_UNSUPPORTED);
                El[j] = e;
                lev[j] = Level{lo, hi - lo + 1};
            }
            res = sternheimer(h, Gv, ncols, El, lev, &iters, s);
            for (int j = 0; j < ncols; ++j) {
                ta.push_back(base + ncols + j); tb.push_back(base + j); tg.push_back(-1.0);
                ta.push_back(base + j); tb.push_back(base + ncols + j); tg.push_back(-1.0);
            }
        }
        const double t_st = std::chrono::duration<double>(std::chrono::steady_clock::now() - tsn).count();
        auto tb0 = std::chrono::steady_clock::now();
        std::vector<c64> Wh;
        adjoint_bin_device(h, V, ld, nv, ta, tb, tg, Wh);
        std::memcpy(W, Wh.data(), sizeof(c64) * Wh.size());
        if (st) {
            st->t_density = 0.0;
            st->t_sternheimer = t_st;
            st->t_bin = std::chrono::duration<double>(std::chrono::steady_clock::now() - tb0).count();
            st->t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            st->sternheimer_iters = iters;
            st->sternheimer_residual = res;
            st->J = 0.0;
        }
        return QD_OK;
    });
}

extern "C" int qd_adjoint_bin(qd_handle *h, int nv, const double *Vh, int nterms, const int *a, const int *b,
                              const double *gamma, double *W) {
    if (!h || !Vh || !a || !b || !gamma || !W || nv < 1 || nterms < 1) { set_error("bad argument"); return QD_E_ARG; }
    if (!h->has_tables) { set_error("no tables uploaded"); return QD_E_ARG; }
    return guarded([&] {
        QD_HIP(hipSetDevice(h->device));
        const long n = h->tab.NH;
        for (int q = 0; q < nterms; ++q)
            if (a[q] < 0 || a[q] >= nv || b[q] < 0 || b[q] >= nv) throw StatusError("term index out of range", QD_E_ARG);
        c64 *V = nullptr;
        QD_HIP(hipMalloc(&V, sizeof(c64) * n * nv));
        struct FreeV { c64 *p; ~FreeV() { (void)hipFree(p); } } fv{V};
        QD_HIP(hipMemcpyAsync(V, Vh, sizeof(c64) * n * nv, hipMemcpyHostToDevice, h->stream));
        std::vector<c64> Wh;
        adjoint_bin_device(h, V, n, nv, std::vector<int>(a, a + nterms), std::vector<int>(b, b + nterms),
                           std::vector<double>(gamma, gamma + nterms), Wh);
        std::memcpy(W, Wh.data(), sizeof(c64) * Wh.size());
        return QD_OK;
    });
}
