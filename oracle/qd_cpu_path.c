/*
 * qd_cpu_path.c -- CPU restatement of the reference's single-particle solve (TEST AND
 * MEASUREMENT INFRASTRUCTURE ONLY: the shipped library never links or calls it).
 *
 * It keeps the reference's loop structure so that its run time is that of the reference's CPU
 * path on the same host:
 *   qc_hamiltonian : hamiltonian_8_band (hamiltonian_8_band.c:23-500): an OpenMP collapse(2)
 *                    guided loop over every (row, column) of H that fills the lower triangle, one
 *                    element-function call per entry; each element function allocates its
 *                    per-shape parameter arrays and sums the shapes per call
 *                    (hamiltonian_8_band_matrix_elements.c:27-969, parameter_matrix_element.c:23-48,
 *                    field_fourier_coefficient.c:23-55).
 *   qc_zheevr      : diagonalise_hamiltonian (diagonalise_hamiltonian.c:55-77), LAPACK zheevr with
 *                    JOBZ='V', RANGE='I', UPLO='L', abstol = dlamch('S'), from Intel MKL (the
 *                    reference's library) loaded at run time with GNU OpenMP threading.
 * Inputs are the coefficient arrays of the numpy oracle (per-shape chi, strain, per-shape
 * convolutions, piezo potential) on the Delta-G grid, index ((dx NT_y + dy) NT_z + dz).
 */
#define _GNU_SOURCE
#include <complex.h>
#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define HBSQM 7.61996   /* hbar^2 / m0, eV Angstrom^2 (8_band.h:51) */

typedef struct {
    double E_cb, E_hh, E_lh, E_so, sc, g1, g2, g3, P, ac, av, b, d;
} qc_par;

/* conv[s] points to 8 consecutive Delta-G arrays: xx yy zz yz xz xy hydrostatic biaxial */
typedef struct {
    int N[3], NT[3];
    const double *G[3];
    double K[3], F[3];
    int N_shape;
    const int *embedded;
    const qc_par *par;
    const double complex *const *chi;
    int strain, piezo;
    const double complex *e[6];
    const double complex *const *conv;
    const double complex *phi;
} qc_problem;

enum { XX = 0, YY, ZZ, YZ, XZ, XY, HYD, BIA };

typedef struct {
    int g1[3], g2[3];      /* plane-wave indices of row and column */
    int b1[3], b2[3];      /* their positions in the basis arrays */
    long didx;             /* Delta-G index */
} qc_pair;

static double complex pme(const qc_problem *p, const qc_pair *q, const double *par) {
    double complex r = (q->g1[0] == q->g2[0] && q->g1[1] == q->g2[1] && q->g1[2] == q->g2[2]) ? par[0] : 0.0;
    for (int s = 1; s < p->N_shape; ++s) r += (par[s] - par[p->embedded[s]]) * p->chi[s][q->didx];
    return r;
}

static double kk(const qc_problem *p, const qc_pair *q, int a) {
    const double G1 = p->G[a][q->b1[a]], G2 = p->G[a][q->b2[a]], K = p->K[a];
    return G1 * G2 + K * (G1 + G2) + K * K;
}
static double ks(const qc_problem *p, const qc_pair *q, int a) {
    return p->G[a][q->b1[a]] + p->G[a][q->b2[a]] + 2.0 * p->K[a];
}
static double kx(const qc_problem *p, const qc_pair *q, int a, int b) {
    return (p->G[a][q->b1[a]] + p->K[a]) * (p->G[b][q->b2[b]] + p->K[b]) +
           (p->G[b][q->b1[b]] + p->K[b]) * (p->G[a][q->b2[a]] + p->K[a]);
}

static double complex field(const qc_problem *p, const qc_pair *q) {
    double complex r = 0.0;
    for (int a = 0; a < 3; ++a) {
        const int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
        if (q->b1[a] != q->b2[a] && q->b1[o1] == q->b2[o1] && q->b1[o2] == q->b2[o2]) {
            const int dm = q->b1[a] - q->b2[a];
            r += I * p->F[a] * ((dm + 1) % 2 == 0 ? 1.0 : -1.0) / (p->G[a][q->b2[a]] - p->G[a][q->b1[a]]);
        }
    }
    return r;
}

/* per-shape sum of (par_s - par_emb) * conv_s[c] */
static double complex shape_conv(const qc_problem *p, const qc_pair *q, const double *par, int c) {
    double complex r = 0.0;
    for (int s = 1; s < p->N_shape; ++s)
        r += (par[s] - par[p->embedded[s]]) * p->conv[s][(long)c * p->NT[0] * p->NT[1] * p->NT[2] + q->didx];
    return r;
}

#define PARAM_ARRAY(name, expr)                                              \
    double *name = (double *)malloc(sizeof(double) * p->N_shape);            \
    for (int s = 0; s < p->N_shape; ++s) { const qc_par *m = &p->par[s]; name[s] = (expr); }

static double complex strain_diag(const qc_problem *p, const qc_pair *q, int band) {
    if (!p->strain) return 0.0;
    const long i = q->didx;
    const double complex hyd = p->e[XX][i] + p->e[YY][i] + p->e[ZZ][i];
    const double complex bia = p->e[XX][i] + p->e[YY][i] - 2.0 * p->e[ZZ][i];
    PARAM_ARRAY(ac, m->ac)
    PARAM_ARRAY(av, m->av)
    PARAM_ARRAY(bb, m->b)
    const qc_par *m0 = &p->par[0];
    double complex r;
    if (band == 0) r = m0->ac * hyd + shape_conv(p, q, ac, HYD);
    else if (band == 1) r = m0->av * hyd + 0.5 * m0->b * bia + shape_conv(p, q, av, HYD) + 0.5 * shape_conv(p, q, bb, BIA);
    else if (band == 2) r = m0->av * hyd - 0.5 * m0->b * bia + shape_conv(p, q, av, HYD) - 0.5 * shape_conv(p, q, bb, BIA);
    else r = m0->av * hyd + shape_conv(p, q, av, HYD);
    if (p->piezo && p->phi) r += p->phi[i];
    free(ac); free(av); free(bb);
    return r;
}

/* E_CB, E_HH, E_LH, E_SO (hamiltonian_8_band_matrix_elements.c:27-397) */
static double complex E_CB(const qc_problem *p, const qc_pair *q) {
    PARAM_ARRAY(E, m->E_cb)
    PARAM_ARRAY(sc, m->sc)
    double complex r = pme(p, q, E) + 0.5 * HBSQM * pme(p, q, sc) * (kk(p, q, 0) + kk(p, q, 1) + kk(p, q, 2));
    r += strain_diag(p, q, 0) + field(p, q);
    free(E); free(sc);
    return r;
}
static double complex E_HH(const qc_problem *p, const qc_pair *q) {
    PARAM_ARRAY(E, m->E_hh)
    PARAM_ARRAY(a, m->g1 + m->g2)
    PARAM_ARRAY(c, m->g1 - 2.0 * m->g2)
    double complex r = pme(p, q, E) - 0.5 * HBSQM * pme(p, q, a) * (kk(p, q, 0) + kk(p, q, 1)) -
                       0.5 * HBSQM * pme(p, q, c) * kk(p, q, 2);
    r += strain_diag(p, q, 1) + field(p, q);
    free(E); free(a); free(c);
    return r;
}
static double complex E_LH(const qc_problem *p, const qc_pair *q) {
    PARAM_ARRAY(E, m->E_lh)
    PARAM_ARRAY(a, m->g1 - m->g2)
    PARAM_ARRAY(c, m->g1 + 2.0 * m->g2)
    double complex r = pme(p, q, E) - 0.5 * HBSQM * pme(p, q, a) * (kk(p, q, 0) + kk(p, q, 1)) -
                       0.5 * HBSQM * pme(p, q, c) * kk(p, q, 2);
    r += strain_diag(p, q, 2) + field(p, q);
    free(E); free(a); free(c);
    return r;
}
static double complex E_SO(const qc_problem *p, const qc_pair *q) {
    PARAM_ARRAY(E, m->E_so)
    PARAM_ARRAY(g1, m->g1)
    double complex r = pme(p, q, E) - 0.5 * HBSQM * pme(p, q, g1) * (kk(p, q, 0) + kk(p, q, 1) + kk(p, q, 2));
    r += strain_diag(p, q, 3) + field(p, q);
    free(E); free(g1);
    return r;
}

/* Q (:404-468) */
static double complex Qe(const qc_problem *p, const qc_pair *q) {
    PARAM_ARRAY(g2, m->g2)
    double complex r = -sqrt(0.5) * HBSQM * pme(p, q, g2) * (kk(p, q, 0) + kk(p, q, 1)) +
                       sqrt(2.0) * HBSQM * pme(p, q, g2) * kk(p, q, 2);
    if (p->strain) {
        PARAM_ARRAY(bb, m->b)
        const long i = q->didx;
        r += sqrt(0.5) * p->par[0].b * (p->e[XX][i] + p->e[YY][i] - 2.0 * p->e[ZZ][i]) +
             sqrt(0.5) * shape_conv(p, q, bb, BIA);
        free(bb);
    }
    free(g2);
    return r;
}

/* R and R_conj (:475-616); sgn = +1 for R, -1 for R_conj */
static double complex Re_(const qc_problem *p, const qc_pair *q, double sgn) {
    PARAM_ARRAY(g2, m->g2)
    PARAM_ARRAY(g3, m->g3)
    double complex r = 0.5 * sqrt(3.0) * HBSQM * pme(p, q, g2) * (kk(p, q, 0) - kk(p, q, 1)) -
                       sgn * 0.5 * I * sqrt(3.0) * HBSQM * pme(p, q, g3) * kx(p, q, 0, 1);
    if (p->strain) {
        PARAM_ARRAY(bb, m->b)
        PARAM_ARRAY(dd, m->d)
        const long i = q->didx;
        const double complex rb = -0.5 * sqrt(3.0) * (p->par[0].b * (p->e[XX][i] - p->e[YY][i]) +
                                                      shape_conv(p, q, bb, XX) - shape_conv(p, q, bb, YY));
        const double complex rd = p->par[0].d * p->e[XY][i] + shape_conv(p, q, dd, XY);
        r += rb + sgn * I * rd;
        free(bb); free(dd);
    }
    free(g2); free(g3);
    return r;
}

/* S and S_conj (:623-758) */
static double complex Se_(const qc_problem *p, const qc_pair *q, double sgn) {
    PARAM_ARRAY(g3, m->g3)
    double complex r = 0.5 * sqrt(1.5) * HBSQM * pme(p, q, g3) * kx(p, q, 0, 2) -
                       sgn * 0.5 * I * sqrt(1.5) * HBSQM * pme(p, q, g3) * kx(p, q, 1, 2);
    if (p->strain) {
        PARAM_ARRAY(dd, m->d)
        const long i = q->didx;
        const double complex sxz = p->par[0].d * p->e[XZ][i] + shape_conv(p, q, dd, XZ);
        const double complex syz = p->par[0].d * p->e[YZ][i] + shape_conv(p, q, dd, YZ);
        r += -sqrt(0.5) * (sxz - sgn * I * syz);
        free(dd);
    }
    free(g3);
    return r;
}

/* P-weighted strain e_ij (matrix term + shape terms) */
static double complex pstrain(const qc_problem *p, const qc_pair *q, const double *P, int c) {
    return p->par[0].P * p->e[c][q->didx] + shape_conv(p, q, P, c);
}

/* T_plus, T_minus (:765-899); sgn = +1 for T_plus */
static double complex Te_(const qc_problem *p, const qc_pair *q, double sgn) {
    PARAM_ARRAY(P, m->P)
    const double complex pm = pme(p, q, P);
    double complex r = pm * ks(p, q, 0) / (2.0 * sqrt(6.0)) + sgn * I * pm * ks(p, q, 1) / (2.0 * sqrt(6.0));
    if (p->strain) {
        const double complex pxx = pstrain(p, q, P, XX), pyy = pstrain(p, q, P, YY), pxy = pstrain(p, q, P, XY);
        const double complex pxz = pstrain(p, q, P, XZ), pyz = pstrain(p, q, P, YZ);
        r += (-0.5 / sqrt(6.0)) * ((pxx + sgn * I * pxy) * ks(p, q, 0) + (pxy + sgn * I * pyy) * ks(p, q, 1) +
                                   (pxz + sgn * I * pyz) * ks(p, q, 2));
    }
    free(P);
    return r;
}

/* U (:906-969) */
static double complex Ue(const qc_problem *p, const qc_pair *q) {
    PARAM_ARRAY(P, m->P)
    double complex r = pme(p, q, P) * ks(p, q, 2) / (2.0 * sqrt(3.0));
    if (p->strain)
        r += (-0.5 / sqrt(3.0)) * (pstrain(p, q, P, XZ) * ks(p, q, 0) + pstrain(p, q, P, YZ) * ks(p, q, 1) +
                                   pstrain(p, q, P, ZZ) * ks(p, q, 2));
    free(P);
    return r;
}

/* Lower-block table of hamiltonian_8_band.c:78-492 (0-based row band, column band). */
static double complex element(const qc_problem *p, const qc_pair *q, int br, int bc) {
    const double r2 = sqrt(2.0), r3 = sqrt(3.0);
    switch (bc) {
    case 0:
        switch (br) {
        case 0: return E_CB(p, q);
        case 1: return -r3 * Te_(p, q, -1.0);
        case 2: return r2 * Ue(p, q);
        case 3: return -Ue(p, q);
        case 6: return -Te_(p, q, 1.0);
        case 7: return -r2 * Te_(p, q, 1.0);
        default: return 0.0;
        }
    case 1:
        switch (br) {
        case 1: return E_HH(p, q);
        case 2: return r2 * Se_(p, q, -1.0);
        case 3: return -Se_(p, q, -1.0);
        case 6: return -Re_(p, q, -1.0);
        case 7: return -r2 * Re_(p, q, -1.0);
        default: return 0.0;
        }
    case 2:
        switch (br) {
        case 2: return E_LH(p, q);
        case 3: return Qe(p, q);
        case 4: return Te_(p, q, 1.0);
        case 5: return Re_(p, q, -1.0);
        case 7: return r3 * Se_(p, q, -1.0);
        default: return 0.0;
        }
    case 3:
        switch (br) {
        case 3: return E_SO(p, q);
        case 4: return r2 * Te_(p, q, 1.0);
        case 5: return r2 * Re_(p, q, -1.0);
        case 6: return -r3 * Se_(p, q, -1.0);
        default: return 0.0;
        }
    case 4:
        switch (br) {
        case 4: return E_CB(p, q);
        case 5: return -r3 * Te_(p, q, 1.0);
        case 6: return r2 * Ue(p, q);
        case 7: return -Ue(p, q);
        default: return 0.0;
        }
    case 5:
        switch (br) {
        case 5: return E_HH(p, q);
        case 6: return r2 * Se_(p, q, 1.0);
        case 7: return -Se_(p, q, 1.0);
        default: return 0.0;
        }
    case 6:
        switch (br) {
        case 6: return E_LH(p, q);
        case 7: return Qe(p, q);
        default: return 0.0;
        }
    default:
        return br == 7 ? E_SO(p, q) : 0.0;
    }
}

/* Lower triangle of H for the columns cols[0..ncols-1] (NULL: every column 0..N_H-1): entry (row,
 * cols[j]) with row >= cols[j] goes to H[j N_H + row].  Every (row, listed column) pair is visited,
 * as the reference's loop visits every (row, column) of its matrix and skips the upper ones. */
int qc_hamiltonian(const qc_problem *p, double complex *H, const int *cols, int ncols, int threads) {
    const int NG = p->N[0] * p->N[1] * p->N[2];
    const int NH = 8 * NG;
    if (!cols) ncols = NH;
    for (int j = 0; cols && j < ncols; ++j)
        if (cols[j] < 0 || cols[j] >= NH) return -1;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for collapse(2) schedule(guided)
    for (int row = 0; row < NH; ++row)
        for (int j = 0; j < ncols; ++j) {
            const int col = cols ? cols[j] : j;
            if (row < col) continue;
            qc_pair q;
            const int m1 = row % NG, m2 = col % NG;
            q.g1[0] = m1 / (p->N[1] * p->N[2]);
            q.g1[1] = (m1 % (p->N[1] * p->N[2])) / p->N[2];
            q.g1[2] = m1 % p->N[2];
            q.g2[0] = m2 / (p->N[1] * p->N[2]);
            q.g2[1] = (m2 % (p->N[1] * p->N[2])) / p->N[2];
            q.g2[2] = m2 % p->N[2];
            for (int a = 0; a < 3; ++a) {
                q.b1[a] = p->NT[a] / 2 - p->N[a] / 2 + q.g1[a];
                q.b2[a] = p->NT[a] / 2 - p->N[a] / 2 + q.g2[a];
            }
            q.didx = (p->NT[2] / 2 + q.g1[2] - q.g2[2]) +
                     (long)p->NT[2] * (p->NT[1] / 2 + q.g1[1] - q.g2[1]) +
                     (long)p->NT[1] * p->NT[2] * (p->NT[0] / 2 + q.g1[0] - q.g2[0]);
            H[(long)j * NH + row] = element(p, &q, row / NG, col / NG);
        }
    return 0;
}

/* Entries H[rows[i], cols[i]] (rows[i] >= cols[i], the stored lower triangle) into out[i]. */
int qc_hamiltonian_entries(const qc_problem *p, const int *rows, const int *cols, long n, double complex *out,
                           int threads) {
    const int NG = p->N[0] * p->N[1] * p->N[2];
    const int NH = 8 * NG;
    for (long i = 0; i < n; ++i)
        if (rows[i] < cols[i] || rows[i] >= NH || cols[i] < 0) return -1;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < n; ++i) {
        const int row = rows[i], col = cols[i];
        qc_pair q;
        const int m1 = row % NG, m2 = col % NG;
        q.g1[0] = m1 / (p->N[1] * p->N[2]);
        q.g1[1] = (m1 % (p->N[1] * p->N[2])) / p->N[2];
        q.g1[2] = m1 % p->N[2];
        q.g2[0] = m2 / (p->N[1] * p->N[2]);
        q.g2[1] = (m2 % (p->N[1] * p->N[2])) / p->N[2];
        q.g2[2] = m2 % p->N[2];
        for (int a = 0; a < 3; ++a) {
            q.b1[a] = p->NT[a] / 2 - p->N[a] / 2 + q.g1[a];
            q.b2[a] = p->NT[a] / 2 - p->N[a] / 2 + q.g2[a];
        }
        q.didx = (p->NT[2] / 2 + q.g1[2] - q.g2[2]) + (long)p->NT[2] * (p->NT[1] / 2 + q.g1[1] - q.g2[1]) +
                 (long)p->NT[1] * p->NT[2] * (p->NT[0] / 2 + q.g1[0] - q.g2[0]);
        out[i] = element(p, &q, row / NG, col / NG);
    }
    return 0;
}

/* ---- LAPACK zheevr from Intel MKL (libmkl_rt, LP64 interface), loaded at run time ---- */
typedef void (*zheevr_fn)(const char *, const char *, const char *, const int *, double complex *, const int *,
                          const double *, const double *, const int *, const int *, const double *, int *, double *,
                          double complex *, const int *, int *, double complex *, const int *, double *, const int *,
                          int *, const int *, int *, size_t, size_t, size_t);
typedef double (*dlamch_fn)(const char *, size_t);
typedef void (*setthreads_fn)(const int *);

static void *g_mkl = NULL;
static char g_mkl_path[512] = "";

/* Loads MKL from the given path (or the usual places); returns 0 on success. */
int qc_load_mkl(const char *path) {
    if (g_mkl) return 0;
    setenv("MKL_THREADING_LAYER", "GNU", 0);   /* share the gcc OpenMP runtime of the H fill */
    setenv("MKL_INTERFACE_LAYER", "LP64", 0);
    const char *cands[] = {path, "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libmkl_rt.so.1", NULL};
    for (int i = 0; i < 4; ++i) {
        if (!cands[i]) continue;
        g_mkl = dlopen(cands[i], RTLD_NOW | RTLD_GLOBAL);
        if (g_mkl) { snprintf(g_mkl_path, sizeof(g_mkl_path), "%s", cands[i]); return 0; }
    }
    return -1;
}

const char *qc_mkl_path(void) { return g_mkl_path; }

/* Eigenpairs il..iu (1-based) of the Hermitian matrix whose lower triangle is in H (column-major,
 * n x n, overwritten), as diagonalise_hamiltonian.c:55-77 calls zheevr.  Returns LAPACK's info, or
 * -1000 when MKL is not loadable. */
int qc_zheevr(int n, double complex *H, int il, int iu, double *w, double complex *Z, int threads) {
    if (qc_load_mkl(NULL) != 0) return -1000;
    zheevr_fn zheevr = (zheevr_fn)dlsym(g_mkl, "zheevr_");
    dlamch_fn dlamch = (dlamch_fn)dlsym(g_mkl, "dlamch_");
    setthreads_fn setth = (setthreads_fn)dlsym(g_mkl, "MKL_Set_Num_Threads");
    if (!zheevr || !dlamch) return -1001;
    if (setth && threads > 0) setth(&threads);
    const double abstol = dlamch("S", 1);
    const double vl = 0.0, vu = 0.0;
    int m = 0, info = 0, lwork = -1, lrwork = -1, liwork = -1;
    double complex wq;
    double rq;
    int iq;
    int *isuppz = (int *)malloc(sizeof(int) * 2 * (size_t)(iu - il + 1));
    double *wall = (double *)malloc(sizeof(double) * (size_t)n);
    zheevr("V", "I", "L", &n, H, &n, &vl, &vu, &il, &iu, &abstol, &m, wall, Z, &n, isuppz, &wq, &lwork, &rq,
           &lrwork, &iq, &liwork, &info, 1, 1, 1);
    if (info == 0) {
        lwork = (int)creal(wq);
        lrwork = (int)rq;
        liwork = iq;
        double complex *work = (double complex *)malloc(sizeof(double complex) * (size_t)lwork);
        double *rwork = (double *)malloc(sizeof(double) * (size_t)lrwork);
        int *iwork = (int *)malloc(sizeof(int) * (size_t)liwork);
        zheevr("V", "I", "L", &n, H, &n, &vl, &vu, &il, &iu, &abstol, &m, wall, Z, &n, isuppz, work, &lwork, rwork,
               &lrwork, iwork, &liwork, &info, 1, 1, 1);
        free(work); free(rwork); free(iwork);
        memcpy(w, wall, sizeof(double) * (size_t)(iu - il + 1));
    }
    free(isuppz);
    free(wall);
    return info;
}
