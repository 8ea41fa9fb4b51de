/*
 * qd_pwem -- drop-in driver of the QUDOS single-particle solve on one MI355X.
 *
 * Mirrors the solve path of the reference executable 8_band_pwem.ex (8_band_main.c): reads
 * parameters.in and supercell.in from the working directory with "-key=value" command-line
 * overrides (read_command_line_input.c), builds the FP64 Delta-G tables, solves the eigen
 * window of eigenstates.c:29-36 on the GPU, and writes
 *   energies_vb.dat, energies_cb.dat      (write_energies.c:23-84)
 *   momentum_matrix_elements.dat          (write_momentum_matrix_elements.c, when MME = 1)
 *   overlaps.dat                          (write_overlaps.c, when Overlaps = 1)
 *   carrier_localisation_{vb,cb}.dat      (write_carrier_localisation.c, when Localisation = 1)
 *   envelope_function_*.dat, probability_density_*.dat
 *                                         (write_envelope_functions*.c, write_probability_density*.c:
 *                                          linescans, cross-sections and the 3-D grid, when enabled)
 *   coulomb_matrix_elements_*.dat         (write_coulomb_matrix_elements.c, when CME >= 1)
 *   calculation_times.dat                 (write_calculation_times.c: 21 rows, one column)
 * into the working directory (or -Output_dir=PATH).
 *
 * Driver-only keys (not passed to the input reader):
 *   -Precision=0   FP64 reduction;  -Precision=1   FP32 reduction + FP64 refinement (default)
 *   -Device=N      GPU ordinal (default 0)
 *   -Output_dir=P  output directory (default .)
 *   -Refine_tol=R  stop of the FP64 refinement, max ||H x - E x|| / ||T||_1 (default 5e-12 when
 *                  Coulomb elements are written, so that they meet 1e-6 relative; else 1e-10)
 * Keys of the reference that select its CPU eigensolver (-Eigensolver=...) are accepted and
 * have no effect.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qudos_gpu.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int fail(const char *what, int rc) {
    fprintf(stderr, "qd_pwem: %s failed (%d): %s\n", what, rc, qd_last_error());
    return 1;
}

static FILE *open_out(const char *dir, const char *name) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "w");
    if (!f) fprintf(stderr, "qd_pwem: cannot write %s\n", path);
    return f;
}

/* Bloch character of state l: weight of each of the 8 bands (bloch_character.c:23-105). */
static void bloch(const double *Z, long nh, long ng, int l, double out[8]) {
    for (int b = 0; b < 8; ++b) {
        double s = 0.0;
        const double *z = Z + 2 * (l * nh + b * ng);
        for (long g = 0; g < ng; ++g) s += z[2 * g] * z[2 * g] + z[2 * g + 1] * z[2 * g + 1];
        out[b] = s;
    }
}

static void energy_row(FILE *f, double E, const double b[8]) {
    fprintf(f, "%17.10e\t%17.10e\t%17.10e\t%17.10e\t%17.10e\n", E, 100.0 * (b[3] + b[7]), 100.0 * (b[2] + b[6]),
            100.0 * (b[1] + b[5]), 100.0 * (b[0] + b[4]));
}

/* ---- real-space grids (Angstrom) and envelope files ---------------------------------------- */

/* n points through the centre of the cell along [100], [010], [001], [110], [101], [011], [111]
   (dir 0..6; real_space_grid_linescan.c:23-284): a diagonal takes the step and end points of its
   axis with the smallest step (the later axis on a tie); points accumulate from the first, the
   last is the end point. */
static void linescan_grid(int dir, int n, const double L[3], double *x, double *y, double *z) {
    static const int use[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    double lo[3], hi[3], d[3], *out[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        lo[a] = -0.5 * L[a];
        hi[a] = 0.5 * L[a];
        d[a] = (hi[a] - lo[a]) / ((double)n - 1.0);
    }
    int m = -1;
    if (dir >= 3 && dir <= 5) {
        const int a0 = dir == 5 ? 1 : 0, a1 = dir == 3 ? 1 : 2;
        m = d[a0] < d[a1] ? a0 : a1;
    } else if (dir == 6) {
        m = (d[0] < d[1] && d[0] < d[2]) ? 0 : (d[1] < d[2] ? 1 : 2);
    }
    for (int a = 0; a < 3; ++a) {
        if (!use[dir][a]) {
            for (int i = 0; i < n; ++i) out[a][i] = 0.0;
            continue;
        }
        const double l0 = m >= 0 ? lo[m] : lo[a], l1 = m >= 0 ? hi[m] : hi[a], da = m >= 0 ? d[m] : d[a];
        out[a][0] = l0;
        for (int i = 1; i < n - 1; ++i) out[a][i] = out[a][i - 1] + da;
        out[a][n - 1] = l1;
    }
}

/* n x n points (row i1, column i2) of the plane through the centre with normal [100], [010],
   [001], [-110], [110], [-101], [0-11] or [111] (dir 0..7; real_space_grid_cross_section.c:23-306);
   the (111) section is the expanded one of the reference. */
static void cross_section_grid(int dir, int n, const double L[3], double *x, double *y, double *z) {
    double lo[3], hi[3], d[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = -0.5 * L[a];
        hi[a] = 0.5 * L[a];
        d[a] = (hi[a] - lo[a]) / ((double)n - 1.0);
    }
    int a0 = -1, a1 = -1;   /* the two axes whose steps are made equal */
    if (dir == 3 || dir == 4 || dir == 7) { a0 = 0; a1 = 1; }
    if (dir == 5) { a0 = 0; a1 = 2; }
    if (dir == 6) { a0 = 1; a1 = 2; }
    if (a0 >= 0) {
        const int from = d[a0] < d[a1] ? a0 : a1, to = from == a0 ? a1 : a0;
        d[to] = d[from];
        lo[to] = lo[from];
        hi[to] = hi[from];
    }
    for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2) {
            const long q = (long)i1 * n + i2;
            double X = 0.0, Y = 0.0, Z = 0.0;
            switch (dir) {
                case 0: Y = lo[1] + i1 * d[1]; Z = lo[2] + i2 * d[2]; break;
                case 1: X = lo[0] + i1 * d[0]; Z = lo[2] + i2 * d[2]; break;
                case 2: X = lo[0] + i1 * d[0]; Y = lo[1] + i2 * d[1]; break;
                case 3: X = lo[0] + i1 * d[0]; Y = lo[1] + i1 * d[1]; Z = lo[2] + i2 * d[2]; break;
                case 4: X = hi[0] - i1 * d[0]; Y = lo[1] + i1 * d[1]; Z = lo[2] + i2 * d[2]; break;
                case 5: X = lo[0] + i1 * d[0]; Y = lo[1] + i2 * d[1]; Z = lo[2] + i1 * d[2]; break;
                case 6: X = lo[0] + i2 * d[0]; Y = lo[1] + i1 * d[1]; Z = lo[2] + i1 * d[2]; break;
                default:
                    Z = 2.0 * lo[2] + 2.0 * i2 * d[2];
                    Y = -2.0 * sqrt(0.25 * L[0] * L[0] + 0.25 * L[1] * L[1]) / sqrt(2.0) - Z / 2.0 + 2.0 * i1 * d[1];
                    X = -Y - Z;
            }
            x[q] = X;
            y[q] = Y;
            z[q] = Z;
        }
}

/* n^3 points, x slowest, z fastest (real_space_grid_3D.c:23-66) */
static void grid_3d(int n, const double L[3], double *x, double *y, double *z) {
    double lo[3], d[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = -0.5 * L[a];
        d[a] = L[a] / ((double)n - 1.0);
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const long q = ((long)i * n + j) * n + k;
                x[q] = lo[0] + i * d[0];
                y[q] = lo[1] + j * d[1];
                z[q] = lo[2] + k * d[2];
            }
}

/* One grid's files for every window state: envelope_function<kind>_{vb,cb}_<n><suffix> (x y z
   [nm], E, the SO, LH, HH, CB components of both spins, re / im, nm^-3/2) and
   probability_density<kind>_... (x y z, E, total, SO, LH, HH, CB densities, nm^-3)
   (write_envelope_functions*.c, write_probability_density*.c).  Valence states are numbered from
   the highest.  A blank line follows every `block` points (0: none).  cb_total: the factor of
   the total density of conduction states, 1000 (nm^-3) except in the 3-D files, where the
   reference writes 10 (write_probability_density_3D.c:51); kept so that the files match. */
static int write_envelope_files(const char *dir, const char *kind, const char *suffix, int block, int npts,
                                const double *x, const double *y, const double *z, const double *w, int nvb,
                                int ncb, const double *F, double cb_total) {
    static const int order[8] = {3, 7, 2, 6, 1, 5, 0, 4};   /* SO, LH, HH, CB; spin 1, spin 2 */
    const double s = sqrt(1000.0);
    for (int st = 0; st < nvb + ncb; ++st) {
        const int vb = st < nvb;
        const int num = vb ? nvb - st : st - nvb + 1;
        char name[256];
        snprintf(name, sizeof(name), "envelope_function%s_%s_%d%s", kind, vb ? "vb" : "cb", num, suffix);
        FILE *fe = open_out(dir, name);
        snprintf(name, sizeof(name), "probability_density%s_%s_%d%s", kind, vb ? "vb" : "cb", num, suffix);
        FILE *fp = open_out(dir, name);
        if (!fe || !fp) return 1;
        const double *Fs = F + 2 * (size_t)st * 8 * npts;
        for (int p = 0; p < npts; ++p) {
            fprintf(fe, "%13e\t%13e\t%13e\t%13e", 0.1 * x[p], 0.1 * y[p], 0.1 * z[p], w[st]);
            fprintf(fp, "%13e\t%13e\t%13e\t%13e", 0.1 * x[p], 0.1 * y[p], 0.1 * z[p], w[st]);
            double P[8];
            for (int b = 0; b < 8; ++b) {
                const double re = Fs[2 * ((size_t)b * npts + p)], im = Fs[2 * ((size_t)b * npts + p) + 1];
                P[b] = 1000.0 * (re * re + im * im);
            }
            for (int k = 0; k < 8; ++k) {
                const int b = order[k];
                fprintf(fe, "\t%13e\t%13e", s * Fs[2 * ((size_t)b * npts + p)], s * Fs[2 * ((size_t)b * npts + p) + 1]);
            }
            const double tot = (vb ? 1.0 : cb_total / 1000.0) * (P[0] + P[1] + P[2] + P[3] + P[4] + P[5] + P[6] + P[7]);
            fprintf(fp, "\t%13e\t%13e\t%13e\t%13e\t%13e", tot, P[3] + P[7], P[2] + P[6], P[1] + P[5], P[0] + P[4]);
            fprintf(fe, "\n");
            fprintf(fp, "\n");
            if (block > 0 && (p + 1) % block == 0) {
                fprintf(fe, "\n");
                fprintf(fp, "\n");
            }
        }
        fclose(fe);
        fclose(fp);
    }
    return 0;
}

/* envelopes of every window state on one grid, then its files */
static int envelope_grid(qd_handle *h, const qd_inputs *in, const char *dir, const char *kind, const char *suffix,
                         int block, int npts, const double *x, const double *y, const double *z, const double *w) {
    const int nvb = 2 * in->N_vb, ncb = 2 * in->N_cb;
    double *F = (double *)malloc(sizeof(double) * 2 * 8 * (size_t)(nvb + ncb) * npts);
    if (!F) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
    const int rc = qd_envelopes(h, in, 0, nvb + ncb, npts, x, y, z, F);
    if (rc != QD_OK) { free(F); return fail("qd_envelopes", rc); }
    const int r = write_envelope_files(dir, kind, suffix, block, npts, x, y, z, w, nvb, ncb, F,
                                       strcmp(kind, "_3D") == 0 ? 10.0 : 1000.0);
    free(F);
    return r;
}

int main(int argc, char **argv) {
    const double t_start = now_s();
    int precision = QD_FP32_REFINE, device = 0;
    double refine_tol = 0.0;
    const char *outdir = ".";
    const char **pass = (const char **)calloc((size_t)argc + 1, sizeof(char *));
    int npass = 0;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (strncmp(a, "-Precision=", 11) == 0) precision = atoi(a + 11) ? QD_FP32_REFINE : QD_FP64;
        else if (strncmp(a, "-Device=", 8) == 0) device = atoi(a + 8);
        else if (strncmp(a, "-Output_dir=", 12) == 0) outdir = a + 12;
        else if (strncmp(a, "-Refine_tol=", 12) == 0) refine_tol = atof(a + 12);
        else if (strcmp(a, "-h") == 0 || strcmp(a, "--help") == 0) {
            printf("usage: qd_pwem [-key=value ...] [-Precision=0|1] [-Device=N] [-Output_dir=PATH]\n"
                   "reads parameters.in and supercell.in from the working directory\n");
            free(pass);
            return 0;
        } else pass[npass++] = a;
    }
    qd_inputs in;
    int rc = qd_read_inputs("parameters.in", "supercell.in", npass, pass, &in);
    free(pass);
    if (rc != QD_OK) return fail("reading parameters.in / supercell.in", rc);

    double t0 = now_s();
    qd_tables tab;
    rc = qd_tables_build(&in, &tab);
    if (rc != QD_OK) return fail("coefficient tables", rc);
    const double t_tables = now_s() - t0;

    int il, iu;
    qd_window(&in, &il, &iu);
    const int k = iu - il + 1;
    if (precision == QD_FP32_REFINE && k > 40) precision = QD_FP64;
    const long ng = (long)in.N[0] * in.N[1] * in.N[2];
    const long nh = 8 * ng;
    printf("qd_pwem: N_H = %ld, eigenvalues %d..%d, %s\n", nh, il, iu,
           precision == QD_FP64 ? "FP64" : "FP32 reduction + FP64 refinement");

    qd_handle *h = NULL;
    if ((rc = qd_create(device, &h)) != QD_OK) return fail("qd_create", rc);
    if ((rc = qd_upload_tables(h, &tab)) != QD_OK) return fail("qd_upload_tables", rc);
    if (refine_tol <= 0.0) refine_tol = in.cme >= 1 ? 5e-12 : 1e-10;
    if ((rc = qd_set_refine_tolerance(h, refine_tol)) != QD_OK) return fail("qd_set_refine_tolerance", rc);
    double *w = (double *)malloc(sizeof(double) * k);
    double *Z = (double *)malloc(sizeof(double) * 2 * nh * k);
    if (!w || !Z) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
    qd_stats st;
    t0 = now_s();
    rc = qd_solve(h, il, iu, precision, w, Z, &st);
    /* the reference aborts when its eigensolver reports info != 0 (diagonalise_hamiltonian.c:279-293) */
    if (rc != QD_OK) return fail("qd_solve", rc);
    const double t_eig = now_s() - t0;

    /* energies: VB highest first, CB lowest first (write_energies.c:40-84) */
    t0 = now_s();
    const int nvb = 2 * in.N_vb, ncb = 2 * in.N_cb;
    FILE *fv = open_out(outdir, "energies_vb.dat");
    FILE *fc = open_out(outdir, "energies_cb.dat");
    if (!fv || !fc) return 1;
    double b[8];
    for (int i = nvb - 1; i >= 0; --i) { bloch(Z, nh, ng, i, b); energy_row(fv, w[i], b); }
    for (int i = nvb; i < nvb + ncb; ++i) { bloch(Z, nh, ng, i, b); energy_row(fc, w[i], b); }
    fclose(fv);
    fclose(fc);
    const double t_bloch = now_s() - t0;

    /* momentum matrix elements of every CB / VB pair, x y z and the [110] / [-110] combinations,
       divided by sqrt(HBSQM / 2) so that the squares are energies (write_momentum_matrix_elements.c:23-116) */
    double t_mme = 0;
    if (in.mme == 1) {
        t0 = now_s();
        double *pm = (double *)malloc(sizeof(double) * 2 * 3 * (size_t)ncb * nvb);
        if (!pm) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
        qd_momentum_stats mst;
        rc = qd_momentum(h, &in, nvb, ncb, pm, &mst);
        if (rc != QD_OK) return fail("qd_momentum", rc);
        FILE *fm = open_out(outdir, "momentum_matrix_elements.dat");
        if (!fm) return 1;
        const double scale = 1.0 / sqrt(7.61996 / 2.0);   /* HBSQM, 8_band.h:51 */
        for (int c = 0; c < ncb; ++c) {
            for (int v = 0; v < nvb; ++v) {
                fprintf(fm, "%d\t%d\t", c + 1, v + 1);
                double re[5], im[5];
                for (int a = 0; a < 3; ++a) {
                    const size_t q = 2 * (((size_t)a * ncb + c) * nvb + v);
                    re[a] = pm[q] * scale;
                    im[a] = pm[q + 1] * scale;
                }
                re[3] = (re[0] + re[1]) / sqrt(2.0);
                im[3] = (im[0] + im[1]) / sqrt(2.0);
                re[4] = (-re[0] + re[1]) / sqrt(2.0);
                im[4] = (-im[0] + im[1]) / sqrt(2.0);
                for (int a = 0; a < 5; ++a) fprintf(fm, "%17.10e\t%17.10e\t%17.10e\t", re[a], im[a], hypot(re[a], im[a]));
                fprintf(fm, "\n");
            }
            fprintf(fm, "\n");
        }
        fclose(fm);
        free(pm);
        t_mme = now_s() - t0;
    }

    /* overlaps <cb|vb> (write_overlaps.c:23-50) */
    double t_ov = 0;
    if (in.overlaps == 1) {
        t0 = now_s();
        double *ov = (double *)malloc(sizeof(double) * 2 * (size_t)ncb * nvb);
        if (!ov) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
        if ((rc = qd_overlaps(h, &in, ov)) != QD_OK) return fail("qd_overlaps", rc);
        FILE *fo = open_out(outdir, "overlaps.dat");
        if (!fo) return 1;
        for (int c = 0; c < ncb; ++c) {
            for (int v = 0; v < nvb; ++v)
                fprintf(fo, "%d\t%d\t%13e\t%13e\n", c, v, ov[2 * ((size_t)c * nvb + v)], ov[2 * ((size_t)c * nvb + v) + 1]);
            fprintf(fo, "\n");
        }
        fclose(fo);
        free(ov);
        t_ov = now_s() - t0;
    }

    /* carrier localisation in percent per shape, column 0 the rest of the cell
       (write_carrier_localisation.c:23-57) */
    double t_loc = 0;
    if (in.localisation == 1) {
        t0 = now_s();
        const int ns = in.N_shape;
        double *lv = (double *)malloc(sizeof(double) * (size_t)nvb * ns), *lc = (double *)malloc(sizeof(double) * (size_t)ncb * ns);
        if (!lv || !lc) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
        if ((rc = qd_localisation(h, &in, lv, lc, NULL)) != QD_OK) return fail("qd_localisation", rc);
        const char *names[2] = {"carrier_localisation_vb.dat", "carrier_localisation_cb.dat"};
        for (int f = 0; f < 2; ++f) {
            FILE *fl = open_out(outdir, names[f]);
            if (!fl) return 1;
            const int rows = f == 0 ? nvb : ncb;
            const double *m = f == 0 ? lv : lc;
            for (int r = 0; r < rows; ++r) {
                for (int s = 0; s < ns; ++s) fprintf(fl, "%13e\t", 100.0 * m[(size_t)r * ns + s]);
                fprintf(fl, "\n");
            }
            fclose(fl);
        }
        free(lv);
        free(lc);
        t_loc = now_s() - t0;
    }

    /* real-space envelope functions and probability densities (8_band_main.c:811-1100) */
    double t_line = 0, t_cross = 0, t_3d = 0;
    {
        const double L[3] = {in.shapes[0].L1, in.shapes[0].L2, in.shapes[0].L3};
        static const char *const line_names[7] = {"_100.dat", "_010.dat", "_001.dat", "_110.dat", "_101.dat", "_011.dat", "_111.dat"};
        static const char *const cross_names[8] = {"_100.dat", "_010.dat", "_001.dat", "_-110.dat", "_110.dat", "_-101.dat",
                                                   "_0-11.dat", "_111.dat"};
        long cap = 0;
        for (int dd = 0; dd < 7; ++dd)
            if (in.linescan[dd] == 1 && in.N_line > cap) cap = in.N_line;
        if (in.cross_sections == 1 && (long)in.N_cross * in.N_cross > cap) cap = (long)in.N_cross * in.N_cross;
        if (in.out3d == 1 && (long)in.N_3D * in.N_3D * in.N_3D > cap) cap = (long)in.N_3D * in.N_3D * in.N_3D;
        double *gx = cap > 0 ? (double *)malloc(sizeof(double) * 3 * cap) : NULL;
        if (cap > 0 && !gx) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
        double *gy = gx + cap, *gz = gx + 2 * cap;
        t0 = now_s();
        for (int dd = 0; dd < 7; ++dd) {
            if (in.linescan[dd] != 1 || in.N_line < 2) continue;
            linescan_grid(dd, in.N_line, L, gx, gy, gz);
            if (envelope_grid(h, &in, outdir, "", line_names[dd], 0, in.N_line, gx, gy, gz, w)) return 1;
        }
        t_line = now_s() - t0;
        t0 = now_s();
        if (in.cross_sections == 1 && in.N_cross >= 2)
            for (int dd = 0; dd < 8; ++dd) {
                if (in.cross_section[dd] != 1) continue;
                cross_section_grid(dd, in.N_cross, L, gx, gy, gz);
                if (envelope_grid(h, &in, outdir, "_cross_section", cross_names[dd], in.N_cross, in.N_cross * in.N_cross,
                                  gx, gy, gz, w))
                    return 1;
            }
        t_cross = now_s() - t0;
        t0 = now_s();
        if (in.out3d == 1 && in.N_3D >= 2) {
            grid_3d(in.N_3D, L, gx, gy, gz);
            if (envelope_grid(h, &in, outdir, "_3D", ".dat", in.N_3D, in.N_3D * in.N_3D * in.N_3D, gx, gy, gz, w)) return 1;
        }
        t_3d = now_s() - t0;
        free(gx);
    }

    double t_cme_b = 0, t_cme_i = 0, t_cme = 0;
    if (in.cme >= 1) {
        const size_t n4h = (size_t)nvb * nvb * nvb * nvb, n4e = (size_t)ncb * ncb * ncb * ncb;
        const size_t neh = (size_t)nvb * nvb * ncb * ncb;
        double *hh = (double *)malloc(sizeof(double) * 2 * n4h), *ee = (double *)malloc(sizeof(double) * 2 * n4e);
        double *ed = (double *)malloc(sizeof(double) * 2 * neh), *ex = (double *)malloc(sizeof(double) * 2 * neh);
        if (!hh || !ee || !ed || !ex) { fprintf(stderr, "qd_pwem: out of host memory\n"); return 1; }
        qd_coulomb_stats cst;
        rc = qd_coulomb(h, &in, &tab, nvb, ncb, NULL, hh, ee, ed, ex, &cst);
        if (rc != QD_OK) return fail("qd_coulomb", rc);
        rc = qd_write_coulomb(outdir, nvb, ncb, hh, ee, ed, ex);
        if (rc != QD_OK) return fail("qd_write_coulomb", rc);
        t_cme_b = cst.t_pairs;
        t_cme_i = cst.t_contract;
        t_cme = cst.t_total;
        free(hh); free(ee); free(ed); free(ex);
    }
    qd_destroy(h);

    /* calculation_times.dat rows (write_calculation_times.c:53-233); the table build is one
       host stage here and is reported in the convolution row */
    const double rows[21] = {0, 0, 0, 0, 0, 0, t_tables, st.t_hamiltonian,
                             st.t_reduce + st.t_tridiag + st.t_backtransform + st.t_refine, t_eig, t_bloch,
                             t_line, t_cross, t_3d, t_ov, t_loc, t_mme, t_cme_b, t_cme_i, t_cme, now_s() - t_start};
    FILE *ft = open_out(outdir, "calculation_times.dat");
    if (!ft) return 1;
    for (int r = 0; r < 21; ++r) fprintf(ft, "%17.10e\n", rows[r]);
    fclose(ft);
    printf("qd_pwem: E_h1 = %.10f eV, E_e1 = %.10f eV, solve %.3f s (refine iterations %d)\n", w[nvb - 1], w[nvb],
           t_eig, st.refine_iters);
    free(w);
    free(Z);
    qd_tables_free(&tab);
    qd_free_inputs(&in);
    return 0;
}
