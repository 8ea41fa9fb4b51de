/*
 * qudos_gpu.h -- C ABI of the MI355X single-particle solve for QUDOS.
 *
 * Plain C types only (pointers, sizes, doubles); complex numbers travel as interleaved
 * (re, im) double pairs.  Every entry point returns QD_OK (0) or a negative QD_E* status and
 * never aborts the process.  A handle binds one GPU and owns device buffers; it is used by one
 * host thread at a time and all of its work is ordered on the handle's stream.
 *
 * Reference interfaces replaced (QUDOS, code/source/8_band):
 *   qd_read_inputs     <- read_parameters (read_parameters.c:23), read_command_line_input
 *                         (read_command_line_input.c:23), read_supercell (read_supercell.c:22)
 *   qd_tables_build    <- material_parameters (material_parameters.c:23), basis_set (basis_set.c:23),
 *                         characteristic_functions_reciprocal_space (..._reciprocal_space.c:24),
 *                         strain_region_reciprocal_space (strain_region_reciprocal_space.c:22),
 *                         piezo_region_reciprocal_space (piezo_region_reciprocal_space.c:22),
 *                         strain_reciprocal_space (strain_reciprocal_space.c:22),
 *                         piezo_reciprocal_space (piezo_reciprocal_space.c:22), convolutions
 *                         (convolutions.c:24); the globals of 8_band.h:218-225 become qd_tables
 *   qd_hamiltonian     <- hamiltonian_8_band (8_band.h:729, hamiltonian_8_band.c:23)
 *   qd_solve           <- eigenstates (8_band.h:720, eigenstates.c:23) with diagonalise_hamiltonian
 *                         (8_band.h:855, diagonalise_hamiltonian.c:24) and collect_eigenstates
 *                         (collect_eigenstates.c:23)
 *   qd_coulomb         <- coulomb_matrix_elements (8_band.h:1059, coulomb_matrix_elements.c:24)
 *   qd_adjoint         <- no reference counterpart (new: gradients of E_n and J_eh)
 *   qd_adjoint_z       <- no reference counterpart (new: gradients of E_n and of the eigenvectors)
 */
#ifndef QUDOS_GPU_H
#define QUDOS_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define QD_OK 0
#define QD_E_ARG -1        /* invalid argument / inconsistent sizes */
#define QD_E_INPUT -2      /* malformed or missing input file / key */
#define QD_E_ALLOC -3      /* host or device allocation failed */
#define QD_E_DEVICE -4     /* HIP runtime error */
#define QD_E_NOCONV -5     /* bisection / inverse iteration / refinement did not converge */
#define QD_E_UNSUPPORTED -6/* option not implemented (message says which) */
#define QD_E_NAN -7        /* non-finite value in inputs or tables */

/* ---- precision of the O(N^3) reduction ---- */
#define QD_FP64 0          /* FP64 Householder reduction + FP64 back-transform */
#define QD_FP32_REFINE 1   /* FP32 MFMA reduction, FP64 bisection/inverse iteration, FP64 refinement */

/* One shape of supercell.in (read_supercell.c:159-303); lengths in Angstrom, angles in degrees.
 * shapes[0] is the supercell itself (L1..L3 = cell lengths). */
typedef struct {
    int shape, embedded, material;
    double x, y;
    double Ox, Oy, Oz;
    double L1, L2, L3, L4, L5;
    double alpha, beta, gamma;
} qd_shape;

/* The solve-path subset of parameters.in / supercell.in / -key=value overrides. */
typedef struct {
    int N_shape;
    qd_shape *shapes;          /* caller- or qd_read_inputs-owned array of N_shape */
    int N[3];                  /* plane waves per axis, N = 2M+1 */
    int N_S[3];                /* strain-cell plane waves */
    int N_C[3];                /* Coulomb-cell plane waves */
    int strain_type;           /* 0 free strain-cell lengths, 1 relative (integer multiples) */
    int coulomb_type;
    int RL[3], RCL[3];
    double SL[3], CL[3];       /* strain / Coulomb cell lengths, Angstrom */
    int N_vb, N_cb;            /* Kramers levels below / above the gap */
    double T;                  /* K */
    double F[3];               /* V/Angstrom */
    double K[3];               /* 1/Angstrom */
    int strain, isotropic, piezo, cme, softening;
    double sigma_softening;    /* Angstrom */
    int madelung_ranks;        /* MPI rank count whose split of the real-space Madelung sum
                                  (coulomb_matrix_elements.c:287-305) is reproduced: rank 0 adds
                                  no real-space term, so the result depends on it; default 4
                                  (the reference example run, run_test_pwem.sh:33); key
                                  "Madelung_ranks" */
    int mme;                   /* momentum matrix elements requested (key "MME"; the driver writes
                                  momentum_matrix_elements.dat) */
    int overlaps;              /* key "Overlaps": the driver writes overlaps.dat */
    int localisation;          /* key "Localisation": the driver writes carrier_localisation_*.dat */
    /* real-space envelope outputs (read_parameters.c:238, 347-476; the driver writes them) */
    int N_line;                /* key "N_line": points per linescan */
    int linescan[7];           /* keys "Linescan_100", _010, _001, _110, _101, _011, _111 */
    int cross_sections;        /* key "Cross_sections" */
    int N_cross;               /* key "N_cross": points per side of a cross-section */
    int cross_section[8];      /* keys "Cross_section_100", _010, _001, _m110, _110, _m101, _0m11, _111 */
    int out3d;                 /* key "3D_output" */
    int N_3D;                  /* key "N_3D": points per axis of the 3-D grid */
} qd_inputs;

/* Delta-G tables feeding the Hamiltonian kernel: QD_NTAB complex tables on the
 * (2N_x-1)(2N_y-1)(2N_z-1) grid, index ((dx*NT_y)+dy)*NT_z+dz with d = m_row - m_col + NT/2. */
enum {
    QD_T_ECB = 0, QD_T_EHH, QD_T_ELH, QD_T_ESO,     /* band edge + strain + piezo + field */
    QD_T_SC, QD_T_G1, QD_T_G2, QD_T_G3, QD_T_P,     /* kinetic / Kane parameter expansions */
    QD_T_Q,                                         /* strain part of Q */
    QD_T_RB, QD_T_RD,                               /* strain part of R = RB + i RD, R* = RB - i RD */
    QD_T_SXZ, QD_T_SYZ,                             /* strain part of S = -(SXZ - i SYZ)/sqrt2 */
    QD_T_PXX, QD_T_PYY, QD_T_PZZ, QD_T_PYZ, QD_T_PXZ, QD_T_PXY, /* P-weighted strain */
    QD_NTAB
};

typedef struct {
    int N[3], NT[3];
    double K[3];
    double L[3];               /* electronic-structure cell, Angstrom */
    double *G[3];              /* NT[a] reciprocal vectors per axis, 1/Angstrom */
    double *tab;               /* QD_NTAB * NT0*NT1*NT2 complex, interleaved re/im */
    int N_vb, N_cb;
    double nr0;                /* matrix refractive index (Coulomb screening) */
    void *state;               /* the builder's intermediates, kept by qd_tables_build_retain for
                                  qd_tables_vjp_retained (NULL otherwise); owned, freed by qd_tables_free */
} qd_tables;

typedef struct {
    double t_hamiltonian;      /* seconds, GPU events */
    double t_reduce;           /* tridiagonal reduction */
    double t_tridiag;          /* bisection + inverse iteration */
    double t_backtransform;
    double t_refine;
    double t_total;
    int refine_iters;
    double max_residual;       /* max_i ||H x_i - lambda_i x_i|| / ||H||_1 estimate, FP64 */
    double max_orth;           /* max |X^H X - I| */
    double t_band;             /* two-stage reduction: the dense -> band part of t_reduce */
    double tri_residual;       /* max_i ||(T - w_i) z_i|| / ||T||_1 of the tridiagonal eigenpairs */
} qd_stats;

typedef struct {
    double t_pairs;            /* pair densities: padding, FFTs, band products, inverse FFTs (s) */
    double t_contract;         /* upsampling, Poisson kernel, contractions (s) */
    double t_total;            /* wall time of qd_coulomb including host assembly (s) */
    int n_pairs;               /* S(S+1)/2 for S = nh + ne states */
    int fft[3];                /* FFT grid per axis (power of two >= 2N-1) */
    long nq;                   /* Coulomb-cell wave vectors (2N_C-1)^3 */
} qd_coulomb_stats;

typedef struct {
    double t_total;            /* wall time of qd_momentum (s) */
} qd_momentum_stats;

typedef struct {
    double t_density;          /* level densities, Coulomb potentials applied to the levels (s) */
    double t_sternheimer;      /* Sternheimer solves (s) */
    double t_bin;              /* Delta-G binning of Re tr(D dH) for all tables (s) */
    double t_total;            /* wall time of qd_adjoint (s) */
    int sternheimer_iters;
    double sternheimer_residual; /* max_i ||P_i g_i - P_i (H - E_i) lambda_i|| / ||P_i g_i|| */
    double J;                  /* J_eh of the two levels (eV) */
} qd_adjoint_stats;

/* Parameters of qd_tables_derivative */
#define QD_DP_FX 0         /* field component, per V/Angstrom (exact: the tables are linear in F) */
#define QD_DP_FY 1
#define QD_DP_FZ 2
#define QD_DP_L1 3         /* Length 1 of the masked shapes, per Angstrom */
#define QD_DP_L2 4
#define QD_DP_L3 5         /* Length 3 (height of a lens / cone, top side of a pyramid), per Angstrom */
#define QD_DP_L12 6        /* Length 1 and Length 2 together (base radius of a circular lens) */
#define QD_DP_X 7          /* composition x of the masked shapes */
#define QD_DP_OX 8         /* centre of the masked shapes, per Angstrom (moves them rigidly) */
#define QD_DP_OY 9
#define QD_DP_OZ 10
#define QD_DP_L4 11        /* Length 4 (height of a truncated pyramid), per Angstrom */

typedef struct qd_handle qd_handle;

/* Human-readable text of the last error on this thread. */
const char *qd_last_error(void);

/* ---- inputs and FP64 tables (host) ---- */
int qd_read_inputs(const char *parameters_path, const char *supercell_path, int argc,
                   const char *const *argv, qd_inputs *out);
void qd_free_inputs(qd_inputs *in);
int qd_tables_build(const qd_inputs *in, qd_tables *out);
/* qd_tables_build that also keeps the builder's intermediates (strain, piezo, shape transforms) in
 * out->state for qd_tables_vjp_retained, which then skips the forward pass it would repeat; the
 * state can be large (N_shape transforms on the strain grid) and is freed by qd_tables_free.  No
 * reference counterpart. */
int qd_tables_build_retain(const qd_inputs *in, qd_tables *out);
void qd_tables_free(qd_tables *t);
/* 1-based eigen window of eigenstates.c:29-36 */
int qd_window(const qd_inputs *in, int *il, int *iu);

/* ---- device ---- */
int qd_create(int device, qd_handle **out);
int qd_destroy(qd_handle *h);
/* Copy tables to device memory (kept resident until the next upload or destroy). */
int qd_upload_tables(qd_handle *h, const qd_tables *t);
/* qd_tables_build computed on the handle's device: the shape transforms, the strain (Green tensor),
 * the piezoelectric polarisation and potential, the transfer to the electronic grid (all three cell
 * modes), the convolutions and the tables; the material interpolation stays on the host.
 * Replaces 8_band_main.c:398-708 (characteristic_functions_reciprocal_space,
 * strain_region_reciprocal_space, piezo_region_reciprocal_space, strain/piezo_reciprocal_space[_fft],
 * convolutions).  out is filled
 * as by qd_tables_build (same layout and status codes; free it with qd_tables_free) and the tables
 * are left uploaded as by qd_upload_tables.  The convolutions are zero-padded FFTs where the host
 * builder uses direct transforms: the two agree to rounding. */
int qd_tables_build_device(qd_handle *h, const qd_inputs *in, qd_tables *out);
/* Build H on the device and copy its lower triangle (column-major, ld = N_H) to host
 * (H_host: N_H*N_H complex; only row >= col is written). precision: QD_FP64 or QD_FP32_REFINE
 * (the latter returns the FP32 matrix widened to double). */
int qd_hamiltonian(qd_handle *h, int precision, double *H_host);
/* Windowed eigensolve: eigenvalues il..iu (1-based, ascending) of the resident tables' H.
 * w: (iu-il+1) doubles; Z: N_H x (iu-il+1) complex column-major (may be NULL). */
int qd_solve(qd_handle *h, int il, int iu, int precision, double *w, double *Z, qd_stats *st);
/* qd_solve, qd_solve_device and qd_solve_matrix return QD_E_NOCONV when a pair misses its
 * tolerance: a bisection that hits its round limit, an inverse-iteration vector of T with
 * ||(T - w) z|| > 1e-11 ||T||_1, or (QD_FP32_REFINE) a refined pair with ||H x - w x|| above the
 * refinement tolerance after the last sweep.  The results are still stored and fetchable; the
 * reference aborts on the corresponding info != 0 of its eigensolver (diagonalise_hamiltonian.c:279-293). */
/* Same solve, results left on the device (bench path); fetch with qd_fetch_results. */
int qd_solve_device(qd_handle *h, int il, int iu, int precision, qd_stats *st);
int qd_fetch_results(qd_handle *h, double *w, double *Z);
/* Solve a caller-supplied Hermitian matrix (lower triangle, column-major, ld = n) with the same
 * kernels: used for tests of the eigensolver on matrices with known spectra. */
int qd_solve_matrix(qd_handle *h, int n, const double *A, int il, int iu, int precision,
                    double *w, double *Z, qd_stats *st);

/* Coulomb matrix elements of nh hole and ne electron (spin-resolved) states; replaces
 * coulomb_matrix_elements (8_band.h:1059, coulomb_matrix_elements.c:24) for CME levels 1-4 of
 * in->cme, free-length (Coulomb_supercell_type 0) or identical (type 1 with all RCL = 1) Coulomb
 * cells.  Z: N_H x (nh+ne) complex column-major host vectors, holes first; NULL uses the
 * device-resident eigenvectors of the last solve on h (holes: the nh highest of the first
 * 2*in->N_vb window states; electrons: the ne lowest of the rest).  Outputs (interleaved
 * complex, the reference's storage of 8_band.h:170-174 and coulomb_integrals.c):
 *   hh[((i*nh + j)*nh + k)*nh + l]            nh^4
 *   ee[((i*ne + j)*ne + k)*ne + l]            ne^4
 *   eh_direct[((i*nh + j)*ne + k)*ne + l]     i, j holes; k, l electrons
 *   eh_exchange[((i*ne + j)*ne + k)*nh + l]   i, l holes; j, k electrons
 * Entries outside the symmetry-limited loops of coulomb_integrals.c:47-349 are zero. */
int qd_coulomb(qd_handle *h, const qd_inputs *in, const qd_tables *t, int nh, int ne, const double *Z,
               double *hh, double *ee, double *eh_direct, double *eh_exchange, qd_coulomb_stats *st);
/* Momentum matrix elements (replaces momentum_matrix_element.c:31-103 with the dH/dk matrices of
 * hamiltonian_derivative_8_band.c:23-500, called from 8_band_main.c:1638-1660):
 * p[((a * nc + c) * nv + v) * 2 + {0,1}] = <psi_v| dH/dk_a |psi_c> (eV Angstrom, complex) for
 * a = x, y, z, the nc lowest conduction states c (c = 0 lowest) and the nv highest valence states
 * v (v = 0 highest) of the last solve on h (device-resident eigenvectors).  The reference's file
 * divides by sqrt(HBSQM / 2) and adds the [110] and [-110] combinations (write_momentum_matrix_elements.c). */
int qd_momentum(qd_handle *h, const qd_inputs *in, int nv, int nc, double *p, qd_momentum_stats *st);
/* Carrier localisation (replaces carrier_localisation.c:23-128 and the loops of 8_band_main.c:1427-1545):
 * the fraction of each window state in each shape, loc_vb[r * N_shape + s] for the 2 N_vb valence
 * states (r = 0 the highest) and loc_cb for the 2 N_cb conduction states (r = 0 the lowest); column
 * 0 is the rest of the cell, 1 - the sum over the shapes embedded in the matrix.  The reference's
 * file holds 100 times these (write_carrier_localisation.c).  seconds may be NULL. */
int qd_localisation(qd_handle *h, const qd_inputs *in, double *loc_vb, double *loc_cb, double *seconds);
/* Overlaps <cb_c|vb_v> of the window states (replaces spatial_overlap.c:23-52, 8_band_main.c:1373-1415):
 * ov[(c * 2 N_vb + v) * 2 + {0,1}], v = 0 the highest valence state. */
int qd_overlaps(qd_handle *h, const qd_inputs *in, double *ov);
/* Real-space envelope functions (replaces envelope_functions.c:23-220, which 8_band_main.c:883-902,
 * :1058-1077 and the 3-D output call on the linescan, cross-section and 3-D grids):
 * F[((s * 8 + b) * npts + p) * 2 + {0,1}] = sum_G c_b(G) exp(i (G + K) . r_p) / sqrt(L1 L2 L3)
 * (Angstrom^-3/2, complex) for the window states first .. first + count - 1 of the last solve
 * (ascending energy) and the points r_p = (x[p], y[p], z[p]) in Angstrom; b in the band order of
 * the Hamiltonian rows.  The reference's files hold sqrt(1000) times these, their probability
 * densities 1000 |F|^2 (write_envelope_functions*.c, write_probability_density*.c). */
int qd_envelopes(qd_handle *h, const qd_inputs *in, int first, int count, int npts, const double *x,
                 const double *y, const double *z, double *F);
/* The four coulomb_matrix_elements_*.dat files into dir; replaces write_coulomb_matrix_elements
 * (write_coulomb_matrix_elements.c:23). */
int qd_write_coulomb(const char *dir, int nh, int ne, const double *hh, const double *ee,
                     const double *eh_direct, const double *eh_exchange);
/* Adjoint of a spectral objective of the last solve on h (no counterpart in the reference):
 *   L = sum_n gE[n] E_n + wJ * J_eh,
 * gE over the solved window (give both members of a Kramers pair the same weight, so that the
 * weighted projector is gauge invariant), J_eh = 1/4 sum_{i in hole level, k in electron level}
 * V^eh_direct_iikk with the CME-1 kernel on the electronic-structure cell (levels: 0-based window
 * index of their first state; two states each; wJ = 0 skips the term).  The eigenvector part
 * of dJ is carried by Sternheimer solutions lambda_i = (H - E_i)^+ P_i dJ/dpsi_i^*.
 * Output W: QD_NTAB x nT complex with dL = Re sum_t sum_d W[t][d] dT_t[d] for any change dT of the
 * tables (so dL/dtheta = Re sum W * qd_tables_derivative(theta)). */
int qd_adjoint(qd_handle *h, const qd_inputs *in, const qd_tables *t, const double *gE, int hole_level,
               int electron_level, double wJ, double *W, qd_adjoint_stats *st);
/* Adjoint of a general objective of the window states of the last solve on h, the interface
 * qd_adjoint(..., dL_dE, dL_dZ, ...) of the adjoint design (no counterpart in the reference):
 *   dL = sum_n gE[n] dE_n + 2 Re sum_j <gZ_j, d psi_{cols[j]}>,  gZ_j = dL/dpsi_{cols[j]}^*
 * gE: k weights over the window (NULL: none); ncols (<= 32) window columns cols[j] with gZ the
 * column-major n x ncols complex gradients (interleaved re/im, n = N_H, rows in the Hamiltonian's
 * order).  L must not change under unitary mixing of the states of a degenerate level (a Kramers
 * pair): each eigenvector term is a Sternheimer solution lambda_j = (H - E_j)^+ P_j gZ_j with P_j
 * the projector onto the complement of that level (window states within 1e-8 eV, at most 4), and
 * D = sum gE psi psi^H - sum_j (lambda_j psi_j^H + psi_j lambda_j^H) is binned into W as in
 * qd_adjoint (dL = Re sum W dT for any table change dT).  st may be NULL. */
int qd_adjoint_z(qd_handle *h, const double *gE, int ncols, const int *cols, const double *gZ, double *W,
                 qd_adjoint_stats *st);
/* dT/dtheta (host FP64, QD_NTAB x nT complex into dT): QD_DP_F* analytically; shape parameters
 * (all shapes whose bit is set in shape_mask) by central differences of the table builder with
 * step rel_step * |value| (rel_step itself for x and the centre).  No reference counterpart; a
 * finite-difference check of qd_tables_vjp. */
int qd_tables_derivative(const qd_inputs *in, int param, unsigned shape_mask, double rel_step, double *dT);
/* dL/dtheta_j = Re sum_t sum_d W[t][d] dT_t[d]/dtheta_j for nparams parameters (params[j] one of
 * QD_DP_*, masks[j] its shapes, bit s = shape s >= 1; the field ignores the mask), from one reverse
 * pass through the FP64 table builder: dL/dchi_s and dL/d(material fields) of every shape, then
 * the exact d chi_s/dtheta of the shape transforms (forward-mode dual numbers through the same
 * formulas and quadrature nodes) and the closed-form d(fields)/dx of the alloy interpolation.  W is
 * what qd_adjoint returns.  Derivatives through the matrix's own composition or elastic constants
 * are not provided (bit 0 is rejected); QD_DP_L1 or QD_DP_L2 alone on a rectangular-base pyramid (shapes 5, 6) returns
 * QD_E_UNSUPPORTED (its faces are one segment turned by 2 atan(L2/L1) k, and the transform's
 * on-axis limiting branches carry no derivative across the axis; QD_DP_L12 is exact).  No
 * reference counterpart. */
/* qd_tables_vjp using the intermediates kept in t->state when t was built by qd_tables_build_retain
 * from inputs equal to in (every field the builder reads); otherwise it runs the forward pass itself. */
int qd_tables_vjp_retained(const qd_tables *t, const qd_inputs *in, const double *W, int nparams, const int *params,
                           const unsigned *masks, double *grad);
int qd_tables_vjp(const qd_inputs *in, const double *W, int nparams, const int *params, const unsigned *masks,
                  double *grad);
/* qd_tables_vjp with the forward fields and the reverse pass through the convolutions, the transfer
 * between the cells, the piezoelectric and strain steps computed on the handle's device, and the
 * gradients contracted there with the shapes' exact derivatives (dual numbers for the lengths, the
 * phase exp(-i G.O) for the centres); only the alloy derivatives are contracted on the host.  Same
 * arguments, statuses and results (to rounding).  No reference counterpart. */
int qd_tables_vjp_device(qd_handle *h, const qd_inputs *in, const double *W, int nparams, const int *params,
                         const unsigned *masks, double *grad);
/* Test support: W of qd_adjoint for an explicit D = sum_q gamma[q] V[:, a[q]] V[:, b[q]]^H, V an
 * N_H x nv complex column-major host array, with the resident tables' geometry. */
int qd_adjoint_bin(qd_handle *h, int nv, const double *V, int nterms, const int *a, const int *b,
                   const double *gamma, double *W);
/* Test support, host only (no GPU): W of qd_adjoint_bin for the resident-table geometry of t,
 * computed twice on the host -- through the binning kernels' own per-element code with a host FFT
 * (W_fft) and by the definition of the binning (W_direct) -- for the CPU tests of the FFT
 * correlation.  V: N_H x nv complex column-major; W_*: QD_NTAB x nT complex. */
int qd_selftest_adjoint_bin_host(const qd_tables *t, int nv, const double *V, int nterms, const int *a,
                                 const int *b, const double *gamma, double *W_fft, double *W_direct);
/* Test support, host only (no GPU): qd_momentum's elements for given valence states V (N_H x nv,
 * ascending energy as in the window, so v = 0 is V's last column) and conduction states C
 * (N_H x nc), computed on the host with the momentum kernel's own element code. */
int qd_selftest_momentum_host(const qd_tables *t, int nv, const double *V, int nc, const double *C, double *p);
/* Test support, host only (no GPU): the upsampling of P pair grids B (QD tables' Delta-G grid,
 * complex) to the Coulomb grid of in (out: P x (2N_C-1)^3 complex) with the matrices the device
 * path uploads, applied in the device's order. */
int qd_selftest_coulomb_upsample_host(const qd_inputs *in, const qd_tables *t, int P, const double *B, double *out);
/* Test support, host only (no GPU): the full Hermitian H (N_H x N_H complex, column-major) from the
 * tables with the assembly kernels' element code, for CPU tests of the tables against the
 * reference's fixtures. */
int qd_selftest_hamiltonian_host(const qd_tables *t, double *H);
/* Test support, host only: chi_idx(G) of one shape of in at n wave vectors G (3n, 1/Angstrom) on
 * the electronic (cell 0), strain (1) or Coulomb (3) cell, or with param one of the QD_DP_L and QD_DP_O
 * parameter its exact derivative; out: n complex. */
int qd_selftest_shape_host(const qd_inputs *in, int idx, int cell, int param, int n, const double *G, double *out);
/* Test support: the same chi_idx(G), or with param one of the QD_DP_L and QD_DP_O parameters its
 * exact derivative, evaluated on the handle's device (every shape type; the 1-D integrals of shapes
 * 3, 4 and 11-15 by the same Gauss-Legendre pair as the host, the length derivatives by dual
 * numbers through them); out: n complex. */
int qd_selftest_shape_device(qd_handle *h, const qd_inputs *in, int idx, int cell, int param, int n, const double *G,
                             double *out);
/* Measurement support (no reference counterpart): when enabled, the dominant kernel of the
 * reduction is bracketed by HIP events on the handle's stream: with QD_REDUCE_TWO_STAGE every
 * panel's trailing-matrix product X = A22 V (k_hemm), with QD_REDUCE_ONE_STAGE every 16th
 * column's Hermitian matrix-vector product (k_hemv).  qd_profile_read returns the number of timed
 * launches, their mean duration (s) and their mean algorithmic work since the last enable:
 * floating-point operations (8 m^2 32 for an m x m trailing matrix) for k_hemm, bytes (the lower
 * triangle of the trailing matrix, read once) for k_hemv. */
int qd_profile_enable(qd_handle *h, int enable);
int qd_profile_read(qd_handle *h, int *launches, double *mean_seconds, double *mean_work);

/* Tridiagonal reduction of later solves on this handle (no reference counterpart; zheevr's own
 * reduction is internal to the library it calls).  QD_REDUCE_TWO_STAGE (default): dense -> band
 * of width 32 by MFMA panel updates, then band -> tridiagonal by bulge chasing.
 * QD_REDUCE_ONE_STAGE: blocked Householder tridiagonalisation (one trailing-matrix
 * matrix-vector product per column). */
#define QD_REDUCE_TWO_STAGE 0
#define QD_REDUCE_ONE_STAGE 1
int qd_set_reduction(qd_handle *h, int method);

/* Stop of the FP64 refinement of QD_FP32_REFINE solves on this handle (no reference counterpart):
 * sweeps end when the Ritz values move by < 1e-10 eV and every ||H x - w x|| <= rel_residual ||T||_1,
 * or when a sweep no longer halves the largest residual (its floor, ~2e-12 ||T||_1, is reached
 * after two sweeps).  Default 1e-10 (one sweep; eigenvalues far inside 1 ueV of the FP64 solve);
 * vectors that feed Coulomb elements or gradients to 1e-6 relative want 5e-12 (two sweeps). */
int qd_set_refine_tolerance(qd_handle *h, double rel_residual);
/* Test support (no reference counterpart): iteration limits of later solves on this handle, a
 * negative value keeping the default: refinement sweeps (6), Sternheimer iterations of qd_adjoint
 * (12), inverse-iteration sweeps (4), bisection rounds (40, each a 65-way multisection).  A limit
 * that stops a solve short makes it return QD_E_NOCONV. */
int qd_debug_set_limits(qd_handle *h, int refine_sweeps, int sternheimer_iters, int invit_sweeps, int bisect_rounds);

/* Test support (no reference counterpart): the tridiagonal (d, e) of the last reduction
 * (n and n-1 doubles), and a device self-test of the 16x16x4 MFMA operand layout used by the
 * trailing update (out: 514 doubles: max |error| for the f32 and f64 instructions, then two raw
 * 64-lane x 4 dumps of the f64 accumulator for one-hot operand probes). */
int qd_debug_tridiagonal(qd_handle *h, double *d, double *e);
/* Test support (no reference counterpart): X <- Q X (adjoint = 0) or Q^H X (adjoint = 1) with the
 * unitary factor Q of the last reduction (Q^H H Q = T); X is N_H x k complex column-major on the
 * host, k <= 64. */
int qd_debug_apply_q(qd_handle *h, int k, int adjoint, double *X);
int qd_selftest_mfma(double *out);
/* Test support (no reference counterpart): record on the device, as a bounded wait between
 * work-groups that gave up would, that the factorisation of the last two-stage reduction is not
 * to be trusted; no kernel is affected, and every later use of the factorisation on this handle
 * (qd_adjoint, qd_debug_apply_q) returns QD_E_DEVICE until the next solve. */
int qd_debug_mark_wait_failure(qd_handle *h);
/* Test support: the DPP lane reductions of the band kernels on 64 lanes (in: 64 doubles, out: 320):
 * out[l], out[64+l], out[128+l] = sum of in[] over the aligned 4-, 8- and 16-lane group of lane l,
 * out[192+l] = in[0] (broadcast of lane 0), out[256+l] = sum of all 64. */
int qd_selftest_quad_sum(const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
