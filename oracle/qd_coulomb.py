"""FP64 CPU oracle of the QUDOS Coulomb matrix elements (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the reference's Coulomb path, used by tests/ to check the MI355X
library.  Nothing in the shipped library imports it.  Paths are relative to
``code/source/8_band`` of the QUDOS repository.

State order follows the reference: the 2*N_VB hole (valence) states in window order, then the
2*N_CB electron states; pair (a <= b) has packed index b*(b+1)/2 + a (B_convolutions.c:427).

Parity pins (tests/test_coulomb_oracle.py):
  * pair densities: brute-force direct correlation sum on tiny grids;
  * upsampling: identity when the Coulomb cell equals the electronic-structure cell, and the
    closed-form box transform against a brute-force real-space integral otherwise;
  * Poisson kernel, Madelung/dipole/quadrupole vectors: closed forms evaluated independently;
  * contractions + writer: the Prism_111 fixture (coulomb_matrix_elements_*.dat) through
    gauge-invariant block sums.
"""
from __future__ import annotations

import numpy as np

ESQEPS = 180.9513    # 8_band.h:52, e^2/eps0 in eV Angstrom
EPS = 1.0e-08        # 8_band.h:59


def pair_index(a, b):
    """Packed pair index for a <= b (B_convolutions.c:427)."""
    return b * (b + 1) // 2 + a


def basis_1d(N, L):
    """basis_set.c:34-86: G_i = (2 pi / L)(i - (NT-1)/2), i in [0, 2N-1)."""
    NT = 2 * N - 1
    return (2.0 * np.pi / L) * (np.arange(NT, dtype=np.float64) - 0.5 * (NT - 1))


def pair_densities(C, N):
    """B_ab(q) = sum_band sum_G conj(c_a,band(G)) c_b,band(G+q) for all pairs a <= b.

    Follows state_redistribution.c:40-58 (zero padding to (2N-1)^3, conj-reversed copy) and
    B_convolutions.c:332-386 (forward FFT of both, product, inverse FFT with 1/size, band sum).
    C: N_H x S complex (rows band*N_G + (ix*Ny + iy)*Nz + iz).  Returns (P, 2Nx-1, 2Ny-1, 2Nz-1)
    with array index n = q + N - 1 per axis.
    """
    S = C.shape[1]
    Nx, Ny, Nz = N
    NG = Nx * Ny * Nz
    NT = (2 * Nx - 1, 2 * Ny - 1, 2 * Nz - 1)
    F = np.zeros((S, 8) + NT, dtype=np.complex128)
    Fc = np.zeros((S, 8) + NT, dtype=np.complex128)
    for s in range(S):
        for b in range(8):
            c = C[b * NG:(b + 1) * NG, s].reshape(Nx, Ny, Nz)
            pad = np.zeros(NT, dtype=np.complex128)
            pad[:Nx, :Ny, :Nz] = c
            padc = np.zeros(NT, dtype=np.complex128)
            padc[:Nx, :Ny, :Nz] = np.conj(c[::-1, ::-1, ::-1])
            F[s, b] = np.fft.fftn(pad)
            Fc[s, b] = np.fft.fftn(padc)
    P = S * (S + 1) // 2
    B = np.empty((P,) + NT, dtype=np.complex128)
    for a in range(S):
        for b in range(a, S):
            B[pair_index(a, b)] = np.fft.ifftn(np.sum(Fc[a] * F[b], axis=0))
    return B


def upsample_free(B, N, NC, L, CL):
    """B_coulomb_supercell.c:23-100 (free Coulomb-cell lengths): separable box transform
    B_C(Q) = (1/Omega) sum_q B(q) I0x(Qx-qx) I0y(Qy-qy) I0z(Qz-qz), I0(g) = 2 sin(g L/2)/g,
    I0(0) = L, I0 = 0 beyond the Coulomb basis cut-off; an axis whose plane-wave count and
    length both equal the electronic-structure ones is copied."""
    mats = []
    for a in range(3):
        G = basis_1d(N[a], L[a])
        GC = basis_1d(NC[a], CL[a])
        if NC[a] == N[a] and L[a] == CL[a]:
            mats.append(L[a] * np.eye(len(GC)))
            continue
        g = GC[:, None] - G[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            m = 2.0 * np.sin(g * L[a] / 2.0) / g
        m = np.where(np.abs(g) < EPS, L[a], m)
        m = np.where(np.abs(g) > GC[-1] + EPS, 0.0, m)
        mats.append(m)
    inv = 1.0 / (L[0] * L[1] * L[2])
    return inv * np.einsum("xi,yj,zk,pijk->pxyz", mats[0], mats[1], mats[2], B, optimize=True)


def upsample_commensurate(B, N, NC, L, CL, RCL, sigma=None):
    """B_coulomb_supercell_fft.c:23-100 with characteristic_function_coulomb_supercell.c (Coulomb
    cell = RCL x the electronic cell): B placed at Coulomb offsets RCL e on a grid of
    U = 2 RCL (N_T/2) + 1 points per axis, zero-padded to U + N_C_T - 1, transformed, multiplied by
    the transform of the (zero-padded) electronic cell on the Coulomb grid (cuboid, 1/Omega_C,
    optional softening Gaussian), transformed back (scale 1/size), the central N_C_T block kept and
    scaled by Omega_C / Omega_es."""
    NT = tuple(2 * n - 1 for n in N)
    NCT = tuple(2 * n - 1 for n in NC)
    U = tuple(2 * RCL[a] * (NT[a] // 2) + 1 for a in range(3))
    sz = tuple(NCT[a] + U[a] - 1 for a in range(3))
    P = B.shape[0]
    Bup = np.zeros((P,) + U, dtype=np.complex128)
    sl = tuple(slice(0, U[a], RCL[a]) for a in range(3))     # offsets RCL e, e = -(N_T/2)..N_T/2
    Bup[(slice(None),) + sl] = B
    gx, gy, gz = coulomb_grid(NC, CL)
    chi = np.ones(NCT, dtype=np.complex128) / (CL[0] * CL[1] * CL[2])
    for g, La in ((gx, L[0]), (gy, L[1]), (gz, L[2])):
        chi = chi * (La * np.sinc(0.5 * g * La / np.pi))
    if sigma is not None:
        chi = chi * np.exp(-0.5 * sigma ** 2 * (gx ** 2 + gy ** 2 + gz ** 2))
    conv = np.fft.ifftn(np.fft.fftn(Bup, sz, axes=(1, 2, 3)) * np.fft.fftn(chi, sz, axes=(0, 1, 2))[None], axes=(1, 2, 3))
    off = tuple((NCT[a] + U[a] - 1) // 2 - NCT[a] // 2 for a in range(3))
    out = conv[(slice(None),) + tuple(slice(off[a], off[a] + NCT[a]) for a in range(3))]
    return out * ((CL[0] * CL[1] * CL[2]) / (L[0] * L[1] * L[2]))


def coulomb_grid(NC, CL):
    Gx, Gy, Gz = (basis_1d(NC[a], CL[a]) for a in range(3))
    gx, gy, gz = np.meshgrid(Gx, Gy, Gz, indexing="ij")
    return gx, gy, gz


def poisson_kernel(NC, CL, nr0):
    """coulomb_matrix_elements.c:384-403: w(q) = ESQEPS / (Omega_C |q|^2 nr0^2), w(0) = 0."""
    gx, gy, gz = coulomb_grid(NC, CL)
    G2 = gx * gx + gy * gy + gz * gz
    c = tuple(2 * n - 1 for n in NC)
    w = np.zeros_like(G2)
    mask = np.ones_like(G2, dtype=bool)
    mask[c[0] // 2, c[1] // 2, c[2] // 2] = False
    w[mask] = ESQEPS / ((CL[0] * CL[1] * CL[2]) * G2[mask] * nr0 ** 2)
    return w


def madelung(NC, CL, nprocs=4):
    """coulomb_matrix_elements.c:250-334, summed over ranks as the reference's MPI_Reduce does.

    Ewald eta = sqrt(pi/(CL1 CL2)); a = -2 eta/sqrt(pi) - pi/(eta^2 Omega) + reciprocal sum over
    q != 0 of (4 pi/Omega) exp(-q^2/(4 eta^2))/q^2 + real-space sum of erfc(eta R)/R.  The
    reference distributes the real-space lattice indices (mx, my, mz) in [0, ...) x [0, NyT/2) x
    [0, NzT/2) so that rank r > 0 takes the index slice of rank r-1 and rank 0 none
    (:287-305): with nprocs ranks the sum covers flat indices [0, sum_{p < nprocs-1} size_p).
    """
    from scipy.special import erfc
    CL1, CL2, CL3 = CL
    om = CL1 * CL2 * CL3
    eta = np.sqrt(np.pi / (CL1 * CL2))
    a = -2.0 * eta / np.sqrt(np.pi) - np.pi / (eta * eta * om)
    gx, gy, gz = coulomb_grid(NC, CL)
    G2 = (gx * gx + gy * gy + gz * gz).ravel()
    c = tuple(2 * n - 1 for n in NC)
    centre = (c[0] // 2) * c[1] * c[2] + (c[1] // 2) * c[2] + c[2] // 2
    m = np.ones(G2.shape, dtype=bool)
    m[centre] = False
    a += np.sum((4.0 * np.pi / om) * np.exp(-G2[m] / (4.0 * eta * eta)) / G2[m])
    ngc = c[0] * c[1] * c[2]
    sizes = [ngc // nprocs + (1 if p < ngc % nprocs else 0) for p in range(nprocs)]
    hi = sum(sizes[:nprocs - 1])
    hy, hz = c[1] // 2, c[2] // 2
    idx = np.arange(hi)
    mx = idx // (hy * hz)
    my = (idx % (hy * hz)) // hz
    mz = (idx % (hy * hz)) % hz
    keep = (mx != 0) | (my != 0) | (mz != 0)
    R = np.sqrt((mx[keep] * CL1) ** 2 + (my[keep] * CL2) ** 2 + (mz[keep] * CL3) ** 2)
    a += np.sum(erfc(R * eta) / R)
    return a


def _I_axis(g, CLa, CLb, CLc):
    """coulomb_matrix_elements.c:408-424 one-axis dipole vector value."""
    return CLb * CLc * ((CLa / (1j * g)) * np.cos(0.5 * g * CLa) + (2.0j / (g * g)) * np.sin(0.5 * g * CLa))


def multipole_vectors(NC, CL):
    """Dipole I_x, I_y, I_z (:406-431) and quadrupole J (:435-478) on the Coulomb grid."""
    gx, gy, gz = coulomb_grid(NC, CL)
    c = tuple(2 * n - 1 for n in NC)
    ix, iy, iz = np.meshgrid(np.arange(c[0]), np.arange(c[1]), np.arange(c[2]), indexing="ij")
    cx, cy, cz = ix == c[0] // 2, iy == c[1] // 2, iz == c[2] // 2
    CL1, CL2, CL3 = CL
    Ix = np.zeros(gx.shape, dtype=np.complex128)
    Iy = np.zeros_like(Ix)
    Iz = np.zeros_like(Ix)
    J = np.zeros_like(Ix)
    mx = ~cx & cy & cz
    my = cx & ~cy & cz
    mz = cx & cy & ~cz
    m0 = cx & cy & cz
    Ix[mx] = _I_axis(gx[mx], CL1, CL2, CL3)
    Iy[my] = _I_axis(gy[my], CL2, CL1, CL3)
    Iz[mz] = _I_axis(gz[mz], CL3, CL2, CL1)

    def jax(g, CLa, CLb, CLc):
        return CLb * CLc * ((0.5 * CLa * CLa / g) * np.sin(0.5 * CLa * g)
                            + (2j / g) * ((CLa / (1j * g)) * np.cos(0.5 * g * CLa)
                                          + (2.0j / (g * g)) * np.sin(0.5 * g * CLa)))
    J[mx] += jax(gx[mx], CL1, CL2, CL3)
    J[my] += jax(gy[my], CL2, CL1, CL3)
    J[mz] += jax(gz[mz], CL3, CL2, CL1)
    J[m0] += (CL1 ** 3 * CL2 * CL3 + CL2 ** 3 * CL1 * CL3 + CL3 ** 3 * CL2 * CL1) / 12.0
    return Ix, Iy, Iz, J


def coulomb_elements(C, nh, ne, N, NC, L, CL, coulomb_type, nr0, cme=1, nprocs=4, RCL=(1, 1, 1), sigma=None):
    """coulomb_matrix_elements.c:24-627 + coulomb_integrals.c:23-356.

    C: N_H x (nh + ne) coefficient vectors (holes first).  Returns dict of the four arrays in the
    reference's storage (8_band.h:170-174), entries outside the symmetry-limited loops zero:
      hh[i, j, k, l] (nh^4), ee (ne^4), eh_direct[i, j, k, l] (i, j holes; k, l electrons),
      eh_exchange[i, j, k, l] (i hole, j, k electrons, l hole).
    """
    S = nh + ne
    B = pair_densities(C, N)
    if coulomb_type == 0:
        B = upsample_free(B, N, NC, L, CL)
    else:
        B = upsample_commensurate(B, N, NC, L, CL, RCL, sigma)
    P = B.shape[0]
    Bf = B.reshape(P, -1)
    w = poisson_kernel(NC, CL, nr0).ravel()
    Brev = Bf[:, ::-1] * w[None, :]          # B(-q) w(q), coulomb_matrix_elements.c:510-512
    # zdotu(x, y) = sum x y; zdotc(x, y) = sum conj(x) y
    M_u = Bf @ Brev.T                        # [pair(i,j), pair(k,l)] -> sum Brev_kl B_ij
    M_c = Bf @ np.conj(Brev).T
    om = CL[0] * CL[1] * CL[2]
    pre = ESQEPS / (4.0 * np.pi * nr0 ** 2)
    a_mad = madelung(NC, CL, nprocs) if cme >= 2 else 0.0
    D = Q = None
    if cme >= 3:
        Ix, Iy, Iz, Jq = multipole_vectors(NC, CL)
        D = np.zeros((3, S, S), dtype=np.complex128)
        for ax, Iv in enumerate((Ix.ravel(), Iy.ravel(), Iz.ravel())):
            for k in range(S):
                for l in range(S):
                    if l >= k:
                        D[ax, l, k] = np.sum(Bf[pair_index(k, l)] * Iv)
                    else:
                        D[ax, l, k] = np.sum(np.conj(Bf[pair_index(l, k)]) * Iv)
        if cme >= 4:
            Q = Bf @ Jq.ravel()

    def V(i, j, k, l):
        """coulomb_integrals.c:75-86 for i <= j."""
        pij = pair_index(i, j)
        if l >= k:
            return M_u[pij, pair_index(k, l)]
        return M_c[pij, pair_index(l, k)]

    def corr(i, j, k, l, monopole, quad):
        v = 0.0
        if cme >= 2 and monopole and i == j and k == l:
            v += -pre * a_mad
        if cme >= 3:
            v += pre * (4.0 * np.pi / (3.0 * om ** 3)) * (
                D[0, j, i] * D[0, l, k] + D[1, j, i] * D[1, l, k] + D[2, j, i] * D[2, l, k])
        if cme >= 4 and quad:
            f = -pre * (2.0 * np.pi / (3.0 * om ** 2))
            if i == j and l >= k:
                v += f * Q[pair_index(k, l)]
            elif i == j and l < k:
                v += f * np.conj(Q[pair_index(l, k)])
            if k == l:
                v += f * Q[pair_index(i, j)]
        return v

    def lstart(i, j, k):
        return k if j == i else (j if k == i else (i + 1 if k < j else i))

    out = {}
    hh = np.zeros((nh,) * 4, dtype=np.complex128)
    for i in range(nh):
        for j in range(i, nh):
            for k in range(i, nh):
                for l in range(lstart(i, j, k), nh):
                    hh[i, j, k, l] = V(i, j, k, l) + corr(i, j, k, l, True, True)
    out["hh"] = hh
    ee = np.zeros((ne,) * 4, dtype=np.complex128)
    o = nh
    for i in range(ne):
        for j in range(i, ne):
            for k in range(i, ne):
                for l in range(lstart(i, j, k), ne):
                    ee[i, j, k, l] = V(o + i, o + j, o + k, o + l) + corr(o + i, o + j, o + k, o + l, True, True)
    out["ee"] = ee
    ed = np.zeros((nh, nh, ne, ne), dtype=np.complex128)
    for i in range(nh):
        for j in range(i, nh):
            for k in range(ne):
                for l in range(k if j == i else 0, ne):
                    ed[i, j, k, l] = V(i, j, o + k, o + l) + corr(i, j, o + k, o + l, True, True)
    out["eh_direct"] = ed
    ex = np.zeros((nh, ne, ne, nh), dtype=np.complex128)
    for i in range(nh):
        for j in range(ne):
            for k in range(j, ne):
                for l in range(i if k == j else 0, nh):
                    # coulomb_integrals.c:326-337: zdotc(Brev[pair(l, o+k)], B[pair(i, o+j)])
                    v = M_c[pair_index(i, o + j), pair_index(l, o + k)]
                    if cme >= 3:
                        v += pre * (4.0 * np.pi / (3.0 * om ** 3)) * (
                            D[0, o + j, i] * D[0, l, o + k] + D[1, o + j, i] * D[1, l, o + k]
                            + D[2, o + j, i] * D[2, l, o + k])
                    ex[i, j, k, l] = v
    out["eh_exchange"] = ex
    return out


def _unique_hh(i, j, k, l):
    return j >= i and k >= i and l >= (k if j == i else (j if k == i else (i + 1 if k < j else i)))


def _fmt(v, conj=False):
    im = -v.imag if conj else v.imag
    return "%17.10e\t%17.10e\t%17.10e\t" % (v.real, im, abs(v))


def write_coulomb_text(cme):
    """write_coulomb_matrix_elements.c:23-395: the four files' text (Patra Eq. 2.86 order,
    symmetry reconstruction V_kijl = V_iklj, V_jlki = V_iklj*).  Returns dict name -> text."""
    hh, ee, ed, ex = cme["hh"], cme["ee"], cme["eh_direct"], cme["eh_exchange"]
    nv, nc = hh.shape[0], ee.shape[0]
    files = {}

    def same_kind(T, n, order_desc, name):
        lines = []
        rng = range(n - 1, -1, -1) if order_desc else range(n)
        lab = (lambda x: n - x) if order_desc else (lambda x: x + 1)
        for i in rng:
            for k in rng:
                for l in rng:
                    for j in rng:
                        s = "%d\t%d\t%d\t%d\t" % (lab(i), lab(k), lab(l), lab(j))
                        if _unique_hh(i, j, k, l):
                            s += _fmt(T[i, j, k, l])
                        elif _unique_hh(k, l, i, j):
                            s += _fmt(T[k, l, i, j])
                        elif _unique_hh(j, i, l, k):
                            s += _fmt(T[j, i, l, k], True)
                        elif _unique_hh(l, k, j, i):
                            s += _fmt(T[l, k, j, i], True)
                        lines.append(s + "\n")
                    lines.append("\n")
        files[name] = "".join(lines)

    same_kind(hh, nv, True, "coulomb_matrix_elements_hh.dat")
    same_kind(ee, nc, False, "coulomb_matrix_elements_ee.dat")
    lines = []
    for k in range(nc):
        for i in range(nv - 1, -1, -1):
            for j in range(nv - 1, -1, -1):
                for l in range(nc):
                    s = "%d\t%d\t%d\t%d\t" % (k + 1, nv - i, nv - j, l + 1)
                    if j >= i and l >= (k if j == i else 0):
                        s += _fmt(ed[i, j, k, l])
                    elif i >= j and k >= (l if i == j else 0):
                        s += _fmt(ed[j, i, l, k], True)
                    lines.append(s + "\n")
                lines.append("\n")
    files["coulomb_matrix_elements_eh_direct.dat"] = "".join(lines)
    lines = []
    for k in range(nc):
        for i in range(nv - 1, -1, -1):
            for j in range(nc):
                for l in range(nv - 1, -1, -1):
                    s = "%d\t%d\t%d\t%d\t" % (k + 1, nv - i, j + 1, nv - l)
                    if k >= j and l >= (i if k == j else 0):
                        s += _fmt(ex[i, j, k, l])
                    elif j >= k and i >= (l if j == k else 0):
                        s += _fmt(ex[l, k, j, i], True)
                    lines.append(s + "\n")
                lines.append("\n")
    files["coulomb_matrix_elements_eh_exchange.dat"] = "".join(lines)
    return files


def read_coulomb_file(path):
    """Parse a coulomb_matrix_elements_*.dat file: rows (i, k, l, j, Re, Im) or (i, k, l, j)
    when the reference printed no value."""
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            idx = tuple(int(x) for x in t[:4])
            val = complex(float(t[4]), float(t[5])) if len(t) >= 6 else None
            rows.append((idx, val))
    return rows
