"""Test-side helpers: compose the oracle's coefficient arrays into the library's Delta-G table
layout (QD_T_* order of include/qudos_gpu.h) and write small problem inputs."""
import functools
import math
import os

import numpy as np

from oracle import qd_oracle as qo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def oracle_tables(t: qo.Tables):
    """20 complex tables from the oracle's chi / strain / convolution / piezo arrays."""
    p, shapes, mp = t.p, t.shapes, t.params
    NT = t.NT
    ns = len(shapes)
    cen = np.zeros(NT, dtype=bool)
    cen[NT[0] // 2, NT[1] // 2, NT[2] // 2] = True

    def pme(f):
        out = f(mp[0]) * cen.astype(complex)
        for s in range(1, ns):
            out = out + (f(mp[s]) - f(mp[shapes[s].embedded])) * t.chi[s]
        return out

    def sh(f, key):
        out = np.zeros(NT, complex)
        for s in range(1, ns):
            out = out + (f(mp[s]) - f(mp[shapes[s].embedded])) * t.conv[s][key]
        return out

    z = np.zeros(NT, complex)
    tabs = [pme(lambda d: d["E_cb"]), pme(lambda d: d["E_hh"]), pme(lambda d: d["E_lh"]), pme(lambda d: d["E_so"]),
            pme(lambda d: d["sc"]), pme(lambda d: d["g1"]), pme(lambda d: d["g2"]), pme(lambda d: d["g3"]),
            pme(lambda d: d["P"]), z, z, z, z, z, z, z, z, z, z, z]
    # field
    Fx, Fy, Fz = p.F
    ix, iy, iz = np.meshgrid(*[np.arange(n) - n // 2 for n in NT], indexing="ij")
    fld = np.zeros(NT, complex)
    L = (shapes[0].L1, shapes[0].L2, shapes[0].L3)
    with np.errstate(all="ignore"):
        for F, d, o1, o2, La in [(Fx, ix, iy, iz, L[0]), (Fy, iy, ix, iz, L[1]), (Fz, iz, ix, iy, L[2])]:
            if F:
                term = 1j * F * np.where(d % 2 == 0, -1.0, 1.0) / (-d * 2 * np.pi / La)
                fld = fld + np.where((d != 0) & (o1 == 0) & (o2 == 0), term, 0)
    for i in range(4):
        tabs[i] = tabs[i] + fld
    if p.strain:
        e = t.strain
        hyd = e["xx"] + e["yy"] + e["zz"]
        bia = e["xx"] + e["yy"] - 2 * e["zz"]
        ph = t.phi if t.phi is not None else 0
        m0 = mp[0]
        tabs[0] = tabs[0] + m0["ac"] * hyd + sh(lambda d: d["ac"], "char_hydrostatic") + ph
        tabs[1] = tabs[1] + m0["av"] * hyd + 0.5 * m0["b"] * bia + sh(lambda d: d["av"], "char_hydrostatic") \
            + 0.5 * sh(lambda d: d["b"], "char_biaxial") + ph
        tabs[2] = tabs[2] + m0["av"] * hyd - 0.5 * m0["b"] * bia + sh(lambda d: d["av"], "char_hydrostatic") \
            - 0.5 * sh(lambda d: d["b"], "char_biaxial") + ph
        tabs[3] = tabs[3] + m0["av"] * hyd + sh(lambda d: d["av"], "char_hydrostatic") + ph
        tabs[9] = math.sqrt(0.5) * (m0["b"] * bia + sh(lambda d: d["b"], "char_biaxial"))
        tabs[10] = -0.5 * math.sqrt(3) * (m0["b"] * (e["xx"] - e["yy"]) + sh(lambda d: d["b"], "char_xx")
                                         - sh(lambda d: d["b"], "char_yy"))
        tabs[11] = m0["d"] * e["xy"] + sh(lambda d: d["d"], "char_xy")
        tabs[12] = m0["d"] * e["xz"] + sh(lambda d: d["d"], "char_xz")
        tabs[13] = m0["d"] * e["yz"] + sh(lambda d: d["d"], "char_yz")
        for i, k in enumerate(["xx", "yy", "zz", "yz", "xz", "xy"]):
            tabs[14 + i] = m0["P"] * e[k] + sh(lambda d: d["P"], "char_" + k)
    return np.stack(tabs)


def oracle_problem(param_path, supercell_path, overrides=()):
    """Oracle Params/shapes with the same '-key=value' overrides the library accepts."""
    text = open(param_path).read()
    for o in overrides:
        k, v = o[1:].split("=", 1)
        text += f"\n{k}: {v}\n"
    p = qo.read_parameters(text)
    shapes = qo.read_supercell(open(supercell_path).read(), p)
    return p, shapes


@functools.lru_cache(maxsize=1)
def prism_oracle_solution():
    """Oracle (params, shapes, tables, eigenvalues, eigenvectors) of the Prism_111 example; the
    FP64 dense solve takes about a minute, so slow tests share it."""
    pp, sp = prism_paths()
    p = qo.read_parameters(open(pp).read())
    sh = qo.read_supercell(open(sp).read(), p)
    t = qo.coefficient_tables(p, sh)
    w, Z = qo.solve(qo.build_hamiltonian(t), p)
    return p, sh, t, w, Z


def prism_paths():
    d = os.path.join(GOLDEN, "prism_111")
    return os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")


def read_momentum_file(path):
    """momentum_matrix_elements.dat (write_momentum_matrix_elements.c:23-116): rows (c, v) with
    re, im, |.| of x, y, z, [110], [-110], in eV^(1/2).  Returns (c, v, 5) complex and the
    magnitudes."""
    rows = [ln.split() for ln in open(path) if ln.strip()]
    a = np.array([[float(x) for x in r] for r in rows])
    nc, nv = int(a[:, 0].max()), int(a[:, 1].max())
    vals = a[:, 2:17].reshape(-1, 5, 3)
    p = (vals[:, :, 0] + 1j * vals[:, :, 1]).reshape(nc, nv, 5)
    mag = vals[:, :, 2].reshape(nc, nv, 5)
    return p, mag


def momentum_pair_sums(mag):
    """sum of |p|^2 over each (conduction Kramers pair, valence Kramers pair) block, per
    polarisation: invariant under any unitary mixing within a Kramers pair."""
    nc, nv, npol = mag.shape
    return (mag ** 2).reshape(nc // 2, 2, nv // 2, 2, npol).sum(axis=(1, 3))


def momentum_columns(p):
    """(c, v, 5) complex in eV^(1/2) from qd_momentum's p[a, c, v] (eV Angstrom)."""
    q = np.moveaxis(p, 0, -1) / math.sqrt(7.61996 / 2.0)
    return np.concatenate([q, ((q[..., 0] + q[..., 1]) / math.sqrt(2.0))[..., None],
                           ((-q[..., 0] + q[..., 1]) / math.sqrt(2.0))[..., None]], axis=-1)
