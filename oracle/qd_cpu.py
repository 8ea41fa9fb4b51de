"""ctypes wrapper of oracle/lib/libqd_cpu.so (TEST AND MEASUREMENT INFRASTRUCTURE ONLY).

The C file restates the reference's CPU path with its own loop structure: the per-element
Hamiltonian fill of hamiltonian_8_band.c:23-500 (OpenMP collapse(2), guided; one element-function
call per entry, per-shape parameter arrays allocated per call) and the zheevr call of
diagonalise_hamiltonian.c:55-77 through Intel MKL, the reference's LAPACK.  Inputs are the numpy
oracle's coefficient arrays (``qd_oracle.Tables``).  Used by tests/ (checked against the numpy
oracle) and by bench.py's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Par(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("E_cb", "E_hh", "E_lh", "E_so", "sc", "g1", "g2", "g3", "P", "ac", "av",
                                           "b", "d")]


class _Problem(C.Structure):
    _fields_ = [("N", C.c_int * 3), ("NT", C.c_int * 3), ("G", C.POINTER(C.c_double) * 3), ("K", C.c_double * 3),
                ("F", C.c_double * 3), ("N_shape", C.c_int), ("embedded", C.POINTER(C.c_int)),
                ("par", C.POINTER(_Par)), ("chi", C.POINTER(C.c_void_p)), ("strain", C.c_int), ("piezo", C.c_int),
                ("e", C.c_void_p * 6), ("conv", C.POINTER(C.c_void_p)), ("phi", C.c_void_p)]


def lib_path():
    return os.path.join(_HERE, "lib", "libqd_cpu.so")


def load():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise OSError(f"{path} is missing: build it with make -C oracle")
        L = C.CDLL(path)
        L.qc_hamiltonian.argtypes = [C.POINTER(_Problem), C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.qc_hamiltonian_entries.argtypes = [C.POINTER(_Problem), C.c_void_p, C.c_void_p, C.c_long, C.c_void_p,
                                             C.c_int]
        L.qc_zheevr.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.qc_load_mkl.argtypes = [C.c_char_p]
        L.qc_mkl_path.restype = C.c_char_p
        _LIB = L
    return _LIB


def mkl_available():
    return load().qc_load_mkl(None) == 0


def mkl_path():
    return load().qc_mkl_path().decode()


class CpuProblem:
    """The reference's global coefficient arrays for one oracle ``Tables``, laid out for C."""

    def __init__(self, t):
        p, shapes, mp = t.p, t.shapes, t.params
        ns = len(shapes)
        self._keep = []
        pr = _Problem()
        nT = int(np.prod(t.NT))
        for a in range(3):
            pr.N[a], pr.NT[a] = t.N[a], t.NT[a]
            g = np.ascontiguousarray(t.G[a], dtype=np.float64)
            self._keep.append(g)
            pr.G[a] = g.ctypes.data_as(C.POINTER(C.c_double))
            pr.K[a], pr.F[a] = p.K[a], p.F[a]
        pr.N_shape = ns
        emb = np.array([0] + [shapes[s].embedded for s in range(1, ns)], dtype=np.int32)
        self._keep.append(emb)
        pr.embedded = emb.ctypes.data_as(C.POINTER(C.c_int))
        pars = (_Par * ns)()
        for s in range(ns):
            for k, _ in _Par._fields_:
                setattr(pars[s], k, float(mp[s][k]))
        self._keep.append(pars)
        pr.par = pars
        chis = (C.c_void_p * ns)()
        for s in range(1, ns):
            c = np.ascontiguousarray(t.chi[s], dtype=np.complex128).reshape(-1)
            self._keep.append(c)
            chis[s] = c.ctypes.data
        self._keep.append(chis)
        pr.chi = chis
        pr.strain = 1 if p.strain else 0
        pr.piezo = 1 if (p.strain and p.piezo and t.phi is not None) else 0
        convs = (C.c_void_p * ns)()
        if p.strain:
            for i, k in enumerate(["xx", "yy", "zz", "yz", "xz", "xy"]):
                e = np.ascontiguousarray(t.strain[k], dtype=np.complex128).reshape(-1)
                self._keep.append(e)
                pr.e[i] = e.ctypes.data
            for s in range(1, ns):
                c = np.empty((8, nT), dtype=np.complex128)
                for i, k in enumerate(["xx", "yy", "zz", "yz", "xz", "xy", "hydrostatic", "biaxial"]):
                    c[i] = t.conv[s]["char_" + k].reshape(-1)
                self._keep.append(c)
                convs[s] = c.ctypes.data
            if pr.piezo:
                ph = np.ascontiguousarray(t.phi, dtype=np.complex128).reshape(-1)
                self._keep.append(ph)
                pr.phi = ph.ctypes.data
        self._keep.append(convs)
        pr.conv = convs
        self.pr = pr
        self.NH = 8 * int(np.prod(t.N))

    def hamiltonian(self, threads=0):
        """Lower triangle (row >= col) of H as an N_H x N_H Fortran-ordered complex array."""
        H = np.zeros((self.NH, self.NH), dtype=np.complex128, order="F")
        if load().qc_hamiltonian(C.byref(self.pr), H.ctypes.data, None, 0, threads) != 0:
            raise ValueError("bad column list")
        return H

    def hamiltonian_columns(self, cols, threads=0):
        """Lower-triangle entries of the listed columns of H: an (N_H, len(cols)) Fortran-ordered
        array whose column j holds H[row, cols[j]] for row >= cols[j]."""
        c = np.ascontiguousarray(cols, dtype=np.int32)
        H = np.zeros((self.NH, len(c)), dtype=np.complex128, order="F")
        if load().qc_hamiltonian(C.byref(self.pr), H.ctypes.data, c.ctypes.data, len(c), threads) != 0:
            raise ValueError("bad column list")
        return H


    def rows(self, rows, threads=0):
        """Full rows H[rows, :] of the Hermitian H, (len(rows), N_H) complex: entries right of the
        diagonal are the conjugates of the stored lower-triangle entries."""
        rows = np.asarray(rows, dtype=np.int64)
        n = self.NH
        R = np.repeat(rows, n)
        Cc = np.tile(np.arange(n, dtype=np.int64), len(rows))
        lo_r = np.maximum(R, Cc).astype(np.int32)
        lo_c = np.minimum(R, Cc).astype(np.int32)
        out = np.zeros(len(R), dtype=np.complex128)
        if load().qc_hamiltonian_entries(C.byref(self.pr), lo_r.ctypes.data, lo_c.ctypes.data, len(R),
                                         out.ctypes.data, threads) != 0:
            raise ValueError("bad entry list")
        out = np.where(Cc > R, np.conj(out), out)
        return out.reshape(len(rows), n)


def zheevr(H, il, iu, threads=0):
    """Eigenpairs il..iu (1-based) of the Hermitian matrix whose lower triangle is in H (an N x N
    Fortran-ordered complex array, overwritten), with MKL's zheevr as the reference calls it."""
    n = H.shape[0]
    if not H.flags.f_contiguous or H.dtype != np.complex128:
        raise ValueError("H must be Fortran-ordered complex128")
    k = iu - il + 1
    w = np.zeros(k)
    Z = np.zeros((n, k), dtype=np.complex128, order="F")   # ldz = n
    info = load().qc_zheevr(n, H.ctypes.data, il, iu, w.ctypes.data, Z.ctypes.data, threads)
    if info == -1000:
        raise OSError("Intel MKL (libmkl_rt.so) is not loadable")
    if info != 0:
        raise RuntimeError(f"zheevr info = {info}")
    return w, Z
