"""Writes tests/golden/prism_111/envelopes.npz from the reference's Prism_111 example outputs
(code/example/Prism_111/Electronic_structure/PWEM_test_ref/probability_density_*.dat): the
probability-density files of every window state (columns x, y, z [nm], E [eV], total, SO, LH, HH,
CB [nm^-3]) for the linescans along [001], [110] and [111], the cross-sections with normals
[-110] and [111] (rows and columns 0, 3, 6, ... of the 40 x 40 grid) and the 3-D grid (points
0, 4, 8, 12, 16 of the 20 per axis).  Data only: the files are read as numbers and subsampled.
Run from the repository root: python tests/golden/make_envelope_fixture.py"""
import os

import numpy as np

SRC = "/root/reference/code/example/Prism_111/Electronic_structure/PWEM_test_ref"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "prism_111", "envelopes.npz")
NVB, NCB = 8, 4          # 2 N_VB and 2 N_CB of the example
N_CROSS, N_3D = 40, 20
CROSS_STEP, STEP_3D = 3, 4


def states():
    return [("vb", i) for i in range(1, NVB + 1)] + [("cb", i) for i in range(1, NCB + 1)]


def load(name):
    return np.loadtxt(os.path.join(SRC, name), ndmin=2)


def main():
    out = {}
    for d in ("001", "110", "111"):
        out[f"line_{d}"] = np.stack([load(f"probability_density_{b}_{i}_{d}.dat") for b, i in states()])
    sel = np.arange(0, N_CROSS, CROSS_STEP)
    for d in ("-110", "111"):
        a = np.stack([load(f"probability_density_cross_section_{b}_{i}_{d}.dat") for b, i in states()])
        a = a.reshape(len(a), N_CROSS, N_CROSS, -1)[:, sel][:, :, sel]
        out[f"cross_{d}"] = a.reshape(len(a), -1, a.shape[-1])
        out[f"cross_{d}_index"] = np.array([(i, j) for i in sel for j in sel])
    sel3 = np.arange(0, N_3D, STEP_3D)
    a = np.stack([load(f"probability_density_3D_{b}_{i}.dat") for b, i in states()])
    a = a.reshape(len(a), N_3D, N_3D, N_3D, -1)[:, sel3][:, :, sel3][:, :, :, sel3]
    out["grid3d"] = a.reshape(len(a), -1, a.shape[-1])
    out["grid3d_index"] = np.array([(i, j, k) for i in sel3 for j in sel3 for k in sel3])
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
