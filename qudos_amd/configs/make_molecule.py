"""Writes qudos_amd/configs/molecule_17x17x21/{parameters.in, supercell.in}: the QD molecule of
two vertically stacked lens dots (base radius 10 nm, height 3 nm) on 0.5 nm In0.4Ga0.6As wetting
layers in a 30 x 30 x 28 nm GaAs supercell, 17 x 17 x 21 plane waves.  Each lens is represented
by K stacked truncated-cone slices (shape 12; the slicing follows the graded truncated-pyramid
example of QUDOS, code/example/Truncated_pyramid/Electronic_structure/supercell.in): slice k
spans heights z_k..z_{k+1} of the lens, with base and top radii R sqrt(1 - (z/h)^2) at its ends,
and takes the In fraction x_In(z) = 0.4 + 0.6 z / h at its mid-height (0.4 at the base, 1.0 at
the apex).  The lower dot's base sits 8 nm below the cell centre; the upper dot's base is the
nominal separation (6 nm) above it.  Separation and field are set per draw at run time.

Shape order: 0 matrix; 1 lower wetting layer; 2..K+1 lower slices; K+2 upper wetting layer;
K+3..2K+2 upper slices.
"""
import math
import os

K = 6                 # slices per dot
R, H = 10.0, 3.0      # nm
WL = 0.5              # nm
BASE_LOWER = -8.0     # nm, from the cell centre
SEPARATION = 6.0      # nm, base to base (nominal)
CELL = (30.0, 30.0, 28.0)
M = (8, 8, 10)


def x_ga(x_in):
    return 1.0 - x_in


def shape_block(shape, x, centre, lengths):
    lines = ["Embedded: 0", f"QD shape: {shape}", "Material: 3", f"Compos x: {x:.6f}", "Compos y: 0.000000",
             f"Centre x: {centre[0]:.6f}", f"Centre y: {centre[1]:.6f}", f"Centre z: {centre[2]:.6f}"]
    lines += [f"Length {i + 1}: {v:.6f}" for i, v in enumerate(lengths)]
    lines += ["Euler alpha angle: 0", "Euler beta angle: 0", "Euler gamma angle: 0", ""]
    return "\n".join(lines)


def dot_blocks(base):
    out = []
    # wetting layer along [001] (Length 1 = 20 -> axis 2), thickness WL, top at the dot base
    out.append(shape_block(17, x_ga(0.4), (0.0, 0.0, base - 0.5 * WL), (2.0, WL, 0.0, 0.0, 0.0)))
    for k in range(K):
        z0, z1 = H * k / K, H * (k + 1) / K
        r0 = R * math.sqrt(max(0.0, 1.0 - (z0 / H) ** 2))
        r1 = R * math.sqrt(max(0.0, 1.0 - (z1 / H) ** 2))
        x_in = 0.4 + 0.6 * 0.5 * (z0 + z1) / H
        out.append(shape_block(12, x_ga(x_in), (0.0, 0.0, base + z0), (r0, r1, z1 - z0, 0.0, 0.0)))
    return out


def main():
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "molecule_17x17x21")
    os.makedirs(d, exist_ok=True)
    blocks = dot_blocks(BASE_LOWER) + dot_blocks(BASE_LOWER + SEPARATION)
    sc = ["# QD molecule: two stacked graded lens dots (written by qudos_amd/configs/make_molecule.py)",
          "QD shape: 0", "Material: 3", "Compos x: 1.0", "Compos y: 0.0",
          f"Length 1: {CELL[0]}", f"Length 2: {CELL[1]}", f"Length 3: {CELL[2]}",
          "Strain length 1: 1", "Strain length 2: 1", "Strain length 3: 1",
          "Coulomb length 1: 1", "Coulomb length 2: 1", "Coulomb length 3: 1", ""] + blocks
    with open(os.path.join(d, "supercell.in"), "w") as f:
        f.write("\n".join(sc))
    n_shape = 1 + len(blocks)
    par = [f"# QD molecule, {2 * M[0] + 1}x{2 * M[1] + 1}x{2 * M[2] + 1} plane waves "
           f"(N_H = {8 * (2 * M[0] + 1) * (2 * M[1] + 1) * (2 * M[2] + 1)}); written by make_molecule.py",
           f"N_shape: {n_shape}", f"M_Gx: {M[0]}", f"M_Gy: {M[1]}", f"M_Gz: {M[2]}", "Strain_supercell_type: 1",
           f"M_S_Gx: {M[0]}", f"M_S_Gy: {M[1]}", f"M_S_Gz: {M[2]}", "N_VB: 10", "N_CB: 10", "T: 300.0",
           "Fx: 0", "Fy: 0", "Fz: 0", "Kx: 0", "Ky: 0", "Kz: 0", "Strain: 1", "Isotropic: 0", "Elastic_const: 0",
           "Piezo: 2", "CME: 0", "Coulomb_supercell_type: 1", f"M_C_Gx: {M[0]}", f"M_C_Gy: {M[1]}",
           f"M_C_Gz: {M[2]}", "Overlaps: 0", "Localisation: 0", "MME: 0", ""]
    with open(os.path.join(d, "parameters.in"), "w") as f:
        f.write("\n".join(par))


if __name__ == "__main__":
    main()
