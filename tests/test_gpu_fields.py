"""Coefficient generation on the device (qd_tables_build_device) against the host FP64 builder.

The host builder (qd_tables_build) is pinned on its own: its tables reproduce the oracle's, whose
strain, piezoelectric and convolution steps are checked against the reference's example files
and closed forms, and its reverse pass matches central differences.  The device builder computes
the same fields, transfers and convolutions with zero-padded FFTs instead of direct transforms,
so the two agree to rounding: every table to 1e-12 of its largest entry.

The shape transforms of all 19 shape types are evaluated on the device as well; they are checked
against the host's transforms (pinned against the oracle and brute-force quadrature) on every
branch of the closed forms, rotated and softened, on both cells, to 1e-12 of |chi(0)|, and their
exact derivatives (dual numbers for the lengths, the phase for the centre) against the host's dual
numbers (themselves pinned by central differences) to 1e-11 of the largest.

The reverse pass on the device (qd_tables_vjp_device) equals the host's (qd_tables_vjp, pinned by
central differences of the builder) to 1e-12 relative for every kind of parameter (shape lengths,
centres, composition and field; relative to 1% of the largest gradient for the small ones, whose
sums cancel).

Cases: the cell modes of the strain transfer (same cell, free lengths, commensurate supercells
with FFT downsampling), first- and second-order piezoelectricity, isotropic elasticity, no strain
with a uniform field, and at full size the bench's lens, the Truncated_pyramid example (100 pyramid
slices, commensurate strain cell, second-order piezo) and the QD molecule (12 truncated-cone slices).  The
resident tables are checked too: the Hamiltonian built from them equals the one built from the
host tables after qd_upload_tables.
"""
import os

import numpy as np
import pytest

import qudos_amd as q
from helpers import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SMALL = ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2")
CASES = [
    ("lens_9x9x9", SMALL),                                                   # same cell, piezo 1 + 2
    ("lens_9x9x9", SMALL + ("-Piezo=1", "-Fz=7")),
    ("lens_9x9x9", SMALL + ("-Piezo=0",)),
    ("lens_9x9x9", SMALL + ("-Isotropic=1",)),
    ("cone_19x19x15", SMALL),
    ("lens_9x9x9", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=2",
                    "-Strain_supercell_type=0")),                            # free lengths
    ("truncated_pyramid", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=4")),  # commensurate
    ("lens_9x9x9", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-Strain=0", "-Fx=7", "-Fz=-3")),
]
FULL = [("lens_11x11x9", ()), ("truncated_pyramid", ()), ("molecule_17x17x21", ())]
# shape type and lengths L1..L4 (nm) of the device forms; the pyramids' base lengths are equal
DEVICE_SHAPES = {
    "cuboid": (0, (6.0, 5.0, 4.0, 0.0)),
    "sphere": (1, (5.0, 0.0, 0.0, 0.0)),
    "ellipsoid": (2, (6.0, 5.0, 4.0, 0.0)),
    "rect_pyramid": (5, (5.5, 5.5, 4.0, 0.0)),
    "trunc_rect_pyramid": (6, (6.0, 6.0, 2.0, 3.5)),
    "hex_pyramid": (7, (4.0, 5.0, 6.0, 0.0)),
    "oct_trunc_pyramid": (8, (4.0, 2.0, 3.0, 8.0)),
    "cylinder": (9, (5.0, 4.0, 0.0, 0.0)),
    "ring": (10, (6.0, 3.0, 4.0, 0.0)),
    "hex_prism": (16, (5.0, 4.0, 6.0, 0.0)),
    "well_z": (17, (2.0, 3.0, 0.0, 0.0)),
    "wire_z": (18, (2.0, 4.1, 3.3, 0.0)),
    "wire_x": (18, (0.0, 3.7, 2.9, 0.0)),
    # 1-D integrals of Bessel functions
    "hemisphere": (3, (5.0, 0.0, 0.0, 0.0)),
    "lens": (4, (6.0, 5.0, 3.0, 0.0)),
    "cone": (11, (5.0, 4.0, 0.0, 0.0)),
    "trunc_cone": (12, (6.0, 3.0, 4.0, 0.0)),
    "conic_ring": (13, (6.0, 3.0, 1.5, 4.0)),
    "elliptic_cone": (14, (6.0, 4.0, 4.5, 0.0)),
    "elliptic_trunc_cone": (15, (6.0, 4.0, 3.0, 4.0)),
}
SHAPE_VARIANTS = [((0.0, 0.0, 0.0), ()), ((30.0, 20.0, 10.0), ()), ((0.0, 0.0, 0.0), ("-Softening=1",))]
DERIV_PARAMS = [q.QD_DP_L1, q.QD_DP_L2, q.QD_DP_L3, q.QD_DP_L4, q.QD_DP_L12, q.QD_DP_OX, q.QD_DP_OZ]
TOL = 1e-12


def cfg(name):
    d = os.path.join(ROOT, "qudos_amd", "configs", name)
    if not os.path.isdir(d):
        d = os.path.join(GOLDEN, name)
    return os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")


def table_array(t):
    nT = t.NT[0] * t.NT[1] * t.NT[2]
    flat = np.ctypeslib.as_array(t.tab, shape=(q.QD_NTAB * nT * 2,)).copy()
    return flat.view(np.complex128).reshape(q.QD_NTAB, nT)


def check_device_tables(solver, name, over, tol=TOL):
    pp, sp = cfg(name)
    inp = q.read_inputs(pp, sp, over)
    host = q.build_tables(inp)
    dev = solver.build_tables(inp)
    assert tuple(dev.NT) == tuple(host.NT) and tuple(dev.N) == tuple(host.N)
    for a in range(3):
        assert dev.L[a] == host.L[a] and dev.K[a] == host.K[a]
        assert np.array_equal(np.ctypeslib.as_array(dev.G[a], shape=(dev.NT[a],)),
                              np.ctypeslib.as_array(host.G[a], shape=(host.NT[a],)))
    assert (dev.N_vb, dev.N_cb, dev.nr0) == (host.N_vb, host.N_cb, host.nr0)
    H, D = table_array(host), table_array(dev)
    scale = np.maximum(np.abs(H).max(axis=1), 1e-300)
    err = np.abs(D - H).max(axis=1) / scale
    assert err.max() <= tol, (name, over, int(err.argmax()), float(err.max()))
    return inp, host, dev


def check_resident(solver, name, over):
    """The tables qd_tables_build_device leaves on the handle are the ones it returned."""
    inp, host, dev = check_device_tables(solver, name, over)
    H_dev = solver.hamiltonian(q.QD_FP64)
    solver.upload(host)
    H_host = solver.hamiltonian(q.QD_FP64)
    assert np.abs(H_dev - H_host).max() <= TOL * np.abs(H_host).max()


@pytest.fixture(scope="module")
def solver():
    s = q.Solver(0)
    yield s
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,over", CASES)
def test_device_tables_match_host_builder(solver, name, over):
    check_device_tables(solver, name, over)


@pytest.mark.gpu
@pytest.mark.parametrize("name,over", FULL)
def test_device_tables_full_size(solver, name, over):
    check_device_tables(solver, name, over)


@pytest.mark.gpu
def test_device_tables_are_resident(solver):
    check_resident(solver, "lens_9x9x9", ("-M_Gx=1", "-M_Gy=1", "-M_Gz=1", "-M_S_Gx=1", "-M_S_Gy=1", "-M_S_Gz=1"))


@pytest.mark.gpu
def test_device_tables_input_errors(solver):
    check_input_errors(solver)


def check_input_errors(solver):
    pp, sp = cfg("truncated_pyramid")
    # a commensurate strain cell too small for the electronic grid is an input error, as on the host
    inp = q.read_inputs(pp, sp, ("-M_Gx=4", "-M_Gy=4", "-M_Gz=4", "-M_S_Gx=2", "-M_S_Gy=2", "-M_S_Gz=2"))
    with pytest.raises(q.QdError) as host_err:
        q.build_tables(inp)
    with pytest.raises(q.QdError) as dev_err:
        solver.build_tables(inp)
    assert dev_err.value.code == host_err.value.code == q.QD_E_INPUT


# the reverse pass: the cases of the host reverse pass's finite-difference test (cell modes, piezo
# orders, isotropic, 100-slice pyramid) and no strain with a field
VJP_CASES = [
    ("lens_9x9x9", SMALL, 2, 1),
    ("lens_9x9x9", SMALL + ("-Piezo=1", "-Fz=7"), 2, 1),
    ("lens_9x9x9", SMALL + ("-Piezo=0",), 2, 1),
    ("cone_19x19x15", SMALL, 2, 1),
    ("lens_9x9x9", SMALL + ("-Isotropic=1",), 2, 1),
    ("lens_9x9x9", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=2", "-Strain_supercell_type=0"), 2, 1),
    ("truncated_pyramid", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-M_S_Gx=3", "-M_S_Gy=3", "-M_S_Gz=4"), 5, 6),
    ("lens_9x9x9", ("-M_Gx=2", "-M_Gy=2", "-M_Gz=2", "-Strain=0", "-Fx=7", "-Fz=-3"), 2, 1),
]


def check_device_vjp(solver, name, over, dot, wl, tol=1e-12):
    """qd_tables_vjp_device against the host reverse pass (itself pinned by central differences of the
    builder) for a random W and every kind of parameter."""
    pp, sp = cfg(name)
    inp = q.read_inputs(pp, sp, over)
    t = q.build_tables(inp)
    NT = tuple(t.NT)
    rng = np.random.default_rng(11)
    W = rng.standard_normal((q.QD_NTAB,) + NT) + 1j * rng.standard_normal((q.QD_NTAB,) + NT)
    md, mx = 1 << dot, (1 << dot) | (1 << wl)
    params = [(q.QD_DP_L3, md), (q.QD_DP_L12, md), (q.QD_DP_X, mx), (q.QD_DP_OZ, md), (q.QD_DP_OX, md),
              (q.QD_DP_FX, 0), (q.QD_DP_FZ, 0)]
    if inp.shapes[dot].shape == 6:
        params.append((q.QD_DP_L4, md))
    want = q.tables_vjp(inp, W, params)
    got = solver.tables_vjp(inp, W, params)
    # the contractions sum many terms of both signs in a different order on each side: relative
    # agreement, with a floor at the rounding of the largest gradient for the small ones
    err = np.abs(got - want) / (np.abs(want) + 0.01 * np.abs(want).max())
    assert err.max() <= tol, (name, over, err, want, got)
    if inp.shapes[dot].shape == 6:   # the host's argument checks hold on the device
        with pytest.raises(q.QdError) as e:
            solver.tables_vjp(inp, W, [(q.QD_DP_L1, md)])
        assert e.value.code == q.QD_E_UNSUPPORTED


@pytest.mark.gpu
@pytest.mark.parametrize("name,over,dot,wl", VJP_CASES)
def test_device_vjp_matches_host(solver, name, over, dot, wl):
    check_device_vjp(solver, name, over, dot, wl)


def check_device_shapes(solver, tmp_path):
    import test_shapes_cpu as ts
    from helpers import prism_paths
    pp, _ = prism_paths()
    G = ts._gvecs()
    for name, (shape, Ls) in DEVICE_SHAPES.items():
        for angles, extra in SHAPE_VARIANTS:
            sp, _ = ts._files(tmp_path, shape, Ls, angles)
            inp = q.read_inputs(pp, sp, ("-N_shape=2",) + extra)
            for cell in (0, 1):
                want = q.selftest_shape_host(inp, 1, G, cell=cell)
                got = solver.shape_transform(inp, 1, G, cell=cell)
                scale = abs(want[0])
                assert scale > 0
                assert np.abs(got - want).max() <= TOL * scale, (name, angles, extra, cell, np.abs(got - want) / scale)
                # exact derivatives: dual numbers on both sides (lengths), the centre's phase (device)
                for param in DERIV_PARAMS:
                    dw = q.selftest_shape_host(inp, 1, G, cell=cell, param=param)
                    dg = solver.shape_transform(inp, 1, G, cell=cell, param=param)
                    dscale = max(np.abs(dw).max(), 1e-300)
                    assert np.abs(dg - dw).max() <= 1e-11 * dscale, (name, angles, extra, cell, param)
    # the host's input checks hold on the device: a regular pyramid needs three sides
    sp, _ = ts._files(tmp_path, 7, (4.0, 5.0, 2.0, 0.0))
    inp = q.read_inputs(pp, sp, ("-N_shape=2",))
    with pytest.raises(q.QdError) as e:
        solver.shape_transform(inp, 1, G)
    assert e.value.code == q.QD_E_INPUT


@pytest.mark.gpu
def test_device_shapes_match_host(solver, tmp_path):
    check_device_shapes(solver, tmp_path)
