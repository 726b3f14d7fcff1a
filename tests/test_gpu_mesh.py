"""GPU parity of mesh import, voxelisation and the static-field precompute:
the library (C ABI) against the oracle on the same inputs, exact: imported
vertices and bounds bit for bit (mesh.rs transforms), flags and B_dyn after
voxelize_mesh, psi (left in E_dyn), B_stat after precompute_B and E_stat after
precompute_E. Geometry: synthetic boxes and spheres, and the reference's own
thruster meshes (stl/deeva_*.stl) in its setup_deeva_test placement
(src/setup.rs:427-447) at reduced resolution."""
import os
import sys

import numpy as np
import pytest

import ionsolver_amd as I
import oracle_lib as O
from ionsolver_amd import Field, ModelType
from meshes import box, icosphere, write_stl

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mesh_ref as MR  # noqa: E402

pytestmark = pytest.mark.gpu

STL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stl")


def make_lbm(shape, mhd=True, lbm_length=None, ecr=False):
    cfg = I.LbmConfig()
    cfg.velocity_set = I.VelocitySet.D3Q19
    cfg.float_type = I.FloatType.FP32
    cfg.n_x, cfg.n_y, cfg.n_z = shape
    cfg.ext_volume_force = mhd
    cfg.ext_magneto_hydro = mhd
    cfg.mhd_lod_depth = 1
    cfg.ext_subgrid_ecr = ecr
    if lbm_length is not None:  # setup_deeva_test units (src/setup.rs:397) at lattice length lbm_length
        cfg.units.set(lbm_length, 1.0, 1.0, 1.0, 1.0, 0.1, 1.0, 10e-8, 1.0, 50000.0)
    return I.Lbm(cfg)


def test_import_matches_the_restatement(tmp_path):
    lbm = make_lbm((64, 128, 64), lbm_length=64.0)
    m = lbm.config.units.m
    for name in ("deeva_disk_magnet", "deeva_quartz_tube", "deeva_e_plate1"):
        p = os.path.join(STL, name + ".stl")
        i = lbm.import_mesh(p, 1.0, 32.0005, 0.0, 32.0)
        v, b = lbm.mesh(i)
        rv, lo, hi = MR.import_mesh(p, 1.0, 32.0005, 0.0, 32.0, 0, 0, 0, units_m=m)
        assert np.array_equal(v, rv) and np.array_equal(b["p_min"], lo) and np.array_equal(b["p_max"], hi)
    p = write_stl(tmp_path / "b.stl", box((0.0, 0.0, 0.0), (1.0, 2.0, 3.0)))
    i = lbm.import_mesh_reposition(p, 30.5, 60.25, 31.0, 0.0, 0.0, 0.0, 20.0)
    v, b = lbm.mesh(i)
    rv, lo, hi = MR.import_mesh_reposition(p, 30.5, 60.25, 31.0, 0, 0, 0, 20.0, (64, 128, 64))
    assert np.array_equal(v, rv) and np.array_equal(b["p_min"], lo)
    # rotations: sin/cos from the C library vs double rounded to float
    i = lbm.import_mesh_reposition(p, 30.5, 60.25, 31.0, 15.0, -40.0, 110.0, 20.0)
    v, _ = lbm.mesh(i)
    rv, _, _ = MR.import_mesh_reposition(p, 30.5, 60.25, 31.0, 15.0, -40.0, 110.0, 20.0, (64, 128, 64))
    np.testing.assert_allclose(v, rv, rtol=0, atol=4e-6 * np.abs(rv).max())


def test_import_errors(tmp_path):
    lbm = make_lbm((16, 16, 16))
    with pytest.raises(I.IonError, match="could not open"):
        lbm.import_mesh(str(tmp_path / "missing.stl"), 1.0, 0, 0, 0)
    with open(tmp_path / "ascii.stl", "wb") as f:
        f.write(b"solid x\nendsolid x\n" + b" " * 100)
    with pytest.raises(I.IonError, match="binary"):
        lbm.import_mesh(str(tmp_path / "ascii.stl"), 1.0, 0, 0, 0)
    with pytest.raises(I.IonError, match="out of range"):
        lbm.voxelise_mesh(3, ModelType.Solid)
    fluid = make_lbm((16, 16, 16), mhd=False)
    with pytest.raises(I.IonError, match="magneto_hydro"):
        fluid.precompute_b()


def run_both(shape, meshes, lbm_length=None, seed=0, precompute=True):
    """meshes: list of (STL path, origin, model type, value), imported with
    import_mesh at scale 1 and voxelised by the library and by the oracle."""
    lbm = make_lbm(shape, lbm_length=lbm_length)
    n = shape[0] * shape[1] * shape[2]
    rng = np.random.default_rng(seed)
    flags0 = rng.choice(np.array([0x00, 0x00, 0x20, 0x02], np.uint8), n)
    b_stat0 = (1e-3 * rng.standard_normal(3 * n)).astype(np.float32)
    e_stat0 = (1e-3 * rng.standard_normal(3 * n)).astype(np.float32)
    lbm.write_field(Field.FLAGS, flags0)
    lbm.write_field(Field.B_STAT, b_stat0)
    lbm.write_field(Field.E_STAT, e_stat0)
    flags, b_dyn = flags0.copy(), np.zeros(3 * n, np.float32)
    k = lbm.config.kernel_constants()
    units = lbm.config.units
    for path, origin, model, value in meshes:
        i = lbm.import_mesh(path, 1.0, *origin)
        v, b = lbm.mesh(i)
        lbm.voxelise_mesh(i, model, value)
        mx = my = mz = 0.0
        if model == ModelType.Magnet:
            conv = np.float32(units.a) / np.float32(units.m)
            mx, my, mz = (np.float32(c) / conv for c in value)
        elif model in (ModelType.Charged, ModelType.ChargedECR):
            mx = np.float32(value[0]) / (np.float32(units.a) * np.float32(units.s))
        O.voxelize(shape, flags, b_dyn, v, b["p_min"], b["p_max"], O.MODEL_FLAG[int(model)], (mx, my, mz))
    assert np.array_equal(lbm.read_field(Field.FLAGS), flags)
    assert np.array_equal(lbm.read_field(Field.B_DYN), b_dyn)
    if not precompute:
        return lbm, flags
    b_stat, e_stat = b_stat0.copy(), e_stat0.copy()
    lbm.precompute_b()
    psi = O.precompute_b(shape, flags, b_dyn, b_stat, k["kmu0"])
    e_dyn = lbm.read_field(Field.E_DYN)
    assert np.array_equal(e_dyn[:psi.size], psi)
    assert np.array_equal(lbm.read_field(Field.B_STAT), b_stat)
    lbm.precompute_e()
    O.precompute_e(shape, flags, b_dyn, e_stat, k["ke"])
    assert np.array_equal(lbm.read_field(Field.E_STAT), e_stat)
    return lbm, flags


def test_voxelize_and_precompute_synthetic(tmp_path):
    shape = (24, 20, 22)
    sphere, _, _ = icosphere((12.3, 9.6, 11.2), 5.7, subdiv=3)
    meshes = [(write_stl(tmp_path / "s.stl", sphere), (0.0, 0.0, 0.0), ModelType.Magnet, (1.0e5, -2.0e5, 3.0e5)),
              (write_stl(tmp_path / "b.stl", box((2.3, 1.6, 3.15), (8.7, 18.4, 6.8))), (0.0, 0.0, 0.0),
               ModelType.Solid, (0.0, 0.0, 0.0)),
              (write_stl(tmp_path / "p.stl", box((15.45, 3.7, 16.2), (21.85, 7.3, 17.6))), (0.0, 0.0, 0.0),
               ModelType.Charged, (2.5e-13, 0.0, 0.0)),
              (write_stl(tmp_path / "q.stl", box((15.45, 12.7, 16.2), (21.85, 16.3, 17.6))), (0.0, 0.0, 0.0),
               ModelType.ChargedECR, (-2.5e-13, 0.0, 0.0))]
    # lattice lengths in metres of one cell = 1 (units default), so the meshes are placed as written
    lbm, flags = run_both(shape, meshes)
    assert np.count_nonzero(flags & 0x10) > 400 and np.count_nonzero((flags & 0x1F) == 0x05) > 10


def test_voxelize_box_with_an_open_sheet(tmp_path):
    """A box and an open square sheet in front of it along the rays: flags and
    B_dyn identical to the oracle's, and the same as for the box alone (the
    reference's parity rule leaves an open sheet out)."""
    shape = (16, 16, 16)
    solid = box((6.3, 3.6, 4.15), (13.7, 10.4, 10.8))
    sheet = np.array([[[2.5, 3.0, 3.5], [2.5, 11.0, 3.5], [2.5, 11.0, 11.5]],
                      [[2.5, 3.0, 3.5], [2.5, 11.0, 11.5], [2.5, 3.0, 11.5]]], np.float32)
    both = write_stl(tmp_path / "bs.stl", np.concatenate([solid, sheet]))
    alone = write_stl(tmp_path / "b.stl", solid)
    _, flags_both = run_both(shape, [(both, (0.0, 0.0, 0.0), ModelType.Solid, (0.0, 0.0, 0.0))], precompute=False)
    _, flags_alone = run_both(shape, [(alone, (0.0, 0.0, 0.0), ModelType.Solid, (0.0, 0.0, 0.0))], precompute=False)
    assert np.count_nonzero((flags_alone & 0x1F) == 0x01) > 100
    assert np.array_equal(flags_both, flags_alone)


def test_precompute_e_ecr_writes_e_var(tmp_path):
    """precompute_E_ECR (src/lbm/mod.rs:319-331, domain.rs:569-578) runs the
    static_e_from_mesh kernel into E_var instead of E_stat: E_var (zero
    after allocation) gains the Coulomb field of the charged cells, E_stat
    is untouched; without ext_subgrid_ecr the call is rejected."""
    shape = (20, 18, 16)
    lbm = make_lbm(shape, ecr=True)
    n = shape[0] * shape[1] * shape[2]
    rng = np.random.default_rng(9)
    e_stat0 = (1e-3 * rng.standard_normal(3 * n)).astype(np.float32)
    lbm.write_field(Field.E_STAT, e_stat0)
    flags = np.zeros(n, np.uint8)
    b_dyn = np.zeros(3 * n, np.float32)
    k = lbm.config.kernel_constants()
    units = lbm.config.units
    p = write_stl(tmp_path / "p.stl", box((4.45, 3.7, 5.2), (14.85, 7.3, 9.6)))
    i = lbm.import_mesh(p, 1.0, 0.0, 0.0, 0.0)
    v, b = lbm.mesh(i)
    lbm.voxelise_mesh(i, ModelType.ChargedECR, (3.0e-13, 0.0, 0.0))
    c = np.float32(3.0e-13) / (np.float32(units.a) * np.float32(units.s))
    O.voxelize(shape, flags, b_dyn, v, b["p_min"], b["p_max"], O.MODEL_FLAG[int(ModelType.ChargedECR)], (c, 0.0, 0.0))
    assert np.array_equal(lbm.read_field(Field.FLAGS), flags) and np.count_nonzero(flags) > 50
    lbm.precompute_e_ecr()
    e_var = np.zeros(3 * n, np.float32)
    O.precompute_e(shape, flags, b_dyn, e_var, k["ke"])
    assert np.array_equal(lbm.read_field(Field.E_VAR), e_var) and np.abs(e_var).max() > 0
    assert np.array_equal(lbm.read_field(Field.E_STAT), e_stat0)
    with pytest.raises(I.IonError, match="subgrid_ecr"):
        make_lbm(shape).precompute_e_ecr()


def test_deeva_thruster_at_reduced_resolution():
    """setup_deeva_test (src/setup.rs:395-453) at a quarter of the lattice
    length per axis: 32 x 64 x 32 cells, the meshes imported at the
    reference's placement scaled by 1/4."""
    from ionsolver_amd.scenarios import deeva
    s = 0.25
    _, meshes = deeva(s)
    meshes = [(os.path.join(STL, f), o, m, v) for f, o, m, v in meshes]
    lbm, flags = run_both((32, 64, 32), meshes, lbm_length=128.0 * s)
    assert np.count_nonzero(flags & 0x10) > 1000


def test_psi_operators_match_the_general_ones():
    """The static-field kernel's specialised square root (every integer in
    [1, 2^24)) and division (sampled over its operand range) give the same
    bits as the compiler's general correctly rounded operators."""
    from ionsolver_amd.lbm import psi_math_check
    sq, dv, samples = psi_math_check()
    assert samples >= 1 << 28
    assert sq == 0 and dv == 0


@pytest.mark.parametrize("m", [(3.0e-13, 0.0, -2.0e-13), (1.0e13, 2.0e13, 0.0)])
def test_precompute_b_outside_the_specialised_range(tmp_path, m):
    """Magnetisations below 2^-40 or above 2^40 (lattice units) take the
    general operators; the result is still exact against the oracle."""
    shape = (18, 16, 20)
    sphere, _, _ = icosphere((9.3, 7.6, 10.2), 4.7, subdiv=2)
    meshes = [(write_stl(tmp_path / "s.stl", sphere), (0.0, 0.0, 0.0), ModelType.Magnet, m)]
    lbm, flags = run_both(shape, meshes)
    assert np.count_nonzero(flags & 0x10) > 100
