"""C ABI of libionsolver_hip.so without a GPU: the library loads, exports
every entry point include/ionsolver.h declares, and its host logic (config
defaults, units, kernel constants, validation, file errors) behaves like the
reference's (src/lbm/mod.rs:105-135, src/lbm/units.rs, src/file.rs)."""
import os
import re

import numpy as np
import pytest

import ionsolver_amd as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ionsolver.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ion_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = I.load_library()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(I.ABI)


def test_config_defaults_match_lbmconfig_new():
    c = I.LbmConfig()
    raw = c._to_c()
    d = I.lbm._Config()
    I.load_library().ion_config_default(d)
    for f, _ in d._fields_:
        if f == "units":
            continue
        assert getattr(d, f) == pytest.approx(getattr(raw, f)), f
    assert d.velocity_set == I.VelocitySet.D2Q9 and d.float_type == I.FloatType.FP16S
    assert d.mhd_lod_depth == 4 and d.graphics_active == 1
    assert d.nu == np.float32(1.0) / np.float32(6.0)


def test_kernel_constants_known_answers():
    """Default units (all 1): KE is Coulomb's constant 1/(4 pi eps0),
    KMU = mu0/(4 pi) = 1e-7, KKGE = -m_e/e, WQ = 1/(2*1 + 1/2) = 0.4,
    DEF_W = 1/(3 nu + 1/2) = 1 for nu = 1/6."""
    k = I.LbmConfig().kernel_constants()
    assert k["ke"] == pytest.approx(8.9875517923e9, rel=1e-6)
    assert k["kmu"] == pytest.approx(1e-7, rel=1e-6)
    assert k["kkge"] == pytest.approx(-9.1093837139e-31 / 1.602176634e-19, rel=1e-6)
    assert k["wq"] == np.float32(0.4)
    assert k["w"] == 1.0
    assert k["c"] == np.float32(0.57735027)


def test_units_set_matches_f32_restatement():
    """Units::set (src/lbm/units.rs:41-59), evaluated in f32."""
    f = np.float32
    args = (128.0, 1.0, 1.0, 1.0, 1.0, 0.1, 1.0, 1.225, 7e-13, 1.0)
    u = I.Units()
    u.set(*args)
    ll, lv, lr, lq, lt, sl, sv, sr, sq, st = [f(a) for a in args]
    m = f(sl / ll)
    kg = f(f(sr / lr) * f(f(m * m) * m))
    s = f(m / f(sv / lv))
    a = f(f(sq / lq) / s)
    assert (f(u.m), f(u.kg), f(u.s), f(u.a), f(u.k)) == (m, kg, s, a, f(st / lt))
    # Coulomb constant in lattice units: 1/(4 pi eps0_lu), eps0_lu = eps0 / (a^2 s^4 / (kg m^3))
    cfg = I.LbmConfig()
    cfg.units = u
    k = cfg.kernel_constants()
    eps0_lu = 8.8541878128e-12 / ((float(a) ** 2 * float(s) ** 4) / (float(kg) * float(m) ** 3))
    assert k["ke"] == pytest.approx(1 / (4 * np.pi * eps0_lu), rel=1e-5)


def test_lod_buffer_size():
    cfg = I.LbmConfig()
    cfg.ext_magneto_hydro = True
    for D, n in [(0, 1), (1, 9), (2, 73), (4, 4681)]:  # sum_k 8^k (src/lbm/domain.rs:110-117)
        cfg.mhd_lod_depth = D
        assert cfg.kernel_constants()["n_lod"] == n


def base_config(**kw):
    c = I.LbmConfig()
    c.velocity_set = I.VelocitySet.D3Q19
    c.n_x = c.n_y = c.n_z = 16
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("kw,code", [
    (dict(velocity_set=I.VelocitySet.D3Q27), -2),
    (dict(d_z=2), -2),
    (dict(n_x=0), -1),
    (dict(ext_magneto_hydro=True), -1),                      # needs ext_volume_force
    (dict(ext_magneto_hydro=True, ext_volume_force=True, n_x=20), -2),  # 20 not divisible by 2^4
    (dict(ext_subgrid_ecr=True), -1),                        # needs ext_magneto_hydro
    (dict(ext_subgrid_ecr=True, ext_magneto_hydro=True, ext_volume_force=True,
          velocity_set=I.VelocitySet.D3Q15), -2),             # built for D3Q19
    (dict(ext_force_field=True), -1),
    (dict(n_x=1024, n_y=1024, n_z=1025, float_type=I.FloatType.FP32), -2),  # 2^30 + 2^20 cells
    (dict(n_x=1024, n_y=1024, n_z=513, float_type=I.FloatType.FP64), -2),  # FP64 above 2^29 cells
    (dict(n_x=1 << 22, n_y=1 << 22, n_z=1 << 22), -2),  # 2^66 cells, 4 after a 64-bit wrap
    (dict(n_x=0xFFFFFFFF, n_y=0xFFFFFFFF, n_z=2), -2),
])
def test_invalid_configs_are_rejected_before_touching_the_device(kw, code):
    with pytest.raises(I.IonError) as e:
        I.Lbm(base_config(**kw))
    assert f"[{code}]" in str(e.value)


def test_load_errors(tmp_path):
    with pytest.raises(I.IonError, match=r"\[-4\]"):
        I.Lbm.load(str(tmp_path / "missing.ion"))
    p = tmp_path / "bad.ion"
    p.write_bytes(b"not an ionsolver file at all....")
    with pytest.raises(I.IonError, match="Invalid Format"):
        I.Lbm.load(str(p))
    p.write_bytes(b"IonSolver setup\n" + bytes(10))
    with pytest.raises(I.IonError, match=r"\[-5\]"):
        I.Lbm.load(str(p))
    # a header whose lattice (2^22 cubed) wraps to 4 cells in 64 bits, followed
    # by the sections of 4 cells: rejected before any section is read
    import struct
    head = (b"IonSolver setup\n" + struct.pack("<3B4f3IB3I4fBB", 2, 0, 2, 1.0, 1.0, 1.0, 1.0, 1 << 22, 1 << 22,
                                                 1 << 22, 0, 1, 1, 1, 0.1, 0.0, 0.0, 0.0, 0, 0))
    p.write_bytes(head + bytes(4 * 17))
    with pytest.raises(I.IonError, match=r"\[-2\].*2\^30"):
        I.Lbm.load(str(p))


def test_no_alloc_field_step_rejects_depth_beyond_its_scratch_array():
    """ion_update_e_b_noalloc holds at most 8^5 LOD sources; the depth check
    runs before any device work, so the buffer pointers are never used."""
    import ctypes
    from harness import make_config
    cfg = make_config(nx=64, ny=64, nz=64, mhd=True, lod_depth=6, float_type=2)
    buf = (ctypes.c_void_p * 8)(*([1] * 8))
    lib = I.load_library()
    assert lib.ion_update_e_b_noalloc(ctypes.byref(cfg._to_c()), buf, 1, None) != 0
    assert "scratch" in lib.ion_last_error().decode()


def test_device_codec_rejects_bad_arguments_before_touching_the_device():
    import ctypes
    lib = I.load_library()
    buf = (ctypes.c_uint8 * 64)()
    assert lib.ion_device_codec(2, 0, buf, buf, 4) != 0 and "FP16S or FP16C" in lib.ion_last_error().decode()
    assert lib.ion_device_codec(1, 2, buf, buf, 4) != 0 and "dir" in lib.ion_last_error().decode()
    assert lib.ion_device_codec(1, 0, None, buf, 4) != 0 and "null" in lib.ion_last_error().decode()
    assert lib.ion_device_codec(1, 0, buf, buf, 1 << 41) != 0 and "2^40" in lib.ion_last_error().decode()
    assert lib.ion_device_codec(1, 0, None, None, 0) == 0  # nothing to do
