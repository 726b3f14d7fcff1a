"""The save format byte for byte (FILE_LAYOUT.txt:1-60; file.rs:42-306),
through ion_load / ion_save, against files written and parsed here with
struct (little endian), independently of the library's own encoder:

* a file built per FILE_LAYOUT.txt (header, 62-byte config, flags, rho,
  u in x/y/z blocks, Q, N_C fixed charges, N_M magnets) loads into exactly
  those fields and that configuration; the float-type byte is read with the
  FloatType discriminants the reference's encoder writes (types.rs:75-83);
* an MHD file as the reference's encoder writes it (no N_C / N_M,
  file.rs:277-303) and a fluid-only file (no Q) load too;
* ion_save's output, parsed field by field, is that layout with N_C = N_M = 0;
* truncated files and trailing bytes are rejected (ION_ERR_FORMAT)."""
import struct

import numpy as np
import pytest

import ionsolver_amd as I
from ionsolver_amd import Field

pytestmark = pytest.mark.gpu

HEADER = b"IonSolver setup\n"
CONFIG = struct.Struct("<BBBffffIIIBIIIffffBB")  # 62 bytes, FILE_LAYOUT.txt:4-25


def config_bytes(vs, rt, ft, units, n, d, nu, f, ext, depth):
    return CONFIG.pack(vs, rt, ft, *units, *n, 0, *d, nu, *f, ext, depth)


def fields(rng, n, mhd):
    N = int(np.prod(n))
    out = dict(flags=rng.choice(np.array([0, 1, 2, 0x11], np.uint8), N),
               rho=(1 + 0.01 * rng.standard_normal(N)).astype(np.float32),
               u=(0.02 * rng.standard_normal(3 * N)).astype(np.float32))
    if mhd:
        out["q"] = (0.05 * rng.standard_normal(N)).astype(np.float32)
    return out


def body(fl):
    b = fl["flags"].tobytes() + fl["rho"].astype("<f4").tobytes() + fl["u"].astype("<f4").tobytes()
    if "q" in fl:
        b += fl["q"].astype("<f4").tobytes()
    return b


def fixed_charges_and_magnets():
    charges = [(5, 1.5e-12), (17, -2.0e-12)]
    magnets = [(3, (0.0, 1e6, 0.0))]
    b = struct.pack("<I", len(charges)) + b"".join(struct.pack("<Qf", i, q) for i, q in charges)
    b += struct.pack("<I", len(magnets)) + b"".join(struct.pack("<Q3f", i, *m) for i, m in magnets)
    return b


def check_loaded(lbm, n, fl, ft, vs, ext, depth, nu, f, units):
    c = lbm.config
    assert (c.n_x, c.n_y, c.n_z) == n and int(c.float_type) == ft and int(c.velocity_set) == vs
    assert (c.d_x, c.d_y, c.d_z) == (1, 1, 1) and c.mhd_lod_depth == depth
    assert c.ext_equilibrium_boudaries == bool(ext & 1) and c.ext_volume_force == bool(ext & 2)
    assert c.ext_force_field == bool(ext & 4) and c.ext_magneto_hydro == bool(ext & 8)
    assert np.float32(c.nu) == np.float32(nu)
    assert np.array_equal(np.float32([c.f_x, c.f_y, c.f_z]), np.float32(f))
    assert np.array_equal(np.float32([c.units.m, c.units.kg, c.units.s, c.units.a]), np.float32(units))
    assert np.array_equal(lbm.read_field(Field.FLAGS), fl["flags"])
    assert np.array_equal(lbm.read_field(Field.RHO), fl["rho"])
    assert np.array_equal(lbm.read_field(Field.U), fl["u"])
    if "q" in fl:
        assert np.array_equal(lbm.read_field(Field.Q), fl["q"])


UNITS = (7.8125e-4, 5.84e-10, 7.8125e-4, 8.96e-10)
NU, FORCE = 0.0333333, (1e-5, -2e-5, 3e-5)


@pytest.mark.parametrize("ft", [0, 1, 2])
def test_load_a_file_built_from_the_layout(tmp_path, ft):
    """FP16S = 0, FP16C = 1, FP32 = 2 as the reference's encoder writes them
    (its decoder swaps 0 and 1, file.rs:68-73; this backend reads what the
    encoder writes)."""
    rng = np.random.default_rng(40 + ft)
    n = (8, 6, 4)
    ext = 1 | 2 | 8
    fl = fields(rng, n, True)
    data = HEADER + config_bytes(2, 1, ft, UNITS, n, (1, 1, 1), NU, FORCE, ext, 1) + body(fl)
    data += fixed_charges_and_magnets()
    p = tmp_path / "layout.ion"
    p.write_bytes(data)
    lbm = I.Lbm.load(str(p))
    try:
        check_loaded(lbm, n, fl, ft, 2, ext, 1, NU, FORCE, UNITS)
        assert int(lbm.config.relaxation_time) == 1
    finally:
        lbm.close()


def test_load_reference_encoder_shapes(tmp_path):
    rng = np.random.default_rng(41)
    n = (8, 4, 4)
    # MHD as file.rs:encode writes it: Q, then nothing (N_C / N_M commented out)
    fl = fields(rng, n, True)
    p = tmp_path / "ref_mhd.ion"
    p.write_bytes(HEADER + config_bytes(2, 0, 2, UNITS, n, (1, 1, 1), NU, FORCE, 2 | 8, 2) + body(fl))
    lbm = I.Lbm.load(str(p))
    try:
        check_loaded(lbm, n, fl, 2, 2, 2 | 8, 2, NU, FORCE, UNITS)
    finally:
        lbm.close()
    # fluid only: no charge section
    fl = fields(rng, n, False)
    p = tmp_path / "ref_fluid.ion"
    p.write_bytes(HEADER + config_bytes(2, 0, 2, UNITS, n, (1, 1, 1), NU, (0.0, 0.0, 0.0), 0, 4) + body(fl))
    lbm = I.Lbm.load(str(p))
    try:
        check_loaded(lbm, n, fl, 2, 2, 0, 4, NU, (0.0, 0.0, 0.0), UNITS)
    finally:
        lbm.close()


def test_save_writes_the_layout(tmp_path):
    rng = np.random.default_rng(42)
    cfg = I.LbmConfig()
    cfg.velocity_set, cfg.float_type = I.VelocitySet.D3Q19, I.FloatType.FP16C
    cfg.relaxation_time = I.RelaxationTime.Trt
    cfg.n_x, cfg.n_y, cfg.n_z = 8, 4, 6
    cfg.nu = NU
    cfg.ext_volume_force = cfg.ext_magneto_hydro = True
    cfg.f_x, cfg.f_y, cfg.f_z = FORCE
    cfg.mhd_lod_depth = 1
    for k, v in zip(("m", "kg", "s", "a"), UNITS):
        setattr(cfg.units, k, v)
    lbm = I.Lbm(cfg)
    fl = fields(rng, (8, 4, 6), True)
    try:
        lbm.write_field(Field.FLAGS, fl["flags"])
        lbm.write_field(Field.RHO, fl["rho"])
        lbm.write_field(Field.U, fl["u"])
        lbm.write_field(Field.Q, fl["q"])
        p = tmp_path / "saved.ion"
        lbm.save(str(p))
    finally:
        lbm.close()
    b = p.read_bytes()
    N = 8 * 4 * 6
    assert b[:16] == HEADER
    vals = CONFIG.unpack_from(b, 16)
    assert vals[:3] == (2, 1, 1)  # D3Q19, TRT, FP16C
    assert np.array_equal(np.float32(vals[3:7]), np.float32(UNITS))
    assert vals[7:10] == (8, 4, 6) and vals[10] == 0 and vals[11:14] == (1, 1, 1)
    assert np.float32(vals[14]) == np.float32(NU) and np.array_equal(np.float32(vals[15:18]), np.float32(FORCE))
    assert vals[18] == 2 | 8 and vals[19] == 1
    o = 16 + CONFIG.size
    assert CONFIG.size == 62
    assert np.array_equal(np.frombuffer(b, np.uint8, N, o), fl["flags"])
    o += N
    assert np.array_equal(np.frombuffer(b, "<f4", N, o), fl["rho"])
    o += 4 * N
    assert np.array_equal(np.frombuffer(b, "<f4", 3 * N, o), fl["u"])
    o += 12 * N
    assert np.array_equal(np.frombuffer(b, "<f4", N, o), fl["q"])
    o += 4 * N
    assert struct.unpack_from("<II", b, o) == (0, 0)  # N_C, N_M (FILE_LAYOUT.txt:46-58)
    assert len(b) == o + 8


def test_malformed_files_are_rejected(tmp_path):
    rng = np.random.default_rng(43)
    n = (4, 4, 4)
    fl = fields(rng, n, True)
    good = HEADER + config_bytes(2, 0, 2, UNITS, n, (1, 1, 1), NU, FORCE, 2 | 8, 1) + body(fl)
    for name, data in (("truncated", good[:-3]), ("trailing", good + fixed_charges_and_magnets() + b"\0"),
                       ("short_header", good[:40])):
        p = tmp_path / f"{name}.ion"
        p.write_bytes(data)
        with pytest.raises(I.IonError, match=r"\[-5\]"):
            I.Lbm.load(str(p))
