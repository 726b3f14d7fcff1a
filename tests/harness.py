"""Builds matched GPU (libionsolver_hip.so through the C ABI) and oracle
simulations from the same configuration and the same seeded inputs, and
compares every buffer.

Comparison rule: exact. Float buffers are compared by value (-0 == +0, NaN
positions must match); DDF storage codes (FP16S/FP16C) are compared as raw
16-bit codes.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import oracle_lib as O  # noqa: E402

import ionsolver_amd as I  # noqa: E402
from ionsolver_amd import Field  # noqa: E402

FIELD_OF = dict(fi=Field.FI, rho=Field.RHO, u=Field.U, flags=Field.FLAGS, F=Field.F, e_stat=Field.E_STAT,
                b_stat=Field.B_STAT, e_dyn=Field.E_DYN, b_dyn=Field.B_DYN, fqi=Field.FQI, q=Field.Q,
                qu_lod=Field.QU_LOD, ei=Field.EI, e_var=Field.E_VAR, eti=Field.ETI, et=Field.ET)


def oracle_for(cfg, float_type=None):
    """OracleSim with the kernel constants of cfg (float_type may be O.FP64)."""
    k = cfg.kernel_constants()
    ext = 0
    if cfg.ext_equilibrium_boudaries:
        ext |= O.EXT_EQUILIBRIUM_BOUNDARIES
    if cfg.ext_volume_force:
        ext |= O.EXT_VOLUME_FORCE
    if cfg.ext_force_field:
        ext |= O.EXT_FORCE_FIELD
    if cfg.ext_magneto_hydro:
        ext |= O.EXT_MAGNETO_HYDRO
    if cfg.graphics_config.graphics_active:
        ext |= O.EXT_UPDATE_FIELDS
    if cfg.ext_charge_inlet:
        ext |= O.EXT_CHARGE_INLET
    if cfg.ext_subgrid_ecr:
        ext |= O.EXT_SUBGRID_ECR
    ft = int(cfg.float_type) if float_type is None else float_type
    return O.OracleSim(cfg.n_x, cfg.n_y, cfg.n_z, velocity_set=int(cfg.velocity_set),
                       relaxation_time=int(cfg.relaxation_time), float_type=ft, lod_depth=cfg.mhd_lod_depth,
                       ext=ext, w=k["w"], fx=cfg.f_x, fy=cfg.f_y, fz=cfg.f_z, ke=k["ke"], kmu=k["kmu"],
                       kkge=k["kkge"], wq=k["wq"], kkbme=k["kkbme"], keabs=k["keabs"],
                       ei_rho0=cfg.electron_density0)


def random_inputs(cfg, rng, solid_frac=0.1, e_frac=0.05, charge=0.05, field=0.05):
    """Seeded random buffers covering the flag classes the kernels branch on."""
    n = cfg.n_x * cfg.n_y * cfg.n_z
    flags = np.zeros(n, np.uint8)
    r = rng.uniform(size=n)
    flags[r < solid_frac] = 0x01                                   # solid
    odd = (r >= solid_frac) & (r < solid_frac + 0.05)
    flags[odd] = rng.choice(np.array([0x11, 0x09, 0x05, 0x20, 0x21], np.uint8), size=int(odd.sum()))  # fluid unless ==1
    if cfg.ext_equilibrium_boudaries:
        e = (r >= solid_frac + 0.05) & (r < solid_frac + 0.05 + e_frac)
        flags[e] = 0x02
    arrays = dict(flags=flags)
    arrays["rho"] = (1.0 + 0.01 * rng.standard_normal(n)).astype(np.float32)
    arrays["u"] = (0.02 * rng.standard_normal(3 * n)).astype(np.float32)
    if cfg.ext_force_field:
        arrays["F"] = (1e-4 * rng.standard_normal(3 * n)).astype(np.float32)
    if cfg.ext_magneto_hydro:
        q = charge * rng.standard_normal(n)
        q[rng.uniform(size=n) < 0.2] = 0.0
        arrays["q"] = q.astype(np.float32)
        arrays["e_stat"] = (field * rng.standard_normal(3 * n)).astype(np.float32)
        arrays["b_stat"] = (field * rng.standard_normal(3 * n)).astype(np.float32)
    if cfg.ext_subgrid_ecr:
        arrays["e_var"] = (field * rng.standard_normal(3 * n)).astype(np.float32)
        arrays["et"] = (0.05 * rng.uniform(size=n)).astype(np.float32)
    return arrays


def write_both(gpu, orc, arrays):
    for k, v in arrays.items():
        gpu.write_field(FIELD_OF[k], v)
        dst = orc.buf(k)
        dst[:] = v.astype(dst.dtype)


def make_config(nx=16, ny=16, nz=16, velocity_set=2, float_type=2, relaxation_time=0, mhd=False, lod_depth=2,
                eqb=True, vforce=True, ffield=False, graphics=True, force=(1e-5, -2e-5, 3e-5), nu=0.02, units=None,
                charge_inlet=False, ecr=False, electron_density0=0.0):
    cfg = I.LbmConfig()
    cfg.velocity_set = I.VelocitySet(velocity_set)
    cfg.float_type = I.FloatType(float_type)
    cfg.relaxation_time = I.RelaxationTime(relaxation_time)
    cfg.n_x, cfg.n_y, cfg.n_z = nx, ny, nz
    cfg.nu = nu
    cfg.ext_equilibrium_boudaries = eqb
    cfg.ext_volume_force = vforce or mhd
    cfg.ext_force_field = ffield
    cfg.ext_magneto_hydro = mhd
    cfg.mhd_lod_depth = lod_depth
    cfg.ext_charge_inlet = charge_inlet
    cfg.ext_subgrid_ecr = ecr
    cfg.graphics_config.graphics_active = graphics
    cfg.electron_density0 = electron_density0
    if vforce or mhd:
        cfg.f_x, cfg.f_y, cfg.f_z = force
    if mhd:
        from ionsolver_amd.scenarios import MHD_UNITS
        cfg.units.set(**(units or MHD_UNITS))
    return cfg


def make_pair(rng, **kw):
    cfg = make_config(**kw)
    gpu = I.Lbm(cfg)
    orc = oracle_for(gpu.config)
    arrays = random_inputs(gpu.config, rng)
    write_both(gpu, orc, arrays)
    return gpu, orc, dict(cfg=gpu.config, arrays=arrays)


def fields_of(cfg):
    names = ["fi", "rho", "u", "flags"]
    if cfg.ext_force_field:
        names.append("F")
    if cfg.ext_magneto_hydro:
        names += ["e_stat", "b_stat", "e_dyn", "b_dyn", "fqi", "q", "qu_lod", "ei"]
        if cfg.ext_subgrid_ecr:
            names += ["e_var", "eti", "et"]
    return names


def same(a, b):
    """Exact comparison (values for floats, codes for integers); returns the
    number of differing elements."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return -1
    if a.dtype.kind == "f":
        eq = (a == b) | (np.isnan(a) & np.isnan(b))
        return int((~eq).sum())
    return int((a != b).sum())


def compare_all(gpu, orc, info, names=None):
    bad = []
    for name in names or fields_of(info["cfg"]):
        g = gpu.read_field(FIELD_OF[name])
        o = orc.buf(name)
        d = same(g, o)
        if d:
            idx = np.flatnonzero(g != o)[:3] if d > 0 else []
            bad.append(f"{name}: {d} differ (e.g. at {list(idx)}: gpu {g[idx]} vs oracle {o[idx]})")
    return bad


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / (den if den > 0 else 1.0))
