"""Synthetic workloads of the benchmark and the tests, as host-side setups.

Each setup returns an ``LbmConfig`` plus the host arrays a caller writes into
the domain buffers before ``initialize()``, the way the reference's setups do
with ``bwrite!`` (src/setup.rs). Arrays are SoA in the reference layout
(n = x + (y + z*Ny)*Nx, vector fields as [x-block, y-block, z-block]).
"""
import os

import numpy as np

from .lbm import FloatType, LbmConfig, ModelType, VelocitySet

# Units of the MHD workloads: the reference's setup_field_vis scaling
# (src/setup.rs:244-246: 128 cells = 0.1 m, 1 m/s, air density) with the
# charge unit chosen so that the Coulomb constant is ~1e-2 in lattice units.
MHD_UNITS = dict(lbm_length=128.0, lbm_velocity=1.0, lbm_rho=1.0, lbm_charge=1.0, lbm_temp=1.0, si_length=0.1,
                 si_velocity=1.0, si_rho=1.2250, si_charge=7e-13, si_temp=1.0)

# Units of the plasma workloads (Units fields set directly): KE = 1.0e-6,
# KMU = 2.8e-5, KKGE = -2.50 in lattice units. With MHD_UNITS KKGE is
# -6.8e-15, so the electron-gas velocity shift F / (2 rho_e KKGE)
# (sim_kernels.cl:651-653) saturates at the clamp +-c for any field and the
# electron gas runs away within a few steps whatever its density; here it
# moves by (E + u_e x B) / 5 per step. The Coulomb constant sets how long the
# state lives: with KE = 1.0e-3 the 256^3 blob workload's velocities reach
# the clamp and its state turns NaN between steps 300 and 400, with
# KE = 1.0e-5 its peak velocity still grows (0.9e-3 at step 500, 2.5e-3 at
# step 3000), with KE = 1.0e-6 it stays below 2e-4 through 3000 steps while
# the charge peaks relax (0.028 -> 0.004).
PLASMA_UNITS = dict(m=1.58e9, kg=2.27e-12, s=1.0, a=1.0)

# Strongly coupled variant (KE = 1.0e-3, KMU = 2.8e-4, KKGE = -2.50): flow and
# dynamic fields rise well above the f32 rounding of the state and of the
# static field within tens of steps, and the state lives a few hundred steps:
# for short comparisons of arithmetic and storage formats.
PLASMA_UNITS_STRONG = dict(m=1.58e8, kg=2.27e-12, s=1.0, a=1.0)


def use_plasma_units(cfg, n0, units=None):
    """PLASMA_UNITS (or the given units) and an initial electron gas of
    density n0 on non-solid cells (LbmConfig.electron_density0, a build
    extension: the reference's initialize starts it at density 0)."""
    for k, v in (PLASMA_UNITS if units is None else units).items():
        setattr(cfg.units, k, v)
    cfg.electron_density0 = n0


def cell_coords(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def fluid_tgv(n=64, tau=0.6, u0=0.05, float_type=FloatType.FP32, velocity_set=VelocitySet.D3Q19):
    """Fluid-only D3Q19 BGK Taylor-Green vortex in a periodic n^3 box.

    Divergence-free 2D form u = u0 cos(kx) sin(ky), v = -u0 sin(kx) cos(ky),
    w = 0, with the matching pressure rho = 1 - 3 u0^2/4 (cos 2kx + cos 2ky),
    on cell-centred coordinates x + 0.5 - n/2 (src/setup.rs:512-514)."""
    cfg = LbmConfig()
    cfg.velocity_set = velocity_set
    cfg.float_type = float_type
    cfg.n_x = cfg.n_y = cfg.n_z = n
    cfg.nu = float(np.float32((np.float32(tau) - np.float32(0.5)) / np.float32(3.0)))
    x, y, z = cell_coords(n, n, n)
    k = 2.0 * np.pi / n
    fx, fy = x + 0.5 - 0.5 * n, y + 0.5 - 0.5 * n
    ux = u0 * np.cos(k * fx) * np.sin(k * fy)
    uy = -u0 * np.sin(k * fx) * np.cos(k * fy)
    uz = np.zeros_like(ux)
    rho = 1.0 - 0.75 * u0 * u0 * (np.cos(2 * k * fx) + np.cos(2 * k * fy))
    u = np.concatenate([ux, uy, uz]).astype(np.float32)
    return cfg, dict(rho=rho.astype(np.float32), u=u)


def mhd_blobs(n=256, seed=1234, float_type=FloatType.FP32, lod_depth=4, n_blobs=8, sigma=6.0, amp=0.05,
              b0=0.05, dipole=3200.0, cyl_r=40.0, nu=None, shape=None, plasma=None):
    """D3Q19 BGK + MHD: Gaussian charge blobs, uniform B0 plus a point dipole,
    Guo Lorentz forcing, a solid z-aligned cylinder (flags 0x01).

    Blob centres are uniform in [2 sigma, n - 2 sigma)^3 outside the
    cylinder; sign +-1 and amplitude U[0.5, 1]*amp come from the same seeded
    stream. With plasma = n0 the blobs sit on a neutral background: the ion
    charge density written to Q (the charge DDFs' density) is n0 + blobs +
    1e-4 normal noise (same stream) and the electron gas starts at density n0
    (use_plasma_units), so that the cell charge Q = rho_q - rho_e is the blobs
    plus the noise; otherwise (the reference's initial state) the charge
    density is the blobs and the electron gas starts at 0.
    Returns (cfg, arrays, blobs)."""
    nx, ny, nz = shape if shape is not None else (n, n, n)
    cfg = LbmConfig()
    cfg.velocity_set = VelocitySet.D3Q19
    cfg.float_type = float_type
    cfg.n_x, cfg.n_y, cfg.n_z = nx, ny, nz
    cfg.units.set(**MHD_UNITS)
    cfg.nu = float(np.float32(0.1) / np.float32(3.0)) if nu is None else nu
    cfg.ext_volume_force = True
    cfg.ext_magneto_hydro = True
    cfg.mhd_lod_depth = lod_depth
    rng = np.random.default_rng(seed)
    x, y, z = cell_coords(nx, ny, nz)
    cx, cy = 0.5 * nx, 0.5 * ny
    r_axis = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    scale = nx / 256.0
    solid = r_axis < cyl_r * scale
    flags = np.where(solid, 1, 0).astype(np.uint8)
    blobs = []
    lo = 2 * sigma
    tries = 0
    while len(blobs) < n_blobs:
        tries += 1
        if tries > 100000:
            raise ValueError(f"no room for {n_blobs} blobs of sigma {sigma} beside the cylinder on {nx}x{ny}x{nz}")
        c = rng.uniform([lo, lo, lo], [nx - lo, ny - lo, nz - lo])
        if np.hypot(c[0] - cx, c[1] - cy) < cyl_r * scale + 2 * sigma:
            continue
        sign = 1.0 if rng.uniform() < 0.5 else -1.0
        a = sign * amp * rng.uniform(0.5, 1.0)
        blobs.append((float(c[0]), float(c[1]), float(c[2]), float(a)))
    q = np.zeros(nx * ny * nz)
    for bx, by, bz, a in blobs:
        q += a * np.exp(-((x - bx) ** 2 + (y - by) ** 2 + (z - bz) ** 2) / (2 * sigma * sigma))
    if plasma is not None:
        use_plasma_units(cfg, plasma)
        # ion density fluctuations of 1e-4 around the background: every LOD
        # block carries charge from the first step, as it does once the
        # electron gas has moved
        q += plasma + 1e-4 * rng.standard_normal(q.size)
    q[solid] = 0.0
    # B_stat = B0 z + point dipole m z at the domain centre: (3 (m.r^) r^ - m) / |r|^3
    rx, ry, rz = x - cx, y - cy, z - 0.5 * nz
    r2 = rx * rx + ry * ry + rz * rz
    r2[r2 == 0] = np.inf
    r = np.sqrt(r2)
    mdotr = dipole * scale ** 3 * rz / r
    bxf = (3 * mdotr * rx / r) / r2 / r
    byf = (3 * mdotr * ry / r) / r2 / r
    bzf = (3 * mdotr * rz / r - dipole * scale ** 3) / r2 / r + b0
    b_stat = np.concatenate([bxf, byf, bzf]).astype(np.float32)
    arrays = dict(flags=flags, q=q.astype(np.float32), b_stat=b_stat)
    return cfg, arrays, blobs


def add_subgrid_ecr(cfg, arrays, et0=1e-3, e_amp=1e-3):
    """Switch on SUBGRID_ECR for a magnetised setup: a uniform initial
    electron temperature et0 and an oscillating field E_var of magnitude
    e_amp along x (lattice units), the direction a microwave feed would give."""
    n = cfg.n_x * cfg.n_y * cfg.n_z
    cfg.ext_subgrid_ecr = True
    e_var = np.zeros(3 * n, np.float32)
    e_var[:n] = e_amp
    arrays["e_var"] = e_var
    arrays["et"] = np.full(n, et0, np.float32)
    return cfg, arrays


def apply(lbm, arrays):
    """Write setup arrays into the domain buffers (the setups' bwrite! calls)."""
    from .lbm import Field
    names = dict(flags=Field.FLAGS, rho=Field.RHO, u=Field.U, q=Field.Q, e_stat=Field.E_STAT, b_stat=Field.B_STAT,
                 F=Field.F, e_var=Field.E_VAR, et=Field.ET)
    for k, v in arrays.items():
        lbm.write_field(names[k], v)


def random_lod(shape, D, rng, n_blobs=8, sigma=(3.0, 16.0), amp=0.05, vel=0.05, noise=0.02):
    """Top-level LOD sums (q, mean v) of a random charged flow, as written by
    stream_collide: 'n_blobs' Gaussian charge blobs (random centre, width,
    sign and amplitude) integrated over each LOD block (separable, exact cell
    sums), plus white block noise, and a smooth random velocity field.
    Returns the [n_top][4] float32 array for the LOD buffer's first entries."""
    nx, ny, nz = shape
    nd = 1 << D
    b = (nx // nd, ny // nd, nz // nd)
    q = np.zeros((nd, nd, nd))
    for _ in range(n_blobs):
        s = rng.uniform(*sigma)
        a = (1.0 if rng.uniform() < 0.5 else -1.0) * amp * rng.uniform(0.5, 1.0)
        prof = []
        for L, bb in zip((nx, ny, nz), b):
            c = rng.uniform(0, L)
            g = np.exp(-(np.arange(L) - c) ** 2 / (2 * s * s))
            prof.append(g.reshape(nd, bb).sum(1))
        q += a * prof[2][:, None, None] * prof[1][None, :, None] * prof[0][None, None, :]
    q += noise * amp * np.prod(b) * rng.standard_normal(q.shape) / np.sqrt(np.prod(b))
    Z, Y, X = np.meshgrid(*(np.arange(nd) + 0.5,) * 3, indexing="ij")
    v = []
    for _ in range(3):
        f = rng.uniform(-1, 1) * 0.5
        for _ in range(3):
            k = rng.integers(1, 4, 3) * 2 * np.pi / nd
            f = f + rng.uniform(-1, 1) * np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 2 * np.pi))
        v.append(vel * f / 2)
    out = np.stack([q.ravel(), v[0].ravel(), v[1].ravel(), v[2].ravel()], 1)
    return out.astype(np.float32)


def _loop_field(r, z, a, wire):
    """(B_r, B_z) of a circular current loop of radius a at z = 0 (unit
    prefactor mu0 I / (2 pi)), closed form with complete elliptic integrals;
    distances to the wire below `wire` are clamped (finite conductor)."""
    from scipy.special import ellipe, ellipk
    d2 = np.maximum((a - r) ** 2 + z * z, wire * wire)
    s2 = (a + r) ** 2 + z * z
    m = np.clip(4.0 * a * r / s2, 0.0, 1.0 - 1e-12)
    K, E = ellipk(m), ellipe(m)
    s = np.sqrt(s2)
    bz = (K + (a * a - r * r - z * z) / d2 * E) / s
    with np.errstate(divide="ignore", invalid="ignore"):
        br = np.where(r > 0, z / (r * s) * (-K + (a * a + r * r + z * z) / d2 * E), 0.0)
    return br, bz


def thruster(n=1024, float_type=None, lod_depth=4, tube_r=300.0, wall=4.0, magnet_r=350.0, b_center=None,
             u_inlet=0.02, q_inlet=0.02, n_blobs=8, sigma=12.0, amp=0.05, seed=1234, log=None, plasma=None,
             outlet=True):
    """Synthetic plasma-thruster workload: a z-aligned quartz tube (solid
    shell, flags 0x01) of radius tube_r, a ring magnet (current loop of
    radius magnet_r at mid-height, analytic field scaled to b_center at its
    centre) as B_stat, an inlet plane (flags 0x02, equilibrium boundary) at
    z = 0 inside the tube with u = (0, 0, u_inlet), with outlet an outlet
    plane of the same kind at z = n - 1, the tube filled with fluid moving
    at u_inlet and seeded with Gaussian charge blobs. With q_inlet set (None
    switches it off) the inlet also injects charge: the charge inlet
    extension holds the inlet cells at charge density q_inlet and the outlet
    cells at the background density. With plasma = n0 the charges sit on a
    neutral background (as in mhd_blobs): PLASMA_UNITS, the electron gas of
    every non-solid cell starts at density n0 and the ion charge density
    written to Q is n0 plus the above. b_center defaults to 0.05 without the
    plasma background and to 0.002 with it: the gas then carries the
    background's charge density n0 = 1, and at 0.05 the Lorentz force
    n0 (u x B) drives the axial velocity at the magnet to the clamp within
    about 300 steps at 256^3 (ion gyration frequency n0 B / rho = 0.05 per
    step), while at 0.002 the flow develops as in the uncharged tube. Lengths
    are given for n = 1024 and scaled with n. Arrays are built one z-layer at
    a time. Returns (cfg, arrays, meta)."""
    from .lbm import FloatType as FT
    ft = FT.FP16C if float_type is None else float_type
    sc = n / 1024.0
    rt, th, rm = tube_r * sc, max(wall * sc, 2.0), magnet_r * sc
    cfg = LbmConfig()
    cfg.velocity_set = VelocitySet.D3Q19
    cfg.float_type = ft
    cfg.n_x = cfg.n_y = cfg.n_z = n
    cfg.units.set(**MHD_UNITS)
    cfg.nu = float(np.float32(0.1) / np.float32(3.0))
    cfg.ext_volume_force = True
    cfg.ext_magneto_hydro = True
    cfg.ext_equilibrium_boudaries = True
    cfg.ext_charge_inlet = q_inlet is not None
    cfg.mhd_lod_depth = lod_depth
    if plasma is not None:
        use_plasma_units(cfg, plasma)
    n0 = 0.0 if plasma is None else plasma
    if b_center is None:
        b_center = 0.05 if plasma is None else 0.002
    N = n ** 3
    c = 0.5 * n
    yy, xx = np.meshgrid(np.arange(n, dtype=np.float64) - c, np.arange(n, dtype=np.float64) - c, indexing="ij")
    r = np.sqrt(xx * xx + yy * yy)
    # the loop field depends on r only: evaluated once per distinct radius of a layer
    r_u, r_inv = np.unique(r, return_inverse=True)
    r_inv = r_inv.reshape(r.shape)
    inside = r < rt
    shell = (r >= rt) & (r < rt + th)
    flags = np.zeros(N, np.uint8)
    u = np.zeros(3 * N, np.float32)
    q = np.zeros(N, np.float32)
    b = np.zeros(3 * N, np.float32)
    br0, bz0 = _loop_field(np.array([0.0]), np.array([0.0]), rm, 2.0)
    scale = b_center / bz0[0]
    rng = np.random.default_rng(seed)
    blobs = []
    while len(blobs) < n_blobs:
        p = rng.uniform([c - rt, c - rt, 2 * sigma * sc], [c + rt, c + rt, n - 2 * sigma * sc])
        if np.hypot(p[0] - c, p[1] - c) > rt - 2 * sigma * sc:
            continue
        blobs.append((p[0], p[1], p[2], (1.0 if rng.uniform() < 0.5 else -1.0) * amp * rng.uniform(0.5, 1.0)))
    s2 = 2.0 * (sigma * sc) ** 2
    gxy = np.stack([a * np.exp(-((xx + c - bx) ** 2 + (yy + c - by) ** 2) / s2) for bx, by, bz, a in blobs])
    gxy = (gxy * inside).reshape(len(blobs), -1)
    lay = n * n
    # per-layer templates: flags (the shell solid; the inlet plane at z = 0),
    # u_z inside the tube, the background charge density off the shell
    f_layer = np.where(shell, 1, 0).astype(np.uint8).ravel()
    f_inlet = np.where(inside, 2, f_layer.reshape(n, n)).astype(np.uint8).ravel()
    uz_layer = np.where(inside, u_inlet, 0.0).astype(np.float32).ravel()
    q_back = np.where(shell, 0.0, n0).ravel()
    q_in = (np.where(inside, q_inlet, 0.0).ravel() + q_back) if q_inlet is not None else None
    for z in range(n):
        sl = slice(z * lay, (z + 1) * lay)
        flags[sl] = f_inlet if z == 0 or (outlet and z == n - 1) else f_layer
        u[2 * N + z * lay:2 * N + (z + 1) * lay] = uz_layer
        if z == 0 and q_in is not None:
            q[sl] = q_in
        elif z == n - 1 and outlet:
            q[sl] = q_back
        else:
            wz = np.array([np.exp(-((z - bz) ** 2) / s2) for bx, by, bz, a in blobs])
            q[sl] = wz @ gxy + q_back
        brl, bzl = _loop_field(r_u, np.full_like(r_u, z - c), rm, 2.0)
        brl, bzl = brl[r_inv], bzl[r_inv]
        with np.errstate(divide="ignore", invalid="ignore"):
            bxl = np.where(r > 0, brl * xx / r, 0.0)
            byl = np.where(r > 0, brl * yy / r, 0.0)
        b[z * lay:(z + 1) * lay] = (scale * bxl).ravel()
        b[N + z * lay:N + (z + 1) * lay] = (scale * byl).ravel()
        b[2 * N + z * lay:2 * N + (z + 1) * lay] = (scale * bzl).ravel()
        if log and z % 128 == 127:
            log(f"thruster setup: layer {z + 1}/{n}")
    arrays = dict(flags=flags, u=u, q=q, b_stat=b)
    meta = dict(tube_r=rt, wall=th, magnet_r=rm, b_center=b_center, u_inlet=u_inlet, q_inlet=q_inlet,
                blobs=len(blobs), seed=seed, plasma=plasma)
    return cfg, arrays, meta


# setup_deeva_test, src/setup.rs:395-453: the thruster meshes in the order the
# reference imports and voxelises them (file, origin, model, magnetisation
# in A/m or charge per cell in C)
DEEVA_MESHES = [
    ("deeva_disk_magnet.stl", (64.001, 0.0, 64.0), ModelType.Magnet, (0.0, 1000000.0, 0.0)),
    ("deeva_inlet.stl", (64.0, 0.0, 64.0), ModelType.Solid, (0.0, 0.0, 0.0)),
    ("deeva_quartz_tube.stl", (64.001, 0.0, 64.0), ModelType.Solid, (0.0, 0.0, 0.0)),
    ("deeva_ring_magnet.stl", (64.001, -0.5, 64.0), ModelType.Magnet, (0.0, 500000.0, 0.0)),
    ("deeva_e_plate1.stl", (64.0, 0.0, 64.0), ModelType.ChargedECR, (0.00000000000021844213 / 2.0, 0.0, 0.0)),
    ("deeva_e_plate2.stl", (64.0, 0.0, 64.0), ModelType.ChargedECR, (-0.00000000000021844213 / 2.0, 0.0, 0.0)),
]


def deeva(scale=1.0, lod_depth=4):
    """LbmConfig of setup_deeva_test (src/setup.rs:395-425) with every length
    scaled by `scale` (1: 128 x 256 x 128). SUBGRID_ECR, which the reference
    switches on there, is left off: it does not reach the static-field
    precompute. Returns (cfg, meshes) with the mesh origins scaled alike."""
    cfg = LbmConfig()
    cfg.units.set(128.0 * scale, 1.0, 1.0, 1.0, 1.0, 0.1, 1.0, 10e-8, 1.0, 50000.0)
    cfg.n_x, cfg.n_y, cfg.n_z = int(128 * scale), int(256 * scale), int(128 * scale)
    cfg.nu = cfg.units.nu_si_lu(0.05)
    cfg.velocity_set = VelocitySet.D3Q19
    cfg.ext_volume_force = True
    cfg.ext_magneto_hydro = True
    cfg.mhd_lod_depth = lod_depth
    meshes = [(f, tuple(scale * c for c in o), m, v) for f, o, m, v in DEEVA_MESHES]
    return cfg, meshes


def voxelise_meshes(lbm, meshes, stl_dir):
    """import_mesh (scale 1) + voxelise_mesh for each (file, origin, model, value)."""
    for f, origin, model, value in meshes:
        i = lbm.import_mesh(os.path.join(stl_dir, f), 1.0, *origin)
        lbm.voxelise_mesh(i, model, value)
