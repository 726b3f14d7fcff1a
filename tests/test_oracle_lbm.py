"""Pins the oracle's stream_collide / initialize / step sequencing
(sim_kernels.cl:234-302, 465-832; src/lbm/mod.rs:214-272) against an
independent textbook implementation: a two-array pull LBM in numpy (float64)
with full-way bounce-back at solid cells, equilibrium (TYPE_E) cells, Guo
forcing and SRT/TRT collision. The oracle's in-place esoteric-pull storage is
invisible in the macroscopic fields, so rho and u must agree to rounding.
Also: Taylor-Green decay, mass and momentum balance."""
import numpy as np
import pytest

import oracle_lib as O
from test_oracle_lattice import SETS

C_LIM = float(np.float32(0.57735027))


class NaiveLBM:
    """Two-array pull scheme over shifted DDFs g_i = f_i - w_i (float64)."""

    def __init__(self, vs, nx, ny, nz, w, trt, force, flags, rho, u, eqb, ffield=None):
        self.c, wts = SETS[vs]
        self.wt = np.array(wts)[:, None]
        self.q = self.c.shape[1]
        self.n = nx * ny * nz
        self.w, self.trt, self.eqb = w, trt, eqb
        self.force = np.asarray(force, np.float64)[:, None]
        self.ffield = ffield
        self.flags = flags
        self.solid = (flags & 0x1F) == 0x01
        self.is_e = eqb & ((flags & 0x1F) == 0x02)
        self.rho_buf = rho.astype(np.float64).copy()
        self.u_buf = u.reshape(3, -1).astype(np.float64).copy()
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        x, y, z = x.ravel(), y.ravel(), z.ravel()
        # pull source of direction i: the cell at x - c_i (periodic)
        self.src = [((x - self.c[0, i]) % nx) + (((y - self.c[1, i]) % ny) + ((z - self.c[2, i]) % nz) * ny) * nx
                    for i in range(self.q)]
        self.opp = np.array([0] + [i + 1 if i % 2 else i - 1 for i in range(1, self.q)])
        self.u_buf[:, self.solid] = 0.0
        self.post = self.feq(self.rho_buf, self.u_buf)

    def feq(self, rho, u):
        cu = self.c.T @ u
        return self.wt * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * np.sum(u * u, 0)) - self.wt

    def step(self):
        fin = np.stack([self.post[i][self.src[i]] for i in range(self.q)])
        rho = fin.sum(0) + 1
        u = (self.c @ fin) / rho
        rho = np.where(self.is_e, self.rho_buf, rho)
        u = np.where(self.is_e, self.u_buf, u)
        F = np.broadcast_to(self.force, (3, self.n)).copy()
        if self.ffield is not None:
            F = F + self.ffield
        u = np.clip(u + F / (2 * rho), -C_LIM, C_LIM)
        cu, cF = self.c.T @ u, self.c.T @ F
        Fi = self.wt * (3 * (cF - np.sum(u * F, 0)) + 9 * cu * cF)
        geq = self.feq(rho, u)
        w = self.w
        if not self.trt:
            post = (1 - w) * fin + w * geq + (1 - w / 2) * Fi
        else:
            wm = 1 / (0.1875 / (1 / w - 0.5) + 0.5)
            fb, eb, Fb = fin[self.opp], geq[self.opp], Fi[self.opp]
            Fc = (0.5 - w / 4) * (Fi + Fb) + (0.5 - wm / 4) * (Fi - Fb)
            post = fin - w * ((fin + fb) - (geq + eb)) / 2 - wm * ((fin - fb) - (geq - eb)) / 2 + Fc
        post = np.where(self.is_e, geq, post)
        reflected = fin[self.opp]  # full-way bounce-back: solid cells reverse what arrived
        self.post = np.where(self.solid, reflected, post)
        upd = ~self.solid & ~self.is_e
        self.rho_buf = np.where(upd, rho, self.rho_buf)
        self.u_buf = np.where(upd, u, self.u_buf)


def make_case(rng, vs, nx, ny, nz, solid_frac=0.15, e_frac=0.05, ffield=False):
    n = nx * ny * nz
    flags = np.zeros(n, np.uint8)
    r = rng.uniform(size=n)
    flags[r < solid_frac] = 0x01
    flags[(r >= solid_frac) & (r < solid_frac + 0.05)] = 0x11  # magnet flag: fluid for the kernels
    flags[(r >= solid_frac + 0.05) & (r < solid_frac + 0.05 + e_frac)] = 0x02
    rho = 1 + 0.01 * rng.standard_normal(n)
    u = 0.03 * rng.standard_normal(3 * n)
    if vs == O.D2Q9:
        u[2 * n:] = 0.0
    F = 2e-4 * rng.standard_normal(3 * n) if ffield else None
    if F is not None and vs == O.D2Q9:
        F[2 * n:] = 0.0
    return flags, rho, u, F


@pytest.mark.parametrize("vs,dims", [(O.D3Q19, (7, 6, 5)), (O.D3Q15, (6, 5, 4)), (O.D2Q9, (9, 7, 1))])
@pytest.mark.parametrize("trt", [0, 1])
def test_oracle_matches_naive_two_array_lbm(vs, dims, trt):
    rng = np.random.default_rng(vs * 10 + trt)
    nx, ny, nz = dims
    w = 1.0 / (3.0 * 0.04 + 0.5)
    force = [2e-4, -1e-4, 3e-4] if vs != O.D2Q9 else [2e-4, -1e-4, 0.0]
    flags, rho, u, F = make_case(rng, vs, nx, ny, nz, ffield=True)
    ext = (O.EXT_EQUILIBRIUM_BOUNDARIES | O.EXT_VOLUME_FORCE | O.EXT_FORCE_FIELD | O.EXT_UPDATE_FIELDS)
    orc = O.OracleSim(nx, ny, nz, velocity_set=vs, relaxation_time=trt, float_type=O.FP64, ext=ext, w=w,
                      fx=force[0], fy=force[1], fz=force[2])
    orc.buf("flags")[:] = flags
    orc.buf("rho")[:] = rho
    orc.buf("u")[:] = u
    orc.buf("F")[:] = F
    force32 = np.float32(force).astype(np.float64)  # kernel arguments are f32 (domain.rs:422-425)
    naive = NaiveLBM(vs, nx, ny, nz, float(np.float32(w)), trt, force32, flags, rho, u, True, F.reshape(3, -1))
    orc.initialize()
    np.testing.assert_allclose(orc.buf("u").reshape(3, -1), naive.u_buf, atol=0)  # solids zeroed
    for step in range(8):
        orc.do_time_step()
        naive.step()
        np.testing.assert_allclose(orc.buf("rho"), naive.rho_buf, rtol=1e-12, atol=1e-13, err_msg=f"step {step}")
        np.testing.assert_allclose(orc.buf("u").reshape(3, -1), naive.u_buf, rtol=1e-11, atol=1e-13,
                                   err_msg=f"step {step}")


def test_step_parity_and_time_counter():
    orc = O.OracleSim(4, 4, 4, ext=O.EXT_UPDATE_FIELDS, w=1.2)
    assert orc.t == 0
    orc.run(3)  # lazily initialises, then three steps (mod.rs:235-246)
    assert orc.t == 3


def test_taylor_green_decay_rate():
    """2D Taylor-Green vortex (an exact Navier-Stokes solution): kinetic
    energy decays as exp(-4 nu k^2 t), nu = (1/w - 1/2)/3."""
    n = 32
    nu = 0.05
    w = 1.0 / (3.0 * nu + 0.5)
    orc = O.OracleSim(n, n, 1, velocity_set=O.D2Q9, float_type=O.FP64, ext=O.EXT_UPDATE_FIELDS, w=w)
    x, y = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="xy")
    k = 2 * np.pi / n
    u0 = 0.01
    ux = u0 * np.cos(k * x) * np.sin(k * y)
    uy = -u0 * np.sin(k * x) * np.cos(k * y)
    orc.buf("u")[:] = np.concatenate([ux.ravel(), uy.ravel(), np.zeros(n * n)])
    orc.buf("rho")[:] = (1 - 0.75 * u0 ** 2 * (np.cos(2 * k * x) + np.cos(2 * k * y))).ravel()
    orc.initialize()
    ke = []
    for t in range(201):
        if t in (20, 200):
            uu = orc.buf("u").reshape(3, -1)
            ke.append(np.sum(uu[0] ** 2 + uu[1] ** 2))
        orc.do_time_step()
    nu_w = (1.0 / float(np.float32(w)) - 0.5) / 3.0
    rate = -np.log(ke[1] / ke[0]) / 180.0
    assert abs(rate / (4 * nu_w * k * k) - 1) < 5e-3


def test_mass_and_momentum_balance_with_volume_force():
    """Periodic box, no solids: sum(rho) is constant and the total momentum
    sum(rho u) grows by N*F per step (Guo forcing)."""
    rng = np.random.default_rng(11)
    n = 8
    F = np.array([3e-5, -2e-5, 1e-5])
    orc = O.OracleSim(n, n, n, float_type=O.FP64, ext=O.EXT_VOLUME_FORCE | O.EXT_UPDATE_FIELDS, w=1.4,
                      fx=F[0], fy=F[1], fz=F[2])
    orc.buf("rho")[:] = 1 + 0.01 * rng.standard_normal(n ** 3)
    orc.buf("u")[:] = 0.01 * rng.standard_normal(3 * n ** 3)
    orc.initialize()
    orc.do_time_step()
    m0 = orc.buf("rho").sum()
    p0 = (orc.buf("rho") * orc.buf("u").reshape(3, -1)).sum(1)
    for _ in range(10):
        orc.do_time_step()
    assert abs(orc.buf("rho").sum() - m0) < 1e-11 * m0
    p1 = (orc.buf("rho") * orc.buf("u").reshape(3, -1)).sum(1)
    np.testing.assert_allclose(p1 - p0, 10 * n ** 3 * np.float32(F).astype(np.float64), rtol=1e-9)


@pytest.mark.parametrize("ft", [O.FP32, O.FP16S, O.FP16C])
def test_storage_formats_stay_close_to_fp64(ft):
    """FP32/FP16S/FP16C storage with FP32 arithmetic vs the FP64 oracle on a
    short fluid run (DDF-shifting keeps the 16-bit formats accurate)."""
    rng = np.random.default_rng(12)
    n = 12
    rho = 1 + 0.01 * rng.standard_normal(n ** 3)
    u = 0.02 * rng.standard_normal(3 * n ** 3)
    sims = []
    for f in (O.FP64, ft):
        s = O.OracleSim(n, n, n, float_type=f, ext=O.EXT_UPDATE_FIELDS, w=1.5)
        s.buf("rho")[:] = rho
        s.buf("u")[:] = u
        s.initialize()
        s.run(20)
        sims.append(s)
    du = sims[1].buf("u").astype(np.float64) - sims[0].buf("u")
    rel = np.linalg.norm(du) / np.linalg.norm(sims[0].buf("u"))
    assert rel < {O.FP32: 1e-5, O.FP16S: 2e-2, O.FP16C: 1e-2}[ft]
