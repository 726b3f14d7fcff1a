"""Pins the oracle's MHD extension (sim_kernels.cl:530-677, 897-1003):
the E/B field update against a float64 brute-force sum over the same source
set (including the reference's LOD index range, own-LOD skip, block centres
and near field), the single-source Coulomb/Biot-Savart closed form, the LOD
block sums, D3Q7 charge diffusion (D = c_s^2 (1/w_q - 1/2) = 1/2) and drift,
charge conservation and the Lorentz force on the gas."""
import numpy as np
import pytest

import oracle_lib as O

MHD_EXT = O.EXT_MAGNETO_HYDRO | O.EXT_VOLUME_FORCE | O.EXT_UPDATE_FIELDS


def n_lod_own(D):
    return sum(8 ** k for k in range(D + 1))


def brute_force_fields(nx, ny, nz, D, flags, q, u, qu, e_stat, b_stat, ke, kmu):
    """float64 restatement of update_e_b_dynamic over the reference's source set."""
    n = nx * ny * nz
    nd = 1 << D
    sh = 1 << (nd & 31)
    ds = [max(v // sh, 1) for v in (nx, ny, nz)]
    hi = n_lod_own(D)
    lo = hi - nd ** 3
    bs = [v // nd for v in (nx, ny, nz)]
    d = np.arange(lo, hi)
    t = d % (nd * nd)
    cen = np.stack([(t % nd) * bs[0] + 0.5 * bs[0], (t // nd) * bs[1] + 0.5 * bs[1], (d // (nd * nd)) * bs[2] + 0.5 * bs[2]], 1)
    qd = qu.reshape(-1, 4)[lo:hi].astype(np.float64)
    E = e_stat.reshape(3, -1).astype(np.float64).copy()
    B = b_stat.reshape(3, -1).astype(np.float64).copy()
    uu = u.reshape(3, -1).astype(np.float64)
    for cell in range(n):
        if (flags[cell] & 0x1F) == 1:
            continue
        z, rem = divmod(cell, nx * ny)
        y, x = divmod(rem, nx)
        p = np.array([x, y, z], np.float64)
        e = np.zeros(3)
        b = np.zeros(3)
        lx, ly, lz = (x // ds[0]) * ds[0], (y // ds[1]) * ds[1], (z // ds[2]) * ds[2]
        for xx in range(lx, min(lx + ds[0], nx)):
            for yy in range(ly, min(ly + ds[1], ny)):
                for zz in range(lz, min(lz + ds[2], nz)):
                    nc = xx + (yy + zz * ny) * nx
                    if nc == cell or q[nc] == 0:
                        continue
                    r = p - np.array([xx, yy, zz])
                    pre = r / np.linalg.norm(r) ** 3
                    e += q[nc] * pre
                    b += q[nc] * np.cross(uu[:, nc], pre)
        own = x // bs[0] + (y // bs[1] + z // bs[2] * nd) * nd
        keep = d != own
        r = p[None, :] - cen[keep]
        pre = r / (np.linalg.norm(r, axis=1) ** 3)[:, None]
        e += (qd[keep, 0:1] * pre).sum(0)
        b += (qd[keep, 0:1] * np.cross(qd[keep, 1:4], pre)).sum(0)
        E[:, cell] += ke * e
        B[:, cell] += kmu * b
    return E, B


@pytest.mark.parametrize("dims,D", [((8, 8, 8), 1), ((16, 8, 8), 2), ((8, 8, 16), 3), ((6, 6, 6), 0)])
def test_update_e_b_matches_float64_brute_force(dims, D):
    nx, ny, nz = dims
    n = nx * ny * nz
    rng = np.random.default_rng(D)
    ke, kmu = 0.7, 0.3
    orc = O.OracleSim(nx, ny, nz, lod_depth=D, ext=MHD_EXT, w=1.2, ke=ke, kmu=kmu, kkge=-1.0)
    flags = np.where(rng.uniform(size=n) < 0.1, 1, 0).astype(np.uint8)
    q = np.where(rng.uniform(size=n) < 0.3, 0.0, rng.standard_normal(n)).astype(np.float32)
    u = (0.05 * rng.standard_normal(3 * n)).astype(np.float32)
    qu = rng.standard_normal(orc.n_lod * 4).astype(np.float32)
    e_stat = rng.standard_normal(3 * n).astype(np.float32)
    b_stat = rng.standard_normal(3 * n).astype(np.float32)
    for k, v in dict(flags=flags, q=q, u=u, qu_lod=qu, e_stat=e_stat, b_stat=b_stat).items():
        orc.buf(k)[:] = v
    orc.kernel("update_e_b_dynamic")
    E, B = brute_force_fields(nx, ny, nz, D, flags, q.astype(np.float64), u, qu, e_stat, b_stat, ke, kmu)
    fl = (flags & 0x1F) != 1
    Eo, Bo = orc.buf("e_dyn").reshape(3, -1), orc.buf("b_dyn").reshape(3, -1)
    np.testing.assert_allclose(Eo[:, fl], E[:, fl], rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(Bo[:, fl], B[:, fl], rtol=2e-5, atol=2e-5)
    assert np.all(Eo[:, ~fl] == 0) and np.all(Bo[:, ~fl] == 0)  # solid cells are not updated


def test_single_source_closed_form():
    """One charged LOD, all other sources empty: E = KE q r/|r|^3 and
    B = KMU q v x r/|r|^3 (Coulomb, Biot-Savart)."""
    nx = ny = nz = 16
    D = 2
    orc = O.OracleSim(nx, ny, nz, lod_depth=D, ext=MHD_EXT, w=1.2, ke=2.0, kmu=0.5, kkge=-1.0)
    qu = np.zeros(orc.n_lod * 4, np.float32)
    L = 40  # block (0, 2, 2) of the 4^3 top level, inside the seen range [9, 73)
    qv = np.array([0.75, 0.01, -0.02, 0.03], np.float32)
    qu[4 * L:4 * L + 4] = qv
    orc.buf("qu_lod")[:] = qu
    orc.kernel("update_e_b_dynamic")
    c = np.array([(L % 4) * 4 + 2, ((L // 4) % 4) * 4 + 2, (L // 16) * 4 + 2], np.float64)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()]).astype(np.float64)
    r = p - c[:, None]
    rn3 = np.linalg.norm(r, axis=0) ** 3
    own = (x.ravel() // 4 + (y.ravel() // 4 + z.ravel() // 4 * 4) * 4) == L
    E = 2.0 * qv[0] * r / np.where(own, 1, rn3)
    B = 0.5 * qv[0] * np.cross(qv[1:].astype(np.float64)[:, None], r, axis=0) / np.where(own, 1, rn3)
    E[:, own] = 0
    B[:, own] = 0
    np.testing.assert_allclose(orc.buf("e_dyn").reshape(3, -1), E, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(orc.buf("b_dyn").reshape(3, -1), B, rtol=1e-6, atol=1e-9)


def test_low_lod_indices_are_never_seen():
    """update_e_b_dynamic reads LODs [n_lod_own - 8^D, n_lod_own)
    (sim_kernels.cl:943): LOD indices below (8^D - 1)/7 carry no field."""
    D = 2
    orc = O.OracleSim(16, 16, 16, lod_depth=D, ext=MHD_EXT, w=1.2, ke=1.0, kmu=1.0, kkge=-1.0)
    qu = np.zeros(orc.n_lod * 4, np.float32)
    qu[0:4 * 9] = 1.0  # LODs 0..8: below the range start 9
    orc.buf("qu_lod")[:] = qu
    orc.buf("e_stat")[:] = 0.25
    orc.kernel("update_e_b_dynamic")
    assert np.all(orc.buf("e_dyn") == 0.25) and np.all(orc.buf("b_dyn") == 0)


def quiet_mhd(nx, ny, nz, D, ft, q0, u0=(0.0, 0.0, 0.0), seed=0):
    """MHD run with no fields (KE = KMU = 0, E = B = 0) and the electron gas
    set to rest at density 1 (all ei slots 0 after initialize), so that the
    charge is a passive scalar advected by the fluid."""
    orc = O.OracleSim(nx, ny, nz, lod_depth=D, float_type=ft, ext=MHD_EXT, w=1.5, ke=0.0, kmu=0.0, kkge=-1.0,
                      wq=0.4)
    n = nx * ny * nz
    orc.buf("q")[:] = q0
    orc.buf("u")[:] = np.repeat(np.asarray(u0, np.float64), n)
    orc.initialize()
    orc.buf("ei")[:] = 0
    return orc


def profile_x(nx, ny, nz, c, s, amp=0.5):
    """Charge profile varying along x only (a plane Gaussian)."""
    x = np.tile(np.arange(nx, dtype=np.float64), ny * nz)
    return amp * np.exp(-(x - c) ** 2 / (2 * s * s)), x


def x_moments(rho_q, x):
    w = rho_q / rho_q.sum()
    m = (w * x).sum()
    return m, (w * (x - m) ** 2).sum()


def test_charge_diffusion_coefficient():
    """D3Q7 BGK with w_q = 0.4 diffuses with D = (1/4)(1/0.4 - 1/2) = 1/2:
    the variance grows by 2 D = 1 per step, once the start-up transient,
    which decays as (1 - w_q)^t, has died out."""
    n = (128, 4, 4)
    q0, x = profile_x(*n, 64.0, 2.0)
    orc = quiet_mhd(*n, 2, O.FP64, q0)
    var = {}
    for t in range(61):
        if t in (30, 60):
            var[t] = x_moments(orc.buf("q") + 1.0, x)[1]  # Q = rho_q - rho_e with rho_e = 1
        orc.do_time_step()
    assert abs((var[60] - var[30]) / 30.0 - 1.0) < 1e-5


def test_charge_drift_and_conservation():
    """At uniform fluid velocity u the charge moves by u per step."""
    n = (128, 4, 4)
    q0, x = profile_x(*n, 60.0, 2.0)
    orc = quiet_mhd(*n, 2, O.FP64, q0, u0=(0.05, 0.0, 0.0))
    orc.do_time_step()
    r0 = orc.buf("q") + 1.0
    tot0, c0 = r0.sum(), x_moments(r0, x)[0]
    for _ in range(40):
        orc.do_time_step()
    r1 = orc.buf("q") + 1.0
    tot1, c1 = r1.sum(), x_moments(r1, x)[0]
    assert abs(tot1 - tot0) < 1e-12 * abs(tot0)
    assert abs((c1 - c0) - 40 * 0.05) < 1e-9


def test_lod_sums_are_block_sums():
    """QU_lod holds, per top-level block, the sum of the cell charges
    Q = rho_q - rho_e and the mean of the (pre-force) fluid velocity."""
    nx, ny, nz, D = 16, 16, 8, 2
    rng = np.random.default_rng(5)
    n = nx * ny * nz
    orc = O.OracleSim(nx, ny, nz, lod_depth=D, float_type=O.FP32, ext=MHD_EXT, w=1.5, ke=0.0, kmu=0.0, kkge=-1.0)
    orc.buf("q")[:] = rng.uniform(-0.5, 0.5, n)
    orc.buf("u")[:] = 0.02 * rng.standard_normal(3 * n)
    flags = np.where(rng.uniform(size=n) < 0.1, 1, 0).astype(np.uint8)
    orc.buf("flags")[:] = flags
    orc.initialize()
    orc.buf("ei")[:] = 0
    orc.do_time_step()
    q = orc.buf("q").astype(np.float64)
    u = orc.buf("u").reshape(3, -1).astype(np.float64)  # no force: stored u is the raw moment
    nd = 1 << D
    bx, by, bz = nx // nd, ny // nd, nz // nd
    z, rem = np.divmod(np.arange(n), nx * ny)
    y, x = np.divmod(rem, nx)
    lod = x // bx + (y // by + z // bz * nd) * nd
    fl = (flags & 0x1F) != 1
    qu = orc.buf("qu_lod").reshape(-1, 4).astype(np.float64)
    for L in range(nd ** 3):
        m = fl & (lod == L)
        assert abs(qu[L, 0] - q[m].sum()) < 1e-5
        np.testing.assert_allclose(qu[L, 1:], u[:, m].sum(1) / (bx * by * bz), atol=1e-7)
    assert np.all(qu[nd ** 3:] == 0)


def test_lorentz_force_on_gas_at_rest():
    """First step from rest with uniform charge and E: the stored velocity is
    the Guo half-force shift u = rho_q E / (2 rho) (sim_kernels.cl:660, 681)."""
    n = 8
    E = np.array([1e-3, -2e-3, 0.5e-3])
    orc = O.OracleSim(n, n, n, lod_depth=1, float_type=O.FP64, ext=MHD_EXT, w=1.5, ke=0.0, kmu=0.0, kkge=-1.0)
    orc.buf("q")[:] = 0.3
    orc.buf("e_stat")[:] = np.repeat(E, n ** 3)
    orc.buf("b_stat")[:] = 0.2
    orc.initialize()
    orc.buf("ei")[:] = 0
    orc.do_time_step()
    u = orc.buf("u").reshape(3, -1)
    np.testing.assert_allclose(u, np.repeat((0.3 * E / 2)[:, None], n ** 3, 1), rtol=1e-12)


@pytest.mark.parametrize("ft,residue", [(O.FP32, -2.3841858e-07), (O.FP16S, 2.4414062e-04), (O.FP16C, -6.1035156e-05)])
def test_electron_gas_starts_at_rounding_residue(ft, residue):
    """ei is initialised to feq(0, u) = -w_i (sim_kernels.cl:823-825): the
    electron density is the rounding residue of the shifted sum, a
    storage-format dependent value (the knife edge of the electron gas)."""
    orc = O.OracleSim(4, 4, 4, lod_depth=1, float_type=ft, ext=MHD_EXT, w=1.2, ke=0.0, kmu=0.0, kkge=-1.0)
    orc.initialize()
    e = orc.ddfs("ei")
    # at a cell with all 19 slots its own after init: sum in index order then + 1
    r = np.float32(e[0, 0])
    for i in range(1, 19):
        r = np.float32(r + e[i, 0])
    assert np.float32(r + np.float32(1.0)) == np.float32(residue)


def inlet_ring(nx, ft, q_in, u0):
    """A ring of nx x 4 x 4 cells, periodic, with the plane x = 0 a TYPE_E
    charge inlet holding q_in; everything else fluid, initially uncharged."""
    n = (nx, 4, 4)
    orc = O.OracleSim(*n, lod_depth=2, float_type=ft, w=1.5, ke=0.0, kmu=0.0, kkge=-1.0, wq=0.4,
                      ext=MHD_EXT | O.EXT_EQUILIBRIUM_BOUNDARIES | O.EXT_CHARGE_INLET)
    cells = nx * 16
    x = np.tile(np.arange(nx), 16)
    inlet = x == 0
    orc.buf("flags")[:] = np.where(inlet, 0x02, 0x00)
    q = np.zeros(cells)
    q[inlet] = q_in
    orc.buf("q")[:] = q
    orc.buf("u")[:] = np.repeat(np.asarray(u0, np.float64), cells)
    orc.initialize()
    orc.buf("ei")[:] = 0  # electron gas at rest with density 1: Q = rho_q - 1 in fluid cells
    return orc, inlet


@pytest.mark.parametrize("u0", [(0.0, 0.0, 0.0), (0.05, 0.0, 0.0)])
def test_charge_inlet_fills_the_ring(u0):
    """With one inlet plane in a periodic ring the only steady state of
    advection-diffusion is the inlet's charge everywhere: the inlet cells keep
    q_in, the total charge grows monotonically and converges to q_in per cell
    (the slowest mode decays by about exp(-0.0066 t) at rest)."""
    nx, q_in = 24, 0.3
    orc, inlet = inlet_ring(nx, O.FP64, q_in, u0)
    tot = []
    for t in range(2200):
        orc.do_time_step()
        if t % 100 == 0:
            rq = np.where(inlet, orc.buf("q"), orc.buf("q") + 1.0)
            tot.append(rq.sum())
    assert np.all(orc.buf("q")[inlet] == q_in)
    assert np.all(np.diff(tot) > 0)
    rq = orc.buf("q")[~inlet] + 1.0
    assert np.max(np.abs(rq - q_in)) < 1e-6


def test_charge_inlet_off_conserves_charge():
    """Planting the same TYPE_E plane without the inlet extension conserves
    the total charge instead (the TYPE_E charge advects like fluid)."""
    nx, q_in = 24, 0.3
    orc = O.OracleSim(nx, 4, 4, lod_depth=2, float_type=O.FP64, w=1.5, ke=0.0, kmu=0.0, kkge=-1.0, wq=0.4,
                      ext=MHD_EXT | O.EXT_EQUILIBRIUM_BOUNDARIES)
    x = np.tile(np.arange(nx), 16)
    orc.buf("flags")[:] = np.where(x == 0, 0x02, 0x00)
    orc.buf("q")[:] = np.where(x == 0, q_in, 0.0)
    orc.initialize()
    orc.buf("ei")[:] = 0
    orc.do_time_step()
    t0 = (orc.buf("q") + 1.0).sum()
    for _ in range(200):
        orc.do_time_step()
    assert abs((orc.buf("q") + 1.0).sum() - t0) < 1e-12 * abs(t0)


@pytest.mark.parametrize("ft,tol", [(O.FP64, 1e-14), (O.FP32, 1e-6), (O.FP16C, 2e-3)])
def test_electron_density0_starts_the_electron_gas(ft, tol):
    """Build extension electron_density0 = n0: initialize writes the electron
    DDFs of non-solid cells as the textbook equilibrium of density n0 at the
    cell's velocity (moments n0, n0 u, n0 (u u + I/3) before the -w_i shift),
    and those of solid cells as density 0 at rest, as the reference does
    for every cell."""
    from test_oracle_lorentz import C19, I3, moments, to_t1_layout
    n0, u0 = 0.75, np.array([0.03, -0.02, 0.01])  # n0 exact in float (the parameter is f32)
    n = 4
    orc = O.OracleSim(n, n, n, lod_depth=1, float_type=ft, ext=MHD_EXT, w=1.2, kkge=-1.0, ei_rho0=n0)
    orc.buf("u")[:] = np.repeat(u0, n ** 3)
    orc.initialize()
    e = orc.ddfs("ei").astype(np.float64)
    assert np.all(e == e[:, :1])  # uniform lattice: every cell's slots alike
    rho, j, pi = moments(to_t1_layout(e[:, 0]))
    assert abs(rho - n0) < tol
    np.testing.assert_allclose(j, n0 * u0, rtol=0, atol=tol)
    np.testing.assert_allclose(pi, n0 * (np.outer(u0, u0) + I3 / 3) - I3 / 3, rtol=0, atol=tol)
    assert C19.shape == (3, 19)
    # solid cells start at density 0: the total over every slot counts fluid cells only
    orc = O.OracleSim(n, n, n, lod_depth=1, float_type=ft, ext=MHD_EXT, w=1.2, kkge=-1.0, ei_rho0=n0)
    flags = orc.buf("flags")
    flags[::5] = 1
    orc.initialize()
    total = orc.ddfs("ei").astype(np.float64).sum() + n ** 3
    assert abs(total - n0 * np.count_nonzero(flags != 1)) < tol * n ** 3


@pytest.mark.parametrize("solid", [True, False])
def test_initialize_solid_cells_rest_and_fields_start_static(solid):
    """initialize (sim_kernels.cl:760-832) on a uniform lattice given a
    velocity and a charge: a solid lattice stores the gas equilibrium at rest
    (momentum 0) and the D3Q7 equilibrium of zero charge, and zeroes u and Q;
    a fluid one keeps rho u and Q. Either way E_dyn and B_dyn start as E_stat
    and B_stat."""
    from test_oracle_lorentz import moments, to_t1_layout
    n = 4
    N = n ** 3
    orc = O.OracleSim(n, n, n, lod_depth=1, ext=MHD_EXT, w=1.2, kkge=-1.0)
    rho0, u0, q0 = 1.05, np.array([0.03, -0.02, 0.01]), 0.125
    orc.buf("rho")[:] = rho0
    orc.buf("u")[:] = np.repeat(u0, N)
    orc.buf("q")[:] = q0
    if solid:
        orc.buf("flags")[:] = 1
    rng = np.random.default_rng(9)
    e_stat = rng.standard_normal(3 * N).astype(np.float32)
    b_stat = rng.standard_normal(3 * N).astype(np.float32)
    orc.buf("e_stat")[:] = e_stat
    orc.buf("b_stat")[:] = b_stat
    orc.initialize()
    f = orc.ddfs("fi").astype(np.float64)
    assert np.all(f == f[:, :1])
    rho, j, _ = moments(to_t1_layout(f[:, 0]))
    assert abs(rho - rho0) < 1e-6
    np.testing.assert_allclose(j, 0.0 if solid else rho0 * u0, rtol=0, atol=1e-7)
    qsum = orc.ddfs("fqi").astype(np.float64)[:, 0].sum() + 1.0  # D3Q7 charge moment
    assert abs(qsum - (0.0 if solid else q0)) < 1e-7
    assert np.all(orc.buf("q") == (0.0 if solid else np.float32(q0)))
    assert np.all(orc.buf("u") == (0.0 if solid else np.repeat(u0, N).astype(np.float32)))
    assert np.array_equal(orc.buf("e_dyn"), e_stat) and np.array_equal(orc.buf("b_dyn"), b_stat)


def test_update_fields_clamps_the_velocity():
    """update_fields (sim_kernels.cl:834-859) recomputes rho and u from the
    DDFs and clamps each velocity component to the lattice speed of sound:
    an equilibrium at u = (0.9, -0.7, 0.1) comes back as (c, -c, 0.1)."""
    n = 4
    orc = O.OracleSim(n, n, n, lod_depth=1, ext=O.EXT_VOLUME_FORCE | O.EXT_UPDATE_FIELDS, w=1.2)
    N = n ** 3
    u0 = np.array([0.9, -0.7, 0.1])
    orc.buf("rho")[:] = 1.0
    orc.buf("u")[:] = np.repeat(u0, N)
    orc.initialize()  # fi = f_eq(1, u0)
    orc.buf("u")[:] = 0.0
    orc.kernel("update_fields")
    c = np.float32(0.57735027)
    u = orc.buf("u").reshape(3, N)
    assert np.all(u[0] == c) and np.all(u[1] == -c)
    np.testing.assert_allclose(u[2], 0.1, rtol=0, atol=1e-6)
    np.testing.assert_allclose(orc.buf("rho"), 1.0, rtol=0, atol=1e-6)
