"""Pins the oracle's SUBGRID_ECR step (sim_kernels.cl:557-629) with closed
forms, on quiet lattices (fluid and electron gas at rest, field update
switched off by KE = KMU = 0, electron density set to 1 by zeroing its
DDFs), one step, float64 oracle:

* heating: Et += KEABS/KKBME (rho_e + 1e-5)/KKGE |E_perp|^2, E_perp the part
  of E_var normal to B, and nothing for E_var parallel to B;
* gyration drift: du_e = KKBME Et grad|B| / |B| with grad_mag_v as the
  reference writes it, mag(c + e) - mag(c - e) / 2 (only the second term
  halved; zero on the lattice faces), and Et -= |du_e| / KKBME; under a
  uniform B this leaves grad_mag_v = |B|/2 on every axis of an interior
  cell, so Et grows by the factor 1 + sqrt(3)/2 there (KKBME < 0);
* ionisation: rho_q and rho_e both grow by 1e-4 Et, so Q is unchanged and
  the charge DDFs, relaxed towards the raised rho_q with rate WQ, gain
  WQ 1e-4 sum(Et)."""
import numpy as np

import oracle_lib as O

EXT = O.EXT_MAGNETO_HYDRO | O.EXT_VOLUME_FORCE | O.EXT_UPDATE_FIELDS
KKGE, KKBME, KEABS = -0.5, -2.0, 3.0


def quiet(n, b, e_var, et0, ecr=True):
    orc = O.OracleSim(n, n, n, lod_depth=0, float_type=O.FP64, ext=EXT | (O.EXT_SUBGRID_ECR if ecr else 0),
                      w=1.5, ke=0.0, kmu=0.0, kkge=KKGE, wq=0.4, kkbme=KKBME, keabs=KEABS)
    N = n ** 3
    orc.buf("b_stat")[:] = b
    if ecr:
        orc.buf("e_var")[:] = e_var
        orc.buf("et")[:] = et0
    orc.initialize()
    orc.buf("ei")[:] = 0.0  # electron density 1, at rest
    assert orc.buf("q").size == N
    return orc


def coords(n):
    x = np.tile(np.arange(n, dtype=np.float64), n * n)
    y = np.tile(np.repeat(np.arange(n, dtype=np.float64), n), n)
    z = np.repeat(np.arange(n, dtype=np.float64), n * n)
    return x, y, z


def interior_mask(n):
    return np.logical_and.reduce([(c > 0) & (c < n - 1) for c in coords(n)])


def uniform(n, v):
    return np.repeat(np.asarray(v, np.float64), n ** 3)


def test_heating_by_the_field_normal_to_b():
    n, et0 = 6, 0.2
    b = uniform(n, (0.0, 0.0, 0.7))
    for ev, perp2 in [((0.3, -0.4, 0.9), 0.25), ((0.0, 0.0, 1.3), 0.0)]:
        orc = quiet(n, b, uniform(n, ev), et0)
        orc.do_time_step()
        heated = et0 + KEABS / KKBME * (1.0 + float(np.float32(1e-5))) / KKGE * perp2
        expect = np.where(interior_mask(n), heated * (1.0 + np.sqrt(3.0) / 2.0), heated)
        np.testing.assert_allclose(orc.buf("et"), expect, rtol=1e-12)


def test_gyration_drift_under_a_linear_field_strength():
    n, et0, b0, g = 8, 0.3, 0.5, 0.05
    x, _, _ = coords(n)
    bz = b0 + g * x
    b = np.concatenate([np.zeros(2 * n ** 3), bz])
    orc = quiet(n, b, np.zeros(3 * n ** 3), et0)
    orc.do_time_step()
    interior = interior_mask(n)
    # grad_mag_v as written: mag(c + e) - mag(c - e) / 2
    gx = (b0 + g * (x + 1)) - (b0 + g * (x - 1)) / 2
    gy = bz - bz / 2
    gz = bz - bz / 2
    du = KKBME * et0 * np.stack([gx, gy, gz]) / bz
    expect = np.where(interior, et0 - np.linalg.norm(du, axis=0) / KKBME, et0)
    np.testing.assert_allclose(orc.buf("et"), expect, rtol=1e-12)


def test_ionisation_adds_charge_and_keeps_q():
    n = 6
    rng = np.random.default_rng(4)
    et0 = 0.1 + 0.2 * rng.uniform(size=n ** 3)
    b = uniform(n, (0.0, 0.0, 0.7))
    ev = uniform(n, (0.0, 0.0, 0.4))  # parallel to B: no heating
    with_ecr, without = quiet(n, b, ev, et0), quiet(n, b, ev, et0, ecr=False)
    for o in (with_ecr, without):
        o.do_time_step()
    et = with_ecr.buf("et")
    assert np.all(et > 0)
    gained = with_ecr.buf("fqi").sum() - without.buf("fqi").sum()
    expect = float(np.float32(0.4)) * float(np.float32(1e-4)) * et.sum()  # WQ and 1e-4 are float literals
    assert abs(gained - expect) < 1e-9 * expect
    np.testing.assert_allclose(with_ecr.buf("q"), without.buf("q"), rtol=0, atol=1e-14)
