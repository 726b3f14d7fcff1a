"""The Lorentz coupling of the HIP stream_collide (sim_kernels.cl:641-677)
against the closed forms of the Guo-forced SRT collision on uniform lattices,
through the C ABI, with no oracle in between (the same statements as
test_oracle_lorentz.py, whose docstring derives them).

The unit factors are chosen so that KKGE is about -2.5 and KE about 1e-3,
and every cell's charge Q = rho_q - rho_e is zero up to rounding, so that the
LOD sources vanish and E_dyn = E_stat, B_dyn = B_stat stay frozen. The LOD
depth is 3 on 8^3 cells: the near-field block (side 8 / 2^(2^3), below one
cell) is then empty, so the field update that initialize runs on the initial
Q buffer (which holds rho_q for the charge DDFs) adds nothing either."""
import numpy as np
import pytest

from harness import FIELD_OF, make_config
from test_oracle_lorentz import B0, E0, I3, feq_shifted, gas_trajectory, moments, to_t1_layout

import ionsolver_amd as I

pytestmark = pytest.mark.gpu

FP64, FP32 = 3, 2


def uniform_gpu(ft, q0, u0, n=8, D=3):
    cfg = make_config(nx=n, ny=n, nz=n, mhd=True, lod_depth=D, float_type=ft, eqb=False, force=(0.0, 0.0, 0.0))
    cfg.units.m, cfg.units.kg, cfg.units.s, cfg.units.a = 1.58e8, 2.27e-12, 1.0, 1.0
    k = cfg.kernel_constants()
    assert -3.0 < k["kkge"] < -2.0 and 0 < k["ke"] < 2e-3
    gpu = I.Lbm(cfg)
    N = n ** 3
    gpu.write_field(FIELD_OF["rho"], np.ones(N))
    gpu.write_field(FIELD_OF["q"], np.full(N, q0))
    gpu.write_field(FIELD_OF["u"], np.repeat(np.asarray(u0, np.float64), N))
    gpu.write_field(FIELD_OF["e_stat"], np.repeat(E0, N))
    gpu.write_field(FIELD_OF["b_stat"], np.repeat(B0, N))
    gpu.initialize()
    assert gpu.get_time_step() == 0
    return gpu, k, N


def stored(ft, v):
    """The value a field buffer of the instantiation holds for v."""
    return np.asarray(v, np.float64) if ft == FP64 else np.asarray(v, np.float32).astype(np.float64)


@pytest.mark.parametrize("ft,tol", [(FP64, 1e-12), (FP32, 2e-6)])
def test_gas_lorentz_force_with_moving_gas(ft, tol):
    u0 = np.array([0.04, -0.025, 0.01])
    gpu, k, N = uniform_gpu(ft, 1.0, u0)  # rho_q = 1 = rho_e: Q = 0
    try:
        gpu.write_field(FIELD_OF["ei"], np.zeros(19 * N))
        expect = gas_trajectory(1.0, stored(ft, u0), stored(ft, E0), stored(ft, B0), 5)
        for s, e in enumerate(expect):
            gpu.do_time_step()
            u = gpu.read_field(FIELD_OF["u"]).reshape(3, -1).astype(np.float64)
            np.testing.assert_allclose(u, np.repeat(e[:, None], N, 1), rtol=tol, atol=0, err_msg=f"step {s}")
            np.testing.assert_array_equal(gpu.read_field(FIELD_OF["e_dyn"]), np.repeat(stored(ft, E0), N))
    finally:
        gpu.close()


@pytest.mark.parametrize("ft,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_electron_gas_lorentz_collision(ft, tol):
    rho_e, u_e = 1.3, np.array([0.02, -0.01, 0.015])
    gpu, k, N = uniform_gpu(ft, rho_e, np.array([0.01, 0.0, -0.02]))
    try:
        e0 = stored(ft, feq_shifted(rho_e, u_e))
        gpu.write_field(FIELD_OF["ei"], np.repeat(to_t1_layout(e0)[:, None], N, 1).reshape(-1))
        rho_in, j_in, _ = moments(e0)
        u_in = j_in / rho_in
        gpu.do_time_step()
        e1 = gpu.read_field(FIELD_OF["ei"]).reshape(19, -1).astype(np.float64)
        assert np.all(e1 == e1[:, :1])
        rho1, j1, pi1 = moments(e1[:, 0])
        w, kk = k["w"], k["kkge"]
        E, B = stored(ft, E0), stored(ft, B0)
        F = -rho_in * (E + np.cross(u_in, B))
        us = u_in + F / (2 * rho_in * kk)
        j_exp = (1 - w) * rho_in * u_in + w * rho_in * us + (1 - w / 2) * F
        pi_exp = ((1 - w) * (rho_in * (np.outer(u_in, u_in) + I3 / 3) - I3 / 3)
                  + w * (rho_in * (np.outer(us, us) + I3 / 3) - I3 / 3)
                  + (1 - w / 2) * (np.outer(us, F) + np.outer(F, us)))
        assert abs(rho1 - rho_in) < tol
        np.testing.assert_allclose(j1, j_exp, rtol=0, atol=tol * np.abs(j_exp).max())
        np.testing.assert_allclose(pi1, pi_exp, rtol=0, atol=tol * np.abs(pi_exp).max())
    finally:
        gpu.close()
