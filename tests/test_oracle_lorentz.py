"""Pins the oracle's Lorentz coupling (sim_kernels.cl:641-677) with values
that the moment equations of the SRT collision with Guo forcing fix, on
lattices whose state is uniform, so that streaming is the identity and one
cell's moments evolve by a closed-form map.

For a collision f' = (1-w) f + w feq(rho, u*) + (1 - w/2) F_i with Guo's
term F_i = w_i [3 (c_i - u*) + 9 (c_i . u*) c_i] . F, the moments of f' are
  sum f'_i            = rho - 1                                (shifted DDFs)
  sum c_i f'_i        = (1-w) rho u + w rho u* + (1 - w/2) F
  sum c_i c_i^T f'_i  = (1-w) [rho (u u^T + I/3) - I/3]
                        + w [rho (u* u*^T + I/3) - I/3] + (1 - w/2) (u* F^T + F u*^T)
where u is the raw moment of f. The reference chooses
  gas:      F = rho_q (E + u x B) with the raw u, u* = u + F / (2 rho)     (:660-662, :681-685)
  electron: F = -rho_e (E + u_e x B), u* = u_e + F / (2 rho_e KKGE)         (:641-656)
and feeds the raw fluid u (not u*) into the LOD velocity mean (:674-676).
The expected values below are computed from these statements with numpy's
cross product; the oracle's own functions are not called for them."""
import numpy as np
import pytest

import oracle_lib as O

MHD_EXT = O.EXT_MAGNETO_HYDRO | O.EXT_VOLUME_FORCE | O.EXT_UPDATE_FIELDS

# D3Q19 velocities and weights (sim_kernels.cl:336-339, domain.rs:763-765)
C19 = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0],
                [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1],
                [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1]], np.float64)
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
I3 = np.eye(3)

E0 = np.array([1.0e-3, 2.0e-3, -1.5e-3])
B0 = np.array([0.3, -0.2, 0.5])


def feq_shifted(rho, u):
    """Textbook second-order D3Q19 equilibrium, DDF-shifted by -w_i."""
    cu = C19.T @ u
    return W19 * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u @ u) - W19


def to_t1_layout(e):
    """Slot values of a uniform lattice such that the step at even t reads
    the DDFs e (initialize stores with t = 1: opposite pairs swapped,
    sim_kernels.cl:234-247)."""
    s = np.empty_like(e)
    s[0] = e[0]
    s[1::2] = e[2::2]
    s[2::2] = e[1::2]
    return s


def moments(e):
    """(rho, j, Pi) of shifted DDFs e."""
    return e.sum() + 1.0, C19 @ e, (C19 * e) @ C19.T


def uniform_mhd(ft, q0, u0, E, B, kkge, w=1.5, n=8, D=1):
    """MHD lattice with every cell in the same state: fields frozen
    (KE = KMU = 0, so E_dyn = E_stat and B_dyn = B_stat every step)."""
    orc = O.OracleSim(n, n, n, lod_depth=D, float_type=ft, ext=MHD_EXT, w=w, ke=0.0, kmu=0.0, kkge=kkge)
    N = n ** 3
    orc.buf("rho")[:] = 1.0
    orc.buf("q")[:] = q0
    orc.buf("u")[:] = np.repeat(np.asarray(u0, np.float64), N)
    orc.buf("e_stat")[:] = np.repeat(np.asarray(E, np.float64), N)
    orc.buf("b_stat")[:] = np.repeat(np.asarray(B, np.float64), N)
    orc.initialize()
    assert orc.t == 0
    return orc


def gas_trajectory(q, u0, E, B, steps):
    """Stored velocities of a uniform gas (rho = 1) under rho_q (E + u x B):
    the raw momentum grows by F each step, the stored u is the Guo half shift."""
    u = np.array(u0, np.float64)
    out = []
    for _ in range(steps):
        F = q * (E + np.cross(u, B))
        out.append(u + F / 2)
        u = u + F
    return out


@pytest.mark.parametrize("ft,tol", [(O.FP64, 1e-12), (O.FP32, 2e-6)])
def test_gas_lorentz_force_with_moving_gas(ft, tol):
    """Moving charged gas in uniform E and B: stored u follows
    u_raw + rho_q (E + u_raw x B)/2 with u_raw advanced by the full force."""
    q, u0 = 0.3, np.array([0.04, -0.025, 0.01])
    orc = uniform_mhd(ft, q, u0, E0, B0, kkge=-1.0)
    orc.buf("ei")[:] = 0  # electron gas at rest with density 1
    for k, expect in enumerate(gas_trajectory(q, u0, E0, B0, 5)):
        orc.do_time_step()
        u = orc.buf("u").reshape(3, -1).astype(np.float64)
        np.testing.assert_allclose(u, np.repeat(expect[:, None], u.shape[1], 1), rtol=tol, atol=0,
                                   err_msg=f"step {k}")


def test_gas_lorentz_force_is_the_magnetic_part():
    """With E = 0 the force is q u x B: perpendicular to u and to B, and
    reversing B reverses the velocity change (fixes the sign of u x B)."""
    q, u0 = 0.3, np.array([0.04, -0.025, 0.01])
    du = {}
    for s in (1.0, -1.0):
        orc = uniform_mhd(O.FP64, q, u0, np.zeros(3), s * B0, kkge=-1.0)
        orc.buf("ei")[:] = 0
        orc.do_time_step()
        du[s] = orc.buf("u").reshape(3, -1)[:, 0] - u0
    expect = q * np.cross(u0, B0) / 2
    np.testing.assert_allclose(du[1.0], expect, rtol=1e-10, atol=0)
    np.testing.assert_allclose(du[-1.0], -expect, rtol=1e-10, atol=0)
    assert abs(du[1.0] @ u0) < 1e-15 and abs(du[1.0] @ B0) < 1e-15


@pytest.mark.parametrize("kkge", [0.37, -2.5, 1.0])
@pytest.mark.parametrize("ft,tol", [(O.FP64, 1e-12), (O.FP32, 1e-5)])
def test_electron_gas_lorentz_collision(kkge, ft, tol):
    """Electron gas at density rho_e and velocity u_e in uniform E and B:
    after one collision its mass is unchanged and its first and second
    moments are those of the Guo-forced SRT collision with
    F = -rho_e (E + u_e x B) and the velocity shift F / (2 rho_e KKGE)."""
    w = 1.5
    rho_e, u_e = 1.3, np.array([0.02, -0.01, 0.015])
    orc = uniform_mhd(ft, 0.2, np.array([0.01, 0.0, -0.02]), E0, B0, kkge=kkge, w=w)
    e0 = feq_shifted(rho_e, u_e)
    raw = orc.buf("ei").reshape(19, -1)
    if ft == O.FP64:
        raw[:] = to_t1_layout(e0)[:, None]
    else:
        raw[:] = to_t1_layout(e0).astype(np.float32)[:, None]
        e0 = e0.astype(np.float32).astype(np.float64)
    rho_in, j_in, _ = moments(e0)
    u_in = j_in / rho_in
    orc.do_time_step()
    e1 = orc.buf("ei").reshape(19, -1).astype(np.float64)
    assert np.all(e1 == e1[:, :1]), "uniform state must stay uniform"
    rho1, j1, pi1 = moments(e1[:, 0])  # slots are in natural order after a step at even t

    kk = float(np.float32(kkge))  # the kernel constant is a float (domain.rs:736-857)
    F = -rho_in * (E0 + np.cross(u_in, B0))
    us = u_in + F / (2 * rho_in * kk)
    assert np.all(np.abs(us) < 0.5), "no clamp in this case"
    j_exp = (1 - w) * rho_in * u_in + w * rho_in * us + (1 - w / 2) * F
    pi_exp = ((1 - w) * (rho_in * (np.outer(u_in, u_in) + I3 / 3) - I3 / 3)
              + w * (rho_in * (np.outer(us, us) + I3 / 3) - I3 / 3)
              + (1 - w / 2) * (np.outer(us, F) + np.outer(F, us)))
    assert abs(rho1 - rho_in) < tol
    np.testing.assert_allclose(j1, j_exp, rtol=0, atol=tol * np.abs(j_exp).max())
    np.testing.assert_allclose(pi1, pi_exp, rtol=0, atol=tol * np.abs(pi_exp).max())
    # the charge field the step writes: Q = rho_q - rho_e
    np.testing.assert_allclose(orc.buf("q").astype(np.float64), 0.2 - rho_in, rtol=0, atol=tol)


def test_electron_velocity_is_clamped():
    """A shift beyond the lattice speed c = 0.57735027 is clamped per
    component (sim_kernels.cl:644-646): with KKGE tiny the shifted velocity
    sits at +-c, its sign that of F."""
    w, c = 1.5, float(np.float32(0.57735027))
    rho_e, u_e = 1.3, np.array([0.02, -0.01, 0.015])
    kkge = 1e-6
    orc = uniform_mhd(O.FP64, 0.2, np.zeros(3), E0, B0, kkge=kkge, w=w)
    e0 = feq_shifted(rho_e, u_e)
    orc.buf("ei").reshape(19, -1)[:] = to_t1_layout(e0)[:, None]
    orc.do_time_step()
    rho1, j1, _ = moments(orc.buf("ei").reshape(19, -1)[:, 0])
    F = -rho_e * (E0 + np.cross(u_e, B0))
    us = np.sign(F) * c
    j_exp = (1 - w) * rho_e * u_e + w * rho_e * us + (1 - w / 2) * F
    np.testing.assert_allclose(j1, j_exp, rtol=0, atol=1e-12)


def test_lod_velocity_is_the_raw_moment():
    """The LOD entries hold the block sums of Q and the block mean of the raw
    fluid velocity, before the Guo shift: with a non-zero force the stored u
    differs from the LOD mean by F / 2."""
    q, u0 = 0.3, np.array([0.04, -0.025, 0.01])
    n, D = 8, 1
    orc = uniform_mhd(O.FP64, q, u0, E0, B0, kkge=-1.0, n=n, D=D)
    orc.buf("ei")[:] = 0  # rho_e = 1
    orc.do_time_step()
    qu = orc.buf("qu_lod").reshape(-1, 4)
    cells = (n >> D) ** 3
    np.testing.assert_allclose(qu[:8, 1:], np.repeat(u0[None, :], 8, 0), rtol=1e-13, atol=0)
    np.testing.assert_allclose(qu[:8, 0], cells * (q - 1.0), rtol=1e-13, atol=0)
    F = q * (E0 + np.cross(u0, B0))
    np.testing.assert_allclose(orc.buf("u").reshape(3, -1)[:, 0], u0 + F / 2, rtol=1e-12, atol=0)
