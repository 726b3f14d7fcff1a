"""Pins the oracle's per-cell lattice functions (sim_kernels.cl:155-233,
366-396) against textbook closed forms: the second-order equilibrium, its
moments, Guo's forcing term and the D3Q7 advection equilibrium."""
import numpy as np
import pytest

import oracle_lib as O

SETS = {
    O.D2Q9: (np.array([[0, 1, -1, 0, 0, 1, -1, 1, -1], [0, 0, 0, 1, -1, 1, -1, -1, 1], [0] * 9]),
             [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4),
    O.D3Q15: (np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1],
                        [0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1],
                        [0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1]]),
              [2 / 9] + [1 / 9] * 6 + [1 / 72] * 8),
    O.D3Q19: (np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0],
                        [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1],
                        [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1]]),
              [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12),
}


def velocity_sets():
    return [O.D2Q9, O.D3Q15, O.D3Q19]


def rand_state(rng, vs):
    rho = 1 + 0.05 * rng.standard_normal()
    u = 0.05 * rng.standard_normal(3)
    if vs == O.D2Q9:
        u[2] = 0.0
    return rho, u


@pytest.mark.parametrize("vs", velocity_sets())
def test_lattice_tables_are_consistent(vs):
    c, w = SETS[vs]
    w = np.array(w)
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(c @ w, 0)
    d = 2 if vs == O.D2Q9 else 3
    assert np.allclose((c[:d] * w) @ c[:d].T, np.eye(d) / 3)  # c_s^2 = 1/3
    # odd/even pairs are opposite directions (the esoteric-pull pairing)
    assert np.array_equal(c[:, 1::2], -c[:, 2::2])


@pytest.mark.parametrize("vs", velocity_sets())
def test_feq_is_the_second_order_equilibrium(vs):
    c, w = SETS[vs]
    w = np.array(w)
    rng = np.random.default_rng(vs)
    for _ in range(20):
        rho, u = rand_state(rng, vs)
        cu = c.T @ u
        ref = w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u) - w  # DDF-shifted by -w_i
        np.testing.assert_allclose(O.f_eq(vs, rho, *u, f64=True), ref, rtol=0, atol=1e-15)
        np.testing.assert_allclose(O.f_eq(vs, rho, *u), ref, rtol=0, atol=5e-8)


@pytest.mark.parametrize("vs", velocity_sets())
def test_feq_moments_and_rho_u_inverse(vs):
    c, w = SETS[vs]
    rng = np.random.default_rng(10 + vs)
    for _ in range(20):
        rho, u = rand_state(rng, vs)
        feq = O.f_eq(vs, rho, *u, f64=True)
        assert abs(feq.sum() - (rho - 1)) < 1e-14
        np.testing.assert_allclose(c @ feq, rho * u, atol=1e-15)
        got = O.rho_u(vs, feq, f64=True)
        np.testing.assert_allclose(got, [rho, *u], atol=1e-14)
        got32 = O.rho_u(vs, O.f_eq(vs, rho, *u))
        np.testing.assert_allclose(got32, [rho, *u], atol=2e-6)


@pytest.mark.parametrize("vs", velocity_sets())
def test_guo_forcing_closed_form_and_moments(vs):
    """Fin_i = w_i [3 (c_i - u) + 9 (c_i.u) c_i] . F (Krueger p.233, without
    the (1 - w/2) factor, which the kernel applies afterwards)."""
    c, w = SETS[vs]
    w = np.array(w)
    rng = np.random.default_rng(20 + vs)
    for _ in range(20):
        _, u = rand_state(rng, vs)
        F = 1e-3 * rng.standard_normal(3)
        if vs == O.D2Q9:
            F[2] = 0.0
        cu, cF = c.T @ u, c.T @ F
        ref = w * (3 * (cF - u @ F) + 9 * cu * cF)
        got = O.forcing(vs, u, F)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9)
        assert abs(got.sum()) < 1e-9
        np.testing.assert_allclose(c @ got.astype(np.float64), F, atol=1e-9)


def test_d3q7_equilibrium():
    """a_eq: w0 = 1/4, w_i = 1/8, c_s^2 = 1/4, shifted by -w_i."""
    c7 = np.array([[0, 1, -1, 0, 0, 0, 0], [0, 0, 0, 1, -1, 0, 0], [0, 0, 0, 0, 0, 1, -1]])
    w7 = np.array([1 / 4] + [1 / 8] * 6)
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.standard_normal()
        u = 0.05 * rng.standard_normal(3)
        ref = w7 * q * (1 + 4 * (c7.T @ u)) - w7
        got = O.a_eq(q, *u)
        np.testing.assert_allclose(got, ref, atol=1e-7)
        assert abs(got.astype(np.float64).sum() - (q - 1)) < 1e-6
        np.testing.assert_allclose(c7 @ got.astype(np.float64), q * u, atol=1e-7)
