"""Host-side workload generators (CPU only)."""
import numpy as np
import pytest

from ionsolver_amd import FloatType
from ionsolver_amd.scenarios import _loop_field, thruster


def test_loop_field_closed_form_on_axis():
    """On the axis of a current loop the field is axial, a^2 / (a^2 + z^2)^1.5
    in units of mu0 I / 2 (here mu0 I / (2 pi) times pi)."""
    a = 7.0
    z = np.array([0.0, 1.0, 3.0, 10.0])
    br, bz = _loop_field(np.zeros_like(z), z, a, 0.1)
    np.testing.assert_allclose(br, 0.0, atol=1e-15)
    np.testing.assert_allclose(bz, np.pi * a * a / (a * a + z * z) ** 1.5, rtol=1e-12)


def test_loop_field_is_divergence_free():
    a, h = 5.0, 1e-4
    r, z = np.array([1.3, 2.2, 7.5]), np.array([0.4, -1.1, 2.0])
    # div B = (1/r) d(r B_r)/dr + dB_z/dz
    brp, _ = _loop_field(r + h, z, a, 0.1)
    brm, _ = _loop_field(r - h, z, a, 0.1)
    _, bzp = _loop_field(r, z + h, a, 0.1)
    _, bzm = _loop_field(r, z - h, a, 0.1)
    div = ((r + h) * brp - (r - h) * brm) / (2 * h * r) + (bzp - bzm) / (2 * h)
    _, bz0 = _loop_field(r, z, a, 0.1)
    assert np.all(np.abs(div) < 1e-6 * np.abs(bz0).max())


def test_thruster_geometry():
    n = 64
    cfg, arr, meta = thruster(n=n, float_type=FloatType.FP16C, lod_depth=3)
    f = arr["flags"].reshape(n, n, n)
    x = np.arange(n) - n / 2
    r = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2)
    assert np.all(f[:, (r >= meta["tube_r"]) & (r < meta["tube_r"] + meta["wall"])] == 1)
    inside = r < meta["tube_r"]
    assert np.all(f[0][inside] == 2) and np.all(f[-1][inside] == 2) and np.all(f[1:-1][:, inside] == 0)
    q = arr["q"].reshape(n, n, n)
    assert np.all(q[:, r >= meta["tube_r"]] == 0) and np.abs(q).max() > 0
    b = arr["b_stat"].reshape(3, n, n, n)
    np.testing.assert_allclose(b[:, n // 2, n // 2, n // 2], [0.0, 0.0, meta["b_center"]], atol=1e-7)
    # mirror symmetry of the loop field about its plane: B_z even, B_x odd in z - n/2
    np.testing.assert_allclose(b[2, n // 2 + 5], b[2, n // 2 - 5], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(b[0, n // 2 + 5], -b[0, n // 2 - 5], rtol=1e-5, atol=1e-9)
    assert cfg.ext_equilibrium_boudaries and cfg.ext_magneto_hydro and cfg.ext_charge_inlet
    assert np.all(q[0][inside] == np.float32(meta["q_inlet"])) and np.all(q[-1][inside] == 0)
    assert meta["b_center"] == 0.05


def test_subgrid_ecr_workload_setup():
    """add_subgrid_ecr (the mhd<N>-<fmt>-ecr bench workloads): the extension
    on, E_var along x with the given magnitude, uniform initial Et, and the
    arrays written to the E_VAR / ET buffers by apply()."""
    from ionsolver_amd import Field
    from ionsolver_amd.scenarios import add_subgrid_ecr, apply, mhd_blobs
    cfg, arrays, _ = mhd_blobs(n=32, float_type=FloatType.FP16C, lod_depth=2, sigma=3.0, cyl_r=6.0)
    cfg, arrays = add_subgrid_ecr(cfg, arrays, et0=2e-3, e_amp=5e-4)
    n = 32 ** 3
    assert cfg.ext_subgrid_ecr and cfg.ext_magneto_hydro
    ev = arrays["e_var"].reshape(3, n)
    assert np.all(ev[0] == np.float32(5e-4)) and not ev[1:].any()
    assert arrays["et"].shape == (n,) and np.all(arrays["et"] == np.float32(2e-3))

    class Sink:
        def __init__(self):
            self.seen = {}

        def write_field(self, field, v):
            self.seen[field] = v

    s = Sink()
    apply(s, arrays)
    assert s.seen[Field.E_VAR] is arrays["e_var"] and s.seen[Field.ET] is arrays["et"]


def test_plasma_thruster_defaults():
    """With the plasma background the ring magnet is weaker (0.002 at its
    centre: the gas carries the background charge density) and the outlet
    plane holds the background charge density."""
    n = 32
    cfg, arr, meta = thruster(n=n, lod_depth=2, plasma=1.0)
    assert meta["b_center"] == 0.002 and cfg.electron_density0 == 1.0
    f = arr["flags"].reshape(n, n, n)
    q = arr["q"].reshape(n, n, n)
    x = np.arange(n) - n / 2
    inside = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2) < meta["tube_r"]
    assert np.all(f[-1][inside] == 2) and np.all(q[-1][inside] == 1.0)
    assert np.all(q[0][inside] == np.float32(1.0 + meta["q_inlet"]))
    b = arr["b_stat"].reshape(3, n, n, n)
    np.testing.assert_allclose(b[2, n // 2, n // 2, n // 2], 0.002, rtol=1e-5)


def test_plasma_unit_constants():
    """The plasma workloads' lattice constants: KE = 1.0e-6, KMU = 2.8e-5,
    KKGE = -2.50 (long-run units) and KE = 1.0e-3, KMU = 2.8e-4 (strongly
    coupled units, for short accuracy comparisons)."""
    from ionsolver_amd import scenarios
    cfg, _, _ = scenarios.mhd_blobs(n=64, lod_depth=2, plasma=1.0)
    k = cfg.kernel_constants()
    assert k["ke"] == pytest.approx(1.0e-6, rel=0.01) and k["kmu"] == pytest.approx(2.8e-5, rel=0.01)
    assert k["kkge"] == pytest.approx(-2.50, rel=0.01) and cfg.electron_density0 == 1.0
    scenarios.use_plasma_units(cfg, 1.0, scenarios.PLASMA_UNITS_STRONG)
    k = cfg.kernel_constants()
    assert k["ke"] == pytest.approx(1.0e-3, rel=0.01) and k["kmu"] == pytest.approx(2.8e-4, rel=0.01)
    assert k["kkge"] == pytest.approx(-2.50, rel=0.01)
