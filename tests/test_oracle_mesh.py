"""Pins the mesh oracle (oracle/mesh_oracle.cpp, oracle/mesh_ref.py): the STL
import transforms against closed forms, voxelize_mesh (sim_kernels.cl:
1150-1231) against the exact cell sets its truncated-distance column walk
gives for an axis-aligned box and against ball bounds for a sphere, and the
static fields (:1234-1299) against float64 brute force, the point-dipole far
field B = mu0/(4 pi) (3 (m.r) r / r^5 - m / r^3) and Coulomb's law."""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from meshes import box, icosphere, inscribed_radius, write_stl

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mesh_ref as MR  # noqa: E402

STL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stl")


def coords(shape):
    nx, ny, nz = shape
    n = np.arange(nx * ny * nz)
    return n % nx, (n // nx) % ny, n // (nx * ny)


def test_rotation_matrix_is_the_product_of_axis_rotations():
    def rx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])

    def ry(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])

    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    for ang in [(0, 0, 0), (30, 0, 0), (0, 45, 0), (0, 0, 60), (15, -40, 110)]:
        m = MR.rotation_matrix(*ang)
        M = np.array([[m[i + j] for j in "xyz"] for i in "xyz"], np.float64)
        r = [math.radians(a) for a in ang]
        np.testing.assert_allclose(M, rx(r[0]) @ ry(r[1]) @ rz(r[2]), atol=2e-7)
    m = MR.rotation_matrix(0, 0, 0)
    assert [m[k] for k in ("xx", "yy", "zz", "xy", "yz", "zx")] == [1, 1, 1, 0, 0, 0]


def test_stl_import_closed_forms(tmp_path):
    p = write_stl(tmp_path / "b.stl", box((0.0, 0.0, 0.0), (0.01, 0.02, 0.04)))
    # import_mesh: scale = len_si_lu(1) = 1/m, placed at the origin argument
    v, lo, hi = MR.import_mesh(p, 1.0, 5.0, 6.0, 7.0, 0, 0, 0, units_m=0.001)
    np.testing.assert_allclose(lo, [5, 6, 7], rtol=1e-6)
    np.testing.assert_allclose(hi, [15, 26, 47], rtol=1e-6)
    # reposition: centred at (c), longest side = size
    v, lo, hi = MR.import_mesh_reposition(p, 20.0, 30.0, 40.0, 0, 0, 0, 16.0, (64, 64, 64))
    np.testing.assert_allclose((lo + hi) / 2, [20, 30, 40], rtol=1e-6)
    np.testing.assert_allclose(hi - lo, [4, 8, 16], rtol=1e-6)
    # size 0: fit into the lattice box
    v, lo, hi = MR.import_mesh_reposition(p, 32.0, 32.0, 32.0, 0, 0, 0, 0.0, (64, 96, 64))
    np.testing.assert_allclose(hi - lo, [16, 32, 64], rtol=1e-6)
    # a 90 degree turn about z maps y extents onto x
    v, lo, hi = MR.import_mesh_reposition(p, 32.0, 32.0, 32.0, 0, 0, 90, 16.0, (64, 64, 64))
    np.testing.assert_allclose(hi - lo, [8, 4, 16], rtol=1e-5)
    with open(tmp_path / "bad.stl", "wb") as f:
        f.write(b"solid ascii\n")
    with pytest.raises(ValueError):
        MR.read_stl_raw(str(tmp_path / "bad.stl"), False, (1, 1, 1), (0, 0, 0), MR.rotation_matrix(0, 0, 0), 1.0)


def test_deeva_meshes_read():
    T = {"deeva_disk_magnet": 252, "deeva_inlet": 252, "deeva_quartz_tube": 512, "deeva_ring_magnet": 512,
         "deeva_e_plate1": 12, "deeva_e_plate2": 12}
    for name, t in T.items():
        v, lo, hi = MR.import_mesh(os.path.join(STL, name + ".stl"), 1.0, 0, 0, 0, 0, 0, 0, units_m=1.0)
        assert v.shape == (t, 3, 3) and np.all(hi > lo)


def box_cells(shape, lo, hi):
    """Cells the reference's voxeliser marks for an axis-aligned box: along
    the ray axis the column walk keeps h with
    floor(lo - h0) < h - h0 < floor(hi - h0) (distances truncated to ushort,
    strict bounds), across it the cells strictly inside the faces."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    w0, w1 = lo - 2, hi + 2
    c = [(w1[1] - w0[1]) * (w1[2] - w0[2]), (w1[2] - w0[2]) * (w1[0] - w0[0]), (w1[0] - w0[0]) * (w1[1] - w0[1])]
    d = 0 if (c[0] < c[1] and c[0] < c[2]) else (1 if c[1] < c[2] else 2)
    xyz = coords(shape)
    inside = np.ones(xyz[0].size, bool)
    for ax in range(3):
        if ax == d:
            h0 = min(max(int(w0[ax]), 0), shape[ax] - 1)
            a = h0 + math.floor(lo[ax] - h0)
            b = h0 + math.floor(hi[ax] - h0)
            inside &= (xyz[ax] > a) & (xyz[ax] < b)
        else:
            inside &= (xyz[ax] > lo[ax]) & (xyz[ax] < hi[ax])
    return inside, d


@pytest.mark.parametrize("lo,hi", [((3.3, 2.6, 4.15), (10.7, 12.4, 7.8)), ((1.45, 3.7, 2.2), (6.85, 5.3, 13.6)),
                                   ((2.2, 1.3, 5.6), (14.1, 11.7, 9.35))])
def test_voxelize_box_cell_set(lo, hi):
    shape = (16, 15, 17)
    n = shape[0] * shape[1] * shape[2]
    inside, d = box_cells(shape, lo, hi)
    # precondition: no lattice point on a face diagonal crossed by the rays
    b, c = [a for a in range(3) if a != d]
    for u in range(shape[b]):
        t = (u - lo[b]) / (hi[b] - lo[b])
        zc = lo[c] + t * (hi[c] - lo[c])
        assert abs(zc - round(zc)) > 1e-3 or not (0 < t < 1)
    rng = np.random.default_rng(3)
    flags0 = rng.choice(np.array([0x00, 0x20, 0x40, 0x02], np.uint8), n)
    flags = flags0.copy()
    b_dyn = rng.standard_normal(3 * n).astype(np.float32)
    b0 = b_dyn.copy()
    tri = box(lo, hi)
    O.voxelize(shape, flags, b_dyn, tri, tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0), 0x11, (0.5, -1.5, 2.5))
    assert np.array_equal(flags[inside], (flags0[inside] & 0xE0) | 0x11)
    assert np.array_equal(flags[~inside], flags0[~inside])
    B = b_dyn.reshape(3, n)
    assert np.all(B[0, inside] == 0.5) and np.all(B[1, inside] == -1.5) and np.all(B[2, inside] == 2.5)
    assert np.array_equal(B[:, ~inside], b0.reshape(3, n)[:, ~inside])
    # charged models write only the x block
    flags = flags0.copy()
    b_dyn = b0.copy()
    O.voxelize(shape, flags, b_dyn, tri, tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0), 0x05, (7.0, 8.0, 9.0))
    B = b_dyn.reshape(3, n)
    assert np.all(B[0, inside] == 7.0) and np.array_equal(B[1:], b0.reshape(3, n)[1:])
    assert np.array_equal(flags[inside], (flags0[inside] & 0xE0) | 0x05)


def test_voxelize_open_sheet_encloses_nothing():
    """A box and, in front of it along the rays, an open square sheet: the
    reference's rule (a cell is inside where the hits ahead and behind the
    ray's start are both odd, then toggled at each hit, sim_kernels.cl:
    1150-1231) leaves the sheet out, so the cells are exactly the box's. A
    rule that took either parity would also fill the cells between the
    ray's start and the box."""
    shape = (16, 15, 17)
    lo, hi = (6.3, 3.6, 4.15), (13.7, 10.4, 10.8)
    inside, d = box_cells(shape, lo, hi)
    assert d == 0  # rays along x, from below the sheet at x = 2.5
    sheet = np.array([[[2.5, 3.0, 3.5], [2.5, 11.0, 3.5], [2.5, 11.0, 11.5]],
                      [[2.5, 3.0, 3.5], [2.5, 11.0, 11.5], [2.5, 3.0, 11.5]]], np.float32)  # diagonal z = y + 1/2
    tri = np.concatenate([box(lo, hi), sheet])
    n = shape[0] * shape[1] * shape[2]
    flags = np.zeros(n, np.uint8)
    O.voxelize(shape, flags, None, tri, tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0), 0x01)
    assert inside.sum() > 0
    assert np.array_equal(flags == 0x01, inside)


@pytest.mark.parametrize("center,r", [((8.2, 8.7, 7.9), 6.3), ((10.4, 7.35, 9.6), 4.9)])
def test_voxelize_sphere_between_its_balls(center, r):
    shape = (20, 18, 19)
    tri, V, f = icosphere(center, r, subdiv=3)
    r_in = inscribed_radius(V, f, center)
    n = shape[0] * shape[1] * shape[2]
    flags = np.zeros(n, np.uint8)
    O.voxelize(shape, flags, None, tri, tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0), 0x01)
    x, y, z = coords(shape)
    dist = np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2)
    marked = flags == 1
    # the column walk truncates distances, so the surface moves by up to one cell
    assert np.all(marked[dist < r_in - 1.0])
    assert not np.any(marked[dist > r])
    vol = 4 / 3 * math.pi * r ** 3
    assert 0.6 * vol < marked.sum() < vol


def brute_psi(shape, flags, M):
    nx, ny, nz = shape
    n = nx * ny * nz
    P = np.stack(coords(shape), 1).astype(np.float64) + 1.0
    sel = (flags & 0x10) != 0
    src, m = P[sel], M.reshape(3, n)[:, sel].T.astype(np.float64)
    pn = np.arange((nx + 2) * (ny + 2) * (nz + 2))
    C = np.stack([pn % (nx + 2), (pn // (nx + 2)) % (ny + 2), pn // ((nx + 2) * (ny + 2))], 1).astype(np.float64)
    psi = np.zeros(len(C))
    for s, mm in zip(src, m):
        d = C - s
        l3 = np.linalg.norm(d, axis=1) ** 3
        ok = l3 > 0
        psi[ok] += (d[ok] @ mm) / l3[ok]
    return psi / (4 * math.pi)


def test_precompute_b_matches_float64():
    shape = (10, 9, 8)
    n = 720
    rng = np.random.default_rng(5)
    flags = np.zeros(n, np.uint8)
    mag = rng.choice(n, 12, replace=False)
    flags[mag] = 0x11
    flags[rng.choice(n, 30, replace=False)] |= 0x01
    b_dyn = (rng.standard_normal(3 * n)).astype(np.float32)
    b_stat = (0.1 * rng.standard_normal(3 * n)).astype(np.float32)
    b0 = b_stat.copy()
    kmu0 = 0.37
    psi = O.precompute_b(shape, flags, b_dyn, b_stat, kmu0)
    ref = brute_psi(shape, flags, b_dyn)
    assert np.max(np.abs(psi - ref)) < 1e-5 * np.max(np.abs(ref))
    # B = B0 - kmu0 * central difference of psi, solid-bit cells untouched
    nx, ny = shape[0] + 2, shape[1] + 2
    x, y, z = coords(shape)
    c = (x + 1) + ((y + 1) + (z + 1) * ny) * nx
    p = psi.astype(np.float64)
    g = np.stack([(p[c + 1] - p[c - 1]) / 2, (p[c + nx] - p[c - nx]) / 2, (p[c + nx * ny] - p[c - nx * ny]) / 2])
    fluid = (flags & 1) == 0
    exp = b0.reshape(3, n).astype(np.float64) - kmu0 * g
    got = b_stat.reshape(3, n)
    np.testing.assert_allclose(got[:, fluid], exp[:, fluid], rtol=1e-5, atol=1e-6)
    assert np.array_equal(got[:, ~fluid], b0.reshape(3, n)[:, ~fluid])


def test_precompute_b_dipole_far_field():
    """A 2^3 magnet with magnetisation M per cell is, far away, a dipole of
    moment m = 8 M: B = mu0/(4 pi) (3 (m.r) r / r^5 - m / r^3)."""
    shape = (32, 32, 32)
    n = 32 ** 3
    x, y, z = coords(shape)
    flags = np.zeros(n, np.uint8)
    blk = (x >= 15) & (x <= 16) & (y >= 15) & (y <= 16) & (z >= 15) & (z <= 16)
    flags[blk] = 0x10
    M = np.array([0.3, 1.0, -0.4])
    b_dyn = np.zeros(3 * n, np.float32)
    b_dyn.reshape(3, n)[:, blk] = M[:, None]
    b_stat = np.zeros(3 * n, np.float32)
    kmu0 = 2.0
    O.precompute_b(shape, flags, b_dyn, b_stat, kmu0)
    m = 8 * M
    r = np.stack([x, y, z], 1) - 15.5
    d = np.linalg.norm(r, axis=1)
    far = (d > 9) & (d < 14)
    rr, dd = r[far], d[far][:, None]
    Bd = kmu0 / (4 * math.pi) * (3 * (rr @ m)[:, None] * rr / dd ** 5 - m / dd ** 3)
    got = b_stat.reshape(3, n)[:, far].T
    err = np.linalg.norm(got - Bd, axis=1) / np.linalg.norm(Bd, axis=1)
    assert np.max(err) < 0.03 and np.median(err) < 0.01


def test_precompute_e_single_charge_and_superposition():
    shape = (12, 10, 9)
    n = 1080
    x, y, z = coords(shape)
    rng = np.random.default_rng(8)
    flags = np.zeros(n, np.uint8)
    src = rng.choice(n, 5, replace=False)
    flags[src[:3]] = 0x09
    flags[src[3:]] = 0x05
    q = rng.uniform(-1, 1, 5).astype(np.float32)
    b_dyn = np.zeros(3 * n, np.float32)
    b_dyn[src] = q
    e_stat = np.full(3 * n, 0.25, np.float32)
    ke = 1.7
    O.precompute_e(shape, flags, b_dyn, e_stat, ke)
    P = np.stack([x, y, z], 1).astype(np.float64)
    E = np.zeros((n, 3))
    for s, qq in zip(src, q):
        d = P - P[s]
        l3 = np.linalg.norm(d, axis=1) ** 3
        ok = l3 > 0
        E[ok] += ke * qq * d[ok] / l3[ok, None]
    fluid = (flags & 1) == 0
    got = e_stat.reshape(3, n).T.astype(np.float64) - 0.25
    np.testing.assert_allclose(got[fluid], E[fluid], rtol=2e-5, atol=2e-6)
    assert np.all(e_stat.reshape(3, n)[:, ~fluid] == 0.25)
