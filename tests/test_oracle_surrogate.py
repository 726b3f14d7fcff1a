"""Pinning of the float64 surrogate network reference (oracle/surrogate_ref.py).

The reference is the checker of the GPU network, so each of its pieces is
checked against something independent: the convolution against torch's
conv3d, the bf16 rounding against torch's bfloat16 cast, pooling and
upsampling against their adjoint identities, the block <-> cell maps against
a brute-force loop, the hand-written backward pass against central finite
differences of the loss, the far-field samples (plus a brute-force near
field) against the C oracle's float64 direct field, and the interpolation
against quadratics and the brute-force far field.
"""
import numpy as np
import pytest
import torch

from oracle import surrogate_ref as S


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-20, 20, 4000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0078125, 65504.0, 3e38]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(S.bf16(x), ref)


@pytest.mark.parametrize("g,cin,cout", [(4, 3, 5), (2, 8, 4), (1, 6, 2)])
def test_conv3_matches_torch_conv3d(g, cin, cout):
    rng = np.random.default_rng(g * 100 + cin)
    x = rng.standard_normal((g, g, g, cin))
    W = np.zeros((cout, S.pad32(27 * cin)))
    W[:, :27 * cin] = rng.standard_normal((cout, 27 * cin))
    # W[o, d*cin + ci] with d = (dx+1) + 3(dy+1) + 9(dz+1)  ->  weight[o, ci, dz+1, dy+1, dx+1]
    wt = W[:, :27 * cin].reshape(cout, 3, 3, 3, cin).transpose(0, 4, 1, 2, 3)
    xt = torch.tensor(x.transpose(3, 0, 1, 2)[None])
    ref = torch.nn.functional.conv3d(xt, torch.tensor(wt), padding=1)[0].numpy().transpose(1, 2, 3, 0)
    np.testing.assert_allclose(S.conv3(x, W, cin), ref, rtol=1e-12, atol=1e-12)


def test_conv3_bwd_matches_autograd():
    rng = np.random.default_rng(5)
    g, cin, cout = 4, 5, 3
    x = rng.standard_normal((g, g, g, cin))
    W = np.zeros((cout, S.pad32(27 * cin)))
    W[:, :27 * cin] = rng.standard_normal((cout, 27 * cin))
    dout = rng.standard_normal((g, g, g, cout))
    dW, dx = S.conv3_bwd(x, W, cin, dout)
    xt = torch.tensor(x, requires_grad=True)
    Wt = torch.tensor(W, requires_grad=True)
    wk = Wt[:, :27 * cin].reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
    y = torch.nn.functional.conv3d(xt.permute(3, 0, 1, 2)[None], wk, padding=1)[0].permute(1, 2, 3, 0)
    (y * torch.tensor(dout)).sum().backward()
    np.testing.assert_allclose(dW, Wt.grad.numpy(), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(dx, xt.grad.numpy(), rtol=1e-11, atol=1e-11)
    assert np.all(dW[:, 27 * cin:] == 0.0)


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((8, 8, 8, 3))
    y = rng.standard_normal((4, 4, 4, 3))
    assert np.isclose(np.sum(S.pool(x) * y), np.sum(x * S.pool_bwd(y)))
    assert np.isclose(np.sum(S.up(y) * x), np.sum(y * S.up_bwd(x)))
    # pool of an upsampled grid is the identity; pooling a constant keeps it
    np.testing.assert_allclose(S.pool(S.up(y)), y, rtol=1e-15)
    assert S.pool(x)[1, 2, 3, 0] == pytest.approx(x[2:4, 4:6, 6:8, 0].mean())


def test_block_cell_maps_brute_force():
    geo = S.Geometry(16, 8, 24, 2, 8)    # nd = 4, blocks of 4 x 2 x 6 cells
    rng = np.random.default_rng(2)
    Y = rng.standard_normal((4, 4, 4, geo.nout))
    F = S.block_to_cells(geo, Y)
    for n in rng.integers(0, 16 * 8 * 24, 200):
        x, y, z = n % 16, (n // 16) % 8, n // 128
        X, Yb, Z = x // geo.bx, y // geo.by, z // geo.bz
        o = (x % geo.bx) + ((y % geo.by) + (z % geo.bz) * geo.by) * geo.bx
        for c in range(6):
            assert F[c, n] == Y[Z, Yb, X, c * geo.cpb + o]
    np.testing.assert_array_equal(S.cells_to_block(geo, F), Y)


def test_net_input_scaling_and_lo_mask():
    geo = S.Geometry(16, 16, 16, 3, 8)
    rng = np.random.default_rng(3)
    qu = rng.standard_normal((geo.nblk, 4)).astype(np.float32)
    qu[:, 1:] *= 1e-4                                   # currents far below the charges
    x, s = S.net_input(geo, qu)
    J = np.stack([qu[:, 0] * qu[:, 1], qu[:, 0] * qu[:, 2], qu[:, 0] * qu[:, 3]], 1).astype(np.float64)
    assert s[0] == np.max(np.abs(qu[:, 0])) and s[1] == pytest.approx(np.max(np.abs(J)), rel=1e-6)
    flat = x.reshape(-1, 8)
    assert np.all(flat[:geo.lo] == 0.0) and np.any(flat[geo.lo:] != 0.0)
    v = flat[:, :4] + flat[:, 4:]                       # hi + lo parts
    ref = np.concatenate([qu[:, :1] / s[0], J / s[1]], 1)
    np.testing.assert_allclose(v[geo.lo:], ref[geo.lo:], rtol=2e-5, atol=1e-7)
    # each group spans [-1, 1] whatever the ratio of the currents to the charges
    assert np.max(np.abs(v[:, 0])) == pytest.approx(1.0, rel=1e-4)
    assert np.max(np.abs(v[:, 1:])) == pytest.approx(1.0, rel=1e-4)
    assert geo.lo == 1 + 8 + 64
    # no current at all: the current scale falls back to 1
    qu[:, 1:] = 0.0
    assert S.net_input(geo, qu)[1][1] == 1.0


def test_output_groups_scale_separately():
    """Scaling the currents by a factor scales the B output by it and leaves
    E unchanged, scaling the charges at fixed velocity scales both (E in q,
    B in J = q v): the two input scales keep each group scale-free."""
    geo, params, _, _, _ = _problem(21)
    rng = np.random.default_rng(22)
    qu = np.zeros((geo.nblk, 4), np.float32)
    qu[geo.lo:, 0] = np.float32(rng.choice([-1.0, 0.5, 1.0, 2.0], geo.nblk - geo.lo))
    qu[geo.lo:, 1:] = np.float32(rng.choice([-0.25, 0.5, 1.0], (geo.nblk - geo.lo, 3)))
    y0, s0, _ = S.forward(geo, params, qu)
    qv = qu.copy()
    qv[:, 1:] *= np.float32(2.0 ** -14)                # slower by an exact power of two
    y1, s1, _ = S.forward(geo, params, qv)
    assert s1[0] == s0[0] and s1[1] == s0[1] * 2.0 ** -14
    # network units are the same; raw units follow the scales
    np.testing.assert_allclose(y1, y0, rtol=1e-12, atol=1e-12)
    # and so do the targets: the raw far field of the slower input has the
    # same E (E does not see the velocities) and B times 2^-14, and both map
    # to the same network units
    raw0 = S.far_interp(geo, S.far_samples(geo, qu))
    raw1 = S.far_interp(geo, S.far_samples(geo, qv))
    assert np.array_equal(raw1[:3], raw0[:3])
    np.testing.assert_allclose(raw1[3:], raw0[3:] * 2.0 ** -14, rtol=1e-12, atol=0)
    assert np.abs(raw0[3:]).max() > 0
    np.testing.assert_allclose(S.targets(geo, s1, raw1), S.targets(geo, s0, raw0), rtol=1e-12, atol=0)


def _problem(seed, c0=8):
    geo = S.Geometry(16, 16, 16, 3, c0)
    rng = np.random.default_rng(seed)
    params = np.concatenate([rng.standard_normal(r * c) / np.sqrt(c) for r, c in geo.shapes])
    qu = rng.standard_normal((geo.nblk, 4)).astype(np.float32)
    T = rng.standard_normal((6, 16 ** 3))
    fluid = (rng.uniform(size=16 ** 3) > 0.1).astype(np.float64)
    return geo, params, qu, T, fluid


def test_loss_grad_central_differences():
    geo, params, qu, T, fluid = _problem(4)
    L, grad, rel, _ = S.loss_and_grad(geo, params, qu, T, fluid)
    assert L == pytest.approx(rel[0] ** 2 + rel[1] ** 2, rel=1e-12)
    rng = np.random.default_rng(7)
    offs = np.cumsum([0] + [r * c for r, c in geo.shapes])
    h = 1e-5
    for li, (r, c) in enumerate(geo.shapes):
        cin_used = 27 * (geo.hcin if li == 7 else sum(list(geo.layer_dims.values())[li][:2]))
        for _ in range(3):
            k = offs[li] + rng.integers(0, r) * c + rng.integers(0, cin_used)
            p1, p2 = params.copy(), params.copy()
            p1[k] += h
            p2[k] -= h
            fd = (S.loss_and_grad(geo, p1, qu, T, fluid)[0] - S.loss_and_grad(geo, p2, qu, T, fluid)[0]) / (2 * h)
            assert grad[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), (li, k)
    # padding columns of every layer get zero gradient
    for li, (r, c) in enumerate(geo.shapes):
        cin = geo.hcin if li == 7 else sum(list(geo.layer_dims.values())[li][:2])
        blk = grad[offs[li]:offs[li + 1]].reshape(r, c)
        assert np.all(blk[:, 27 * cin:] == 0.0)


def test_loss_grad_directional_derivative():
    """Whole-vector check: <grad, v> equals the central difference along a random v."""
    geo, params, qu, T, fluid = _problem(9)
    _, grad, _, _ = S.loss_and_grad(geo, params, qu, T, fluid)
    v = np.random.default_rng(11).standard_normal(params.size)
    h = 1e-6
    fd = (S.loss_and_grad(geo, params + h * v, qu, T, fluid)[0] -
          S.loss_and_grad(geo, params - h * v, qu, T, fluid)[0]) / (2 * h)
    assert np.dot(grad, v) == pytest.approx(fd, rel=1e-6)


def test_linearity_in_the_sources():
    """The network is linear: the output of a + b equals output(a) + output(b)
    when the inputs are exactly representable (no hi/lo rounding)."""
    geo, params, _, _, _ = _problem(12)
    rng = np.random.default_rng(13)
    qa = np.zeros((geo.nblk, 4), np.float32)
    qb = np.zeros((geo.nblk, 4), np.float32)
    qa[300:310, 0] = 1.0
    qb[400:410, 0] = 1.0
    qa[0, 0] = qb[0, 0] = 2.0      # same scale s = 2 for both and for the sum (index < lo: masked)
    qs = qa + qb
    qs[0, 0] = 2.0
    ya, _, _ = S.forward(geo, params, qa)
    yb, _, _ = S.forward(geo, params, qb)
    ys, _, _ = S.forward(geo, params, qs)
    np.testing.assert_allclose(ys, ya + yb, rtol=1e-10, atol=1e-12)
    del rng


# ---------------------------------------------------------------------------
# Global path: far-field samples and their interpolation
# ---------------------------------------------------------------------------
def test_lagrange3_nodes_and_quadratics():
    for b in (2, 3, 8, 16):
        h = b / 3.0
        nodes = (np.arange(3) + 0.5) * h - 0.5
        np.testing.assert_allclose(S.lagrange3(nodes, b), np.eye(3), atol=1e-12)
        u = np.linspace(-0.5, b - 0.5, 17)
        f = lambda t: 0.3 - 1.7 * t + 0.45 * t * t  # noqa: E731
        np.testing.assert_allclose(np.sum(S.lagrange3(u, b) * f(nodes)[:, None], 0), f(u), rtol=1e-12, atol=1e-12)


def _near_brute(geo, qu):
    """Field at every cell from the 26 neighbouring blocks' sources (q != 0,
    index >= lo), r / |r|^3, sources at the block centres."""
    nd = geo.nd
    n = np.arange(geo.nx * geo.ny * geo.nz)
    x, y, z = n % geo.nx, (n // geo.nx) % geo.ny, n // (geo.nx * geo.ny)
    cb = np.stack([x // geo.bx, y // geo.by, z // geo.bz], 1)
    out = np.zeros((6, n.size))
    for d in range(geo.lo, geo.nblk):
        q = float(qu[d, 0])
        if q == 0.0:
            continue
        sc = np.array([d % nd, (d // nd) % nd, d // (nd * nd)])
        dd = np.abs(cb - sc)
        near = np.all(dd <= 1, axis=1) & np.any(dd > 0, axis=1)
        r = np.stack([x, y, z], 1) - (sc + 0.5) * np.array([geo.bx, geo.by, geo.bz])
        with np.errstate(divide="ignore"):  # r = 0 only at the own block's centre, masked out
            t = np.where(near, q / np.sum(r * r, 1) ** 1.5, 0.0)
        out[:3] += (t[:, None] * r).T
        out[3:] += (t[:, None] * np.cross(qu[d, 1:4].astype(np.float64), r)).T
    return out


def test_near_plus_far_samples_equal_the_oracle_direct_field():
    """With 3-cell blocks the sample lattice is the cell lattice, so near (26
    neighbours) + far samples must add up to the C oracle's float64
    update_e_b_dynamic (KE = KMU = 1, zero static fields) at every cell."""
    import oracle_lib as O
    n, D = 24, 3
    geo = S.Geometry(n, n, n, D, 8)
    rng = np.random.default_rng(21)
    qu = np.zeros((geo.nblk, 4))
    qu[:, 0] = rng.standard_normal(geo.nblk) * (rng.uniform(size=geo.nblk) < 0.7)
    qu[:, 1:] = 0.05 * rng.standard_normal((geo.nblk, 3))
    orc = O.OracleSim(n, n, n, velocity_set=O.D3Q19, float_type=O.FP64, lod_depth=D,
                      ext=O.EXT_VOLUME_FORCE | O.EXT_MAGNETO_HYDRO, ke=1.0, kmu=1.0)
    lod = orc.buf("qu_lod").reshape(-1, 4)
    lod[:geo.nblk] = qu
    orc.kernel("update_e_b_dynamic")
    direct = np.concatenate([orc.buf("e_dyn").reshape(3, -1), orc.buf("b_dyn").reshape(3, -1)])
    far = S.far_samples(geo, qu)                       # sample p of block B sits on a cell
    cells = S.far_interp(geo, far)                     # interpolation is exact at nodes
    mine = _near_brute(geo, qu) + cells
    np.testing.assert_allclose(mine, direct, rtol=1e-9, atol=1e-12 * np.abs(direct).max())


def test_far_interpolation_error_is_small():
    """6-cell blocks: tri-quadratic interpolation of the samples against the
    brute-force far field at every cell (the error the surrogate starts from)."""
    n, D = 48, 3
    geo = S.Geometry(n, n, n, D, 8)
    rng = np.random.default_rng(22)
    from ionsolver_amd.scenarios import random_lod
    qu = random_lod((n, n, n), D, rng).astype(np.float64)
    far = S.far_samples(geo, qu)
    approx = S.far_interp(geo, far)
    # brute-force far field at the cells (sources outside the 3^3 neighbourhood)
    nd = geo.nd
    nn = np.arange(n ** 3)
    x, y, z = nn % n, (nn // n) % n, nn // (n * n)
    cb = np.stack([x // geo.bx, y // geo.by, z // geo.bz], 1)
    exact = np.zeros((6, nn.size))
    for d in range(geo.lo, geo.nblk):
        if qu[d, 0] == 0.0:
            continue
        sc = np.array([d % nd, (d // nd) % nd, d // (nd * nd)])
        far_m = np.any(np.abs(cb - sc) > 1, axis=1)
        r = np.stack([x, y, z], 1) - (sc + 0.5) * geo.bx
        with np.errstate(divide="ignore"):
            t = np.where(far_m, qu[d, 0] / np.sum(r * r, 1) ** 1.5, 0.0)
        exact[:3] += (t[:, None] * r).T
        exact[3:] += (t[:, None] * np.cross(qu[d, 1:4], r)).T
    total = exact + _near_brute(geo, qu)
    for sl in (slice(0, 3), slice(3, 6)):
        err = np.linalg.norm(approx[sl] - exact[sl]) / np.linalg.norm(total[sl])
        assert err < 3e-3, err
