"""Pinning of the float64 surrogate network reference (oracle/surrogate_ref.py).

The reference is the checker of the GPU network, so each of its pieces is
checked against something independent: the convolution against torch's
conv3d, the bf16 rounding against torch's bfloat16 cast, pooling and
upsampling against their adjoint identities, the block <-> cell maps against
a brute-force loop, and the hand-written backward pass against central finite
differences of the loss.
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
    x, s = S.net_input(geo, qu)
    assert s == np.max(np.abs(qu[:, 0]))
    flat = x.reshape(-1, 8)
    assert np.all(flat[:geo.lo] == 0.0) and np.any(flat[geo.lo:] != 0.0)
    v = flat[:, :4] + flat[:, 4:]                       # hi + lo parts
    ref = np.stack([qu[:, 0], qu[:, 0] * qu[:, 1], qu[:, 0] * qu[:, 2], qu[:, 0] * qu[:, 3]], 1) / s
    np.testing.assert_allclose(v[geo.lo:], ref[geo.lo:], rtol=2e-5, atol=1e-7)
    assert geo.lo == 1 + 8 + 64


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
