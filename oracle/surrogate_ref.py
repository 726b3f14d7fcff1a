"""Float64 reference of the learned field surrogate (test infrastructure only:
the checker of the GPU network's forward output and gradients).

Restates the network of ionsolver_amd/csrc/surrogate_host.cpp in numpy:
input (q / s_q, J / s_J, J = q v) of the top-level LOD blocks as bf16 hi +
lo parts (s_q = max |q|, s_J = max |J| over blocks and components),
a linear 4-level 3D conv U-Net over the LOD grid (3^3 convolutions, 2^3
average pooling, nearest upsampling, skip concatenation), and the sub-pixel
output head that maps the 3^3 neighbourhood of [d0, x] to the 6 field
components of every cell of a LOD block; plus the global path: far-field
samples of every block (direct sums over the sources outside its 3^3
neighbourhood at a 3^3 point lattice) and their tri-quadratic
interpolation onto the cells. Loss: relative squared error of E and of B
over fluid cells. The backward pass is written out by hand; the tests check
it against central finite differences, the far field against the direct
field of the C oracle.

Parameter layout (flat): layers e0, e1, e2, e3, d2, d1, d0, head, each a
[cout][kpad] matrix with column k = delta * cin + ci, delta = (dx+1) +
3 (dy+1) + 9 (dz+1), kpad = 27 cin rounded up to a multiple of 32.
"""
import numpy as np

LAYERS = ["e0", "e1", "e2", "e3", "d2", "d1", "d0"]


def pad32(v):
    return (v + 31) // 32 * 32


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even), as float64."""
    f = np.asarray(x, np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    rounded = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


class Geometry:
    def __init__(self, nx, ny, nz, D, c0):
        self.nx, self.ny, self.nz, self.D, self.c0 = nx, ny, nz, D, c0
        self.nd = 1 << D
        self.bx, self.by, self.bz = nx // self.nd, ny // self.nd, nz // self.nd
        self.cpb = self.bx * self.by * self.bz
        self.nout = 6 * self.cpb
        self.gain = float(np.float32(self.cpb ** (2.0 / 3.0)))
        self.nblk = self.nd ** 3
        self.lo = sum(8 ** k for k in range(D + 1)) - self.nblk
        C = [c0, 2 * c0, 3 * c0, 4 * c0]
        g0 = self.nd
        # (cin_a, cin_b, cout, grid)
        self.layer_dims = {
            "e0": (8, 0, C[0], g0), "e1": (C[0], 0, C[1], g0 // 2), "e2": (C[1], 0, C[2], g0 // 4),
            "e3": (C[2], 0, C[3], g0 // 8), "d2": (C[3], C[2], C[2], g0 // 4), "d1": (C[2], C[1], C[1], g0 // 2),
            "d0": (C[1], C[0], C[0], g0),
        }
        self.hcin = c0 + 8
        self.hk = pad32(27 * self.hcin)
        self.shapes = [(d[2], pad32(27 * (d[0] + d[1]))) for d in (self.layer_dims[n] for n in LAYERS)]
        self.shapes.append((self.nout, self.hk))

    def split(self, flat):
        out, off = [], 0
        for r, c in self.shapes:
            out.append(flat[off:off + r * c].reshape(r, c))
            off += r * c
        assert off == flat.size
        return out

    def n_params(self):
        return sum(r * c for r, c in self.shapes)


def conv3(x, W, cin):
    """x: (g, g, g, cin) [z, y, x, c]; W: (cout, kpad) -> (g, g, g, cout)."""
    g = x.shape[0]
    p = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((g, g, g, W.shape[0]))
    for d in range(27):
        dx, dy, dz = d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1
        sl = p[1 + dz:1 + dz + g, 1 + dy:1 + dy + g, 1 + dx:1 + dx + g, :]
        out += sl @ W[:, d * cin:(d + 1) * cin].T
    return out


def conv3_bwd(x, W, cin, dout):
    """Gradients of conv3 w.r.t. W (full kpad width) and x."""
    g = x.shape[0]
    p = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros_like(W)
    dp = np.zeros_like(p)
    for d in range(27):
        dx, dy, dz = d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1
        sl = (slice(1 + dz, 1 + dz + g), slice(1 + dy, 1 + dy + g), slice(1 + dx, 1 + dx + g), slice(None))
        dW[:, d * cin:(d + 1) * cin] = np.einsum("zyxo,zyxc->oc", dout, p[sl])
        dp[sl] += dout @ W[:, d * cin:(d + 1) * cin]
    return dW, dp[1:-1, 1:-1, 1:-1, :]


def pool(x):
    g = x.shape[0]
    return x.reshape(g // 2, 2, g // 2, 2, g // 2, 2, -1).mean(axis=(1, 3, 5))


def pool_bwd(d):
    return np.repeat(np.repeat(np.repeat(d, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_bwd(d):
    g = d.shape[0]
    return d.reshape(g // 2, 2, g // 2, 2, g // 2, 2, -1).sum(axis=(1, 3, 5))


def net_input(geo, qu):
    """(nblk, 4) LOD sums -> (g, g, g, 8) input, scales (s_q, s_J): the charge
    channel over s_q = max |q|, the three current channels J = q v (float32
    products) over s_J = max |J|; a zero maximum scales by 1. E is linear in
    q and B in J, so each group of outputs is scale-free on its own however
    small J is next to q."""
    q = qu[:geo.nblk].astype(np.float32)
    J = q[:, 0:1] * q[:, 1:4]                                   # float32 products, as the GPU forms them
    s_q = float(np.max(np.abs(q[:, 0])))
    s_J = float(np.max(np.abs(J)))
    s_q = s_q if s_q > 0 else 1.0
    s_J = s_J if s_J > 0 else 1.0
    inv = np.float32([1.0 / np.float32(s_q)] + [1.0 / np.float32(s_J)] * 3)
    v = np.concatenate([q[:, 0:1], J], 1)
    v = np.float32(v * inv).astype(np.float64)
    v[:geo.lo] = 0.0
    hi = bf16(v)
    lo = bf16(v - hi)
    x = np.concatenate([hi, lo], 1)
    g = geo.nd
    return x.reshape(g, g, g, 8), (s_q, s_J)


def forward(geo, params, qu):
    Ws = geo.split(params)
    W = dict(zip(LAYERS, Ws[:7]))
    Wh = Ws[7]
    x, s = net_input(geo, qu)
    a = {}
    a["x"] = x
    a["e0"] = conv3(x, W["e0"], 8)
    a["p0"] = pool(a["e0"])
    a["e1"] = conv3(a["p0"], W["e1"], geo.layer_dims["e1"][0])
    a["p1"] = pool(a["e1"])
    a["e2"] = conv3(a["p1"], W["e2"], geo.layer_dims["e2"][0])
    a["p2"] = pool(a["e2"])
    a["e3"] = conv3(a["p2"], W["e3"], geo.layer_dims["e3"][0])
    a["u2"] = np.concatenate([up(a["e3"]), a["e2"]], 3)
    a["d2"] = conv3(a["u2"], W["d2"], a["u2"].shape[3])
    a["u1"] = np.concatenate([up(a["d2"]), a["e1"]], 3)
    a["d1"] = conv3(a["u1"], W["d1"], a["u1"].shape[3])
    a["u0"] = np.concatenate([up(a["d1"]), a["e0"]], 3)
    a["d0"] = conv3(a["u0"], W["d0"], a["u0"].shape[3])
    a["h"] = np.concatenate([a["d0"], x], 3)
    Y = conv3(a["h"], Wh, geo.hcin)  # (g, g, g, nout) per block
    return Y, s, a


def block_to_cells(geo, Y):
    """(g, g, g, 6*cpb) block outputs -> (6, ncell) fields."""
    g = geo.nd
    Y = Y.reshape(g, g, g, 6, geo.bz, geo.by, geo.bx)           # Z, Y, X, c, oz, oy, ox
    Y = Y.transpose(3, 0, 4, 1, 5, 2, 6)                         # c, Z, oz, Y, oy, X, ox
    return Y.reshape(6, geo.nx * geo.ny * geo.nz)


def cells_to_block(geo, F):
    g = geo.nd
    F = F.reshape(6, g, geo.bz, g, geo.by, g, geo.bx).transpose(1, 3, 5, 0, 2, 4, 6)
    return F.reshape(g, g, g, 6 * geo.cpb)


def targets(geo, s, raw):
    """Network-unit targets from the direct kernel's raw source sums
    (E then B, 6 x N): E rows raw * gain / s_q, B rows raw * gain / s_J,
    gain = (cells per block)^(2/3)."""
    t = raw.reshape(6, -1).astype(np.float64).copy()
    t[:3] *= geo.gain / s[0]
    t[3:] *= geo.gain / s[1]
    return t


def net_units(geo, s, f):
    """(6, N) raw-unit field -> network units (the scaling of targets)."""
    return targets(geo, s, f)


def far_samples(geo, qu):
    """Far field of every top-level block: the in-range sources (LOD index
    >= lo, q != 0) outside the block's 3^3 neighbourhood, summed directly
    (E from q, B from q v x r, r / |r|^3) at the 3^3 lattice of points
    ((k + 1/2)/3) b - 1/2 of each axis. Sources sit at the LOD centres
    ((i + 1/2) b) in cell coordinates (sim_kernels.cl:442-447). Returns
    [nblk][27][6] float64 (E then B)."""
    nd, D = geo.nd, geo.D
    q = qu[:geo.nblk].astype(np.float64)
    src = np.arange(geo.lo, geo.nblk)
    src = src[q[src, 0] != 0.0]
    b = np.array([geo.bx, geo.by, geo.bz], np.float64)
    sc = np.stack([src % nd, (src // nd) % nd, src // (nd * nd)], 1)
    spos = (sc + 0.5) * b
    blk = np.arange(geo.nblk)
    bc = np.stack([blk % nd, (blk // nd) % nd, blk // (nd * nd)], 1)
    k = np.arange(27)
    pk = np.stack([k % 3, (k // 3) % 3, k // 9], 1)
    pts = bc[:, None, :] * b + ((pk[None, :, :] + 0.5) / 3.0) * b - 0.5       # [nblk][27][3]
    out = np.zeros((geo.nblk, 27, 6))
    for j, sidx in enumerate(src):
        far = np.any(np.abs(bc - sc[j]) > 1, axis=1)                          # [nblk]
        r = pts - spos[j]
        r3 = np.sum(r * r, axis=2) ** 1.5
        t = np.where(far[:, None], q[sidx, 0] / r3, 0.0)
        v = q[sidx, 1:4]
        out[:, :, 0:3] += t[..., None] * r
        out[:, :, 3:6] += t[..., None] * np.cross(v, r)
    return out


def lagrange3(u, b):
    """Quadratic Lagrange basis of the nodes ((k + 1/2)/3) b - 1/2, k = 0, 1, 2."""
    h = b / 3.0
    s = (np.asarray(u, np.float64) - (0.5 * h - 0.5)) / h
    return np.stack([0.5 * (s - 1) * (s - 2), -s * (s - 2), 0.5 * s * (s - 1)])


def far_interp(geo, samples):
    """Tri-quadratic interpolation of far samples onto every cell: (6, N)."""
    n = np.arange(geo.nx * geo.ny * geo.nz)
    x, y, z = n % geo.nx, (n // geo.nx) % geo.ny, n // (geo.nx * geo.ny)
    blk = x // geo.bx + geo.nd * (y // geo.by + geo.nd * (z // geo.bz))
    Lx, Ly, Lz = lagrange3(x % geo.bx, geo.bx), lagrange3(y % geo.by, geo.by), lagrange3(z % geo.bz, geo.bz)
    out = np.zeros((6, n.size))
    for p in range(27):
        w = Lz[p // 9] * Ly[(p // 3) % 3] * Lx[p % 3]
        out += w[None, :] * samples[blk, p, :].T
    return out


def loss_and_grad(geo, params, qu, T, fluid, y_add=None):
    """Loss sum_E (y-t)^2/sum_E t^2 + sum_B (y-t)^2/sum_B t^2 over fluid cells,
    its gradient w.r.t. the flat parameters, and (rel_e, rel_b). A group whose
    sum t^2 is at most 1e-8 of the other's is left out of the loss and the
    gradient (its weight would overflow the GPU's bf16 gradients). y_add: a
    constant (6, N) term added to the network output (the far field in
    network units)."""
    Y, s, a = forward(geo, params, qu)
    y = block_to_cells(geo, Y)
    if y_add is not None:
        y = y + y_add
    m = fluid[None, :]
    d = (y - T) * m
    nE, nB = np.sum((T[:3] * m) ** 2), np.sum((T[3:] * m) ** 2)
    wE = 1.0 / nE if nE > 1e-8 * nB else 0.0
    wB = 1.0 / nB if nB > 1e-8 * nE else 0.0
    L = np.sum(d[:3] ** 2) * wE + np.sum(d[3:] ** 2) * wB
    dy = np.concatenate([2 * d[:3] * wE, 2 * d[3:] * wB], 0)
    dY = cells_to_block(geo, dy)
    Ws = geo.split(params)
    W = dict(zip(LAYERS, Ws[:7]))
    grads = {}
    grads["head"], dh = conv3_bwd(a["h"], Ws[7], geo.hcin, dY)
    c0, C1, C2 = geo.c0, 2 * geo.c0, 3 * geo.c0
    dd0 = dh[..., :c0]
    grads["d0"], du0 = conv3_bwd(a["u0"], W["d0"], a["u0"].shape[3], dd0)
    dd1, de0 = up_bwd(du0[..., :C1]), du0[..., C1:]
    grads["d1"], du1 = conv3_bwd(a["u1"], W["d1"], a["u1"].shape[3], dd1)
    dd2, de1 = up_bwd(du1[..., :C2]), du1[..., C2:]
    grads["d2"], du2 = conv3_bwd(a["u2"], W["d2"], a["u2"].shape[3], dd2)
    de3, de2 = up_bwd(du2[..., :4 * c0]), du2[..., 4 * c0:]
    grads["e3"], dp2 = conv3_bwd(a["p2"], W["e3"], C2, de3)
    de2 = de2 + pool_bwd(dp2)
    grads["e2"], dp1 = conv3_bwd(a["p1"], W["e2"], C1, de2)
    de1 = de1 + pool_bwd(dp1)
    grads["e1"], dp0 = conv3_bwd(a["p0"], W["e1"], c0, de1)
    de0 = de0 + pool_bwd(dp0)
    grads["e0"], _ = conv3_bwd(a["x"], W["e0"], 8, de0)
    flat = np.concatenate([grads[n].ravel() for n in LAYERS + ["head"]])
    rel = (np.sqrt(np.sum(d[:3] ** 2) / nE), np.sqrt(np.sum(d[3:] ** 2) / nB))
    return L, flat, rel, y
