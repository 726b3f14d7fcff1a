"""float32 restatement of the reference's STL import (src/mesh.rs).

TEST INFRASTRUCTURE ONLY: the tests use it to check the library's host-side
mesh import; nothing in the product imports it.

Every operation is a float32 operation in the order the Rust source writes
it (Rust does not contract multiply-adds): matrix-vector products
(m.xx v.x + m.xy v.y) + m.xz v.z (mesh.rs:397-405), matrix products likewise
(:411-421), f32::min / f32::max bounds (:97-113), and the placement
center + scale (offset + p) (:207-211). sin and cos are taken in double and
rounded to float32 (the reference calls the C library's sinf / cosf, which
can differ from that by one unit in the last place for angles other than 0).
"""
import math

import numpy as np

F = np.float32
STL_RECORD = np.dtype([("normal", "<f4", (3,)), ("v", "<f4", (3, 3)), ("attr", "<u2")])


def rotm_around_v(v, r):
    """F32_3_3::rotm_around_v, mesh.rs:58-68."""
    sr, cr = F(math.sin(float(r))), F(math.cos(float(r)))
    x, y, z = (F(c) for c in v)
    one = F(1.0)
    return dict(xx=x * x + (one - x * x) * cr, xy=x * y * (one - cr) - z * sr, xz=x * z * (one - cr) + y * sr,
                yx=x * y * (one - cr) + z * sr, yy=y * y + (one - y * y) * cr, yz=y * z * (one - cr) - x * sr,
                zx=x * z * (one - cr) - y * sr, zy=y * z * (one - cr) + x * sr, zz=z * z + (one - z * z) * cr)


def matmul(a, m):
    """Mul<F32_3_3> for F32_3_3, mesh.rs:411-421."""
    r = {}
    for i in "xyz":
        for j in "xyz":
            r[i + j] = a[i + "x"] * m["x" + j] + a[i + "y"] * m["y" + j] + a[i + "z"] * m["z" + j]
    return r


def rotation_matrix(rx, ry, rz):
    """construct_rotation_matrix with degrees converted as mesh.rs:237 does."""
    pi = F(math.pi)
    rad = [F(a) * pi / F(180.0) for a in (rx, ry, rz)]
    return matmul(matmul(rotm_around_v((1, 0, 0), rad[0]), rotm_around_v((0, 1, 0), rad[1])),
                  rotm_around_v((0, 0, 1), rad[2]))


def apply(m, p):
    """Mul<F32_3> for F32_3_3 on an [..., 3] float32 array."""
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return np.stack([m["xx"] * x + m["xy"] * y + m["xz"] * z,
                     m["yx"] * x + m["yy"] * y + m["yz"] * z,
                     m["zx"] * x + m["zy"] * y + m["zz"] * z], -1).astype(F)


def bounds(v):
    flat = v.reshape(-1, 3)
    return flat.min(0), flat.max(0)


def read_stl_raw(path, reposition, box, center, rot, size):
    """Mesh::read_stl_raw, mesh.rs:175-217: returns (vertices [T, 3, 3], p_min, p_max)."""
    b = open(path, "rb").read()
    T = int(np.frombuffer(b[80:84], "<u4")[0]) if len(b) >= 84 else 0
    if not (T > 0 and len(b) == 84 + 50 * T):
        raise ValueError("corrupted or unsupported STL")
    v = np.frombuffer(b, STL_RECORD, count=T, offset=84)["v"].astype(F)
    v = apply(rot, v)
    lo, hi = bounds(v)
    size = F(size)
    box = np.asarray(box, F)
    if size == 0:
        scale = min(box[0] / (hi[0] - lo[0]), min(box[1] / (hi[1] - lo[1]), box[2] / (hi[2] - lo[2])))
    elif size > 0:
        scale = size / max(hi[0] - lo[0], max(hi[1] - lo[1], hi[2] - lo[2]))
    else:
        scale = -size
    scale = F(scale)
    offset = (F(-0.5) * (lo + hi)).astype(F) if reposition else np.zeros(3, F)
    c = np.asarray(center, F)
    v = (c + scale * (offset + v)).astype(F)
    lo, hi = bounds(v)
    return v, lo, hi


def import_mesh(path, scale, ox, oy, oz, rx, ry, rz, units_m):
    """Lbm::import_mesh, mesh.rs:234-240 (units_m: Units::m)."""
    scale_lu = F(scale) / F(units_m)
    return read_stl_raw(path, False, (1.0, 1.0, 1.0), (ox, oy, oz), rotation_matrix(rx, ry, rz), -abs(scale_lu))


def import_mesh_reposition(path, cx, cy, cz, rx, ry, rz, size, shape):
    """Lbm::import_mesh_reposition, mesh.rs:248-254."""
    return read_stl_raw(path, True, tuple(float(s) for s in shape), (cx, cy, cz), rotation_matrix(rx, ry, rz), size)
