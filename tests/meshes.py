"""Synthetic closed triangle meshes written as binary STL (80-byte header,
uint32 count, 50-byte records: normal, 3 vertices, attribute), the format
the reference reads (src/mesh.rs:175-199)."""
import numpy as np


def write_stl(path, tri):
    tri = np.asarray(tri, np.float32).reshape(-1, 3, 3)
    rec = np.zeros(len(tri), np.dtype([("normal", "<f4", (3,)), ("v", "<f4", (3, 3)), ("attr", "<u2")]))
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    rec["normal"] = n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    rec["v"] = tri
    with open(path, "wb") as f:
        f.write(b"synthetic mesh".ljust(80, b" "))
        f.write(np.uint32(len(tri)).tobytes())
        f.write(rec.tobytes())
    return str(path)


def box(lo, hi):
    """Axis-aligned box, 12 outward-facing triangles."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    tri = []
    for a, b, cc, d in quads:
        tri += [[c[a], c[b], c[cc]], [c[a], c[cc], c[d]]]
    return np.array(tri, np.float32)


def icosphere(center, r, subdiv=2):
    """Geodesic sphere of radius r (vertices on the sphere), outward facing."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(v) * r + np.asarray(center, np.float64)
    return np.array([[V[a], V[b], V[c]] for a, b, c in f], np.float32), V, f


def inscribed_radius(V, faces, center):
    """Distance from center to the nearest face plane (the largest ball inside)."""
    c = np.asarray(center, np.float64)
    d = []
    for a, b, cc in faces:
        n = np.cross(V[b] - V[a], V[cc] - V[a])
        d.append(abs(np.dot(n / np.linalg.norm(n), V[a] - c)))
    return min(d)
