"""Procedural STL geometry (float32 triangles, closed 2-manifold meshes) and a
binary STL writer.  Vertices are computed once per lattice point of the mesh
and shared, so every edge is shared by exactly two triangles."""
import struct

import numpy as np


def uv_sphere(radius, center, n_lat=120, n_lon=240):
    """Closed UV sphere with n_lat latitude bands and n_lon longitude segments:
    2 * n_lon * (n_lat - 1) triangles, outward orientation."""
    cx, cy, cz = center
    theta = np.pi * np.arange(n_lat + 1) / n_lat
    phi = 2.0 * np.pi * np.arange(n_lon) / n_lon
    st, ct = np.sin(theta), np.cos(theta)
    verts = np.zeros((n_lat + 1, n_lon, 3), np.float32)
    verts[..., 0] = (cx + radius * st[:, None] * np.cos(phi)[None, :]).astype(np.float32)
    verts[..., 1] = (cy + radius * st[:, None] * np.sin(phi)[None, :]).astype(np.float32)
    verts[..., 2] = (cz + radius * ct[:, None] * np.ones(n_lon)[None, :]).astype(np.float32)
    verts[0, :, :] = verts[0, 0, :]            # single north pole vertex
    verts[n_lat, :, :] = verts[n_lat, 0, :]    # single south pole vertex
    tris = []
    for a in range(n_lat):
        for b in range(n_lon):
            b1 = (b + 1) % n_lon
            p00, p01 = verts[a, b], verts[a, b1]
            p10, p11 = verts[a + 1, b], verts[a + 1, b1]
            if a != 0:
                tris.append(np.concatenate([p00, p10, p01]))
            if a != n_lat - 1:
                tris.append(np.concatenate([p01, p10, p11]))
    return np.array(tris, np.float32)


def box(lo, hi):
    """Closed axis-aligned box, 12 triangles, outward orientation."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    faces = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
             (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return np.array([np.concatenate([v[a], v[b], v[c]]) for a, b, c in faces], np.float32)


def write_binary_stl(path, tris):
    tris = np.asarray(tris, np.float32).reshape(-1, 9)
    with open(path, "wb") as fh:
        fh.write(b"lbm3 procedural".ljust(80, b" "))
        fh.write(struct.pack("<I", len(tris)))
        for t in tris:
            a, b, c = t[0:3], t[3:6], t[6:9]
            n = np.cross(b - a, c - a)
            nn = np.linalg.norm(n)
            n = n / nn if nn > 0 else n
            fh.write(struct.pack("<12fH", *n.astype(np.float32), *t, 0))


def read_binary_stl(path):
    with open(path, "rb") as fh:
        fh.read(80)
        n = struct.unpack("<I", fh.read(4))[0]
        rec = np.frombuffer(fh.read(50 * n), dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")]))
    return np.array(rec["v"], np.float32)
