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


def _axis_frame(axis):
    """(a, b, c): the axis index and the two in-plane indices, right-handed."""
    return axis, (axis + 1) % 3, (axis + 2) % 3


def rotor(n_blades=4, radius=10.0, hub_ratio=0.2, twist_deg=25.0, length=6.0, center=(0.0, 0.0, 0.0), axis=2,
          n_theta=1024, n_axial=48, sharpness=24.0):
    """Closed N-blade rotor: a star-shaped section r(theta) = r_hub + (R - r_hub)
    |cos(N theta / 2)|^p (N blades of hub radius hub_ratio * R), extruded
    along `axis` over `length` with the section turned by twist_deg from one
    end to the other.  2 * n_theta * n_axial side triangles plus two fans of
    n_theta, outward orientation, every edge shared by two triangles."""
    R, rh = float(radius), float(hub_ratio) * float(radius)
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    r = rh + (R - rh) * np.abs(np.cos(0.5 * n_blades * th)) ** sharpness
    z = -0.5 * length + length * np.arange(n_axial + 1) / n_axial
    tw = np.deg2rad(twist_deg) * (z / length + 0.5)
    a, b, c = _axis_frame(axis)
    ring = np.zeros((n_axial + 1, n_theta, 3))
    ring[..., a] = z[:, None] + center[a]
    ring[..., b] = center[b] + r[None, :] * np.cos(th[None, :] + tw[:, None])
    ring[..., c] = center[c] + r[None, :] * np.sin(th[None, :] + tw[:, None])
    ring = ring.astype(np.float32)
    ends = np.zeros((2, 3), np.float32)
    ends[:, a] = (z[0] + center[a], z[-1] + center[a])
    ends[:, b], ends[:, c] = center[b], center[c]
    i = np.arange(n_theta)
    i1 = (i + 1) % n_theta
    tris = []
    for k in range(n_axial):
        p00, p01, p10, p11 = ring[k, i], ring[k, i1], ring[k + 1, i], ring[k + 1, i1]
        tris.append(np.concatenate([p00, p01, p11], 1))
        tris.append(np.concatenate([p00, p11, p10], 1))
    lo = np.repeat(ends[:1], n_theta, 0)
    hi = np.repeat(ends[1:], n_theta, 0)
    tris.append(np.concatenate([lo, ring[0, i1], ring[0, i]], 1))
    tris.append(np.concatenate([hi, ring[-1, i], ring[-1, i1]], 1))
    return np.concatenate(tris).astype(np.float32)


def tube(r_in, r_out, length, center=(0.0, 0.0, 0.0), axis=0, n_theta=256):
    """Closed thick-walled cylinder (an annulus extruded along `axis`), the
    duct wall around a rotor: 8 * n_theta triangles, outward orientation."""
    a, b, c = _axis_frame(axis)
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    i = np.arange(n_theta)
    i1 = (i + 1) % n_theta

    def ring(rad, zz):
        p = np.zeros((n_theta, 3))
        p[:, a] = center[a] + zz
        p[:, b] = center[b] + rad * np.cos(th)
        p[:, c] = center[c] + rad * np.sin(th)
        return p.astype(np.float32)

    z0, z1 = -0.5 * length, 0.5 * length
    o0, o1, n0, n1 = ring(r_out, z0), ring(r_out, z1), ring(r_in, z0), ring(r_in, z1)
    t = [np.concatenate([o0[i], o0[i1], o1[i1]], 1), np.concatenate([o0[i], o1[i1], o1[i]], 1),      # outer wall
         np.concatenate([n0[i], n1[i1], n0[i1]], 1), np.concatenate([n0[i], n1[i], n1[i1]], 1),      # inner wall
         np.concatenate([o1[i], o1[i1], n1[i1]], 1), np.concatenate([o1[i], n1[i1], n1[i]], 1),      # end z1
         np.concatenate([o0[i], n0[i1], o0[i1]], 1), np.concatenate([o0[i], n0[i], n0[i1]], 1)]      # end z0
    return np.concatenate(t).astype(np.float32)


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
