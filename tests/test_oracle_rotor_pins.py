"""Pins of the oracle's rotors (rotorFunctions.h) against what the geometry
and the mechanics fix.  CPU only.

* buildRotors on an axis-aligned box: every interpolation cell's eight corner
  counts equal the number of solid voxels among the 2x2x2 voxels around each
  corner, which for a box factorises into per-axis counts (0, 1 or 2) of the
  voxel layers inside the box; blocks are marked exactly where their 6^3 voxel
  neighbourhood meets the box, with increasing compact indices;
* the solid fraction is the trilinear interpolation of count / 8: 1 deep
  inside, 0 outside, and a rotor turned by a quarter about z sees the point
  turned back by a quarter;
* forcing: inside a fully solid rotor a cell leaves the collision with
  momentum rho * (Omega x r), the rigid rotation; in a partly solid cell the
  momentum moves toward it by s^2 (3 - 2 s) of the gap, s the solid fraction;
* the angle advances by dt * Omega per step modulo 2 pi;
* the torque sum of the force cells matches sum over cells of r x g."""
import numpy as np
import pytest

from lbm3_amd import cases, geometry
from lbm3_amd import ctypes_defs as T
from tests.oracle_lib import OracleSim
from tests.test_oracle_pins import _dirs

LO, HI = (-5.3, -4.1, -3.7), (4.9, 3.3, 5.2)


def _box_rotor(double=True, lo=LO, hi=HI, n=24):
    case, dom, _ = cases.periodic_box(n)
    o = OracleSim(case, dom, [], double=double)
    o.build_rotors([geometry.box(lo, hi)])
    return o


def _axis_inside(v, length, half_open_low):
    """Voxel layers of one axis inside the box, by the voxelizer's sampling
    rule: layer v sits at v * res from the box's low face; x and y rays lie
    inside for 0 < p <= L (the snap puts faces 1/1000 of a voxel above their
    lattice position), z hits count 0 <= p < L (ceil of the hit height)."""
    p = v * 0.5
    return (p >= 0) & (p < length) if half_open_low else (p > 0) & (p <= length)


def test_box_rotor_blocks_and_corner_counts_by_geometry():
    o = _box_rotor()
    info = o.rotor_info(0, 0)
    res, cx, cy, cz, nz = info[0], int(info[7]), int(info[8]), int(info[9]), int(info[11])
    assert res == 0.5 and cx % 4 == 0 and cy % 4 == 0 and cz % 4 == 0
    L = np.array(HI, np.float32) - np.array(LO, np.float32)
    rmap, index, interp = o.array("rotor0Map"), o.array("rotor0Index"), o.array("rotor0Interp")
    bx, by, bz = cx // 4, cy // 4, cz // 4
    assert len(rmap) == bx * by * bz and len(index) == nz and len(interp) == 64 * nz
    vcount = [int(L[a] / np.float32(0.5)) + 2 for a in range(3)]
    # per-axis solid layers of the voxel grid
    inside = [_axis_inside(np.arange(vcount[a]), L[a], a == 2) for a in range(3)]

    def axis_solid(a, v):
        v = np.asarray(v)
        ok = (v >= 0) & (v < vcount[a])
        out = np.zeros(v.shape, bool)
        out[ok] = inside[a][v[ok]]
        return out

    # blocks: marked iff the 6^3 neighbourhood (x, y clamped to the grid) meets the box
    marked = np.zeros(len(rmap), bool)
    for b in range(len(rmap)):
        kb, jb, ib = b // (bx * by), (b % (bx * by)) // bx, b % bx
        hit = [axis_solid(a, np.arange(4 * c - 2, 4 * c + 4)).any() for a, c in ((0, ib), (1, jb), (2, kb))]
        marked[b] = all(hit)
    assert np.array_equal(rmap >= 0, marked)
    assert np.array_equal(np.flatnonzero(marked), index)
    assert np.array_equal(rmap[index], np.arange(nz))
    # corner counts: corner P (interpolation-vertex index) sits between voxels P-2 and P-1 per axis
    for x, b in enumerate(index):
        kb, jb, ib = b // (bx * by), (b % (bx * by)) // bx, b % bx
        for q in range(64):
            ki, ji, ii = q // 16, (q // 4) % 4, q % 4
            packed = int(interp[64 * x + q])
            for corner in range(8):
                ci, cj, ck = corner & 1, (corner >> 1) & 1, corner >> 2
                P = (4 * ib + ii + ci, 4 * jb + ji + cj, 4 * kb + ki + ck)
                want = 1
                for a in range(3):
                    want *= int(axis_solid(a, [P[a] - 2, P[a] - 1]).sum())
                assert (packed >> (4 * corner)) & 15 == want, (b, q, corner)


def test_solid_fraction_inside_outside_and_turned():
    o = _box_rotor(double=False)
    o.set_rotor(0, 0, 0.0, T.AXIS_Z, (0.0, 0.0, 0.0))
    assert o.rotor_fraction(0, 0, 0.0, 0.0, 0.0, 0.5) == 1.0
    assert o.rotor_fraction(0, 0, 0.0, 8.0, 0.0, 0.0) == 0.0
    assert o.rotor_fraction(0, 0, 0.0, 0.0, 0.0, 7.9) == 0.0
    # across a face the fraction falls monotonically from 1 to 0 within about a voxel
    xs = np.linspace(3.5, 6.0, 26)
    fr = np.array([o.rotor_fraction(0, 0, 0.0, x, 0.0, 0.5) for x in xs])
    assert fr[0] == 1.0 and fr[-1] == 0.0 and np.all(np.diff(fr) <= 0)
    # a quarter turn: the rotor frame sees (x, y) at (y, -x)
    rng = np.random.default_rng(5)
    for _ in range(40):
        x, y, z = rng.uniform(-7, 7), rng.uniform(-7, 7), rng.uniform(-4, 6)
        a = o.rotor_fraction(0, 0, np.float32(np.pi / 2), x, y, z)
        b = o.rotor_fraction(0, 0, 0.0, y, -x, z)
        assert abs(a - b) < 1e-4, (x, y, z)


def _logical_f(o):
    """Pre-collision populations f[cell, q] of the current parity from the
    stored slots (esotwist: direction q is read from the neighbour that lies
    in +x/+y/+z where c_q is negative, in slot q or inv(q) by the flipper)."""
    info = o.info()
    n = info.cellCount
    raw = o.array("f").reshape(27, n)
    sh = o.array("shifter")
    rjp, rkp, rjkp = o.array("rowJPlus"), o.array("rowKPlus"), o.array("rowJKPlus")
    c = np.arange(n)
    head = np.where(sh >= 0, c, c + sh)
    row = sh[head]
    back = c - head
    wrap = lambda v: np.where(v + 1 >= n, 0, v + 1)
    j, k, jk = rjp[row] + back, rkp[row] + back, rjkp[row] + back
    nb = {(0, 0, 0): c, (1, 0, 0): wrap(c), (0, 1, 0): j, (0, 0, 1): k, (1, 1, 0): wrap(j), (1, 0, 1): wrap(k),
          (0, 1, 1): jk, (1, 1, 1): wrap(jk)}
    cx, cy, cz, inv = _dirs()
    f = np.zeros((n, 27))
    for q in range(27):
        plane = inv[q] if info.esotwistFlipper else q
        f[:, q] = raw[plane, nb[(int(cx[q] < 0), int(cy[q] < 0), int(cz[q] < 0))]]
    return f


def _one_step_post_momentum(o):
    """Run one step; post-collision momentum of every interior cell
    (sum_q c_q f*_q(y) = sum_q c_q f_q(y + c_q) after streaming), with the
    positions and pre-step densities."""
    info = o.info()
    n1 = info.cellCountX
    m0 = o.macros().astype(np.float64)
    o.run(1)
    f = _logical_f(o).reshape(n1, n1, n1, 27)
    cx, cy, cz, _ = _dirs()
    P = np.zeros((n1, n1, n1, 3))
    for q in range(27):
        fq = np.roll(f[..., q], shift=(-cz[q], -cy[q], -cx[q]), axis=(0, 1, 2))
        P += fq[..., None] * np.array([cx[q], cy[q], cz[q]], float)
    inner = (slice(2, -2),) * 3
    k, j, i = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    pos = np.stack([i * info.res + info.ox, j * info.res + info.oy, k * info.res + info.oz], -1)
    return P[inner].reshape(-1, 3), pos[inner].reshape(-1, 3), m0.reshape(n1, n1, n1, 4)[inner].reshape(-1, 4)


@pytest.mark.parametrize("axis", [T.AXIS_X, T.AXIS_Y, T.AXIS_Z])
def test_fully_solid_rotor_sets_rigid_rotation_momentum(axis):
    """A rotor box larger than the domain (solid fraction 1 everywhere): each
    cell leaves the collision with momentum rho * u_target, u_target =
    (Omega dt / res) e_axis x (r - r0)."""
    case, dom, _ = cases.periodic_box(16)
    o = OracleSim(case, dom, [], double=True)
    o.build_rotors([geometry.box((-30, -30, -30), (30, 30, 30))])
    omega = 2.0e-3 / (o.info().dtPhys / o.info().res)       # lattice scale 2e-3 per mm
    origin = (0.7, -1.1, 0.4)
    o.set_rotor(0, 0, omega, axis, origin)
    o.init()
    P, pos, m0 = _one_step_post_momentum(o)
    info = o.info()
    scale = np.float64(np.float32(omega) * np.float32(info.dtPhys) / np.float32(info.res))
    r = pos - np.array(origin, np.float32).astype(np.float64)
    e = np.zeros(3)
    e[axis] = 1.0
    ut = scale * np.cross(e, r)
    rho = 1.0 + m0[:, 0]
    assert np.max(np.abs(P - rho[:, None] * ut)) < 1e-13


def test_partly_solid_cells_move_by_the_smoothstep_share():
    """Box rotor inside the domain, no rotation, uniform start u:
    post-collision momentum = rho u + s^2 (3 - 2 s) rho (0 - u), s the solid
    fraction at the cell."""
    case, dom, _ = cases.periodic_box(16, u=(0.03, -0.02, 0.01))
    o = OracleSim(case, dom, [], double=True)
    o.build_rotors([geometry.box(LO, HI)])
    o.set_rotor(0, 0, 0.0, T.AXIS_NONE, (0.0, 0.0, 0.0))
    o.init()
    P, pos, m0 = _one_step_post_momentum(o)
    o32 = _box_rotor(double=False, n=16)
    s = np.array([o32.rotor_fraction(0, 0, 0.0, *p) for p in pos.astype(np.float32)], np.float64)
    ff = s * s * (3 - 2 * s)
    rho = 1.0 + m0[:, 0]
    want = rho[:, None] * m0[:, 1:] * (1 - ff)[:, None]
    assert (s > 0).sum() > 50 and ((s > 0) & (s < 1)).sum() > 20
    assert np.max(np.abs(P - want)) < 1e-6


def test_angle_advances_by_dt_omega_modulo_two_pi():
    case, dom, _ = cases.periodic_box(8)
    o = OracleSim(case, dom, [], double=False)
    o.build_rotors([geometry.box((-1, -1, -1), (1, 1, 1))])
    omega = np.float32(2.5) / np.float32(o.info().dtPhys)
    o.set_rotor(0, 0, omega, T.AXIS_Z)
    o.init()
    o.run(7)
    want = np.fmod(7 * np.float64(np.float32(o.info().dtPhys)) * np.float64(omega), 2 * np.pi)
    assert abs(o.rotor_info(0, 0)[10] - want) < 1e-12


def test_torque_of_the_force_cells_matches_the_cell_sum():
    """From rest, with track: every fully solid cell gets g = u_target; the
    force cells' torque about the axis approaches sum over cells of
    (r x g)_z = scale * sum r_perp^2 (lever arms at force-cell centres)."""
    case, dom, _ = cases.periodic_box(16)
    case.trackRotorForce = 1
    o = OracleSim(case, dom, [], double=True)
    o.build_rotors([geometry.box((-5.1, -5.1, -5.1), (5.1, 5.1, 5.1))])
    omega = 1.0e-3 / (o.info().dtPhys / o.info().res)
    o.set_rotor(0, 0, omega, T.AXIS_Z)
    o.init()
    o.run(1)
    F = o.rotor_forces(0, 0)
    info = o.info()
    n1 = info.cellCountX
    k, j, i = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    x, y, z = i * info.res + info.ox, j * info.res + info.oy, k * info.res + info.oz
    angle = o.rotor_info(0, 0)[10]               # the step's rotor angle
    o32 = _box_rotor(double=False, lo=(-5.1, -5.1, -5.1), hi=(5.1, 5.1, 5.1), n=16)
    o32.set_rotor(0, 0, 0.0, T.AXIS_Z)
    s = np.array([o32.rotor_fraction(0, 0, angle, a, b, c) for a, b, c in zip(x.ravel(), y.ravel(), z.ravel())])
    ff = s * s * (3 - 2 * s)
    ff[o.array("wallMap") == -2] = 0.0          # periodic ghost layers do not track force (updateGrid.h:57)
    scale = np.float64(np.float32(omega) * np.float32(info.dtPhys) / np.float32(info.res))
    gx, gy = -scale * y.ravel() * ff, scale * x.ravel() * ff
    tz = np.sum(x.ravel() * gy - y.ravel() * gx)
    # the force sums are the cell sums turned into the rotor frame (binning moves only the lever arms)
    c, sn = np.cos(angle), np.sin(angle)
    fx, fy = c * gx.sum() + sn * gy.sum(), -sn * gx.sum() + c * gy.sum()
    assert abs(F[0] - fx) < 1e-6 * np.abs(gx).sum() and abs(F[1] - fy) < 1e-6 * np.abs(gy).sum()
    assert abs(F[2]) < 1e-15
    assert abs(F[5] / tz - 1.0) < 0.02
