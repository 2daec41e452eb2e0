"""Pins of the CPU oracle against what the mathematics and the geometry fix.

The reference ships no tests and no golden vectors for the hot path, and it
cannot be built here (TNL and nvcc are absent), so the oracle is checked
against closed forms, invariants, symmetries, brute-force restatements of the
streaming and analytic geometry.  Each check is chosen so that a plausible
slip in the oracle (a wrong weight, a transposed axis, a dropped force term, a
wrong esotwist slot, an off-by-one in the grid origin) fails it.  CPU only."""
import ctypes as C
import itertools

import numpy as np
import pytest

from lbm3_amd import cases, geometry
from lbm3_amd import ctypes_defs as T
from tests.oracle_lib import OracleSim, lib as orc

NAMES = ["OOO", "POO", "MOO", "OOM", "OOP", "OMO", "OPO", "POM", "MOP", "POP", "MOM", "MMO", "PPO", "OPM", "OMP",
         "MPO", "PMO", "OPP", "OMM", "MPM", "PMP", "MMP", "PPM", "PMM", "MPP", "MMM", "PPP"]
SIGN = {"P": 1, "O": 0, "M": -1}


def _dirs():
    L = orc()
    cx = np.array((C.c_int * 27).in_dll(L, "ORC_CX")[:])
    cy = np.array((C.c_int * 27).in_dll(L, "ORC_CY")[:])
    cz = np.array((C.c_int * 27).in_dll(L, "ORC_CZ")[:])
    inv = np.array((C.c_int * 27).in_dll(L, "ORC_INV")[:])
    return cx, cy, cz, inv


def _weights(cx, cy, cz):
    n = np.abs(cx) + np.abs(cy) + np.abs(cz)
    return np.array([8 / 27, 2 / 27, 1 / 54, 1 / 216])[n]


# ---------------------------------------------------------------- A1: lattice
def test_direction_table_matches_the_reference_names():
    """D3Q27Directions.h:8-34 names each direction by its (x, y, z) signs."""
    cx, cy, cz, inv = _dirs()
    for q, name in enumerate(NAMES):
        assert (cx[q], cy[q], cz[q]) == tuple(SIGN[ch] for ch in name), name
    # inverse pairs are (2k-1, 2k) and really are opposite vectors
    for q in range(1, 27):
        assert inv[q] == (q + 1 if q % 2 else q - 1)
        assert (cx[inv[q]], cy[inv[q]], cz[inv[q]]) == (-cx[q], -cy[q], -cz[q])
    assert inv[0] == 0


# ---------------------------------------------------------------- A5: equilibrium and moments
def _feq(m4):
    m4 = np.ascontiguousarray(m4, np.float64)
    out = np.zeros((len(m4), 27))
    orc().orc_feq_batch_f64(m4.ctypes.data, len(m4), out.ctypes.data)
    return out


def _moments(f):
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros((len(f), 4))
    orc().orc_moments_batch_f64(f.ctypes.data, len(f), out.ctypes.data)
    return out


def test_equilibrium_moments_closed_form():
    """Sum of the shifted feq = dRho; first moment = rho u; second moment of the
    unshifted feq = rho (delta/3 + u u) (factorised D3Q27 equilibrium)."""
    cx, cy, cz, _ = _dirs()
    w = _weights(cx, cy, cz)
    rng = np.random.default_rng(1)
    m4 = np.column_stack([rng.uniform(-0.05, 0.05, 200), rng.uniform(-0.1, 0.1, (200, 3))])
    fb = _feq(m4)
    f = fb + w
    rho = 1.0 + m4[:, 0]
    c = np.stack([cx, cy, cz], 1).astype(np.float64)
    assert np.allclose(fb.sum(1), m4[:, 0], atol=1e-15, rtol=0)
    assert np.allclose(f @ c, rho[:, None] * m4[:, 1:], atol=1e-15, rtol=0)
    for a, b in itertools.product(range(3), range(3)):
        second = f @ (c[:, a] * c[:, b])
        expect = rho * ((a == b) / 3.0 + m4[:, 1 + a] * m4[:, 1 + b])
        assert np.allclose(second, expect, atol=1e-15, rtol=0)
    # getDRhoUxUyUz inverts it
    assert np.allclose(_moments(fb), m4, atol=1e-15, rtol=0)


# ---------------------------------------------------------------- A6: collision
def _collide(f, nu, **bc):
    f = np.ascontiguousarray(f, np.float64).copy()
    orc().orc_collide_batch_f64(f.ctypes.data, len(f), T.default_bc(**bc), nu)
    return f


def _random_state(n, seed, amp=0.01):
    cx, cy, cz, _ = _dirs()
    w = _weights(cx, cy, cz)
    rng = np.random.default_rng(seed)
    m4 = np.column_stack([rng.uniform(-0.03, 0.03, n), rng.uniform(-0.08, 0.08, (n, 3))])
    return _feq(m4) + amp * w * rng.standard_normal((n, 27))


MODELS = [dict(collisionLimiter=0.0), dict(collisionLimiter=0.05), dict(collisionLimiter=1e-3, nuMultiplier=0.7)]


@pytest.mark.parametrize("bc", MODELS)
def test_collision_conserves_mass_and_adds_the_force(bc):
    cx, cy, cz, _ = _dirs()
    c = np.stack([cx, cy, cz], 1).astype(np.float64)
    f = _random_state(300, 2)
    g = np.array([2e-4, -1e-4, 5e-5], np.float32).astype(np.float64)   # BCStruct holds floats
    post = _collide(f, 0.01, gx=g[0], gy=g[1], gz=g[2], **bc)
    assert np.allclose(post.sum(1), f.sum(1), atol=1e-15, rtol=0)
    assert np.allclose(post @ c, f @ c + g, atol=2e-15, rtol=0)


@pytest.mark.parametrize("bc", MODELS)
@pytest.mark.parametrize("nu", [3.2e-4, 0.02, 0.1])
def test_equilibrium_is_a_fixed_point(bc, nu):
    rng = np.random.default_rng(3)
    m4 = np.column_stack([rng.uniform(-0.03, 0.03, 100), rng.uniform(-0.08, 0.08, (100, 3))])
    feq = _feq(m4)
    assert np.max(np.abs(_collide(feq, nu, **bc) - feq)) < 1e-16 * 50


def test_k15_at_omega_one_returns_the_equilibrium():
    """nu = 1/6 gives omega1 = omega2 = 1: at rest (u = 0, where the Galilean
    correction of the second-order relaxation vanishes) K15 relaxes every moment
    to its equilibrium and returns feq(dRho, 0)."""
    cx, cy, cz, _ = _dirs()
    c = np.stack([cx, cy, cz], 1).astype(np.float64)
    w = _weights(cx, cy, cz)
    f = _random_state(100, 4)
    j = f @ c
    f = f - 3.0 * w * (j @ c.T)            # remove the momentum: sum w c c = I/3
    m = _moments(f)
    assert np.max(np.abs(m[:, 1:])) < 1e-16
    post = _collide(f, 1.0 / 6.0, collisionLimiter=0.0)
    assert np.max(np.abs(post - _feq(m))) < 1e-15


def test_galilean_correction_at_omega_one():
    """Moving fluid, omega1 = 1: the result differs from feq only by the
    Galilean correction, which is of order u^2 times the non-equilibrium part
    and vanishes with it."""
    f = _random_state(100, 4, amp=1e-3)
    eq = _feq(_moments(f))
    d_small = np.max(np.abs(_collide(f, 1.0 / 6.0, collisionLimiter=0.0) - eq))
    f2 = eq + 10.0 * (f - eq)
    d_big = np.max(np.abs(_collide(f2, 1.0 / 6.0, collisionLimiter=0.0) - _feq(_moments(f2))))
    assert d_small < 1e-5
    assert 5.0 < d_big / d_small < 15.0


def _permute(f, perm, flips):
    """Populations of the lattice image under x_a -> flips[a] * x_perm[a]."""
    cx, cy, cz, _ = _dirs()
    c = np.stack([cx, cy, cz], 1)
    index = {tuple(v): q for q, v in enumerate(c)}
    target = [index[tuple(flips[a] * c[q][perm[a]] for a in range(3))] for q in range(27)]
    out = np.empty_like(f)
    out[:, target] = f
    return out


@pytest.mark.parametrize("bc", MODELS)
@pytest.mark.parametrize("perm,flips", [((1, 0, 2), (1, 1, 1)), ((2, 0, 1), (1, 1, 1)), ((0, 1, 2), (-1, 1, 1)),
                                        ((0, 2, 1), (1, -1, -1))])
def test_collision_is_cubic_symmetric(bc, perm, flips):
    """collide(R f) = R collide(f) for every lattice rotation/reflection R: catches a
    transposed axis or a wrong index in the moment chains."""
    f = _random_state(200, 5, amp=0.02)
    a = _collide(_permute(f, perm, flips), 0.01, **bc)
    b = _permute(_collide(f, 0.01, **bc), perm, flips)
    assert np.max(np.abs(a - b)) < 1e-15


def test_collision_relaxes_shear_stress_at_omega1():
    """The deviatoric second moment decays by (1 - omega1) in one collision
    (omega1 = 1 / (3 nu + 1/2)); K15 and K17 agree on this order."""
    cx, cy, cz, _ = _dirs()
    c = np.stack([cx, cy, cz], 1).astype(np.float64)
    w = _weights(cx, cy, cz)
    nu = 0.03
    w1 = 1.0 / (3.0 * nu + 0.5)
    m4 = np.array([[0.0, 0.0, 0.0, 0.0]])
    feq = _feq(m4)
    noneq = np.zeros(27)
    noneq[c[:, 0] * c[:, 1] == 1] = 1.0
    noneq[c[:, 0] * c[:, 1] == -1] = -1.0
    noneq *= w * 1e-3 * (np.abs(c[:, 2]) == 0)
    f = feq + noneq
    for bc in MODELS[:2]:
        post = _collide(f, nu, **bc)
        pre_xy = (f + w) @ (c[:, 0] * c[:, 1])
        post_xy = (post + w) @ (c[:, 0] * c[:, 1])
        assert abs(post_xy - (1 - w1) * pre_xy) < 1e-15


# ---------------------------------------------------------------- A3, A4, A7, A10: streaming on a periodic box
def _stored_ijk(o, level=0):
    """(i, j, k) of every stored cell from the compressed row format."""
    sh = o.array("shifter", level)
    ri, rj, rk = o.array("rowI", level), o.array("rowJ", level), o.array("rowK", level)
    n = len(sh)
    head = np.where(sh >= 0, np.arange(n), np.arange(n) + sh)
    row = sh[head]
    back = np.arange(n) - head
    return np.stack([ri[row] + back, rj[row], rk[row]], 1)


@pytest.mark.parametrize("steps", [1, 2, 3])
def test_esotwist_steps_equal_collide_then_pull_stream(steps):
    """TGV 16^3 (+1 ghost layer per face) in FP64: each oracle step equals a
    brute-force step on the logical populations -- collide every stored cell,
    then f_q(x) = f*_q(x - c_q) on the 18^3 box, then the ghost layers take the
    pre-collision populations of the opposite interior planes."""
    case, dom, _ = cases.taylor_green(16, u0=0.05, reynolds=100.0)
    o = OracleSim(case, dom, [], double=True)
    o.init()
    info = o.info()
    n = info.cellCountX
    assert info.cellCount == n ** 3
    ijk = _stored_ijk(o)
    # the stored cells are the whole box, cell order k, j, i
    assert np.array_equal(ijk[:, 0] + n * (ijk[:, 1] + n * ijk[:, 2]), np.arange(n ** 3))
    cx, cy, cz, _ = _dirs()
    f = _feq(o.macros()).reshape(n, n, n, 27)          # [k][j][i][q]
    nu = info.nu
    for _ in range(steps):
        post = _collide(f.reshape(-1, 27), nu).reshape(n, n, n, 27)
        for q in range(27):
            f[..., q] = np.roll(post[..., q], shift=(cz[q], cy[q], cx[q]), axis=(0, 1, 2))
        # periodic ghosts: index 0 <- n-2, n-1 <- 1 along each axis, edges and corners included
        for ax in range(3):
            sl_dst0 = [slice(None)] * 3
            sl_src0 = [slice(None)] * 3
            sl_dst0[ax], sl_src0[ax] = 0, n - 2
            f[tuple(sl_dst0)] = f[tuple(sl_src0)]
            sl_dst1 = [slice(None)] * 3
            sl_src1 = [slice(None)] * 3
            sl_dst1[ax], sl_src1[ax] = n - 1, 1
            f[tuple(sl_dst1)] = f[tuple(sl_src1)]
    o.run(steps)
    got = o.macros()
    want = _moments(f.reshape(-1, 27))
    assert np.max(np.abs(got - want)) < 1e-15


# ---------------------------------------------------------------- A7/A9: physics
def test_plane_poiseuille_profile():
    """Body force g between moment-based no-slip walls at the first and last node
    rows, steady state: zero velocity on the wall nodes, a symmetric profile,
    and the momentum balance nu u'' = -g in the bulk of the channel (second
    differences).  Near the walls the scheme is not the parabola: the
    moment-based wall sets the wall-normal shear moment to its equilibrium, and
    where the periodic ghost layers meet the wall faces the cells take the
    wall condition with an edge normal; both perturb a layer of a few nodes."""
    nu, g, ny = 0.1, 1e-5, 24
    case, dom, _ = cases.channel(4, ny, 4, force=g, nu=nu, periodic_z=True)
    o = OracleSim(case, dom, [], double=True)
    o.init()
    o.run(8000)
    nu32 = o.info().nu
    ijk = _stored_ijk(o)
    m = o.macros()
    H = ny - 1
    sel = (ijk[:, 0] == 2) & (ijk[:, 2] == 2)
    y = ijk[sel, 1]
    u = m[sel, 1][np.argsort(y)]
    assert abs(u[0]) < 1e-15 and abs(u[H]) < 1e-15
    assert np.max(np.abs(u - u[::-1])) < 1e-12
    second = u[:-2] - 2 * u[1:-1] + u[2:]          # centred on y = 1 .. H-1
    g32 = float(np.float32(g))
    err = np.abs(second + g32 / nu32) / (g32 / nu32)
    assert np.max(err[7:H - 8]) < 2e-3            # y = 8 .. H-8: the bulk of the channel
    assert np.max(u) > 0.9 * g32 * H * H / (8 * nu32)


def test_uniform_flow_is_kept_by_inlet_outlet_and_far_field():
    """The sphere case without the sphere, started at its inlet state: velocity
    inlet, non-reflective pressure outlet and far-field faces leave the exact
    uniform solution unchanged."""
    case, dom, _ = cases.sphere(24, 12, 12, diameter=4.0, center=(8.0, 6.0, 6.0))
    o = OracleSim(case, dom, [], double=True)
    o.init()
    o.run(50)
    m = o.macros()
    u0 = case.init.ux
    assert np.max(np.abs(m[:, 0])) < 1e-13
    assert np.max(np.abs(m[:, 1] - u0)) < 1e-13
    assert np.max(np.abs(m[:, 2:])) < 1e-13


def test_taylor_green_viscous_energy_decay_rate():
    """Creeping TGV (Re = 0.2, so the nonlinear transfer time 1/(u0 k) is far
    beyond the run): once the start-up transient of the equilibrium
    initialisation has died out (|1 - omega1|^t, omega1 ~ 1.98), the energy
    decays as exp(-6 nu k^2 t), k = 2 pi / n."""
    n = 24
    case, dom, _ = cases.taylor_green(n, u0=1e-4, reynolds=0.2)
    o = OracleSim(case, dom, [], double=True)
    o.init()
    nu = o.info().nu
    o.run(400)
    e0 = o.kinetic_energy()
    o.run(400)
    e1 = o.kinetic_energy()
    k = 2 * np.pi / n
    rate = -np.log(e1 / e0) / 400
    assert abs(rate / (6 * nu * k * k) - 1.0) < 0.01


# ---------------------------------------------------------------- A16-A19: voxelizer and builder
def _box_case(lo, hi, n=16):
    case, _, _ = cases.channel(n, n, n, force=0.0)
    case.nOpenRules = 0
    cases.add_open_rule(case, dirichletU=1, openBCID=0, nonReflective=0)
    dom = cases.bounds(0, 0, 0, n, n, n)
    return case, dom, [geometry.box(lo, hi)]


def test_box_voxelization_and_link_codes_known_answer():
    """Axis-aligned box with faces 0.3 cell spacings from the cell centres:
    stored wall cells are exactly the stored centres inside the box; every
    axis link that crosses a face has q = 0.3, code int(0.3 * 254 + 0.5) = 76;
    link bits are exactly the directions whose neighbour centre is inside."""
    # find the cell-centre lattice first
    case, dom, tris = _box_case((5.3, 5.3, 5.3), (10.7, 10.7, 10.7))
    probe = OracleSim(case, dom, tris)
    info = probe.info()
    ox, oy, oz, res = info.ox, info.oy, info.oz, info.res
    lo = np.array([ox, oy, oz]) + res * 5.3
    hi = np.array([ox, oy, oz]) + res * 10.7
    case, dom, tris = _box_case(tuple(lo), tuple(hi))
    o = OracleSim(case, dom, tris)
    ijk = _stored_ijk(o)
    centre = np.array([ox, oy, oz]) + res * ijk
    inside = np.all((centre > lo) & (centre < hi), axis=1)
    wm = o.array("wallMap")
    assert np.array_equal(wm == -3, inside)
    # deep interior deleted, one wall halo layer kept: stored inside cells are 6 and 10 planes
    assert inside.sum() == 5 ** 3 - 3 ** 3
    cx, cy, cz, _ = _dirs()
    lookup = {tuple(v): c for c, v in enumerate(ijk)}
    wi, wd, ll = o.array("wallIndex"), o.array("wallData"), o.array("linkLength").reshape(7, -1)
    W = len(wi)
    assert W > 0
    for w in range(W):
        c = wi[w]
        assert wm[c] == w and not inside[c]
        links = [d for d in range(1, 27) if (wd[w] >> d) & 1]
        want = [d for d in range(1, 27)
                if (lambda nb: nb is not None and inside[nb])(lookup.get((ijk[c, 0] + cx[d], ijk[c, 1] + cy[d],
                                                                           ijk[c, 2] + cz[d])))]
        assert links == want
        codes = [(ll[k // 4, w] >> (8 * (k % 4))) & 255 for k in range(len(links))]
        for d, code in zip(links, codes):
            assert 1 <= code <= 254
            if abs(cx[d]) + abs(cy[d]) + abs(cz[d]) == 1:
                assert code == 76, (w, d, code)
    # wall id bits: the single body is wall 0
    assert np.all((wd >> 27) == 0)


def test_sphere_voxelization_against_the_analytic_ball():
    """Centres well inside the inscribed polyhedron are wall, centres outside
    the circumscribed ball are not; the UV sphere lies between the two."""
    R, ctr = 6.0, np.array([12.3, 12.1, 11.8])
    nlat, nlon = 24, 48
    tri = geometry.uv_sphere(R, tuple(ctr), nlat, nlon)
    case, _, _ = cases.channel(24, 24, 24, force=0.0)
    case.nOpenRules = 0
    cases.add_open_rule(case, dirichletU=1, openBCID=0, nonReflective=0)
    o = OracleSim(case, cases.bounds(0, 0, 0, 24, 24, 24), [tri])
    info = o.info()
    ijk = _stored_ijk(o)
    centre = np.array([info.ox, info.oy, info.oz]) + info.res * ijk
    r = np.linalg.norm(centre - ctr, axis=1)
    wall = o.array("wallMap") == -3
    r_in = R * np.cos(np.pi / nlat) * np.cos(np.pi / nlon)   # inscribed ball of the faceted sphere
    assert np.all(wall[r < r_in - 1e-3])
    assert not np.any(wall[r > R + 1e-3])
    # every fluid cell is stored; deleted cells are deep inside
    assert info.cellCount == len(ijk)
