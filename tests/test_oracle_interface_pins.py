"""Pins of the oracle's coarse <-> fine interface transfer (updateInterface.h)
against what the mathematics fixes.  CPU only.

* the reconstruction of populations from (dRho, u, second moments) returns the
  equilibrium when the second moments are at equilibrium, and the stencil
  quantities computed back from those populations are the inputs;
* the interpolation coefficients reproduce every field of the interpolant's
  own family exactly (trilinear density, trilinear + axis-quadratic velocity)
  when fed that field's corner values and corner rates of strain;
* on multilevel grids a uniform flow is kept exactly, and the creeping
  Taylor-Green vortex on two levels decays at 6 nu k^2 like the uniform grid."""
import numpy as np
import pytest

from lbm3_amd import cases
from tests.oracle_lib import OracleSim, lib as orc
from tests.test_oracle_pins import _feq

CORNERS = np.array([[(n & 1) - 0.5, ((n >> 1) & 1) - 0.5, ((n >> 2) & 1) - 0.5] for n in range(8)])


def _reconstruct(dRho, u, k011, k101, k110, k200, k020, k002):
    inp = np.array([dRho, *u, k011, k101, k110, k200, k020, k002], np.float64)
    f = np.zeros(27)
    orc().orc_if_reconstruct_f64(inp.ctypes.data, f.ctypes.data)
    return f


def _values(f, w1):
    f = np.ascontiguousarray(np.atleast_2d(f), np.float64)
    out = np.zeros((len(f), 9))
    orc().orc_if_values_batch_f64(f.ctypes.data, len(f), w1, out.ctypes.data)
    return out


def _coefficients(corner9x8):
    c = np.ascontiguousarray(corner9x8, np.float64)
    out = np.zeros(41)
    orc().orc_if_coefficients_f64(c.ctypes.data, out.ctypes.data)
    return out


def test_reconstruction_at_equilibrium_is_feq():
    rng = np.random.default_rng(1)
    for _ in range(50):
        d = rng.uniform(-0.03, 0.03)
        u = rng.uniform(-0.08, 0.08, 3)
        f = _reconstruct(d, u, 0.0, 0.0, 0.0, d / 3, d / 3, d / 3)
        assert np.max(np.abs(f - _feq(np.array([[d, *u]]))[0])) < 1e-16


def test_stencil_values_invert_the_reconstruction():
    """Second moments put in by the reconstruction come back as rates of strain
    -3 w1 k_ab / rho (off-diagonal) and -3/2 w1 (k_aa - k_bb) / rho."""
    rng = np.random.default_rng(2)
    w1 = 1.7
    for _ in range(50):
        d = rng.uniform(-0.03, 0.03)
        u = rng.uniform(-0.08, 0.08, 3)
        k = rng.uniform(-1e-3, 1e-3, 6)
        k011, k101, k110 = k[:3]
        k200, k020, k002 = d / 3 + k[3], d / 3 + k[4], d / 3 - k[3] - k[4] + k[5]
        f = _reconstruct(d, u, k011, k101, k110, k200, k020, k002)
        v = _values(f, w1)[0]
        rho = 1 + d
        assert abs(v[0] - d) < 1e-16
        assert np.max(np.abs(v[1:4] - u)) < 1e-16
        assert abs(v[4] - (-3 * w1 * k110 / rho)) < 1e-14
        assert abs(v[5] - (-3 * w1 * k011 / rho)) < 1e-14
        assert abs(v[6] - (-3 * w1 * k101 / rho)) < 1e-14
        assert abs(v[7] - (-1.5 * w1 * (k200 - k020) / rho)) < 1e-14
        assert abs(v[8] - (-1.5 * w1 * (k200 - k002) / rho)) < 1e-14


def _poly_field(rng):
    d = rng.standard_normal(8)          # 1 x y z xy xz yz xyz
    a = rng.standard_normal((3, 11))    # 1 x y z xy xz yz xyz x2 y2 z2
    return d, a


def _eval(d, a, p):
    x, y, z = p
    tri = np.array([1, x, y, z, x * y, x * z, y * z, x * y * z])
    full = np.concatenate([tri, [x * x, y * y, z * z]])
    grad = np.zeros((3, 3))
    for c in range(3):
        k = a[c]
        grad[c] = [k[1] + k[4] * y + k[5] * z + k[7] * y * z + 2 * k[8] * x,
                   k[2] + k[4] * x + k[6] * z + k[7] * x * z + 2 * k[9] * y,
                   k[3] + k[5] * x + k[6] * y + k[7] * x * y + 2 * k[10] * z]
    return d @ tri, a @ full, grad


def test_coefficients_reproduce_their_interpolant_family():
    rng = np.random.default_rng(3)
    for _ in range(20):
        d, a = _poly_field(rng)
        corner = np.zeros((8, 9))
        for n, p in enumerate(CORNERS):
            rho, u, g = _eval(d, a, p)
            corner[n] = [rho, *u, g[0, 1] + g[1, 0], g[1, 2] + g[2, 1], g[0, 2] + g[2, 0], g[0, 0] - g[1, 1],
                         g[0, 0] - g[2, 2]]
        co = _coefficients(corner)
        assert np.max(np.abs(co[:8] - d)) < 1e-12
        for c in range(3):
            assert np.max(np.abs(co[8 + 11 * c:19 + 11 * c] - a[c])) < 1e-12


@pytest.mark.parametrize("make,scheme", [("tgv2", 0), ("sphere3", 0), ("tgv2", 1), ("sphere3", 1)])
def test_uniform_flow_is_kept_across_levels(make, scheme):
    """A uniform flow stays exact on every level with the quadratic (scheme 0)
    and the linear interface (scheme 1, updateInterfaceLINEAR.h)."""
    if make == "tgv2":
        case, dom, tris = cases.taylor_green(16, u0=0.05, reynolds=100.0, levels=2)
        case.init.kind = 0
        case.init.ux, case.init.uy, case.init.uz = 0.03, -0.01, 0.02
        u = (0.03, -0.01, 0.02)
    else:
        case, dom, tris = cases.multilevel_sphere(48, 24, 24, level1_box=(8, 6, 6, 40, 18, 18),
                                                  level2_box=(12, 8, 8, 24, 16, 16), center=(18.0, 12.0, 12.0),
                                                  diameter_fine=8.0, segments=(12, 24))
        tris = []                 # the sphere would break uniformity; keep the grid and boundaries
        u = (case.init.ux, 0.0, 0.0)
    case.interfaceScheme = scheme
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    o.run(6)
    for L in range(o.levels):
        m = o.macros(L)
        assert np.max(np.abs(m[:, 0])) < 1e-13, L
        assert np.max(np.abs(m[:, 1:] - np.array(u, np.float32).astype(np.float64))) < 1e-13, L


def test_two_level_creeping_taylor_green_decay_rate():
    """Creeping TGV refined in the central cube: the energy of both levels
    (fine cells weighted 1/8) decays at 6 nu k^2 after the start-up transient."""
    n = 24
    case, dom, _ = cases.taylor_green(n, u0=1e-4, reynolds=0.2, levels=2)
    o = OracleSim(case, dom, [], double=True)
    o.init()
    nu = o.info(0).nu
    o.run(200)
    e0 = o.kinetic_energy()
    o.run(200)
    e1 = o.kinetic_energy()
    k = 2 * np.pi / n
    rate = -np.log(e1 / e0) / 200
    assert abs(rate / (6 * nu * k * k) - 1.0) < 0.02

