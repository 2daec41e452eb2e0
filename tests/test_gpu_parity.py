"""GPU (HIP, gfx950) vs CPU oracle parity of the per-step hot path.

Bars (per the hot-path contract): FP32 macroscopic fields rho, u within 1e-5
relative max-norm after 100 steps; the FP64 build within 1e-10.  Grids are built
by the oracle and imported, so these tests isolate the per-step kernels;
builder parity is covered separately."""
import numpy as np
import pytest

from lbm3_amd import cases
from lbm3_amd import ctypes_defs as T
from lbm3_amd import lib
from tests.helpers import level_arrays, rel_max
from tests.oracle_lib import OracleSim, lib as orc

pytestmark = pytest.mark.gpu


def _random_cells(n, seed, dtype):
    rng = np.random.default_rng(seed)
    m4 = np.zeros((n, 4), dtype)
    m4[:, 0] = rng.uniform(-0.02, 0.02, n)
    m4[:, 1:] = rng.uniform(-0.06, 0.06, (n, 3))
    feq = np.zeros((n, 27), dtype)
    fn = orc().orc_feq_batch_f32 if dtype == np.float32 else orc().orc_feq_batch_f64
    fn(m4.ctypes.data, n, feq.ctypes.data)
    w = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
    return np.ascontiguousarray(feq + (0.01 * w * rng.standard_normal((n, 27))).astype(dtype))


BCS = [dict(), dict(collisionLimiter=0.0), dict(collisionLimiter=0.1, gx=1e-4, gy=-2e-5, gz=3e-5),
       dict(collisionLimiter=0.05, nuMultiplier=2.0)]


@pytest.mark.parametrize("bcargs", BCS)
@pytest.mark.parametrize("nu", [3.2e-4, 0.02, 0.16])
def test_collision_f32_matches_oracle(bcargs, nu):
    f = _random_cells(4096, 7, np.float32)
    bc = T.default_bc(**bcargs)
    ref = f.copy()
    orc().orc_collide_batch_f32(ref.ctypes.data, len(ref), bc, nu)
    got = lib.collide_cells(f.copy(), bc, nu)
    # one collision: a few float ulps of the population scale
    assert rel_max(got, ref) < 2e-6


@pytest.mark.parametrize("bcargs", BCS)
def test_collision_f64_matches_oracle(bcargs):
    f = _random_cells(2048, 11, np.float64)
    bc = T.default_bc(**bcargs)
    ref = f.copy()
    orc().orc_collide_batch_f64(ref.ctypes.data, len(ref), bc, 0.01)
    got = lib.collide_cells(f.copy(), bc, 0.01)
    assert rel_max(got, ref) < 1e-13


def _pair(case, domain, tris, storage):
    o = OracleSim(case, domain, tris, double=(storage == "f64"))
    g = lib.Sim(case, storage)
    for L in range(o.levels):
        g.import_level(L, level_arrays(o, L), o.info(L))
    o.init()
    g.init()
    if storage == "f64":
        # the initial-condition hooks evaluate sin/cos in float, whose last bit differs
        # between libm and the device; start the FP64 comparison from identical populations
        for L in range(o.levels):
            g.set_populations(L, o.array("f", L).reshape(27, -1))
    return o, g


def _compare(o, g, level=0):
    mo, mg = o.macros(level), g.macros(level)
    rho_err = rel_max(1.0 + mg[:, 0], 1.0 + mo[:, 0])
    u_err = rel_max(mg[:, 1:], mo[:, 1:])
    return rho_err, u_err


@pytest.mark.parametrize("n", [16, 32])
def test_taylor_green_100_steps_f32(n):
    case, dom, tris = cases.taylor_green(n, u0=0.05, reynolds=1600.0)
    o, g = _pair(case, dom, tris, "f32")
    r0, u0 = _compare(o, g)
    assert r0 < 1e-6 and u0 < 1e-6, (r0, u0)
    o.run(100)
    g.run(100)
    g.sync()
    r, u = _compare(o, g)
    print("TGV %d^3 after 100 steps: rel max rho %.2e, u %.2e" % (n, r, u))
    assert r < 1e-5 and u < 1e-5, (r, u)


def test_taylor_green_100_steps_f64():
    case, dom, tris = cases.taylor_green(16, u0=0.05, reynolds=1600.0)
    o, g = _pair(case, dom, tris, "f64")
    o.run(100)
    g.run(100)
    g.sync()
    r, u = _compare(o, g)
    assert r < 1e-10 and u < 1e-10, (r, u)


def _small_sphere(**kw):
    return cases.sphere(64, 32, 32, diameter=8.0, center=(16.0, 16.0, 16.0), segments=(24, 48), **kw)


def test_sphere_100_steps_f32():
    """Reference sphere-case collision (K15 where the open boundaries are).  The
    oracle's own FP32 run sits ~8e-6 from its FP64 run after 100 steps here, so
    two FP32 implementations with different (equally valid) rounding can be
    ~1e-5 apart: the bar is 1e-5 or twice that sensitivity, and the GPU FP32 run
    must be as close to the FP64 solution as the oracle's FP32 run is."""
    case, dom, tris = _small_sphere()
    o, g = _pair(case, dom, tris, "f32")
    o64 = OracleSim(case, dom, tris, double=True)
    o64.init()
    for x in (o, g, o64):
        x.run(100)
    g.sync()
    r, u = _compare(o, g)
    m64 = o64.macros()
    sens = rel_max(o.macros()[:, 1:], m64[:, 1:])
    err64 = rel_max(g.macros()[:, 1:], m64[:, 1:])
    print("sphere 64x32x32 after 100 steps: GPU vs oracle FP32 rho %.2e u %.2e; FP32 vs FP64: oracle %.2e, GPU %.2e"
          % (r, u, sens, err64))
    assert r < 1e-5 and u < max(1e-5, 2 * sens), (r, u, sens)
    assert err64 < 1.5 * sens + 1e-6, (err64, sens)


def test_sphere_k17_100_steps_f32_within_fp32_sensitivity():
    """K17 next to the impulsively started sphere and the outlet corners is
    ill-conditioned in FP32: the oracle's own FP32 and FP64 runs differ by
    ~1e-4 after 100 steps.  The GPU FP32 run must stay within a small multiple
    of that sensitivity; exact agreement is checked by the FP64 test below."""
    case, dom, tris = _small_sphere(k15_on_level0=False)
    o, g = _pair(case, dom, tris, "f32")
    o64 = OracleSim(case, dom, tris, double=True)
    o64.init()
    for x in (o, g, o64):
        x.run(100)
    g.sync()
    r, u = _compare(o, g)
    m64, mo = o64.macros(), o.macros()
    sens = rel_max(mo[:, 1:], m64[:, 1:])
    err64 = rel_max(g.macros()[:, 1:], m64[:, 1:])
    print("K17 sphere: GPU vs oracle FP32 %.2e; FP32 vs FP64: oracle %.2e, GPU %.2e" % (u, sens, err64))
    assert r < 1e-5 and u < 3 * sens + 1e-6, (r, u, sens)
    assert err64 < 1.5 * sens + 1e-6, (err64, sens)


@pytest.mark.parametrize("k15", [True, False])
def test_sphere_100_steps_f64(k15):
    case, dom, tris = _small_sphere(k15_on_level0=k15)
    o, g = _pair(case, dom, tris, "f64")
    o.run(100)
    g.run(100)
    g.sync()
    r, u = _compare(o, g)
    assert r < 1e-10 and u < 1e-10, (r, u)


def test_channel_periodic_and_moment_walls_f64():
    """Body-force channel (periodic in x, moment-based no-slip walls): exact
    arithmetic parity in FP64."""
    case, dom, tris = cases.channel(32, 16, 16, force=1e-4)
    o, g = _pair(case, dom, tris, "f64")
    o.run(100)
    g.run(100)
    g.sync()
    r, u = _compare(o, g)
    assert r < 1e-10 and u < 1e-10, (r, u)


def test_channel_periodic_and_moment_walls_f32_within_fp32_sensitivity():
    """The slow start-up of the channel sits at the FP32 floor of the shifted
    populations: the oracle's own FP32 and FP64 runs differ by ~3e-5 relative
    (an absolute ~3e-7 in u) after 100 steps, so the 1e-5 bar is replaced by a
    small multiple of that sensitivity; exact agreement is the FP64 test."""
    case, dom, tris = cases.channel(32, 16, 16, force=1e-4)
    o, g = _pair(case, dom, tris, "f32")
    o64 = OracleSim(case, dom, tris, double=True)
    o64.init()
    for x in (o, g, o64):
        x.run(100)
    g.sync()
    r, u = _compare(o, g)
    m64 = o64.macros()
    sens = rel_max(o.macros()[:, 1:], m64[:, 1:])
    err64 = rel_max(g.macros()[:, 1:], m64[:, 1:])
    print("channel: GPU vs oracle FP32 %.2e; FP32 vs FP64: oracle %.2e, GPU %.2e" % (u, sens, err64))
    assert r < 1e-5 and u < 3 * sens + 1e-6, (r, u, sens)
    assert err64 < 1.5 * sens + 1e-6, (err64, sens)


# ---------------------------------------------------------------- multilevel (interface transfer)
def _compare_all(o, g):
    r = u = 0.0
    for L in range(o.levels):
        rl, ul = _compare(o, g, L)
        r, u = max(r, rl), max(u, ul)
    return r, u


def _cut_sphere():
    case, dom, tris = cases.sphere(48, 24, 24, diameter=7.0, center=(16.0, 12.0, 12.0), segments=(16, 32))
    case.gridLevelCount = 2
    cases.add_refine_rule(case, 0, (4, 4, 4, 16, 20, 20))
    return case, dom, tris


def _small_sphere3():
    return cases.multilevel_sphere(64, 32, 32, level1_box=(12, 8, 8, 36, 24, 24), level2_box=(14, 10, 10, 26, 22, 22),
                                   center=(20.0, 16.0, 16.0), diameter_fine=12.0, segments=(16, 32))


def _linear(make):
    """The same case with the linear interface (updateInterfaceLINEAR.h)."""
    def mk():
        case, dom, tris = make()
        case.interfaceScheme = 1
        return case, dom, tris
    return mk


MULTILEVEL = {"tgv2": lambda: cases.taylor_green(16, u0=0.05, reynolds=400.0, levels=2),
              "sphere3": _small_sphere3, "cut_sphere": _cut_sphere}
MULTILEVEL["tgv2_linear"] = _linear(MULTILEVEL["tgv2"])
MULTILEVEL["cut_sphere_linear"] = _linear(_cut_sphere)


@pytest.mark.parametrize("name", sorted(MULTILEVEL))
def test_multilevel_f64(name):
    """FP64: coarse/fine sub-cycling with the F->C, C->F and leftover transfers,
    20 coarse steps, within 1e-10 of the oracle on every level."""
    case, dom, tris = MULTILEVEL[name]()
    o, g = _pair(case, dom, tris, "f64")
    assert o.levels > 1
    o.run(20)
    g.run(20)
    g.sync()
    r, u = _compare_all(o, g)
    print("%s FP64 after 20 coarse steps: rel max rho %.2e, u %.2e" % (name, r, u))
    assert r < 1e-10 and u < 1e-10, (r, u)


@pytest.mark.parametrize("name", sorted(MULTILEVEL))
def test_multilevel_f32(name):
    """FP32, 50 coarse steps: the 1e-5 bar, or 3x the oracle's own FP32-vs-FP64
    difference where that is larger (the K17 levels next to the sphere)."""
    case, dom, tris = MULTILEVEL[name]()
    o, g = _pair(case, dom, tris, "f32")
    o64 = OracleSim(case, dom, tris, double=True)
    o64.init()
    for x in (o, g, o64):
        x.run(50)
    g.sync()
    r, u = _compare_all(o, g)
    sens = max(rel_max(o.macros(L)[:, 1:], o64.macros(L)[:, 1:]) for L in range(o.levels))
    err64 = max(rel_max(g.macros(L)[:, 1:], o64.macros(L)[:, 1:]) for L in range(o.levels))
    print("%s FP32 after 50 coarse steps: rel max rho %.2e, u %.2e; FP32 vs FP64: oracle %.2e, GPU %.2e"
          % (name, r, u, sens, err64))
    assert r < 1e-5 and u < max(1e-5, 3 * sens + 1e-6), (r, u, sens)
    assert err64 < 1.5 * sens + 1e-6, (err64, sens)


# ---------------------------------------------------------------- 16-bit storage accuracy
@pytest.mark.parametrize("storage,bar", [("f16", 1e-3), ("bf16", 1e-2), ("f16c", 1e-4)])
def test_16bit_storage_velocity_error_vs_fp64(storage, bar):
    """Shifted 16-bit populations (deviation from the free-stream equilibrium,
    in units of the weight): relative L2 velocity error against the FP64 run of
    the same kernels after 1000 steps of the sphere case.  f16c (11 mantissa
    bits) holds half of FP16's error."""
    case, dom, tris = _small_sphere()
    runs = {}
    for st in (storage, "f64"):
        g = lib.Sim(case, st)
        g.build(dom, tris)
        g.init()
        g.run(1000)
        g.sync()
        runs[st] = g.macros()[:, 1:].astype(np.float64)
        fluid = g.array("wallMap") != -3      # solid cells hold no flow
        g.close()
    a, b = runs[storage][fluid], runs["f64"][fluid]
    err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print("%s sphere 64x32x32, 1000 steps: rel L2 u error vs FP64 %.2e" % (storage, err))
    assert err < bar, err


@pytest.mark.parametrize("storage", ["f16", "f16c"])
def test_16bit_multilevel_velocity_error_vs_fp64(storage):
    """FP16 (and f16c) shifted storage on the three-level sphere: relative L2
    velocity error of every level against the FP64 run of the same kernels
    after 400 coarse steps (1600 finest sub-steps)."""
    case, dom, tris = _small_sphere3()
    runs = {}
    for st in (storage, "f64"):
        g = lib.Sim(case, st)
        g.build(dom, tris)
        g.init()
        g.run(400)
        g.sync()
        runs[st] = [(g.macros(L)[:, 1:].astype(np.float64), g.array("wallMap", L) != -3) for L in range(3)]
        g.close()
    errs = []
    for (u16, fluid), (u64, _) in zip(runs[storage], runs["f64"]):
        errs.append(float(np.linalg.norm((u16 - u64)[fluid]) / np.linalg.norm(u64[fluid])))
    print("%s three-level sphere, 400 coarse steps: rel L2 u vs FP64 per level %s" % (storage, ["%.2e" % e for e in errs]))
    assert max(errs) < (1e-3 if storage == "f16" else 5e-4), errs


def _f16c_value(bits):
    """The f16c format by its definition: sign bit 15, exponent bits 11-14
    (bias 15; 0: subnormal, 2^-14 * m / 2^11), mantissa bits 0-10."""
    b = bits.astype(np.int64)
    sign = np.where(b & 0x8000, -1.0, 1.0)
    e = (b >> 11) & 0xF
    m = (b & 0x7FF).astype(np.float64)
    mag = np.where(e == 0, np.ldexp(m / 2048.0, -14), np.ldexp(1.0 + m / 2048.0, e - 15))
    return sign * mag


def test_f16c_format_against_its_definition():
    """The initial Taylor-Green populations (reference velocity 0: the stored
    value is (f - w) / w) in f16c decode, by the format's definition, to the
    FP32 run's values rounded to nearest at 11 mantissa bits: within half a
    unit in the last place, and about half of them not exact (the rounding is
    live); every code lies in the format's range."""
    case, dom, tris = cases.taylor_green(24)
    pops = {}
    for st in ("f32", "f16c"):
        g = lib.Sim(case, st)
        g.build(dom, tris)
        g.init()
        pops[st] = g.populations(0)
        g.close()
    w = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)[:, None]
    v32 = pops["f32"].astype(np.float64) / w
    v = _f16c_value(pops["f16c"])
    e = np.floor(np.log2(np.maximum(np.abs(v32), 2.0 ** -14)))
    half_ulp = np.ldexp(1.0, (e - 12).astype(np.int64))
    # (the library divides in FP32 before it rounds: allow that rounding too)
    assert np.all(np.abs(v - v32) <= half_ulp * (1 + 2.0 ** -10))
    inexact = np.mean(v != v32)
    print("f16c initial populations: %d values, %.2f not exact, max |error| / half ulp %.3f"
          % (v.size, inexact, np.max(np.abs(v - v32) / half_ulp)))
    assert 0.2 < inexact and np.max(np.abs(v)) < 2.0
