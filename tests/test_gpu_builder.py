"""GPU builder vs the oracle builder: every array the per-step kernels read must
be bit-identical (voxelizer flags, wall map and link codes, compressed rows,
open-boundary and periodic lists, interface lists)."""
import numpy as np
import pytest

from lbm3_amd import cases
from lbm3_amd import lib
from tests.helpers import level_arrays
from tests.oracle_lib import OracleSim

pytestmark = pytest.mark.gpu

ARRAYS = ("shifter", "rowI", "rowJ", "rowK", "rowJPlus", "rowKPlus", "rowJKPlus", "wallMap", "wallIndex", "wallData",
          "linkLength", "periodicIndex", "periodicSource")


def _check(case, dom, tris):
    o = OracleSim(case, dom, tris)
    g = lib.Sim(case, "f32")
    fu = g.build(dom, tris)
    assert fu == o.fluid_updates
    for L in range(o.levels):
        io, ig = o.info(L), g.info(L)
        for f in ("cellCount", "cellCountX", "cellCountY", "cellCountZ"):
            assert getattr(io, f) == getattr(ig, f), f
        for f in ("res", "ox", "oy", "oz", "nu", "dtPhys"):
            assert getattr(io, f) == getattr(ig, f), f
        ref = level_arrays(o, L)
        for name in ARRAYS:
            a, b = g.array(name, L), ref[name]
            assert a.shape == b.shape and np.array_equal(a.astype(np.int64), b.astype(np.int64)), (L, name)
        assert np.array_equal(g.array("openCount", L), ref["openCount"]), L
        assert np.array_equal(g.array("openIndex", L), ref["openIndex"]), L
        assert np.array_equal(g.array("openSource", L), ref["openSource"]), L
        assert np.array_equal(g.array("rayHitScan", L), o.array("rayHitScan", L)), L
        assert np.array_equal(g.array("rayMap", L), o.array("rayMap", L)), L
        for name in ("f2cIndex", "f2cChild", "c2fIndex", "c2fChild", "leftoverIndex", "leftoverParent", "leftoverNbrI",
                     "leftoverNbrJ", "leftoverNbrK"):
            assert np.array_equal(g.array(name, L), o.array(name, L)), (L, name)
    return o, g


def test_builder_taylor_green():
    _check(*cases.taylor_green(16))


def test_builder_channel():
    _check(*cases.channel(32, 16, 16))


def test_builder_small_sphere():
    o, g = _check(*cases.sphere(64, 32, 32, diameter=8.0, center=(16.0, 16.0, 16.0), segments=(24, 48)))
    assert g.info().cellCount < 64 * 32 * 32   # the sphere interior is removed


def test_builder_offcentre_sphere_fine_mesh():
    _check(*cases.sphere(48, 40, 36, diameter=13.3, center=(17.3, 19.1, 18.2), segments=(60, 120)))


def _cut_sphere():
    """Two levels whose refinement box ends inside the sphere: leftover cells with
    opposite-side neighbours and wall-adjacent coarse parents."""
    case, dom, tris = cases.sphere(48, 24, 24, diameter=7.0, center=(16.0, 12.0, 12.0), segments=(16, 32))
    case.gridLevelCount = 2
    cases.add_refine_rule(case, 0, (4, 4, 4, 16, 20, 20))
    return case, dom, tris


def _small_sphere3():
    return cases.multilevel_sphere(64, 32, 32, level1_box=(12, 8, 8, 36, 24, 24), level2_box=(14, 10, 10, 26, 22, 22),
                                   center=(20.0, 16.0, 16.0), diameter_fine=12.0, segments=(16, 32))


def test_builder_two_level_taylor_green():
    _check(*cases.taylor_green(16, levels=2))


def test_builder_three_level_sphere():
    o, g = _check(*_small_sphere3())
    assert o.levels == 3 and len(o.array("f2cIndex", 1)) > 0


def test_builder_refinement_cutting_the_sphere():
    o, g = _check(*_cut_sphere())
    assert len(o.array("leftoverIndex", 0)) > 0
