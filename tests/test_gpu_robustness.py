"""Setup refusals and in-loop checks on the GPU: the closed-mesh check of
readSTL (checkSTLEdges, STLFunctions.h:130-315) with its counts bit-exact
against the oracle and an open STL refused; the FP64 global mass balance
(lbm3_mass) exact on a periodic box and equal to the oracle's fields; the
divergence watchdog of lbm3_run on populations with a NaN uploaded from the
host."""
import numpy as np
import pytest

from lbm3_amd import cases, geometry, lib
from tests.oracle_lib import OracleSim, stl_edge_counts

pytestmark = pytest.mark.gpu


def test_edge_counts_bit_exact_and_open_stl_refused():
    case, dom, tris = cases.sphere(40, 20, 20, diameter=6.0, center=(12.0, 10.0, 10.0), segments=(16, 32))
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    got = s.array("stl0EdgeCount").reshape(-1, 3)
    want = stl_edge_counts(tris[0], case.resGlobal)
    assert np.array_equal(got, want) and np.all(got == 2)
    holed = np.delete(np.asarray(tris[0], np.float32).reshape(-1, 9), 17, axis=0)
    with pytest.raises(lib.LBM3Error, match="not closed"):
        s.build(dom, [holed])
    got = s.array("stl0EdgeCount").reshape(-1, 3)           # the refused body's counts
    assert np.array_equal(got, stl_edge_counts(holed, case.resGlobal))
    assert int((got != 2).sum()) == 3
    s.build(dom, tris)                                      # the simulation builds again afterwards
    s.close()


def test_rotor_stl_passes_the_same_check():
    """The seeded 6-blade rotor of the rotor-duct configuration (200,704
    triangles) is closed: its counts on the GPU equal the oracle's."""
    case, dom, statics, rotor, ri = cases.rotor_duct(seed=0, levels=1)
    cx = ri["origin"][0]
    s = lib.Sim(case, "f32")
    s.build(cases.bounds(cx - 16, -36, -36, cx + 16, 36, 36), [rotor])   # the rotor alone as a body
    got = s.array("stl0EdgeCount").reshape(-1, 3)
    assert len(got) == 200704
    assert np.array_equal(got, stl_edge_counts(rotor, case.resGlobal))
    s.close()


def _counted_drho(osim, level):
    m = osim.macros(level)
    wm = osim.array("wallMap", level)
    wd = osim.array("wallData", level)
    keep = (wm != -3) & (wm != -2)
    over = np.zeros_like(keep)
    idx = np.nonzero(wm >= 0)[0]
    over[idx] = (wd[wm[idx]] & 1) != 0
    keep &= ~over
    return float(np.sum(m[keep, 0], dtype=np.float64)), int(keep.sum())


def test_mass_is_conserved_on_a_periodic_box_and_equals_the_oracle():
    case, dom, tris = cases.taylor_green(24, u0=0.05, reynolds=200.0, pressure=True)
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    g = lib.Sim(case, "f64")
    g.build(dom, tris)
    g.init()
    g.set_populations(0, o.array("f", 0).reshape(27, -1))
    d0, n0 = g.mass()
    assert n0 == 24 ** 3
    od, on = _counted_drho(o, 0)
    assert on == n0 and abs(d0 - od) < 1e-12
    g.run(150)
    o.run(150)
    d1, n1 = g.mass()
    od1, _ = _counted_drho(o, 0)
    print("periodic FP64 box: sum dRho %.3e -> %.3e over %d cells; oracle %.3e" % (d0, d1, n1, od1))
    assert n1 == n0
    assert abs(d1 - d0) < 1e-11 * n0
    assert abs(d1 - od1) < 1e-11 * n0
    g.close()


def test_multilevel_mass_weights_levels():
    """Two-level Taylor-Green: the count is sum over levels of counted cells / 8^L."""
    case, dom, tris = cases.taylor_green(16, levels=2)
    o = OracleSim(case, dom, tris)
    g = lib.Sim(case, "f32")
    g.build(dom, tris)
    g.init()
    o.init()
    d, n = g.mass()
    want = sum(_counted_drho(o, L)[1] / 8.0 ** L for L in range(2))
    assert n == want
    dwant = sum(_counted_drho(o, L)[0] / 8.0 ** L for L in range(2))
    assert abs(d - dwant) < 1e-6 * n


def test_watchdog_reports_a_nan():
    case, dom, tris = cases.taylor_green(16)
    g = lib.Sim(case, "f32")
    g.build(dom, tris)
    g.init()
    g.set_watchdog(4)
    g.run(10)                                    # finite: no error
    f = g.populations(0)
    interior = int(np.nonzero(g.array("wallMap") == -1)[0][100])
    f[5, interior] = np.nan
    g.set_populations(0, f)
    with pytest.raises(lib.LBM3Error, match="non-finite populations after coarse step 12"):
        g.run(10)
    g.set_watchdog(0)
    g.run(2)                                     # off: the run goes on
    g.close()
