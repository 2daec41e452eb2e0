"""Section cuts (lbm3_section_cut, the image exportSectionCutPlot{XY,ZY,ZX}
writes, include/plotter/exportSectionCutPlot.h:11-330) on the GPU: against a
restatement of the exporter on the FP64 oracle's fields over two levels and
three planes, against the one-level centre-line sampling, and the reference's
two-level Poiseuille centre-line pressure (results/INTERFACE/
QUADRATIC_centerlinePressure.txt) through it."""
import os

import numpy as np
import pytest

from lbm3_amd import cases, geometry, lib
from tests.helpers import GOLDEN, centreline_pressure, flow_report_reference, section_cut_reference
from tests.oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def _pipe(levels, res=0.4):
    tri = geometry.read_binary_stl(os.path.join(GOLDEN, "pipeD20L200STL.STL"))
    return cases.pipe(tri, length=20.0, res=res, reynolds=100.0, u_mean=0.01, inlet_non_reflective=False,
                      profile_init=True, levels=levels)


def test_section_cut_matches_the_restatement_f64():
    """Two levels (the pipe core refined), FP64, 30 coarse steps: the solid
    marker and cell size of every pixel equal, pressure and velocity within
    1e-6 of their largest values (both sides round to FP32 at the end)."""
    case, dom, tris = _pipe(2)
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    o.run(30)
    g = lib.Sim(case, "f64")
    g.build(dom, tris)
    g.init()
    g.run(30)
    for plane, cut in (("zy", 0.0), ("xy", 1.3), ("zx", -2.1)):
        a, b = g.section_cut(plane, cut), section_cut_reference(o, case, plane, cut)
        assert a.shape == b.shape, plane
        assert np.array_equal(a[..., 4], b[..., 4]) and np.array_equal(a[..., 5], b[..., 5]), plane
        assert np.allclose(np.unique(a[..., 5]), [0.2, 0.4]), plane        # both levels paint pixels
        for k in range(4):
            scale = max(np.max(np.abs(b[..., k])), 1e-30)
            assert np.max(np.abs(a[..., k] - b[..., k])) <= 1e-6 * scale, (plane, k)
    g.close()


def test_centre_rows_of_a_one_level_cut_are_the_centreline():
    case, dom, tris = _pipe(1)
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    s.init()
    s.run(200)
    img = s.section_cut("zy", 0.0)
    ny = img.shape[0]
    p = img[(ny - 1) // 2: ny // 2 + 1, :, 0].mean(axis=0)
    _, pc = centreline_pressure(s, case)
    s.close()
    assert np.array_equal(p, pc)


def test_two_level_poiseuille_centreline_matches_the_reference():
    """0.4 mm cells with the core (r < 5 mm, |z| < 5 mm) refined, 10000 coarse
    steps (the physical time of the uniform runs): the centre-line pressure of
    the x = 0 cut within 3e-3 of the inlet pressure of the reference's
    (measured 8.9e-4; curve level, as the uniform-grid comparison)."""
    case, dom, tris = _pipe(2)
    ref = np.loadtxt(os.path.join(GOLDEN, "poiseuille_quadratic_centerline_pressure.txt"), delimiter=";", skiprows=1)
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    s.init()
    s.run(10000)
    img = s.section_cut("zy", 0.0)
    s.close()
    ny = img.shape[0]
    p = img[(ny - 1) // 2: ny // 2 + 1, :, 0].mean(axis=0)
    assert len(p) == len(ref)
    dev = np.max(np.abs(p - ref[:, 1])) / np.max(np.abs(ref[:, 1]))
    print("two-level poiseuille: centre-line pressure max deviation %.2e of the inlet pressure" % dev)
    assert dev < 3e-3


def test_flow_report_matches_the_restatement_f64():
    """lbm3_flow_report (getFlowReport*) on two levels, FP64: every quantity
    within 1e-6 of the reduction restated on the oracle's section image, for a
    cross-section of the pipe and a cropped plane along it."""
    case, dom, tris = _pipe(2)
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    o.run(30)
    g = lib.Sim(case, "f64")
    g.build(dom, tris)
    g.init()
    g.run(30)
    for plane, cut in (("xy", 0.3), ("xy", -7.9), ("zy", 0.0)):
        a, b = g.flow_report(plane, cut), flow_report_reference(o, case, plane, cut)
        for k, v in b.items():
            assert abs(a[k] - v) <= 1e-6 * max(abs(v), 1e-30), (plane, cut, k, a[k], v)
    g.close()


def test_flow_report_conserves_the_inlet_mass_flow():
    """One level, 10000 steps: the mass flow through the mid-pipe section equals
    the inlet's (open-boundary statistics, outer normal) within 1%, over the
    bore's area (pi (10 mm)^2 within 1%)."""
    case, dom, tris = _pipe(1)
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    s.init()
    s.run(10000)
    s.reset_open_boundary_stats()
    s.run(1)
    inlet = s.open_boundary_stats(0)
    rep = s.flow_report("xy", 0.0)
    s.close()
    print("mid-pipe mass flow %.6e kg/s, inlet %.6e kg/s, area %.2f mm^2" % (rep["mass_flow"], -inlet["mass_flow"], rep["area_mm2"]))
    assert abs(rep["mass_flow"] / -inlet["mass_flow"] - 1.0) < 0.01
    assert abs(rep["area_mm2"] / (np.pi * 100.0) - 1.0) < 0.01
