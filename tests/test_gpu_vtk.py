"""VTK snapshots (lbm3_export_vtk, the file exportVTK writes,
include/exportVTK.h:17-158) on the GPU: the file's layout, and its nine fields
per cell against a restatement of the exporter on the FP64 oracle's fields
over two levels."""
import os

import numpy as np
import pytest

from lbm3_amd import cases, geometry, lib
from tests.helpers import GOLDEN, read_vtk, vtk_reference
from tests.oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def _pipe(levels, res=0.4):
    tri = geometry.read_binary_stl(os.path.join(GOLDEN, "pipeD20L200STL.STL"))
    return cases.pipe(tri, length=20.0, res=res, reynolds=100.0, u_mean=0.01, inlet_non_reflective=False,
                      profile_init=True, levels=levels)


def test_vtk_matches_the_restatement_f64(tmp_path):
    """Two levels (the pipe core refined), FP64, 30 coarse steps: markers, grid
    ids and resolutions equal, coordinates within an FP32 rounding, velocity
    and pressure within 1e-6 of their largest values."""
    case, dom, tris = _pipe(2)
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    o.run(30)
    g = lib.Sim(case, "f64")
    g.build(dom, tris)
    g.init()
    g.run(30)
    path = tmp_path / "snap.vtk"
    g.export_vtk(path)
    cells = [g.info(L).cellCount for L in range(2)]
    g.close()
    v, r = read_vtk(path), vtk_reference(o, case)
    n = sum(cells)
    assert v["title"] == "LBM snapshot; coordinates in mm; grid count 2"
    assert len(v["points"]) == n and np.array_equal(v["vertices"], np.stack([np.ones(n), np.arange(n)], 1))
    for k in ("wall", "grid_id", "grid_resolution_mm", "interface_overlap"):
        assert np.array_equal(v[k], r[k]), k
    assert np.allclose(v["points"], r["points"], rtol=2.5e-7, atol=1e-6)
    assert v["interface_overlap"].sum() > 0 and v["wall"].sum() > 0
    assert np.all(v["velocity_m_s"][v["wall"] == 1] == 0) and np.all(v["pressure_Pa"][v["wall"] == 1] == 0)
    for k in ("velocity_m_s", "pressure_Pa"):
        scale = np.max(np.abs(r[k]))
        assert scale > 0 and np.max(np.abs(v[k] - r[k])) <= 1e-6 * scale, k


def test_vtk_fields_are_the_downloaded_macros_f32(tmp_path):
    """One level, FP32, 100 steps: the exported velocity and pressure are the
    level's downloaded macroscopic fields in physical units, exactly."""
    case, dom, tris = _pipe(1)
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    s.init()
    s.run(100)
    path = tmp_path / "snap.vtk"
    s.export_vtk(path)
    I = s.info(0)
    m = s.macros(0)
    wm = s.array("wallMap", 0)
    s.close()
    v = read_vtk(path)
    f32 = np.float32
    res, dt = f32(I.res), f32(I.dtPhys)
    fluid = wm != -3
    ss = f32(0.57735026918962576) * (res / f32(1000.0)) / dt
    assert np.array_equal(v["velocity_m_s"][fluid], (m[:, 1:4] * (res / f32(1000.0)) / dt)[fluid])
    assert np.array_equal(v["pressure_Pa"][fluid], (m[:, 0] * f32(case.rhoPhys) * ss * ss)[fluid])
