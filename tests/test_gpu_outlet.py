"""GPU non-reflective pressure outlet and strict edge/corner cells against the
oracle (updateGrid.h:259-280): a plane acoustic pulse leaving a duct through
the outlet, and a randomly perturbed box whose every boundary cell is a
pressure outlet (all 6 face, 12 edge and 8 corner normals)."""
import numpy as np
import pytest

from lbm3_amd import lib
from tests.helpers import cell_ijk, level_arrays, rel_max
from tests.oracle_lib import OracleSim
from tests.test_oracle_outlet_pins import _duct, _lattice_case, reflection, pulse_fields
from lbm3_amd import cases

pytestmark = pytest.mark.gpu


def _gpu_from(o, case, storage):
    g = lib.Sim(case, storage)
    g.import_level(0, level_arrays(o, 0), o.info(0))
    return g


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_pulse_leaves_through_the_outlet(storage):
    case, dom = _duct(1)
    o = OracleSim(case, dom, [])
    g = _gpu_from(o, case, storage)
    inc_o, ref_o = reflection(o)
    inc_g, ref_g = reflection(g)
    g.close()
    print("%s: incident %.3e (oracle %.3e), reflected ratio %.4f (oracle %.4f)"
          % (storage, inc_g, inc_o, ref_g / inc_g, ref_o / inc_o))
    assert abs(inc_g - inc_o) < (1e-6 if storage == "f32" else 2e-5)
    assert ref_g / inc_g < 0.1


def test_pulse_through_the_outlet_f64_equals_oracle():
    case, dom = _duct(1)
    o = OracleSim(case, dom, [], double=True)
    g = _gpu_from(o, case, "f64")
    fields = pulse_fields(cell_ijk(o))
    o.init_fields(0, fields)
    g.init_fields(0, fields)
    g.set_populations(0, o.array("f", 0).reshape(27, -1))   # feq from identical bits
    o.run(150)
    g.run(150)
    g.sync()
    mo, mg = o.macros(), g.macros()
    err = max(rel_max(1 + mg[:, 0], 1 + mo[:, 0]), np.abs(mg[:, 1:] - mo[:, 1:]).max() / 1e-3)
    g.close()
    print("duct pulse FP64, 150 steps: GPU vs oracle %.2e" % err)
    assert err < 1e-10


def _outlet_box(storage, n=12):
    c = _lattice_case(0.02)
    cases.add_open_rule(c, openBCID=0, dirichletRho=1, dRho=0.004, nonReflective=1)
    return c, cases.bounds(0, 0, 0, n, n, n)


def _box_fields(N):
    rng = np.random.default_rng(3)
    fields = np.zeros((N, 4), np.float32)
    fields[:, 0] = 0.01 * rng.standard_normal(N)
    fields[:, 1:] = 0.02 + 0.01 * rng.standard_normal((N, 3))
    return fields


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_outlet_box_every_normal(storage):
    """Every boundary cell a non-reflective pressure outlet, random start, 5
    steps: FP64 equals the oracle to 1e-10; FP32 lies as close to the FP64
    solution as the oracle's FP32 run does (1.5x + 1e-6, relative max-norm);
    the edge and corner cells hold the target density exactly."""
    n = 12
    c, dom = _outlet_box(storage, n)
    o = OracleSim(c, dom, [], double=(storage == "f64"))
    g = _gpu_from(o, c, storage)
    fields = _box_fields(o.info().cellCount)
    o.init_fields(0, fields)
    g.init_fields(0, fields)
    if storage == "f64":
        g.set_populations(0, o.array("f", 0).reshape(27, -1))
    o.run(5)
    g.run(5)
    g.sync()
    mo, mg = o.macros().astype(np.float64), g.macros().astype(np.float64)
    g.close()
    idx = o.array("open0Index")
    ijk = cell_ijk(o)[idx]
    edge = ((ijk == 0) | (ijk == n - 1)).sum(axis=1) > 1
    target = float(np.float32(0.004))
    edge_dev = np.abs(mg[idx[edge], 0] - target).max()
    err = max(rel_max(1 + mg[:, 0], 1 + mo[:, 0]), rel_max(mg[:, 1:], mo[:, 1:]))
    if storage == "f64":
        print("f64 outlet box, 5 steps: GPU vs oracle %.2e, edge/corner |dRho - target| %.2e" % (err, edge_dev))
        assert err < 1e-10 and edge_dev < 1e-12
        return
    o64 = OracleSim(c, dom, [], double=True)
    o64.init_fields(0, fields)
    o64.run(5)
    m64 = o64.macros()
    sens = max(rel_max(1 + mo[:, 0], 1 + m64[:, 0]), rel_max(mo[:, 1:], m64[:, 1:]))
    err64 = max(rel_max(1 + mg[:, 0], 1 + m64[:, 0]), rel_max(mg[:, 1:], m64[:, 1:]))
    print("f32 outlet box, 5 steps: GPU vs oracle %.2e; FP32 vs FP64: oracle %.2e, GPU %.2e; edge/corner %.2e"
          % (err, sens, err64, edge_dev))
    assert err64 < 1.5 * sens + 1e-6 and edge_dev < 5e-7
