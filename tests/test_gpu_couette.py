"""Rotating walls on the GPU (getLocalBC rigid rotation, lbm3_local_rule
SET_ROTATION, in the wall kernel) on circular Couette flow: against the FP64
oracle step for step, and in FP32 against the closed-form torque
(torqueTestMain.cu:51-60) at steady state."""
import numpy as np
import pytest

from lbm3_amd import cases, lib
from tests.helpers import rel_max, wall_force_sums
from tests.oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def test_couette_matches_the_oracle_f64():
    case, dom, tris, _ = cases.couette()
    o = OracleSim(case, dom, tris, double=True)
    g = lib.Sim(case, "f64")
    g.build(dom, tris)
    for name in ("shifter", "wallMap", "wallData", "linkLength"):
        assert np.array_equal(g.array(name), o.array(name)), name
    o.init()
    g.init()
    o.run(300)
    g.run(300)
    g.sync()
    mo, mg = o.macros(), g.macros()
    solid = o.array("wallMap") == -3
    r = rel_max(1.0 + mg[~solid, 0], 1.0 + mo[~solid, 0])
    u = rel_max(mg[~solid, 1:], mo[~solid, 1:])
    print("couette f64 after 300 steps: rho %.2e u %.2e" % (r, u))
    assert r < 1e-10 and u < 1e-10
    fo, fg = wall_force_sums(o, 2), g.wall_forces(0, (0.0, 0.0, 0.0), 2)
    assert np.max(np.abs(fg - fo)) < 1e-9 * np.max(np.abs(fo))
    g.close()


def test_couette_torque_f32_matches_the_closed_form():
    case, dom, tris, ci = cases.couette()
    g = lib.Sim(case, "f32")
    g.build(dom, tris)
    g.init()
    g.run(6000)
    g.reset_wall_forces()
    g.run(2000)
    T = g.wall_forces(0, (0.0, 0.0, 0.0), 2)[:, 5] / 2000.0
    R1, R2, L, nu, om = ci["r_inner"], ci["r_outer"], ci["length"], ci["nu"], ci["omega"]
    Ta = 4 * np.pi * nu * om * R1 ** 2 * R2 ** 2 / (R2 ** 2 - R1 ** 2) * L
    print("couette f32 torque on the fluid: rotor %.6e stator %.6e closed form %.6e" % (T[1], T[0], Ta))
    assert abs(T[1] / Ta - 1.0) < 0.02 and abs(-T[0] / Ta - 1.0) < 0.02
    assert abs(T[0] + T[1]) < 0.005 * Ta
    g.close()


def test_update_tracker_torque_in_physical_units():
    """lbm3_update_tracker (updateTracker's wall part, trackerFunctions.h:356-474):
    the period-averaged torque on the fluid in N m equals the closed form in
    physical units, 4 pi rho nu L Omega R1^2 R2^2 / (R2^2 - R1^2) with the
    case's rhoPhys, nuPhys [m^2/s], Omega [rad/s] and lengths in m, within
    2%; a second call reads the next period only (the first one reset)."""
    case, dom, tris, ci = cases.couette()
    g = lib.Sim(case, "f32")
    g.build(dom, tris)
    g.init()
    g.run(6000)
    g.update_tracker(walls=2)                 # the start-up period, then the reset
    g.run(2000)
    t = g.update_tracker(walls=2)["walls"]
    g.close()
    R1, R2, L = ci["r_inner"] * 1e-3, ci["r_outer"] * 1e-3, ci["length"] * 1e-3
    Ta = 4 * np.pi * case.rhoPhys * case.nuPhys * L * ci["omega"] * R1 ** 2 * R2 ** 2 / (R2 ** 2 - R1 ** 2)
    print("tracker torque on the fluid: rotor %.6e stator %.6e N m, closed form %.6e" % (t[1, 5], t[0, 5], Ta))
    assert abs(t[1, 5] / Ta - 1.0) < 0.02 and abs(-t[0, 5] / Ta - 1.0) < 0.02
