"""Circular Couette flow against its closed form: the pin of a rotating wall
(getLocalBC's rigid-rotation wall velocity, lbm3_local_rule SET_ROTATION),
interpolated bounceback on curved walls and the per-body torque sums, after
the reference's torque test (torqueTestMain.cu:51-60: T = 4 pi mu L Omega
R1^2 R2^2 / (R2^2 - R1^2) on a periodic annulus) at a small size: R1 = 12,
R2 = 20, L = 8 cells, Re = 10, FP64 oracle, steady state.

The wall forces are the forces on the fluid (momentum exchange,
interpolatedBouncebackFunctions.h:79-100), so the turning inner cylinder puts
+T into the fluid about z and the fixed tube takes it out.  The measured
torques lie within 0.8% of the closed form and the azimuthal velocity within
1.2% of u = A r + B / r (the wall velocity is set at the cell centres next to
the wall, and q is quantised to 1/254): bars of 2%.  A wrong sign or lever-arm
unit of the rotation, or a wrong body split, fails by far more."""
import numpy as np

from lbm3_amd import cases
from tests.helpers import cell_ijk, wall_force_sums
from tests.oracle_lib import OracleSim


def test_couette_torque_and_profile_match_the_closed_form():
    case, dom, tris, ci = cases.couette()
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    o.run(6000)
    t1 = wall_force_sums(o, 2)[:, 5]
    o.run(2000)
    T = (wall_force_sums(o, 2)[:, 5] - t1) / 2000.0          # per step, lattice force x mm (1 mm cells)
    R1, R2, L, nu, om = ci["r_inner"], ci["r_outer"], ci["length"], ci["nu"], ci["omega"]
    Ta = 4 * np.pi * nu * om * R1 ** 2 * R2 ** 2 / (R2 ** 2 - R1 ** 2) * L
    print("torque on the fluid: rotor %.6e stator %.6e closed form %.6e" % (T[1], T[0], Ta))
    assert abs(T[1] / Ta - 1.0) < 0.02
    assert abs(-T[0] / Ta - 1.0) < 0.02
    assert abs(T[0] + T[1]) < 0.005 * Ta                     # steady: what goes in comes out
    m = o.macros()
    ijk = cell_ijk(o)
    info = o.info()
    x = ijk[:, 0] * info.res + info.ox
    y = ijk[:, 1] * info.res + info.oy
    r = np.hypot(x, y)
    ut = (-y * m[:, 1] + x * m[:, 2]) / np.maximum(r, 1e-9)
    wm = o.array("wallMap")
    A = -om * R1 ** 2 / (R2 ** 2 - R1 ** 2)
    B = om * R1 ** 2 * R2 ** 2 / (R2 ** 2 - R1 ** 2)
    for rr in (14.0, 16.0, 18.0):
        sel = (np.abs(r - rr) < 0.5) & (wm != -3) & (wm != -2)
        ratio = ut[sel].mean() / (A * r[sel] + B / r[sel]).mean()
        assert abs(ratio - 1.0) < 0.02, (rr, ratio)
    assert np.max(np.abs(m[(wm != -3), 3])) < 1e-12             # no axial flow
