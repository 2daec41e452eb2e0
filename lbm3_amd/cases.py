"""Case descriptions: the reference's case files expressed as lbm3_case data.

Each builder returns (case, domain_bounds, static_triangles) where
static_triangles is a list of float32 arrays of shape (n, 9) (one per STL
body).  Float arithmetic that sets up the grid follows the reference case
files in float32, so the builder sees the same numbers:

* taylor_green  cases/taylorGreenVortex/taylorGreenVortexMain.cu:1-44, 82-102, 182-198
* sphere        cases/flowAroundSphere/flowAroundSphereMain.cu:96-151, 183-197
* channel       body-force driven periodic duct (serving mix; walls by moment BC)
* multilevel_sphere  three-level sphere (coarse 128x64x64, nested boxes, finest patch 96^3)
"""
import numpy as np

from . import ctypes_defs as T
from . import geometry

f32 = np.float32


def _case(levels=1, res=1.0, dt=1.0, nu=1.0, rho=1.0):
    c = T.Case()
    c.gridLevelCount = levels
    c.wallRefinementCount = 0
    c.resGlobal = float(f32(res))
    c.dtPhysGlobal = float(f32(dt))
    c.nuPhys = float(f32(nu))
    c.rhoPhys = float(f32(rho))
    c.trackWallForce = 1
    c.trackRotorForce = 0
    c.trackOpenBoundaries = 0
    c.openDefault = T.default_bc()
    return c


def add_open_rule(c, x=T.FACE_ANY, y=T.FACE_ANY, z=T.FACE_ANY, level_mask=0, **bc):
    r = c.openRules[c.nOpenRules]
    r.levelMask = level_mask
    r.faceMask[0], r.faceMask[1], r.faceMask[2] = x, y, z
    r.bc = T.default_bc(**bc)
    c.nOpenRules += 1


def add_local_rule(c, wall_id=-2, level_mask=0, **kw):
    r = c.localRules[c.nLocalRules]
    r.levelMask = level_mask
    r.wallID = wall_id
    r.set = 0
    if "velocity" in kw:
        r.set |= T.SET_VELOCITY
        r.ux, r.uy, r.uz = kw["velocity"]
    if "omega" in kw:
        r.set |= T.SET_ROTATION
        for a in range(3):
            r.omega[a] = kw["omega"][a]
            r.origin[a] = kw.get("origin", (0.0, 0.0, 0.0))[a]
    if "force" in kw:
        r.set |= T.SET_FORCE
        r.gx, r.gy, r.gz = kw["force"]
    if "limiter" in kw:
        r.set |= T.SET_LIMITER
        r.collisionLimiter = kw["limiter"]
    if "nu_multiplier" in kw:
        r.set |= T.SET_NU_MULTIPLIER
        r.nuMultiplier = kw["nu_multiplier"]
    if "use_ibb" in kw:
        r.set |= T.SET_USE_IBB
        r.useIBB = int(kw["use_ibb"])
    c.nLocalRules += 1


def add_refine_rule(c, level, box, mode=T.REFINE_SET, cylinder=None):
    """Refinement box (open, physical mm); cylinder=(axis, origin, radius)
    further limits it to r < radius about that line."""
    r = c.refineRules[c.nRefineRules]
    r.level, r.mode = level, mode
    r.xMin, r.yMin, r.zMin, r.xMax, r.yMax, r.zMax = [float(f32(v)) for v in box]
    if cylinder is not None:
        set_profile(r.cylinder, *cylinder)
    c.nRefineRules += 1


def set_profile(p, axis, origin, radius):
    """Fill an lbm3_profile: max(0, 1 - r^2 / radius^2) about the line through origin along axis."""
    p.axis = axis
    for a in range(3):
        p.origin[a] = float(f32(origin[a]))
    p.radius = float(f32(radius))


def bounds(xmin, ymin, zmin, xmax, ymax, zmax):
    b = T.Bounds()
    b.xMin, b.yMin, b.zMin, b.xMax, b.yMax, b.zMax = [float(f32(v)) for v in (xmin, ymin, zmin, xmax, ymax, zmax)]
    return b


def taylor_green(n=64, u0=0.05, reynolds=1600.0, pressure=False, limiter=None, levels=1):
    """Periodic Taylor-Green vortex on n^3 active cells (plus one ghost layer per face).

    Lattice viscosity nu = u0 * (n / 2 pi) / Re, i.e. Re on the length n/(2 pi).
    levels=2 refines the central cube [-pi L/2, pi L/2]^3 once, as the
    reference's refined TGV run (taylorGreenVortexMain.cu:52-66)."""
    PI = f32(3.14159265358979323846)
    L = f32(1000.0)                       # reference length [mm]
    domain = f32(2.0) * PI * L
    res = domain / f32(n)
    nu_phys = f32(1.0) * (L / f32(1000.0)) / f32(reynolds)
    dt = (f32(u0) / f32(1.0)) * (res / f32(1000.0))
    c = _case(levels, res, dt, nu_phys)
    c.trackWallForce = 0
    c.init.kind = T.INIT_TAYLOR_GREEN
    c.init.u0 = float(f32(u0))
    c.init.length = float(L)
    c.init.pressure = int(bool(pressure))
    add_open_rule(c, periodic=1)
    if limiter is not None:
        add_local_rule(c, limiter=limiter)
    half = f32(0.5) * domain
    eps = f32(1e-6) * domain
    lo = -half - res - eps
    hi = half + res + eps
    if levels > 1:
        w = f32(0.25) * domain
        add_refine_rule(c, 0, (-w, -w, -w, w, w, w))
    return c, bounds(lo, lo, lo, hi, hi, hi), []


def multilevel_sphere(nx=128, ny=64, nz=64, diameter_fine=24.0, u_inlet=0.04, reynolds=100.0, seed=None,
                      level1_box=(24, 16, 16, 72, 48, 48), level2_box=(28, 20, 20, 52, 44, 44),
                      center=(40.0, 32.0, 32.0), segments=(60, 120), k15_on_level0=True):
    """Flow around a sphere on three levels: a coarse nx*ny*nz grid of 1 mm cells
    and two nested refinement boxes (factor 2 each), the finest around a sphere
    of `diameter_fine` finest-level cells.  Boundaries as the sphere case.
    With `seed`, the inlet velocity is drawn uniformly from [0.02, 0.06]."""
    if seed is not None:
        u_inlet = float(np.random.default_rng(seed).uniform(0.02, 0.06))
    diameter = diameter_fine / 4.0                      # in coarse cells (1 mm)
    c, dom, _ = sphere(nx, ny, nz, diameter=diameter, center=center, u_inlet=u_inlet, reynolds=reynolds,
                       k15_on_level0=k15_on_level0, segments=segments)
    c.gridLevelCount = 3
    add_refine_rule(c, 0, level1_box)
    add_refine_rule(c, 1, level2_box)
    tri = geometry.uv_sphere(diameter / 2.0, center, segments[0], segments[1])
    return c, dom, [tri]


def sphere(nx=256, ny=128, nz=128, diameter=32.0, center=(64.0, 64.0, 64.0), u_inlet=0.05, reynolds=100.0,
           limiter=0.1, k15_on_level0=True, segments=(120, 240)):
    """Flow around a sphere on a uniform nx*ny*nz grid of 1 mm cells.

    Boundaries and collision follow the reference sphere case
    (flowAroundSphereMain.cu:108-151): velocity inlet at i=0, non-reflective
    pressure outlet at i=nx-1, far-field velocity on every other boundary cell,
    K17 with limiter 0.1 except K15 on level 0 (where the open boundaries are);
    the sphere (wall 0) is a no-slip interpolated-bounceback wall.  On a single
    uniform level, k15_on_level0=False selects K17 everywhere."""
    res = f32(1.0)
    u_phys = f32(1.0)
    nu_phys = u_phys * (f32(diameter) / f32(1000.0)) / f32(reynolds)
    dt = (f32(u_inlet) / u_phys) * (res / f32(1000.0))
    c = _case(1, res, dt, nu_phys)
    c.init.kind = T.INIT_UNIFORM
    c.init.ux = float(f32(u_inlet))
    U = float(f32(u_inlet))
    add_open_rule(c, x=T.FACE_HIGH, y=T.FACE_MID, z=T.FACE_MID, dirichletRho=1, dRho=0.0, openBCID=1, nonReflective=1)
    add_open_rule(c, x=T.FACE_LOW, y=T.FACE_MID, z=T.FACE_MID, dirichletU=1, ux=U, openBCID=0, nonReflective=0)
    add_open_rule(c, dirichletU=1, ux=U, openBCID=2, nonReflective=0)
    add_local_rule(c, limiter=limiter)
    add_local_rule(c, wall_id=0, velocity=(0.0, 0.0, 0.0))
    if k15_on_level0:
        add_local_rule(c, level_mask=1, limiter=0.0)
    c.storeRef[0] = U                     # 16-bit stores: encode the deviation from the free stream
    tri = geometry.uv_sphere(diameter / 2.0, center, segments[0], segments[1])
    return c, bounds(0, 0, 0, nx, ny, nz), [tri]


def pipe(triangles, length=100.0, res=0.2, diameter=20.0, reynolds=1000.0, u_mean=0.03, inlet_non_reflective=True,
         profile_init=False, levels=1, nu_phys=1.5e-5, rho_phys=1.225, impedance_tolerance=None):
    """Flow through the 20 mm bore of the reference's pipe STL
    (pipeD20L200STL.STL: a 25 mm square block with a 200 mm bore along z),
    cut to `length` mm about z = 0, as the reference's pipe cases set it up
    (nonReflectiveTestMain.cu:1-105: Re 1000, mean inlet velocity 0.03, fluid at
    rest at the start, 0.2 mm cells; poiseuilleFlowMain.cu:1-95: Re 100, mean
    inlet velocity 0.01, Poiseuille profile in the whole bore at the start):
    parabolic velocity inlet at k = 0 (centre-line velocity 2 u_mean, open
    boundary 0, non-reflective or strict Dirichlet), pressure outlet dRho = 0 at
    k = nz - 1 (open boundary 1), K15 everywhere, the bore a no-slip wall
    without interpolated bounceback, open-boundary flow tracked.  With
    levels = 2 the core r < D/4, |z| < length/4 is refined once
    (poiseuilleFlowMain.cu:32-46).  impedance_tolerance selects the impedance
    non-reflective scheme with that density tolerance
    (updateGrid_SCHLAFFER_IMPEDANCE_SCHEME.h, RHO_REFLECTION_TOLERANCE) in place
    of the characteristic one.  Returns (case, domain, triangles)."""
    R = f32(0.5) * f32(diameter)
    u_phys = f32(reynolds) * f32(nu_phys) / (f32(diameter) / f32(1000.0))
    dt = (f32(u_mean) / u_phys) * (f32(res) / f32(1000.0))
    c = _case(levels, res, dt, nu_phys, rho_phys)
    c.trackWallForce = 0
    c.trackOpenBoundaries = 1
    if impedance_tolerance is not None:
        c.openScheme = T.OPEN_IMPEDANCE
        c.rhoReflectionTolerance = float(f32(impedance_tolerance))
    U = float(f32(2.0) * f32(u_mean))
    add_open_rule(c, z=T.FACE_LOW, dirichletU=1, ux=0.0, uy=0.0, uz=U, nonReflective=int(bool(inlet_non_reflective)),
                  openBCID=0)
    set_profile(c.openRules[c.nOpenRules - 1].profile, T.AXIS_Z, (0.0, 0.0, 0.0), R)
    add_open_rule(c, z=T.FACE_HIGH, dirichletRho=1, dRho=0.0, nonReflective=0, openBCID=1)
    add_local_rule(c, limiter=0.0, use_ibb=0)
    add_local_rule(c, wall_id=0, velocity=(0.0, 0.0, 0.0))
    c.init.kind = T.INIT_UNIFORM
    if profile_init:
        c.init.uz = U
        set_profile(c.init.profile, T.AXIS_Z, (0.0, 0.0, 0.0), R)
    if levels > 1:
        zr = f32(length) * f32(0.25)
        add_refine_rule(c, 0, (-1e30, -1e30, -zr, 1e30, 1e30, zr), cylinder=(T.AXIS_Z, (0.0, 0.0, 0.0), f32(0.5) * R))
    tri = np.ascontiguousarray(triangles, np.float32).reshape(-1, 9)
    v = tri.reshape(-1, 3)
    lo, hi = v.min(axis=0), v.max(axis=0)
    half = f32(length) * f32(0.5)
    return c, bounds(lo[0], lo[1], -half, hi[0], hi[1], half), [tri]


def couette(r_inner=12.0, r_outer=20.0, length=8.0, reynolds=10.0, u_wall=0.02, n_theta=512):
    """Circular Couette flow, the set-up of the reference's torque test
    (torqueTestMain.cu:1-112) at a small size: a solid inner cylinder (wall 1)
    turning at the angular velocity that gives it the surface speed u_wall
    (lattice units), inside a fixed outer tube (wall 0) that also covers the
    domain corners; periodic along the axis z, fluid at rest at the start, K15,
    interpolated bounceback on both walls.  Lengths in cells of 1 mm;
    Re = u_wall (r_outer - r_inner) / nu.  Both meshes extend beyond the
    periodic ends.  Returns (case, domain, [outer tube, inner cylinder], info)
    with info: omega (rad/s), nu (lattice), the radii and the length."""
    res = f32(1.0)
    nu = f32(u_wall) * f32(r_outer - r_inner) / f32(reynolds)
    dt = f32(1.0)
    c = _case(1, res, dt, f32(1.0))
    c.nuPhys = float(nu * (res / f32(1000.0)) * (res / f32(1000.0)) / dt)
    c.trackWallForce = 1
    c.init.kind = T.INIT_UNIFORM
    omega = float(f32(u_wall) / f32(r_inner) / dt)     # rad/s, lattice speed u_wall at r_inner
    add_open_rule(c, z=T.FACE_LOW | T.FACE_HIGH, periodic=1)
    add_local_rule(c, limiter=0.0)
    add_local_rule(c, wall_id=0, velocity=(0.0, 0.0, 0.0))
    add_local_rule(c, wall_id=1, omega=(0.0, 0.0, omega), origin=(0.0, 0.0, 0.0))
    h = float(r_outer) + 3.0
    half = 0.5 * float(length)
    eps = 1e-4
    dom = bounds(-h, -h, -half - 1.0 - eps, h, h, half + 1.0 + eps)
    outer = geometry.tube(r_outer, 1.5 * h, length + 8.0, axis=2, n_theta=n_theta)
    inner = geometry.rotor(1, r_inner, 1.0, 0.0, length + 8.0, axis=2, n_theta=n_theta, n_axial=1)
    info = {"omega": omega, "nu": float(nu), "r_inner": float(r_inner), "r_outer": float(r_outer), "length": float(length)}
    return c, dom, [outer, inner], info


def rotor_duct(nx=192, ny=96, nz=96, seed=0, levels=2, n_theta=2048, n_axial=48, tip_speed=0.08, u_inlet=0.02,
               nu=0.004):
    """Rotating rotor in a cylindrical duct on a coarse nx*ny*nz grid of 1 mm
    cells (axis along x): the duct is a thick tube wall of inner radius 0.42 ny
    whose outside covers the domain corners; the rotor (N blades, N from the
    seed in 3..6, twist from the seed in 10..40 deg, hub radius 0.2 R, R = 0.8
    times the duct radius, ~2 n_theta (n_axial + 1) triangles) sits at mid
    length.  Velocity inlet at i = 0, non-reflective pressure outlet at
    i = nx - 1, K15, the tube a no-slip IBB wall; with levels = 2 a box around
    the rotor is refined once.  The rotor turns about x at the angular
    velocity that gives tip_speed lattice units (the same on every level under
    acoustic scaling).

    Returns (case, domain, static triangles, rotor triangles, rotor dict)."""
    rng = np.random.default_rng(seed)
    n_blades = int(rng.integers(3, 7))
    twist = float(rng.uniform(10.0, 40.0))
    res, dt = f32(1.0), f32(1.0)
    c = _case(levels, res, dt, f32(1.0))
    c.nuPhys = float(f32(nu) * (res / f32(1000.0)) * (res / f32(1000.0)))
    c.trackWallForce = 1
    c.trackRotorForce = 1
    c.init.kind = T.INIT_UNIFORM
    c.init.ux = float(f32(u_inlet))
    U = float(f32(u_inlet))
    add_open_rule(c, x=T.FACE_HIGH, y=T.FACE_MID, z=T.FACE_MID, dirichletRho=1, dRho=0.0, openBCID=1, nonReflective=1)
    add_open_rule(c, x=T.FACE_LOW, y=T.FACE_MID, z=T.FACE_MID, dirichletU=1, ux=U, openBCID=0, nonReflective=0)
    add_open_rule(c, dirichletU=1, ux=U, openBCID=2, nonReflective=0)
    add_local_rule(c, limiter=0.0)
    add_local_rule(c, wall_id=0, velocity=(0.0, 0.0, 0.0))
    r_duct = 0.42 * ny
    R = 0.8 * r_duct
    length = 0.25 * ny
    cx = 0.5 * nx
    if levels > 1:
        h = 0.5 * length + 6.0
        add_refine_rule(c, 0, (cx - h, -R - 4.0, -R - 4.0, cx + h, R + 4.0, R + 4.0))
    statics = [geometry.tube(r_duct, 0.75 * ny, nx + 20.0, center=(cx, 0.0, 0.0), axis=0, n_theta=512)]
    rotor = geometry.rotor(n_blades, R, 0.2, twist, length, (cx, 0.0, 0.0), axis=0, n_theta=n_theta, n_axial=n_axial)
    omega = float(f32(tip_speed) / (f32(R) * dt / res))
    info = {"blades": n_blades, "twist_deg": twist, "radius": R, "omega": omega, "axis": T.AXIS_X,
            "origin": (cx, 0.0, 0.0), "triangles": len(rotor)}
    return c, bounds(0, -0.5 * ny, -0.5 * nz, nx, 0.5 * ny, 0.5 * nz), statics, rotor, info


def periodic_box(n=24, nu=0.05, u=(0.0, 0.0, 0.0), limiter=0.0):
    """Fully periodic cube of n^3 active cells of 1 mm (one ghost layer per
    face), centred on the origin, uniform initial state, lattice viscosity nu."""
    res = f32(1.0)
    c = _case(1, res, f32(1.0), f32(1.0))
    c.nuPhys = float(f32(nu) * (res / f32(1000.0)) * (res / f32(1000.0)))
    c.trackWallForce = 0
    c.init.kind = T.INIT_UNIFORM
    c.init.ux, c.init.uy, c.init.uz = [float(f32(v)) for v in u]
    add_open_rule(c, periodic=1)
    if limiter is not None:
        add_local_rule(c, limiter=limiter)
    h = f32(0.5) * f32(n)
    eps = f32(1e-6) * f32(n)
    lo, hi = -h - res - eps, h + res + eps
    return c, bounds(lo, lo, lo, hi, hi, hi), []


def channel(nx=64, ny=32, nz=32, force=1e-6, nu=0.1, limiter=0.0, periodic_z=False):
    """Body-force driven duct: periodic in x, no-slip (moment-based u=0) on y and z faces
    (with periodic_z: periodic in z too, i.e. plane Poiseuille flow between the y walls).

    Lattice viscosity `nu` (K17 is singular at nu = 1/6, where omega1 = omega2 = 1).
    K15 by default: like the reference cases, K17 is not used next to moment-based
    boundaries."""
    res = f32(1.0)
    c = _case(1, res, f32(1.0), f32(1.0))
    # lattice viscosity: nu = dt * nu_phys / (res/1000)^2 with dt = 1
    c.dtPhysGlobal = float(f32(1.0))
    c.nuPhys = float(f32(nu) * (res / f32(1000.0)) * (res / f32(1000.0)))
    c.trackWallForce = 0
    add_open_rule(c, y=T.FACE_LOW | T.FACE_HIGH, dirichletU=1, openBCID=0, nonReflective=0)
    if not periodic_z:
        add_open_rule(c, z=T.FACE_LOW | T.FACE_HIGH, dirichletU=1, openBCID=0, nonReflective=0)
    add_open_rule(c, periodic=1)
    add_local_rule(c, force=(force, 0.0, 0.0))
    if limiter is not None:
        add_local_rule(c, limiter=limiter)
    # one ghost layer per periodic direction, walls are the first/last cell layers; where a
    # ghost layer meets a wall face the cell takes the wall rule (edge normal), as the
    # reference's getOpenBC gives the first matching condition
    if periodic_z:
        return c, bounds(-1, 0, -1, nx + 1, ny, nz + 1), []
    return c, bounds(-1, 0, 0, nx + 1, ny, nz), []
