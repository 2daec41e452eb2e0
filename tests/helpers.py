"""Shared helpers: move an oracle-built level into the GPU import layout, error
norms, the reference-held Taylor-Green kinetic-energy curves."""
import os

import numpy as np


def level_arrays(osim, level=0):
    """Arrays of one oracle-built level in the lbm3_import_level layout."""
    a = {n: osim.array(n, level) for n in ("shifter", "rowI", "rowJ", "rowK", "rowJPlus", "rowKPlus", "rowJKPlus",
                                          "wallMap", "wallIndex", "wallData", "linkLength", "periodicIndex",
                                          "periodicSource")}
    n_open = osim.open_count(level)
    idx = [osim.array("open%dIndex" % b, level) for b in range(n_open)]
    src = [osim.array("open%dSource" % b, level) for b in range(n_open)]
    a["openCount"] = np.array([len(x) for x in idx], np.int32)
    a["openIndex"] = np.concatenate(idx).astype(np.int32) if idx else np.zeros(0, np.int32)
    a["openSource"] = np.concatenate(src).astype(np.int32) if src else np.zeros(0, np.int32)
    a["wallCount"] = np.array(int((a["wallMap"] == -3).sum()))
    if level < osim.levels - 1:
        for n in ("f2cIndex", "f2cChild", "c2fIndex", "c2fChild", "leftoverIndex", "leftoverParent", "leftoverNbrI",
                  "leftoverNbrJ", "leftoverNbrK"):
            a[n] = osim.array(n, level).astype(np.int32)
    return a


def rel_max(a, b, scale=None):
    """max |a - b| / max |b| (relative max-norm)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / s)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- Taylor-Green kinetic-energy curves
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# fixture name -> (active coarse cells per axis, grid levels, iterations between samples)
# (taylorGreenVortexMain.cu:8-13,39-40)
TGV_CURVES = {"coarse": (128, 1, 10), "quadratic": (128, 2, 10), "fine": (256, 1, 20), "linear": (128, 2, 10)}


def tgv_curve(name):
    """(t*, E*) of a reference-held Taylor-Green kinetic-energy curve."""
    d = np.loadtxt(os.path.join(GOLDEN, "tgv_%s_kinetic_energy.txt" % name), delimiter=";", skiprows=1)
    return d[:, 0], d[:, 1]


def tgv_reference_case(name):
    """The reference's Taylor-Green run for a curve: Re = 1600, Ma = 0.1 (lattice
    u0 = Ma / sqrt(3) in FP32), NASA pressure perturbation, periodic, default
    collision (K17, limiter 0.05).  Returns (case, domain, triangles, sample
    iterations, energy scale) with E* = kinetic_energy() * scale."""
    from lbm3_amd import cases
    f32 = np.float32
    n, levels, period = TGV_CURVES[name]
    u0 = float(f32(0.1) * f32(0.5773502691896258))
    case, dom, tris = cases.taylor_green(n, u0=u0, reynolds=1600.0, pressure=True, levels=levels)
    if name == "linear":
        case.interfaceScheme = 1          # INTERFACE_VERSIONS_BACKUP/updateInterfaceLINEAR.h
    dt = f32(case.dtPhysGlobal)
    iterations = int(f32(20.0) / dt + f32(0.5))        # ITERATION_COUNT, taylorGreenVortexMain.cu:34-35
    samples = list(range(0, iterations + 1, period))
    if samples[-1] != iterations:
        samples.append(iterations)
    return case, dom, tris, np.array(samples), 1.0 / (n ** 3 * u0 * u0)


# ---------------------------------------------------------------- cell positions, pipe inlet histories
def cell_ijk(sim, level=0):
    """(i, j, k) of every cell from the compressed IJK arrays (cellFunctions.h:21-94);
    `sim` is a lib.Sim or an OracleSim."""
    sh = sim.array("shifter", level).astype(np.int64)
    ri, rj, rk = (sim.array(n, level).astype(np.int64) for n in ("rowI", "rowJ", "rowK"))
    cell = np.arange(len(sh))
    head = np.where(sh >= 0, cell, cell + sh)
    row = sh[head]
    back = cell - head
    return np.stack([ri[row] + back, rj[row], rk[row]], axis=1)


def open_cells(sim, bcid, level=0):
    """Cells of open boundary `bcid` (the oracle names them open<b>Index, the
    library concatenates every boundary in openIndex)."""
    if hasattr(sim, "open_count"):
        return sim.array("open%dIndex" % bcid, level)
    counts = sim.array("openCount", level)
    start = int(np.sum(counts[:bcid]))
    return sim.array("openIndex", level)[start:start + int(counts[bcid])]


def pipe_inlet_history(name):
    """Rows (iteration, inlet velocity m/s, inlet pressure Pa, relative L2 error of
    the inlet velocity profile) of a reference-held pipe inlet history."""
    return np.loadtxt(os.path.join(GOLDEN, "pipe_inlet_%s.txt" % name), delimiter=";", skiprows=1)


def pipe_reference_case(name):
    """The reference's non-reflective pipe test (nonReflectiveTestMain.cu) with the
    characteristic non-reflective inlet ("characteristic"), the strict Dirichlet
    inlet ("dirichlet") or the impedance scheme at a density tolerance
    ("impedance2e-6")."""
    from lbm3_amd import cases, geometry
    tri = geometry.read_binary_stl(os.path.join(GOLDEN, "pipeD20L200STL.STL"))
    tol = float(name[len("impedance"):]) if name.startswith("impedance") else None
    return cases.pipe(tri, inlet_non_reflective=(name != "dirichlet"), impedance_tolerance=tol)


def pipe_inlet_profile_error(sim, case):
    """getInletProfileError (nonReflectiveTestMain.cu:114-196): relative L2 error of
    uz against the prescribed profile over the counted cells of open boundary 0
    on the k = 0 face, from the current macroscopic fields."""
    f32 = np.float32
    m = sim.macros(0)
    cells = open_cells(sim, 0)
    wm = sim.array("wallMap", 0)[cells]
    wd = sim.array("wallData", 0)
    keep = (wm != -3) & (wm != -2)
    keep &= ~((wm >= 0) & ((wd[np.maximum(wm, 0)] & 1) != 0))
    ijk = cell_ijk(sim, 0)[cells]
    keep &= ijk[:, 2] == 0
    info = sim.info(0)
    x = ijk[:, 0].astype(f32) * f32(info.res) + f32(info.ox)
    y = ijk[:, 1].astype(f32) * f32(info.res) + f32(info.oy)
    rule = case.openRules[0]
    R = f32(rule.profile.radius)
    r2 = x * x + y * y
    target = f32(rule.bc.uz) * np.maximum(f32(0.0), f32(1.0) - r2 / (R * R))
    err = m[cells, 3].astype(f32) - target
    e2 = np.sum((err[keep].astype(np.float64)) ** 2)
    t2 = np.sum((target[keep].astype(np.float64)) ** 2)
    return float(np.sqrt(e2 / t2))


def wall_force_sums(sim, bodies):
    """Per-body sums of the accumulated wall force (on the fluid) and its torque
    about the origin, from the oracle's arrays: (bodies, 6) as lib.Sim.wall_forces
    (lever arms: the wall-adjacent cell centres in mm)."""
    info = sim.info()
    idx = sim.array("wallIndex")
    W = len(idx)
    g = sim.array("wallForce").reshape(3, -1)[:, :W].astype(np.float64)
    ijk = cell_ijk(sim)[idx]
    f32 = np.float32
    p = [(ijk[:, a].astype(f32) * f32(info.res) + f32(o)).astype(np.float64) for a, o in enumerate((info.ox, info.oy, info.oz))]
    body = (sim.array("wallData").astype(np.int64) >> 27) & 31
    t = [p[1] * g[2] - p[2] * g[1], p[2] * g[0] - p[0] * g[2], p[0] * g[1] - p[1] * g[0]]
    out = np.zeros((bodies, 6))
    for b in range(bodies):
        sel = body == b
        out[b] = [g[0][sel].sum(), g[1][sel].sum(), g[2][sel].sum(), t[0][sel].sum(), t[1][sel].sum(), t[2][sel].sum()]
    return out


def centreline_pressure(sim, case):
    """The pipe cases' centre-line pressure on a one-level grid, as
    exportSectionCutPlotZY at x = 0 (exportSectionCutPlot.h:11-330) and
    exportCenterlinePressure.py sample it: the section plane i = (int)((0 - ox)
    / res + 0.5), the two middle rows j = (ny - 1) / 2 and ny / 2 averaged, every
    k, dRho converted to Pa by convertToPhysicalPressure (cellFunctions.h:288-294).
    Returns (z [mm], p [Pa])."""
    f32 = np.float32
    info = sim.info(0)
    ox, res = f32(info.ox), f32(info.res)
    icut = int((f32(0.0) - ox) / res + f32(0.5))
    ny, nz = info.cellCountY, info.cellCountZ
    ijk = cell_ijk(sim, 0)
    sel = np.where(ijk[:, 0] == icut)[0]
    where = {(int(ijk[c, 1]), int(ijk[c, 2])): c for c in sel}
    m = sim.macros(0)
    cs = f32(0.57735026918962576) * (res / f32(1000.0)) / f32(info.dtPhys)
    p = np.zeros(nz)
    for k in range(nz):
        v = [f32(f32(m[where[(j, k)], 0]) * f32(case.rhoPhys) * cs * cs) for j in ((ny - 1) // 2, ny // 2)]
        p[k] = np.mean(np.array(v, np.float32))
    return float(info.oz) + np.arange(nz) * float(res), p


def section_cut_reference(sim, case, plane, cut, lattice=False):
    """exportSectionCutPlotGeneral (include/plotter/exportSectionCutPlot.h:11-330)
    restated on an oracle's macroscopic fields, all levels: returns the
    (V, H, 6) image {p, u horizontal, u vertical, u normal, marker, res} of
    the plane ("xy", "zy", "zx") at `cut` mm, without rotor fractions; with
    lattice=True the image before conversion {dRho, ux, uy, uz, marker, level}
    and the image level."""
    f32 = np.float32
    P = {"xy": 0, "zy": 1, "zx": 2}[plane]
    levels = sim.levels
    infos = [sim.info(L) for L in range(levels)]
    F = infos[-1]
    org = lambda I: f32((I.oz, I.ox, I.oy)[P])
    cnt = lambda I: (I.cellCountZ, I.cellCountX, I.cellCountY)[P]
    cut_index = int((f32(cut) - org(F)) / f32(F.res) + f32(0.5))
    assert 0 <= cut_index < cnt(F)
    cut_pos = float(org(F)) + float(cut_index) * float(f32(F.res))
    H = F.cellCountX if P == 0 else F.cellCountZ
    V = F.cellCountX if P == 2 else F.cellCountY
    image_level = levels - 1
    while H * V > 16000000 and image_level > 0:
        image_level -= 1
        H //= 2
        V //= 2
    img = np.zeros((V, H, 6), np.float32)
    img[..., 4] = 1.0
    for L, I in enumerate(infos):
        ni = (cut_pos - float(org(I))) / float(f32(I.res))
        ni = min(max(ni, 0.0), cnt(I) - 1)
        lower = int(np.floor(ni))
        alpha = f32(ni - lower)
        down = 2 ** (L - image_level) if L > image_level else 1
        up = 2 ** (image_level - L) if L < image_level else 1
        ijk = cell_ijk(sim, L)
        m = sim.macros(L).astype(np.float32)
        wm = sim.array("wallMap", L)
        where = {tuple(c): k for k, c in enumerate(ijk.tolist())}
        unit = [(0, 0, 1), (1, 0, 0), (0, 1, 0)][P]
        for c in np.where(ijk[:, (2, 0, 1)[P]] == lower)[0]:
            i, j, k = ijk[c]
            hc, vc = (i if P == 0 else k), (i if P == 2 else j)
            if hc % down or vc % down:
                continue
            ih, iv = (hc // down) * up, (vc // down) * up
            if ih + up <= 0 or ih >= H or iv + up <= 0 or iv >= V:
                continue
            cu = c
            if alpha > 0:
                cu = where.get((i + unit[0], j + unit[1], k + unit[2]))
                if cu is None:
                    continue
            lower_solid, upper_solid = wm[c] == -3, wm[cu] == -3
            solid = lower_solid if alpha <= 0.5 else upper_solid
            val = np.zeros(4, np.float32)
            if not solid:
                val = m[cu if lower_solid else c].copy()
                if alpha > 0 and not lower_solid and not upper_solid:
                    val = val + alpha * (m[cu] - val)
            img[iv:iv + up, ih:ih + up, :4] = val
            img[iv:iv + up, ih:ih + up, 4] = 1.0 if solid else 0.0
            img[iv:iv + up, ih:ih + up, 5] = L
    if lattice:
        return img, image_level
    out = img.copy()
    for L, I in enumerate(infos):
        sel = img[..., 5] == L
        res, dt = f32(I.res), f32(I.dtPhys)
        ss = f32(0.57735026918962576) * (res / f32(1000.0)) / dt
        vel = lambda v: v * (res / f32(1000.0)) / dt
        ux, uy, uz = vel(img[..., 1]), vel(img[..., 2]), vel(img[..., 3])
        comp = [(ux, uy, uz), (uz, uy, ux), (uz, ux, uy)][P]
        out[..., 0] = np.where(sel, img[..., 0] * f32(case.rhoPhys) * ss * ss, out[..., 0])
        for a in range(3):
            out[..., 1 + a] = np.where(sel, comp[a], out[..., 1 + a])
        out[..., 5] = np.where(sel, res, out[..., 5])
    return out


def flow_report_reference(sim, case, plane, cut):
    """getFlowReportGeneral's reduction (flowReportFunctions.h:266-335) over the
    restated section image: dict as lib.Sim.flow_report."""
    img, level = section_cut_reference(sim, case, plane, cut, lattice=True)
    fluid = img[..., 4] < 0.5
    d = img[..., 0][fluid].astype(np.float64)
    u = img[..., {"xy": 3, "zy": 1, "zx": 2}[plane]][fluid].astype(np.float64)
    n = len(d)
    I = sim.info(level)
    px = float(np.float32(I.res)) / 1000.0
    area, vel = px * px, px / float(np.float32(I.dtPhys))
    pres = float(np.float32(case.rhoPhys)) * vel * vel / 3.0
    mf = float(np.float32(case.rhoPhys)) * area * vel
    rho_u = (1.0 + d) * u
    return {"normal_velocity": u.sum() / n * vel, "mass_flow": (u.sum() + (d * u).sum()) * mf,
            "momentum_thrust": (rho_u * np.abs(u)).sum() * mf * vel, "pressure": d.sum() / n * pres,
            "pressure_power": (d * u).sum() * area * pres * vel, "normal_kinetic_power": (0.5 * rho_u * u * u).sum() * mf * vel * vel,
            "area_mm2": n * float(np.float32(I.res)) ** 2}


def read_vtk(path):
    """Parse the legacy ASCII polydata exportVTK writes (include/exportVTK.h:118-156):
    dict with the header line, points (N,3), vertex pairs (N,2) and every
    point-data array by name."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[2] == "ASCII" and lines[3] == "DATASET POLYDATA"
    out = {"title": lines[1]}
    n = int(lines[4].split()[1])
    assert lines[4] == "POINTS %d float" % n
    at = 5
    take = lambda k, dt: np.array([r.split() for r in lines[at:at + k]], dt).reshape(k, -1)
    out["points"] = take(n, np.float32)
    at += n
    assert lines[at] == "VERTICES %d %d" % (n, 2 * n)
    at += 1
    out["vertices"] = take(n, np.int64)
    at += n
    assert lines[at] == "POINT_DATA %d" % n
    at += 1
    while at < len(lines) and lines[at]:
        head = lines[at].split()
        at += 1
        if head[0] == "SCALARS":
            assert lines[at] == "LOOKUP_TABLE default"
            at += 1
        dt = np.int64 if head[2] == "int" else np.float32
        out[head[1]] = take(n, dt).squeeze(1) if head[0] == "SCALARS" else take(n, dt)
        at += n
    assert at == len(lines) - 1
    return out


def vtk_reference(sim, case):
    """exportVTK's nine fields per cell (include/exportVTK.h:64-112) restated on
    an oracle's macroscopic fields, levels in order: dict of arrays."""
    f32 = np.float32
    parts = {k: [] for k in ("points", "wall", "velocity_m_s", "pressure_Pa", "grid_id", "grid_resolution_mm",
                             "interface_overlap")}
    for L in range(sim.levels):
        I = sim.info(L)
        ijk = cell_ijk(sim, L)
        wm = sim.array("wallMap", L)
        wd = sim.array("wallData", L)
        m = sim.macros(L).astype(f32)
        res, dt = f32(I.res), f32(I.dtPhys)
        parts["points"].append(np.stack([ijk[:, 0].astype(f32) * res + f32(I.ox), ijk[:, 1].astype(f32) * res + f32(I.oy),
                                         ijk[:, 2].astype(f32) * res + f32(I.oz)], 1))
        wall = wm == -3
        wd = np.concatenate([wd, np.zeros(1, wd.dtype)])                   # a level without walls has none
        overlap = ~wall & ((wm == -2) | ((wm >= 0) & ((wd[np.maximum(wm, 0)] & 1) != 0)))
        vel = m[:, 1:4] * (res / f32(1000.0)) / dt
        ss = f32(0.57735026918962576) * (res / f32(1000.0)) / dt
        p = m[:, 0] * f32(case.rhoPhys) * ss * ss
        parts["wall"].append(wall.astype(np.int64))
        parts["velocity_m_s"].append(np.where(wall[:, None], f32(0), vel))
        parts["pressure_Pa"].append(np.where(wall, f32(0), p))
        parts["grid_id"].append(np.full(len(wm), I.gridID, np.int64))
        parts["grid_resolution_mm"].append(np.full(len(wm), res, f32))
        parts["interface_overlap"].append(overlap.astype(np.int64))
    return {k: np.concatenate(v) for k, v in parts.items()}
