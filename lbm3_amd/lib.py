"""ctypes binding of liblbm3mi.so (include/lbm3mi.h).

The library is loaded from the package directory (or from the path in the
LBM3MI_LIB environment variable, e.g. a build with other tuning defines);
there is no fallback: if the HIP library is missing or no GPU is present the
calls fail loudly."""
import ctypes as C
import os

import numpy as np

from . import ctypes_defs as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LBM3MI_LIB") or os.path.join(HERE, "liblbm3mi.so")
HEADER = os.path.join(HERE, "..", "include", "lbm3mi.h")

_lib = None


class LBM3Error(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LBM3Error("%s is not built (run __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32, f64 = C.c_void_p, C.c_int, C.c_longlong, C.c_float, C.c_double
    sig = {
        "lbm3_last_error": (C.c_char_p, []),
        "lbm3_device_info": (i32, [C.c_char_p, i32, C.POINTER(i64), C.POINTER(i32)]),
        "lbm3_sim_create": (i32, [C.POINTER(T.Case), i32, C.POINTER(vp)]),
        "lbm3_sim_destroy": (i32, [vp]),
        "lbm3_build_grids": (i32, [vp, vp, vp, i32, T.Bounds, C.POINTER(i64)]),
        "lbm3_import_level": (i32, [vp, i32, vp]),
        "lbm3_validate_level": (i32, [vp]),
        "lbm3_apply_initial_condition": (i32, [vp, i32]),
        "lbm3_apply_initial_fields": (i32, [vp, i32, vp]),
        "lbm3_update_single_grid": (i32, [vp, i32]),
        "lbm3_update_all_grids": (i32, [vp, i32]),
        "lbm3_run": (i32, [vp, i32]),
        "lbm3_set_watchdog": (i32, [vp, i32]),
        "lbm3_synchronize": (i32, [vp]),
        "lbm3_run_timed": (i32, [vp, i32, i32, C.POINTER(f64)]),
        "lbm3_get_info": (i32, [vp, i32, C.POINTER(T.Info)]),
        "lbm3_level_count": (i32, [vp]),
        "lbm3_download_macros": (i32, [vp, i32, vp]),
        "lbm3_download_populations": (i32, [vp, i32, vp, i64]),
        "lbm3_upload_populations": (i32, [vp, i32, vp, i64]),
        "lbm3_download_array": (i32, [vp, i32, C.c_char_p, vp, i64, C.POINTER(i64)]),
        "lbm3_kinetic_energy": (i32, [vp, C.POINTER(f64)]),
        "lbm3_mass": (i32, [vp, C.POINTER(f64)]),
        "lbm3_wall_forces": (i32, [vp, i32, C.POINTER(f64), i32, C.POINTER(f64)]),
        "lbm3_wall_forces_noalloc": (i32, [vp, i32, C.POINTER(f64), i32, C.POINTER(f64)]),
        "lbm3_autotune": (i32, [vp, i32, vp, i32, C.POINTER(i32)]),
        "lbm3_get_tune": (i32, [vp, vp]),
        "lbm3_set_tune": (i32, [vp, vp]),
        "lbm3_set_hooks": (i32, [vp, vp]),
        "lbm3_set_info_controls": (i32, [vp, i32, C.c_float, C.c_float, i32]),
        "lbm3_refresh_hooks": (i32, [vp]),
        "lbm3_device_array": (i32, [vp, i32, C.c_char_p, C.POINTER(vp), C.POINTER(C.c_longlong)]),
        "lbm3_population_layout": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
        "lbm3_reset_wall_forces": (i32, [vp]),
        "lbm3_open_boundary_stats": (i32, [vp, i32, C.POINTER(f64)]),
        "lbm3_reset_open_boundary_stats": (i32, [vp]),
        "lbm3_section_cut": (i32, [vp, i32, C.c_float, C.POINTER(T.Bounds), vp, i64, C.POINTER(i32)]),
        "lbm3_flow_report": (i32, [vp, i32, C.c_float, C.POINTER(T.Bounds), C.POINTER(f64)]),
        "lbm3_export_vtk": (i32, [vp, C.c_char_p]),
        "lbm3_run_graph": (i32, [vp, i32]),
        "lbm3_save_checkpoint": (i32, [vp, C.c_char_p]),
        "lbm3_load_checkpoint": (i32, [vp, C.c_char_p]),
        "lbm3_update_tracker": (i32, [vp, vp, i32, vp, i32, vp, i32]),
        "lbm3_build_rotors": (i32, [vp, vp, vp, i32, i32]),
        "lbm3_set_rotor": (i32, [vp, i32, i32, C.c_float, i32, C.c_float, C.c_float, C.c_float]),
        "lbm3_rotor_info": (i32, [vp, i32, i32, C.POINTER(f64)]),
        "lbm3_rotor_forces": (i32, [vp, i32, i32, C.POINTER(f64)]),
        "lbm3_copy_probe": (i32, [i64, i32, C.POINTER(f64)]),
        "lbm3_server_create": (i32, [i64, i32, C.POINTER(vp)]),
        "lbm3_server_destroy": (i32, [vp]),
        "lbm3_server_submit": (i32, [vp, C.POINTER(T.Case), i32, vp, vp, i32, T.Bounds, i32, i32, C.POINTER(i64)]),
        "lbm3_server_step": (i32, [vp, C.POINTER(i64), i32, C.POINTER(i32)]),
        "lbm3_server_result": (i32, [vp, i64, C.POINTER(f64), i32]),
        "lbm3_server_macros": (i32, [vp, i64, vp, i64]),
        "lbm3_server_stats": (i32, [vp, C.POINTER(f64)]),
        "lbm3_collide_cells_f32": (i32, [vp, i32, C.POINTER(T.BC), f32]),
        "lbm3_collide_cells_f64": (i32, [vp, i32, C.POINTER(T.BC), f64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise LBM3Error("liblbm3mi error %d: %s" % (rc, load().lbm3_last_error().decode()))


class LevelArrays(C.Structure):
    _fields_ = [("info", T.Info), ("nRows", C.c_int)] + \
               [(n, C.c_void_p) for n in ("shifter", "rowI", "rowJ", "rowK", "rowJPlus", "rowKPlus", "rowJKPlus")] + \
               [("W", C.c_int)] + [(n, C.c_void_p) for n in ("wallMap", "wallIndex", "wallData", "linkLength")] + \
               [("nOpenIDs", C.c_int)] + [(n, C.c_void_p) for n in ("openCount", "openIndex", "openSource")] + \
               [("nPeriodic", C.c_int), ("periodicIndex", C.c_void_p), ("periodicSource", C.c_void_p),
                ("wallCount", C.c_int)] + \
               [("f2cCount", C.c_int), ("f2cIndex", C.c_void_p), ("f2cChild", C.c_void_p),
                ("c2fCount", C.c_int), ("c2fIndex", C.c_void_p), ("c2fChild", C.c_void_p),
                ("leftoverCount", C.c_int)] + \
               [(n, C.c_void_p) for n in ("leftoverIndex", "leftoverParent", "leftoverNbrI", "leftoverNbrJ",
                                          "leftoverNbrK")]

INTERFACE_ARRAYS = ("f2cIndex", "f2cChild", "c2fIndex", "c2fChild", "leftoverIndex", "leftoverParent",
                    "leftoverNbrI", "leftoverNbrJ", "leftoverNbrK")


def level_struct(arrays, info, counts=None):
    """(LevelArrays, kept numpy arrays) from named host arrays; `counts`
    overrides the sizes taken from the arrays (W, nOpenIDs, nPeriodic, ...)."""
    keep = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
    a = LevelArrays()
    a.info = info
    a.nRows = len(keep["rowI"])
    for n in ("shifter", "rowI", "rowJ", "rowK", "rowJPlus", "rowKPlus", "rowJKPlus", "wallMap", "wallIndex",
              "wallData", "linkLength", "openCount", "openIndex", "openSource", "periodicIndex", "periodicSource"):
        v = keep.get(n)
        setattr(a, n, v.ctypes.data if v is not None and v.size else None)
    a.W = len(keep["wallIndex"])
    a.nOpenIDs = len(keep["openCount"])
    a.nPeriodic = len(keep["periodicIndex"])
    a.wallCount = int(np.asarray(keep.get("wallCount", 0)).reshape(-1)[0])
    for n in INTERFACE_ARRAYS:
        v = keep.get(n)
        if v is not None and v.size:
            setattr(a, n, v.ctypes.data)
    a.f2cCount = len(keep.get("f2cIndex", ()))
    a.c2fCount = len(keep.get("c2fIndex", ()))
    a.leftoverCount = len(keep.get("leftoverIndex", ()))
    for k, v in (counts or {}).items():
        setattr(a, k, v)
    return a, keep


def validate_level(arrays, info, counts=None):
    """lbm3_validate_level: the import checks, on the host (no GPU needed)."""
    a, keep = level_struct(arrays, info, counts)
    rc = load().lbm3_validate_level(C.byref(a))
    del keep
    check(rc)


STORE = {"f32": T.STORE_F32, "f16": T.STORE_F16, "bf16": T.STORE_BF16, "f64": T.STORE_F64, "f16c": T.STORE_F16C}
_UINT_ARRAYS = ("wallData", "linkLength", "openMeta")


class Sim:
    """A simulation (all grid levels) resident on the GPU."""

    def __init__(self, case, storage="f32"):
        L = load()
        self.case = case
        self.storage = storage
        h = C.c_void_p()
        check(L.lbm3_sim_create(C.byref(case), STORE[storage], C.byref(h)))
        self.h = h
        self.levels = case.gridLevelCount
        self.fluid_updates = None
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            load().lbm3_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # --- build ---
    def build(self, domain, triangles=()):
        tris = [np.ascontiguousarray(t, np.float32).reshape(-1, 9) for t in triangles]
        allt = np.concatenate(tris) if tris else np.zeros((1, 9), np.float32)
        counts = np.array([len(t) for t in tris] or [0], np.int32)
        out = C.c_longlong()
        check(load().lbm3_build_grids(self.h, allt.ctypes.data, counts.ctypes.data, len(tris), domain, C.byref(out)))
        self.fluid_updates = out.value
        return out.value

    def import_level(self, level, arrays, info):
        """Upload a level from named host arrays (the gridBuilderToGrid layout)."""
        a, keep = level_struct(arrays, info)
        self._keep.append(keep)
        check(load().lbm3_import_level(self.h, level, C.byref(a)))

    # --- run ---
    def init(self, level=None):
        for L in range(self.levels) if level is None else [level]:
            check(load().lbm3_apply_initial_condition(self.h, L))
        self.sync()

    def init_fields(self, level, fields):
        f = np.ascontiguousarray(fields, np.float32)
        check(load().lbm3_apply_initial_fields(self.h, level, f.ctypes.data))

    def run(self, steps):
        check(load().lbm3_run(self.h, steps))

    def set_watchdog(self, period):
        check(load().lbm3_set_watchdog(self.h, period))

    def step_level(self, level):
        check(load().lbm3_update_single_grid(self.h, level))

    def sync(self):
        check(load().lbm3_synchronize(self.h))

    def run_timed(self, steps, kernel_events=False):
        """(total ms, summed level-0 bulk collide-stream kernel ms, launches,
        summed level-0 wall-adjacent collide kernel ms) over `steps` coarse steps."""
        out = (C.c_double * 4)()
        check(load().lbm3_run_timed(self.h, steps, int(kernel_events), out))
        return out[0], out[1], int(out[2]), out[3]

    # --- state ---
    def info(self, level=0):
        i = T.Info()
        check(load().lbm3_get_info(self.h, level, C.byref(i)))
        return i

    def macros(self, level=0):
        n = self.info(level).cellCount
        out = np.zeros((n, 4), np.float64 if self.storage == "f64" else np.float32)
        check(load().lbm3_download_macros(self.h, level, out.ctypes.data))
        return out

    def populations(self, level=0):
        n = self.info(level).cellCount
        dt = {"f32": np.float32, "f64": np.float64, "f16": np.float16, "bf16": np.uint16, "f16c": np.uint16}[self.storage]
        out = np.zeros((27, n), dt)
        check(load().lbm3_download_populations(self.h, level, out.ctypes.data, out.nbytes))
        return out

    def set_populations(self, level, f):
        f = np.ascontiguousarray(f)
        check(load().lbm3_upload_populations(self.h, level, f.ctypes.data, f.nbytes))

    def array(self, name, level=0):
        cnt = C.c_longlong()
        check(load().lbm3_download_array(self.h, level, name.encode(), None, 0, C.byref(cnt)))
        n = max(cnt.value, 0)
        if name in ("rayHitScan", "cellRow"):
            dt = np.int64
        elif name in ("bulkSkip", "openBits"):
            dt = np.uint64
        elif name.startswith("rotor") and name.endswith("Force"):
            dt = np.float64
        elif name in ("wallForce",):
            dt = np.float64
        elif name == "openRefs":
            dt = np.float64 if self.storage == "f64" else np.float32
        elif name in _UINT_ARRAYS or (name.startswith("rotor") and name.endswith("Interp")):
            dt = np.uint32
        else:
            dt = np.int32
        out = np.zeros(n, dt)
        if n:
            check(load().lbm3_download_array(self.h, level, name.encode(), out.ctypes.data, n, C.byref(cnt)))
        return out

    def cell_kinds(self, level=0):
        """Counts of the cells each per-step kernel updates on a level:
        {"bulk", "wall", "open", "rotor", "open_group", "solid", "open_in_bulk"}
        ("bulk": the cells the bulk kernel updates, from its skip mask; open,
        rotor and the fluid cells grouped with open cells have kernels of their
        own, but with open groups of 1 cell the bulk kernel steps the open cells
        it marks too, counted in "open_in_bulk" and in "bulk")."""
        code = (self.array("cellRow", level) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
        n = code.size

        def bits(name):
            w = self.array(name, level)
            return np.unpackbits(w.view(np.uint8), bitorder="little")[:n] if w.size else np.zeros(n, np.uint8)
        return {"bulk": int((bits("bulkSkip") == 0).sum()), "wall": int((code >= 0).sum()),
                "open": int((code <= -8).sum()), "rotor": int((code == -4).sum()), "open_group": int((code == -5).sum()),
                "solid": int((code == -3).sum()), "open_in_bulk": int(bits("openBits").sum())}

    def kinetic_energy(self):
        v = C.c_double()
        check(load().lbm3_kinetic_energy(self.h, C.byref(v)))
        return v.value

    def mass(self):
        """(sum of dRho, counted cells) in FP64, level L weighted by 8^-L."""
        out = (C.c_double * 2)()
        check(load().lbm3_mass(self.h, out))
        return out[0], out[1]

    def autotune(self, steps=10):
        """lbm3_autotune: time every launch choice over `steps` coarse steps,
        keep the fastest, leave the state as it was.  Returns (chosen
        {bulk_lds, xcd_remap, overlap, open_group, open_split, rotor_in_bulk},
        [(lds, remap, overlap, group, split, rotor, ms per step), ...])."""
        out = np.zeros((1024, 7))
        n = C.c_int()
        check(load().lbm3_autotune(self.h, steps, out.ctypes.data, 1024, C.byref(n)))
        return self.tune(), [tuple(r) for r in out[:min(n.value, 1024)]]

    TUNE = ("bulk_lds", "xcd_remap", "overlap", "open_group", "open_split", "rotor_in_bulk")

    def tune(self):
        t = np.zeros(6, np.int32)
        check(load().lbm3_get_tune(self.h, t.ctypes.data))
        return dict(zip(self.TUNE, (int(x) for x in t)))

    def set_tune(self, bulk_lds, xcd_remap, overlap, open_group, open_split=0, rotor_in_bulk=0):
        t = np.array([bulk_lds, xcd_remap, overlap, open_group, open_split, rotor_in_bulk], np.int32)
        check(load().lbm3_set_tune(self.h, t.ctypes.data))

    def wall_forces(self, level=0, origin=(0.0, 0.0, 0.0), bodies=1):
        o = (C.c_double * 3)(*origin)
        out = (C.c_double * (6 * bodies))()
        check(load().lbm3_wall_forces(self.h, level, o, bodies, out))
        return np.array(out[:]).reshape(bodies, 6)

    OPEN_STATS = ("area", "normal_velocity", "mass_flow", "momentum_thrust", "pressure", "pressure_power",
                  "normal_kinetic_power")

    def open_boundary_stats(self, open_bc_id):
        """updateTracker's open-boundary statistics of one boundary since the last
        reset, physical units (area m^2, velocity m/s, pressure Pa, ...)."""
        out = (C.c_double * 7)()
        check(load().lbm3_open_boundary_stats(self.h, open_bc_id, out))
        return dict(zip(self.OPEN_STATS, out[:]))

    def reset_open_boundary_stats(self):
        check(load().lbm3_reset_open_boundary_stats(self.h))

    def update_tracker(self, walls=0, rotors=0, open_boundaries=0):
        """updateTracker for one period, physical units, then the reset:
        {"walls": (walls, 6), "rotors": (rotors, 6), "open": (open_boundaries, 7)}."""
        w = np.zeros((walls, 6))
        r = np.zeros((rotors, 6))
        o = np.zeros((open_boundaries, 7))
        check(load().lbm3_update_tracker(self.h, w.ctypes.data if walls else None, walls,
                                         r.ctypes.data if rotors else None, rotors,
                                         o.ctypes.data if open_boundaries else None, open_boundaries))
        return {"walls": w, "rotors": r, "open": o}

    PLANES = {"xy": 0, "zy": 1, "zx": 2}

    def section_cut(self, plane, cut, bounds=None):
        """exportSectionCutPlot's image of the plane ("xy", "zy", "zx") at `cut`
        mm, cropped to bounds (an lbm3_bounds or None): (vertical, horizontal, 6)
        float32 of {p [Pa], u horizontal, u vertical, u normal [m/s], marker, res [mm]}."""
        pl = self.PLANES[plane]
        b = C.byref(bounds) if bounds is not None else None
        dims = (C.c_int * 2)()
        check(load().lbm3_section_cut(self.h, pl, float(cut), b, None, 0, dims))
        out = np.zeros((dims[0], dims[1], 6), np.float32)
        check(load().lbm3_section_cut(self.h, pl, float(cut), b, out.ctypes.data, out.size, dims))
        return out

    def run_graph(self, steps):
        """run(steps) through a captured HIP graph of two coarse steps; returns
        when the steps are done."""
        check(load().lbm3_run_graph(self.h, int(steps)))

    def save_checkpoint(self, path):
        """The state a time step changes, to a file (restart with load_checkpoint)."""
        check(load().lbm3_save_checkpoint(self.h, str(path).encode()))

    def load_checkpoint(self, path):
        """Restore a save_checkpoint file into a simulation built from the same case and grid."""
        check(load().lbm3_load_checkpoint(self.h, str(path).encode()))

    def export_vtk(self, path):
        """exportVTK: every stored cell of every level to a legacy ASCII VTK file."""
        check(load().lbm3_export_vtk(self.h, str(path).encode()))

    FLOW_REPORT = ("normal_velocity", "mass_flow", "momentum_thrust", "pressure", "pressure_power",
                   "normal_kinetic_power", "area_mm2")

    def flow_report(self, plane, cut, bounds=None):
        """getFlowReport{XY,ZY,ZX}: the section's fluxes in physical units (dict)."""
        out = (C.c_double * 7)()
        b = C.byref(bounds) if bounds is not None else None
        check(load().lbm3_flow_report(self.h, self.PLANES[plane], float(cut), b, out))
        return dict(zip(self.FLOW_REPORT, out[:]))

    def build_rotors(self, rotors, level_mask=0xFF):
        """buildRotors on the levels of level_mask; rotors: list of (n, 9) triangle arrays."""
        tris = [np.ascontiguousarray(t, np.float32).reshape(-1, 9) for t in rotors]
        allt = np.concatenate(tris) if tris else np.zeros((1, 9), np.float32)
        counts = np.array([len(t) for t in tris] or [0], np.int32)
        check(load().lbm3_build_rotors(self.h, allt.ctypes.data, counts.ctypes.data, len(tris), int(level_mask)))

    def set_rotor(self, level, rotor, radians_per_second, axis, origin=(0.0, 0.0, 0.0)):
        check(load().lbm3_set_rotor(self.h, level, rotor, float(radians_per_second), int(axis), *[float(v) for v in origin]))

    def rotor_info(self, level, rotor):
        out = (C.c_double * 12)()
        check(load().lbm3_rotor_info(self.h, level, rotor, out))
        return np.array(out[:])

    def rotor_forces(self, level, rotor):
        out = (C.c_double * 6)()
        check(load().lbm3_rotor_forces(self.h, level, rotor, out))
        return np.array(out[:])

    def reset_wall_forces(self):
        check(load().lbm3_reset_wall_forces(self.h))


def collide_cells(f, bc, nu):
    """applyCollision on the GPU for cell-major populations f[n, 27] (float32 or float64)."""
    f = np.ascontiguousarray(f)
    n = f.shape[0]
    if f.dtype == np.float32:
        check(load().lbm3_collide_cells_f32(f.ctypes.data, n, C.byref(bc), nu))
    else:
        check(load().lbm3_collide_cells_f64(f.ctypes.data, n, C.byref(bc), nu))
    return f


def device_info():
    name = C.create_string_buffer(256)
    mem = C.c_longlong()
    cus = C.c_int()
    check(load().lbm3_device_info(name, 256, C.byref(mem), C.byref(cus)))
    return name.value.decode(), mem.value, cus.value


def exported_symbols_in_header():
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:const char \*|int |void )\s*(lbm3_\w+)\s*\(", text, re.M)))


def copy_probe(nbytes, reps=20):
    """Milliseconds per device-to-device copy of nbytes (nbytes read + nbytes written)."""
    ms = C.c_double()
    check(load().lbm3_copy_probe(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


class Server:
    """Continuous-batching server of single-level simulations (lbm3_server_*)."""

    RESULT_FIELDS = ("wait_s", "time_to_solution_s", "fluid_updates", "kinetic_energy", "fx", "fy", "fz", "tx", "ty",
                     "tz", "status", "steps")

    def __init__(self, hbm_budget_bytes, max_active=64):
        h = C.c_void_p()
        check(load().lbm3_server_create(int(hbm_budget_bytes), int(max_active), C.byref(h)))
        self.h = h
        self._cells = {}

    def close(self):
        if getattr(self, "h", None):
            load().lbm3_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, case, domain, triangles=(), steps=100, storage="f32", keep_macros=False):
        tris = [np.ascontiguousarray(t, np.float32).reshape(-1, 9) for t in triangles]
        allt = np.concatenate(tris) if tris else np.zeros((1, 9), np.float32)
        counts = np.array([len(t) for t in tris] or [0], np.int32)
        rid = C.c_longlong()
        check(load().lbm3_server_submit(self.h, C.byref(case), STORE[storage], allt.ctypes.data, counts.ctypes.data,
                                         len(tris), domain, int(steps), int(keep_macros), C.byref(rid)))
        self._cells[rid.value] = storage
        return rid.value

    def step(self, cap=1024):
        ids = (C.c_longlong * cap)()
        n = C.c_int()
        check(load().lbm3_server_step(self.h, ids, cap, C.byref(n)))
        return [ids[i] for i in range(min(n.value, cap))]

    def result(self, rid):
        out = (C.c_double * 12)()
        check(load().lbm3_server_result(self.h, int(rid), out, 12))
        return dict(zip(self.RESULT_FIELDS, list(out)))

    def macros(self, rid, cells):
        dt = np.float64 if self._cells.get(rid) == "f64" else np.float32
        out = np.zeros((cells, 4), dt)
        check(load().lbm3_server_macros(self.h, int(rid), out.ctypes.data, out.nbytes))
        return out

    def stats(self):
        out = (C.c_double * 5)()
        check(load().lbm3_server_stats(self.h, out))
        return {"active": int(out[0]), "queued": int(out[1]), "bytes": out[2], "steps": int(out[3]),
                "updates": out[4]}
