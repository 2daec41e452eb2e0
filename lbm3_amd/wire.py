"""Serving wire format and results store.

A request is one JSON object (one line of a JSON-lines stream):

    {"id": "s1", "kind": "sphere", "dims": [128, 64, 64], "diameter": 16.0,
     "reynolds": 100.0, "u": 0.05, "steps": 500, "precision": "f32", "seed": 7}
    {"id": "t1", "kind": "tgv", "n": 64, "reynolds": 400.0, "steps": 1000}
    {"id": "c1", "kind": "channel", "dims": [64, 32, 32], "nu": 0.05,
     "force": 1e-6, "steps": 800, "precision": "f16"}

`parse_request` checks it (kind, required numbers and their ranges, unknown
keys) and `to_case` maps it to the case descriptors the builder and the step
read (open-boundary, periodic and body-force rules, the STL sphere).  The
server (lbm3_server_*) solves the requests; `record` turns what it returns
into a result record -- status, wait and time to solution, fluid updates,
final kinetic energy, time-averaged force and torque on the body and, for a
sphere, its drag coefficient -- and `ResultsStore` appends the records to a
JSON-lines file and summarises them (p50 / p99 time to solution, MLUPS).  The
reference has no counterpart: its batch runner (run_intakes_automated.py)
solves case programs one after another.  `serve` is the loop that ties them
together: requests in, records out."""
import json
import math
import os
import time

import numpy as np

from . import cases
from . import lib

KINDS = ("tgv", "channel", "sphere")
PRECISIONS = ("f32", "f16", "bf16", "f64", "f16c")
MAX_CELLS = 1 << 30           # one request's grid, cells (fits one device with room to spare)

# kind -> {field: (type, required, low, high)}; low / high inclusive, None: open
_FIELDS = {
    "tgv": {"n": (int, True, 8, 1024), "reynolds": (float, True, 1.0, 1e6), "u0": (float, False, 1e-4, 0.2)},
    "channel": {"dims": (list, True, 4, 4096), "nu": (float, True, 1e-4, 1.0), "force": (float, True, 0.0, 1e-2)},
    "sphere": {"dims": (list, True, 8, 4096), "diameter": (float, False, 2.0, None), "reynolds": (float, True, 1.0, 1e5),
               "u": (float, True, 1e-4, 0.2)},
}
_COMMON = {"id", "kind", "steps", "precision", "seed"}


class RequestError(ValueError):
    """A request that does not follow the wire format."""


def _number(spec, key, typ, low, high):
    v = spec[key]
    # a Python int is an integer at any size; a float only if it has no fraction
    integral = isinstance(v, int) or (isinstance(v, float) and math.isfinite(v) and v.is_integer())
    if isinstance(v, bool) or not isinstance(v, (int, float)) or (typ is int and not integral):
        raise RequestError("%s: %r is not %s" % (key, v, "an integer" if typ is int else "a number"))
    v = typ(v)
    if (typ is float and not math.isfinite(v)) or (low is not None and v < low) or (high is not None and v > high):
        raise RequestError("%s: %r outside [%s, %s]" % (key, v, low, high))
    return v


def parse_request(obj):
    """Check one request (a dict or its JSON text) and return it normalised:
    defaults filled in, numbers typed, dims a tuple of three integers."""
    if isinstance(obj, (str, bytes)):
        try:
            obj = json.loads(obj)
        except json.JSONDecodeError as e:
            raise RequestError("not JSON: %s" % e) from None
    if not isinstance(obj, dict):
        raise RequestError("a request is a JSON object")
    kind = obj.get("kind")
    if kind not in KINDS:
        raise RequestError("kind: %r is not one of %s" % (kind, ", ".join(KINDS)))
    fields = _FIELDS[kind]
    unknown = set(obj) - _COMMON - set(fields)
    if unknown:
        raise RequestError("unknown keys for kind %s: %s" % (kind, ", ".join(sorted(unknown))))
    out = {"id": str(obj.get("id", "")), "kind": kind}
    if "steps" not in obj:
        raise RequestError("steps: missing")
    out["steps"] = _number(obj, "steps", int, 1, 10 ** 8)
    out["precision"] = obj.get("precision", "f32")
    if out["precision"] not in PRECISIONS:
        raise RequestError("precision: %r is not one of %s" % (out["precision"], ", ".join(PRECISIONS)))
    if "seed" in obj:
        out["seed"] = _number(obj, "seed", int, 0, 2 ** 63 - 1)
    for key, (typ, required, low, high) in fields.items():
        if key not in obj:
            if required:
                raise RequestError("%s: missing for kind %s" % (key, kind))
            continue
        if typ is list:
            d = obj[key]
            if not isinstance(d, (list, tuple)) or len(d) != 3:
                raise RequestError("%s: three integers expected" % key)
            out[key] = tuple(_number({key: v}, key, int, low, high) for v in d)
        else:
            out[key] = _number(obj, key, typ, low, high)
    if kind == "sphere":
        nx, ny, nz = out["dims"]
        d = out.get("diameter", ny / 4.0)
        if d > min(ny, nz) - 4:
            raise RequestError("diameter: %r does not fit a %dx%d cross-section" % (d, ny, nz))
        out["diameter"] = float(d)
    return out


def cells_of(req):
    """Stored cells of a request's grid (the count the server budgets)."""
    if req["kind"] == "tgv":
        return (req["n"] + 2) ** 3
    nx, ny, nz = req["dims"]
    return (nx + 2) * ny * nz if req["kind"] == "channel" else nx * ny * nz


def to_case(req):
    """(case, domain, triangles) of a parsed request: the descriptors the
    builder and the step read."""
    if cells_of(req) > MAX_CELLS:
        raise RequestError("grid of %d cells exceeds %d" % (cells_of(req), MAX_CELLS))
    if req["kind"] == "tgv":
        return cases.taylor_green(req["n"], u0=req.get("u0", 0.05), reynolds=req["reynolds"])
    nx, ny, nz = req["dims"]
    if req["kind"] == "channel":
        return cases.channel(nx, ny, nz, force=req["force"], nu=req["nu"])
    return cases.sphere(nx, ny, nz, diameter=req["diameter"], center=(nx / 4.0, ny / 2.0, nz / 2.0), u_inlet=req["u"],
                        reynolds=req["reynolds"], segments=(60, 120))


def record(req, result):
    """Result record of a request from the server's result fields
    (lib.Server.RESULT_FIELDS): forces and torques averaged over the steps,
    lattice units; a sphere's drag coefficient 2 D / (u^2 pi d^2 / 4) with the
    drag D = -<Fx>, the momentum-exchange sum being the force on the fluid
    (flowAroundSphereMain.cu:212-213)."""
    steps = max(int(result["steps"]), 1)
    force = [result[k] / steps for k in ("fx", "fy", "fz")]
    torque = [result[k] / steps for k in ("tx", "ty", "tz")]
    rec = {"id": req["id"], "kind": req["kind"], "precision": req["precision"], "status": int(result["status"]),
           "steps": int(result["steps"]), "fluid_updates": float(result["fluid_updates"]),
           "wait_s": float(result["wait_s"]), "time_to_solution_s": float(result["time_to_solution_s"]),
           "kinetic_energy": float(result["kinetic_energy"]), "force": force, "torque": torque}
    if "seed" in req:
        rec["seed"] = req["seed"]
    if req["kind"] == "sphere":
        area = math.pi * req["diameter"] ** 2 / 4.0
        rec["drag_coefficient"] = -2.0 * force[0] / (req["u"] ** 2 * area)
    return rec


class ResultsStore:
    """Append-only JSON-lines file of result records."""

    def __init__(self, path):
        self.path = os.fspath(path)

    def append(self, rec):
        with open(self.path, "a") as fh:
            fh.write(json.dumps(rec, sort_keys=True) + "\n")

    def records(self):
        if not os.path.exists(self.path):
            return []
        with open(self.path) as fh:
            return [json.loads(line) for line in fh if line.strip()]

    def get(self, rid):
        hits = [r for r in self.records() if r["id"] == str(rid)]
        return hits[-1] if hits else None

    def summary(self, span_s=None):
        """p50 / p99 / mean time to solution over the successful records, their
        fluid updates and, given the wall time they span, MLUPS."""
        ok = [r for r in self.records() if r["status"] == 0]
        out = {"records": len(self.records()), "ok": len(ok)}
        if ok:
            t = np.array([r["time_to_solution_s"] for r in ok])
            out.update(p50_s=float(np.percentile(t, 50)), p99_s=float(np.percentile(t, 99)), mean_s=float(t.mean()),
                       fluid_updates=float(sum(r["fluid_updates"] for r in ok)))
            if span_s:
                out["mlups"] = out["fluid_updates"] / span_s / 1e6
        return out


def serve(requests, store, budget=200 << 30, max_active=64):
    """Solve parsed requests on one server, appending one record per request to
    `store` as each finishes; returns the wall time from the first submit to
    the last record.  Requests are submitted together (the server admits them
    in order as memory and its active limit allow)."""
    srv = lib.Server(budget, max_active)
    try:
        ids = {}
        t0 = time.perf_counter()
        for req in requests:
            case, dom, tris = to_case(req)
            ids[srv.submit(case, dom, tris, req["steps"], req["precision"])] = req
        done = 0
        while done < len(ids):
            for rid in srv.step():
                store.append(record(ids[rid], srv.result(rid)))
                done += 1
        return time.perf_counter() - t0
    finally:
        srv.close()
