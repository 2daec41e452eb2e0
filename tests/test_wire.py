"""Serving wire format (lbm3_amd.wire) on the host: request checking, the
mapping to case descriptors, result records and the results store."""
import json
import math

import numpy as np
import pytest

from lbm3_amd import ctypes_defs as T
from lbm3_amd import wire

SPHERE = {"id": "s1", "kind": "sphere", "dims": [96, 48, 48], "diameter": 12.0, "reynolds": 100.0, "u": 0.05,
          "steps": 300, "precision": "f32", "seed": 7}


def test_parse_fills_defaults_and_types():
    r = wire.parse_request(json.dumps({"kind": "tgv", "n": 32, "reynolds": 400, "steps": 100}))
    assert r == {"id": "", "kind": "tgv", "steps": 100, "precision": "f32", "n": 32, "reynolds": 400.0}
    s = wire.parse_request(SPHERE)
    assert s["dims"] == (96, 48, 48) and s["seed"] == 7 and isinstance(s["reynolds"], float)
    d = wire.parse_request({k: v for k, v in SPHERE.items() if k != "diameter"})
    assert d["diameter"] == 12.0                        # a quarter of the cross-section


@pytest.mark.parametrize("bad, why", [
    ({"kind": "cube", "steps": 1}, "kind"),
    ({"kind": "tgv", "n": 32, "steps": 10}, "reynolds: missing"),
    ({"kind": "tgv", "n": 32.5, "reynolds": 1.0, "steps": 10}, "integer"),
    ({"kind": "tgv", "n": 32, "reynolds": 1.0, "steps": 0}, "steps"),
    ({"kind": "tgv", "n": 32, "reynolds": 1.0}, "steps: missing"),
    ({"kind": "tgv", "n": 32, "reynolds": float("nan"), "steps": 5}, "reynolds"),
    ({"kind": "tgv", "n": 32, "reynolds": 1.0, "steps": 5, "precision": "f8"}, "precision"),
    ({"kind": "tgv", "n": 32, "reynolds": 1.0, "steps": 5, "colour": 1}, "unknown keys"),
    ({"kind": "channel", "dims": [64, 32], "nu": 0.1, "force": 1e-6, "steps": 5}, "three integers"),
    ({"kind": "channel", "dims": [64, 32, True], "nu": 0.1, "force": 1e-6, "steps": 5}, "integer"),
    (dict(SPHERE, diameter=46.0), "does not fit"),
    (dict(SPHERE, u=0.5), "u:"),
    ("{not json", "not JSON"),
    ("[1, 2]", "JSON object"),
])
def test_parse_refuses(bad, why):
    with pytest.raises(wire.RequestError, match=why):
        wire.parse_request(bad)


def test_to_case_maps_the_descriptors():
    case, dom, tris = wire.to_case(wire.parse_request(SPHERE))
    assert len(tris) == 1 and len(tris[0]) > 1000          # the STL sphere
    assert case.trackWallForce == 1 and case.nOpenRules > 0
    rules = [case.openRules[k] for k in range(case.nOpenRules)]
    assert any(r.bc.dirichletU and abs(r.bc.ux - 0.05) < 1e-7 for r in rules)  # the inlet velocity
    c, _, t = wire.to_case(wire.parse_request({"kind": "channel", "dims": [64, 32, 32], "nu": 0.05, "force": 2e-6,
                                               "steps": 10}))
    assert len(t) == 0
    t2, _, _ = wire.to_case(wire.parse_request({"kind": "tgv", "n": 16, "reynolds": 100.0, "steps": 10}))
    assert t2.init.kind != 0                                # the vortex initial field
    with pytest.raises(wire.RequestError, match="exceeds"):
        wire.to_case(wire.parse_request({"kind": "tgv", "n": 1024, "reynolds": 1.0, "steps": 1}))


def test_record_and_store(tmp_path):
    req = wire.parse_request(SPHERE)
    res = {"wait_s": 0.5, "time_to_solution_s": 2.0, "fluid_updates": 1e9, "kinetic_energy": 3.0, "fx": 600.0,
           "fy": 30.0, "fz": -3.0, "tx": 0.0, "ty": 3.0, "tz": 6.0, "status": 0, "steps": 300}
    rec = wire.record(req, res)
    assert rec["force"] == [2.0, 0.1, -0.01] and rec["torque"] == [0.0, 0.01, 0.02]
    assert math.isclose(rec["drag_coefficient"], -2.0 * 2.0 / (0.05 ** 2 * math.pi * 36.0))   # drag = -Fx
    assert rec["seed"] == 7 and rec["status"] == 0
    store = wire.ResultsStore(tmp_path / "results.jsonl")
    assert store.records() == [] and store.summary() == {"records": 0, "ok": 0}
    times = [1.0, 2.0, 3.0, 4.0]
    for k, t in enumerate(times):
        store.append(dict(rec, id="r%d" % k, time_to_solution_s=t))
    store.append(dict(rec, id="bad", status=-3))
    assert store.get("r2")["time_to_solution_s"] == 3.0 and store.get("none") is None
    sm = store.summary(span_s=10.0)
    assert sm["records"] == 5 and sm["ok"] == 4
    assert sm["p50_s"] == float(np.percentile(times, 50)) and sm["p99_s"] == float(np.percentile(times, 99))
    assert sm["mlups"] == 4e9 / 10.0 / 1e6


def test_the_serving_mix_is_wire_format():
    """Every request of the seeded serving mix carries a spec the wire format
    accepts, and builds from it."""
    from lbm3_amd import serving
    for r in serving.make_requests(n=256, seed=1, steps_scale=0.02):
        spec = wire.parse_request(json.dumps(r["spec"]))
        assert spec["kind"] == r["kind"] and spec["steps"] == r["steps"] and wire.cells_of(spec) == r["cells"]


def test_integer_fields_take_any_python_int_in_range():
    """A seed above 2**53 (not exact in a double) is still an integer; a float
    with no fraction is accepted, one with a fraction is not."""
    big = 2 ** 63 - 1
    assert wire.parse_request(dict(SPHERE, seed=big))["seed"] == big
    assert wire.parse_request(dict(SPHERE, seed=2 ** 53 + 1))["seed"] == 2 ** 53 + 1
    assert wire.parse_request(dict(SPHERE, seed=7.0))["seed"] == 7
    with pytest.raises(wire.RequestError, match="outside"):
        wire.parse_request(dict(SPHERE, seed=2 ** 63))
    with pytest.raises(wire.RequestError, match="integer"):
        wire.parse_request(dict(SPHERE, seed=7.5))
    with pytest.raises(wire.RequestError, match="integer"):
        wire.parse_request(dict(SPHERE, seed=float("inf")))
