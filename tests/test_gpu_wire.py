"""Requests in the serving wire format solved on the GPU server (wire.serve):
one result record per request in the results store, each equal to the same
case solved alone with the single-simulation path (kinetic energy, mean force
and torque, drag coefficient) within the served-vs-alone tolerance."""
import json
import math

import numpy as np
import pytest

from lbm3_amd import lib, wire

pytestmark = pytest.mark.gpu

REQUESTS = [
    {"id": "tgv", "kind": "tgv", "n": 32, "reynolds": 400.0, "steps": 120, "precision": "f32"},
    {"id": "chan", "kind": "channel", "dims": [64, 32, 32], "nu": 0.05, "force": 2e-6, "steps": 150, "precision": "f16"},
    {"id": "sph", "kind": "sphere", "dims": [96, 48, 48], "diameter": 12.0, "reynolds": 100.0, "u": 0.05, "steps": 200,
     "precision": "f32", "seed": 3},
]


def test_served_records_match_the_cases_alone(tmp_path):
    reqs = [wire.parse_request(json.dumps(r)) for r in REQUESTS]
    store = wire.ResultsStore(tmp_path / "results.jsonl")
    span = wire.serve(reqs, store)
    recs = {r["id"]: r for r in store.records()}
    assert sorted(recs) == sorted(r["id"] for r in REQUESTS)
    for req in reqs:
        rec = recs[req["id"]]
        assert rec["status"] == 0 and rec["steps"] == req["steps"] and rec["precision"] == req["precision"]
        case, dom, tris = wire.to_case(req)
        s = lib.Sim(case, req["precision"])
        fu = s.build(dom, tris)
        s.init()
        s.run(req["steps"])
        ke = s.kinetic_energy()
        f = s.wall_forces(0)[0] / req["steps"] if case.trackWallForce else np.zeros(6)
        s.close()
        assert rec["fluid_updates"] == fu * req["steps"]
        assert abs(rec["kinetic_energy"] / ke - 1.0) < 1e-5, req["id"]
        got = np.array(rec["force"] + rec["torque"])
        assert np.max(np.abs(got - f)) <= 1e-5 * max(np.max(np.abs(f)), 1e-30), req["id"]
        if req["kind"] == "sphere":
            cd = -2.0 * f[0] / (req["u"] ** 2 * math.pi * req["diameter"] ** 2 / 4.0)
            assert rec["drag_coefficient"] > 0 and abs(rec["drag_coefficient"] / cd - 1.0) < 1e-5
    sm = store.summary(span_s=span)
    assert sm["ok"] == 3 and sm["p99_s"] >= sm["p50_s"] > 0 and sm["mlups"] > 0


FP64_REQUESTS = [
    {"id": "chan64", "kind": "channel", "dims": [32, 16, 16], "nu": 0.05, "force": 2e-5, "steps": 60, "precision": "f64"},
    {"id": "sph64", "kind": "sphere", "dims": [48, 24, 24], "diameter": 6.0, "reynolds": 100.0, "u": 0.05, "steps": 40,
     "precision": "f64", "seed": 5},
]


def test_served_fp64_records_match_the_oracle(tmp_path):
    """FP64 requests served in a mix: the record's kinetic energy and mean wall
    force equal the CPU oracle's FP64 run of the same case (uniform initial
    states, so both start from the same populations) within 1e-9."""
    from tests.helpers import wall_force_sums
    from tests.oracle_lib import OracleSim
    reqs = [wire.parse_request(json.dumps(r)) for r in FP64_REQUESTS]
    store = wire.ResultsStore(tmp_path / "results.jsonl")
    wire.serve(reqs, store)
    recs = {r["id"]: r for r in store.records()}
    for req in reqs:
        rec = recs[req["id"]]
        assert rec["status"] == 0
        case, dom, tris = wire.to_case(req)
        o = OracleSim(case, dom, tris, double=True)
        o.init()
        o.run(req["steps"])
        ke = o.kinetic_energy()
        print("%s: served KE %.12e, oracle %.12e" % (req["id"], rec["kinetic_energy"], ke))
        assert abs(rec["kinetic_energy"] / ke - 1.0) < 1e-9, req["id"]
        if case.trackWallForce and req["kind"] == "sphere":
            f = wall_force_sums(o, 1)[0] / req["steps"]
            got = np.array(rec["force"] + rec["torque"])
            print("%s: served force %s, oracle %s" % (req["id"], got[:3], f[:3]))
            assert np.max(np.abs(got - f)) <= 1e-9 * np.max(np.abs(f)), req["id"]
