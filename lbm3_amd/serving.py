"""Serving workload: Poisson arrivals of a seeded mix of simulation requests,
solved by the continuous-batching server (lbm3_server_*) and, for comparison,
one after another with the single-simulation path.  Reports time to solution
(p50/p99, from arrival to result) and aggregate MLUPS (fluid updates over the
time from the first arrival to the last result).

The mix follows the serving configuration of the hot-path contract:
Taylor-Green vortices of 32^3 to 128^3, body-force channels of 64x32x32 to
256x64x64 and spheres in a 128x64x64 to 192x96x96 box, Reynolds numbers and
500 to 4000 steps drawn from the seed.  Each request carries its wire-format
spec (lbm3_amd.wire), from which it is built."""
import time

import numpy as np

from . import lib
from . import wire

TGV_SIZES = (32, 48, 64, 96, 128)
CHANNEL_SIZES = ((64, 32, 32), (128, 32, 32), (128, 64, 64), (256, 64, 64))
SPHERE_SIZES = ((128, 64, 64), (160, 80, 80), (192, 96, 96))


def make_requests(n=256, seed=0, steps=(500, 4000), steps_scale=1.0, load=0.9, glups=15.0):
    """n request descriptions with arrival times (seconds from the start).

    The mean inter-arrival time is set so that the offered work is `load`
    times what `glups` billion cell updates per second would clear."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        kind = rng.choice(["tgv", "channel", "sphere"])
        st = int(max(10, round(rng.integers(steps[0], steps[1] + 1) * steps_scale)))
        if kind == "tgv":
            size = int(rng.choice(TGV_SIZES))
            p = {"n": size, "reynolds": float(rng.uniform(100.0, 1600.0))}
            cells = (size + 2) ** 3
        elif kind == "channel":
            size = CHANNEL_SIZES[rng.integers(len(CHANNEL_SIZES))]
            p = {"dims": tuple(int(v) for v in size), "nu": float(rng.uniform(0.02, 0.1)),
                 "force": float(rng.uniform(1e-6, 1e-5))}
            cells = (size[0] + 2) * size[1] * size[2]
        else:
            size = SPHERE_SIZES[rng.integers(len(SPHERE_SIZES))]
            d = size[1] / 4.0
            p = {"dims": tuple(int(v) for v in size), "diameter": d, "reynolds": float(rng.uniform(50.0, 300.0)),
                 "u": float(rng.uniform(0.03, 0.06))}
            cells = size[0] * size[1] * size[2]
        spec = dict(p, id=str(i), kind=str(kind), steps=st, precision="f32", seed=int(seed))
        if "dims" in spec:
            spec["dims"] = list(spec["dims"])
        reqs.append({"id": i, "kind": kind, "params": p, "steps": st, "cells": cells, "spec": spec})
    work = sum(r["cells"] * r["steps"] for r in reqs)
    mean_gap = work / n / (glups * 1e9) / load
    t = np.cumsum(rng.exponential(mean_gap, n))
    for r, a in zip(reqs, t - t[0]):
        r["arrival"] = float(a)
    return reqs


def build(req):
    """(case, domain, triangles) of a request, through its wire-format spec
    (lbm3_amd.wire: checked, then mapped to the case descriptors)."""
    return wire.to_case(wire.parse_request(req["spec"]))


def _summary(tts, updates, span):
    tts = np.asarray(tts)
    return {"requests": int(len(tts)), "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
            "mean_s": float(tts.mean()), "makespan_s": float(span), "mlups": float(updates / span / 1e6)}


def run_served(reqs, storage="f32", budget=200 << 30, max_active=64):
    """Solve on the continuous-batching server in real time: each request is
    submitted when its arrival time has passed."""
    geo = [build(r) for r in reqs]          # host-side case data, prepared before the clock starts
    srv = lib.Server(budget, max_active)
    pending = list(range(len(reqs)))
    submitted = {}
    tts, updates = {}, 0.0
    t0 = time.perf_counter()
    while len(tts) < len(reqs):
        now = time.perf_counter() - t0
        while pending and reqs[pending[0]]["arrival"] <= now:
            i = pending.pop(0)
            case, dom, tris = geo[i]
            rid = srv.submit(case, dom, tris, reqs[i]["steps"], storage)
            submitted[rid] = (i, time.perf_counter() - t0)
        st = srv.stats()
        if st["active"] == 0 and st["queued"] == 0:
            if pending:
                time.sleep(max(0.0, reqs[pending[0]]["arrival"] - (time.perf_counter() - t0)))
            continue
        for rid in srv.step():
            i, tsub = submitted[rid]
            res = srv.result(rid)
            if res["status"] != 0:
                raise RuntimeError("request %d failed with status %d" % (i, res["status"]))
            # from the arrival time: the submit happens when the loop next looks
            tts[i] = res["time_to_solution_s"] + (tsub - reqs[i]["arrival"])
            updates += res["fluid_updates"]
    span = time.perf_counter() - t0
    srv.close()
    return _summary([tts[i] for i in range(len(reqs))], updates, span)


def run_sequential(reqs, storage="f32"):
    """One request at a time in arrival order with the single-simulation path."""
    geo = [build(r) for r in reqs]
    tts, updates = [], 0.0
    t0 = time.perf_counter()
    for r, (case, dom, tris) in zip(reqs, geo):
        now = time.perf_counter() - t0
        if r["arrival"] > now:
            time.sleep(r["arrival"] - now)
        sim = lib.Sim(case, storage)
        fu = sim.build(dom, tris)
        sim.init()
        sim.run(r["steps"])
        sim.kinetic_energy()
        sim.close()
        tts.append(time.perf_counter() - t0 - r["arrival"])
        updates += fu * r["steps"]
    span = time.perf_counter() - t0
    return _summary(tts, updates, span)
