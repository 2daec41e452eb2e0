#!/usr/bin/env python3
"""Headline benchmark of the LBM3 hot path on one MI355X.

Workload: flow around a sphere on a uniform 256x128x128 grid (D = 32 cells at
(64,64,64), inlet 0.05, Re = 100, velocity inlet, non-reflective pressure
outlet, far-field sides, interpolated bounceback on the STL sphere), synthetic
data.  The collision follows the reference sphere case: K15 on level 0 (the
only level here); the K17 (limiter 0.1) kernel is reported as a variant.  A step is one time step of updateAllGrids over the
whole grid.  `value` is MLUPS in the reference's own count
(sum over levels of (cells - walls) * 2^level per step) for FP32 population
storage; FP16 / BF16 shifted storage are reported as variants with their
velocity error against an FP64 run of the same steps.

The roofline entry is the fused bulk collide-stream kernel (every fluid cell
that is not next to a wall): algorithmic bytes per launch (bulk cells x 2 x 27
x bytes per population) over its average duration, measured with HIP events
around each launch on the simulation's stream.  The wall-adjacent cells run in
a second, much smaller kernel whose time is reported beside it.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
SBYTES = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8, "f16c": 2}
# cells of the sphere workload the bulk kernel updates with the open cells inside
BULK_CELLS_SPHERE = 4171784


def workload(k17=False):
    from lbm3_amd import cases
    return cases.sphere(256, 128, 128, diameter=32.0, center=(64.0, 64.0, 64.0), u_inlet=0.05, reynolds=100.0,
                        limiter=0.1, k15_on_level0=not k17)


def measure(storage, steps, warmup, k17=False):
    from lbm3_amd import lib
    case, dom, tris = workload(k17)
    sim = lib.Sim(case, storage)
    t0 = time.time()
    fluid_updates = sim.build(dom, tris)
    build_s = time.time() - t0
    info = sim.info(0)
    walls = int((sim.array("wallMap") == -3).sum())
    wall_adjacent = int(sim.array("wallIndex").size)
    sim.init()
    # launch choices picked by timing them on this box (lbm3_autotune; the state is kept)
    tuned, table = sim.autotune(steps=10)
    sim.run(warmup)
    sim.sync()
    total_ms, _, _, _ = sim.run_timed(steps, kernel_events=False)
    _, kernel_ms, launches, wall_ms = sim.run_timed(steps, kernel_events=True)
    kinds = sim.cell_kinds(0)          # after the first run: open cells may have their own kernel
    bulk_cells = kinds["bulk"]
    bytes_per_launch = bulk_cells * 2 * 27 * SBYTES[storage]
    avg_kernel_s = kernel_ms / launches / 1e3
    achieved = bytes_per_launch / avg_kernel_s / 1e9
    out = {
        "storage": storage,
        "mlups": fluid_updates * steps / (total_ms / 1e3) / 1e6,
        "ms_per_step": total_ms / steps,
        "kernel_us": avg_kernel_s * 1e6,
        "boundary_kernels_us": wall_ms / launches * 1e3,
        "kernel_mlups": bulk_cells / avg_kernel_s / 1e6,
        "bulk_cells": bulk_cells, "wall_adjacent": wall_adjacent, "open_cells": kinds["open"],
        "open_group_cells": kinds["open_group"], "open_cells_in_bulk_kernel": kinds["open_in_bulk"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None},
        "cells": info.cellCount, "walls": walls, "fluid_updates_per_step": fluid_updates,
        "bytes_per_launch": bytes_per_launch, "build_s": round(build_s, 3),
        "steps_done": warmup + 2 * steps,
        "tune": {"chosen": tuned, "candidates": len(table),
                 "ms_per_step_best": round(min(r[-1] for r in table), 4),
                 "ms_per_step_worst": round(max(r[-1] for r in table), 4),
                 # fastest time of each layout of the open and rotor cells (group, x-face choice, rotor)
                 "layouts": {"%d/%d/%d" % (g, sp, ro): round(min(r[-1] for r in table if (r[3], r[4], r[5]) == (g, sp, ro)), 4)
                             for g, sp, ro in sorted({(int(r[3]), int(r[4]), int(r[5])) for r in table})}},
    }
    u = sim.macros(0)[:, 1:].astype(np.float64)
    u[sim.array("wallMap") == -3] = np.nan     # solid cells hold no flow
    sim.close()
    return out, u


def sphere_run(storage, steps, reynolds=100.0, window=1000):
    """The sphere workload for `steps` steps: velocity field and the drag
    coefficient C_D = 2 F_drag / (rho U^2 pi D^2 / 4) from the FP64
    momentum-exchange force averaged over the last `window` steps
    (flowAroundSphereMain.cu:212-213)."""
    from lbm3_amd import cases, lib
    case, dom, tris = cases.sphere(256, 128, 128, diameter=32.0, center=(64.0, 64.0, 64.0), u_inlet=0.05,
                                   reynolds=reynolds, limiter=0.1, k15_on_level0=True)
    s = lib.Sim(case, storage)
    s.build(dom, tris)
    s.init()
    s.run(steps - window)
    s.reset_wall_forces()
    s.run(window)
    s.sync()
    fx = s.wall_forces(0)[0][0] / window
    u = s.macros(0)[:, 1:].astype(np.float64)
    u[s.array("wallMap") == -3] = np.nan
    s.close()
    cd = -2.0 * fx / (0.05 ** 2 * np.pi * 32.0 ** 2 / 4.0)    # the force on the fluid is minus the drag
    return u, float(cd)


def accuracy_runs(steps, storages=("f32", "f16", "bf16", "f16c")):
    """Relative L2 velocity error of each storage precision against the FP64
    run of the same kernels after `steps` steps of the sphere workload, and the
    FP64 drag coefficient at Re 100 and 300 (Schiller-Naumann: 1.09 and 0.69
    in free stream; here the sides hold the inlet velocity and the sphere
    blocks 4.9% of the section)."""
    u64, cd100 = sphere_run("f64", steps)
    mask = np.isfinite(u64).all(axis=1)
    out = {"steps": steps, "drag_coefficient_f64": {"re100": round(cd100, 5)}}
    for st in storages:
        u, cd = sphere_run(st, steps)
        m = mask & np.isfinite(u).all(axis=1)
        out["%s_rel_l2_u_vs_f64" % st] = float(np.linalg.norm((u - u64)[m]) / np.linalg.norm(u64[m]))
        out["drag_coefficient_%s" % st] = round(cd, 5)
    _, cd300 = sphere_run("f64", steps, reynolds=300.0)
    out["drag_coefficient_f64"]["re300"] = round(cd300, 5)
    print("accuracy: %s" % out, file=sys.stderr, flush=True)
    return out


def measure_multilevel(storage, steps=200, warmup=20):
    """Three-level sphere (coarse 128x64x64, two nested factor-2 boxes, D = 24
    finest cells, seeded inlet velocity): MLUPS of whole coarse steps, in the
    reference's count (each level's cells times 2^level)."""
    from lbm3_amd import cases, lib
    case, dom, tris = cases.multilevel_sphere(128, 64, 64, diameter_fine=24.0, seed=0)
    sim = lib.Sim(case, storage)
    t0 = time.time()
    fu = sim.build(dom, tris)
    build_s = time.time() - t0
    sim.init()
    sim.run(warmup)
    tuned, _ = sim.autotune(steps=5)
    sim.sync()
    total_ms = sim.run_timed(steps)[0]
    out = {"storage": storage, "mlups": round(fu * steps / (total_ms / 1e3) / 1e6, 1),
           "ms_per_coarse_step": round(total_ms / steps, 4), "fluid_updates_per_coarse_step": fu,
           "cells_per_level": [sim.info(L).cellCount for L in range(3)], "build_s": round(build_s, 3), "tune": tuned}
    sim.close()
    return out


def measure_graph(storage, n=48, steps=400, warmup=20):
    """Launch-bound case: two-level periodic Taylor-Green vortex (n^3 coarse,
    the inner half refined), whole coarse steps launched one by one (lbm3_run)
    and replayed from a captured HIP graph (lbm3_run_graph); wall time around
    each, synchronized."""
    from lbm3_amd import cases, lib
    case, dom, tris = cases.taylor_green(n, levels=2)
    sim = lib.Sim(case, storage)
    fu = sim.build(dom, tris)
    sim.init()
    sim.run(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    eager = time.perf_counter() - t0
    sim.run_graph(2)
    t0 = time.perf_counter()
    sim.run_graph(steps)
    graph = time.perf_counter() - t0
    out = {"case": "taylor_green %d^3, 2 levels" % n, "storage": storage, "fluid_updates_per_coarse_step": fu,
           "eager_ms_per_coarse_step": round(eager / steps * 1e3, 4), "graph_ms_per_coarse_step": round(graph / steps * 1e3, 4),
           "eager_mlups": round(fu * steps / eager / 1e6, 1), "graph_mlups": round(fu * steps / graph / 1e6, 1),
           "speedup": round(eager / graph, 3)}
    sim.close()
    return out


def measure_rotor(storage, windows=2, window=100, warmup=20):
    """Rotor in a duct (coarse 192x96x96, two levels, seeded N-blade rotor of
    ~200k triangles voxelized on the GPU, tip speed 0.08): MLUPS of whole
    coarse steps and the FP64 rotor torque about the axis, averaged over each
    window of coarse steps (per fine step of the rotor's level)."""
    from lbm3_amd import cases, lib
    case, dom, statics, rotor, ri = cases.rotor_duct(192, 96, 96, seed=0, levels=2)
    sim = lib.Sim(case, storage)
    t0 = time.time()
    fu = sim.build(dom, statics)
    sim.build_rotors([rotor], 3)
    for L in range(2):
        sim.set_rotor(L, 0, ri["omega"], ri["axis"], ri["origin"])
    build_s = time.time() - t0
    sim.init()
    sim.run(warmup)
    tuned, _ = sim.autotune(steps=5)
    sim.sync()
    torque, ms = [], 0.0
    for _ in range(windows):
        sim.reset_wall_forces()
        ms += sim.run_timed(window)[0]
        torque.append(float(sim.rotor_forces(1, 0)[3] / (2 * window)))
    kinds = sim.cell_kinds(1)
    out = {"storage": storage, "mlups": round(fu * windows * window / (ms / 1e3) / 1e6, 1),
           "ms_per_coarse_step": round(ms / (windows * window), 4), "fluid_updates_per_coarse_step": fu,
           "rotor": {"blades": ri["blades"], "twist_deg": round(ri["twist_deg"], 2), "triangles": ri["triangles"],
                     "rotor_cells_fine": kinds["rotor"]},
           "torque_x_per_window": torque, "window_coarse_steps": window, "build_s": round(build_s, 3), "tune": tuned}
    sim.close()
    return out


def _cpu_sample(seconds, batch=10, max_steps=400):
    """Steps of the CPU oracle on the sphere workload until `seconds` of timed
    work: (MLUPS, steps, seconds)."""
    from tests.oracle_lib import OracleSim
    case, dom, tris = workload()
    o = OracleSim(case, dom, tris)
    o.init()
    o.run(1)
    steps = 0
    t0 = time.time()
    while steps < max_steps and (steps == 0 or time.time() - t0 < seconds):
        o.run(batch)
        steps += batch
    dt = time.time() - t0
    return o.fluid_updates * steps / dt / 1e6, steps, dt


def _host_cpu():
    """lscpu's model name, sockets, cores per socket, threads per core, NUMA nodes."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    for line in out.splitlines():
        k, _, v = line.partition(":")
        info[k.strip()] = v.strip()
    res = {"model": info.get("Model name")}
    try:
        res["physical_cores"] = int(info["Socket(s)"]) * int(info["Core(s) per socket"])
        res["threads_per_core"] = int(info["Thread(s) per core"])
    except (KeyError, ValueError):
        pass
    if "NUMA node(s)" in info:
        res["numa_nodes"] = info["NUMA node(s)"]
    return res


def native_oracle():
    """The oracle built for this host (-O3 -march=native, contraction off as in
    the parity build): oracle/_build/liboracle_native.so, made here because the
    build container's CPU differs from the GPU box's host."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                   capture_output=True, text=True)
    return os.path.join(ROOT, "oracle", "_build", "liboracle_native.so")


def cpu_baseline(threads, seconds=4.0, runs=3, single_seconds=6.0):
    """The CPU oracle (C, OpenMP over cells, -O3 -march=native) on the same
    sphere workload: the median of `runs` bounded samples with `threads`
    threads pinned close, and one single-thread sample in a child process."""
    lib_path = native_oracle()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from tests import oracle_lib
    oracle_lib.LIB = lib_path
    samples = [_cpu_sample(seconds) for _ in range(runs)]
    mlups = sorted(s[0] for s in samples)[len(samples) // 2]
    env = dict(os.environ, OMP_NUM_THREADS="1", LBM3_ORACLE_LIB=lib_path)
    code = ("import sys, json; sys.path.insert(0, %r); from tests import oracle_lib; import os; "
            "oracle_lib.LIB = os.environ['LBM3_ORACLE_LIB']; import bench; "
            "print(json.dumps(bench._cpu_sample(%r, batch=1, max_steps=50)))" % (ROOT, single_seconds))
    single = None
    try:
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
        if r.returncode == 0:
            v, st, dt = json.loads(r.stdout.strip().splitlines()[-1])
            single = {"value": round(v, 3), "steps": st, "seconds": round(dt, 1)}
    except (subprocess.SubprocessError, ValueError, IndexError):
        single = None
    steps = sum(s[1] for s in samples)
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "oracle (C restatement, -O3 -march=native, OpenMP over cells, %d threads, OMP_PROC_BIND=close) "
                      "on the full 256x128x128 sphere grid: median of %d runs of 10-step batches for about %.0f s each "
                      "(%d steps in all) after 1 untimed step" % (threads, runs, seconds, steps),
            "runs_mlups": [round(s[0], 3) for s in samples], "single_thread": single, "host": _host_cpu()}


def traffic_from_profiles(storage, bulk_cells, tune):
    """Measured HBM bytes per launch of the kernel the tuned step runs, if the
    counters were taken on the same cell split and kernel: open cells in groups
    (key `storage`), in the bulk kernel (`storage`_in_bulk), with the block's
    lanes rotated (`_in_bulk_rotated`), or with the wall cells' blocks in the
    same launch (`_in_bulk_walls`: side-stream choice 3 with the open cells
    inside and the lanes in order)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_sphere.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    if tune["open_group"] != 1:
        key = storage
    elif tune["open_split"] == 3:
        key = storage + "_in_bulk_rotated"
    elif tune["overlap"] == 3:
        key = storage + "_in_bulk_walls"
    else:
        key = storage + "_in_bulk"
    e = d.get(key)
    if e is None or e.get("cells") != bulk_cells:
        return None
    b = e.get("by_block_order", {}).get(str(tune["xcd_remap"]))
    return b["bytes_per_launch"] if b else e.get("bytes_per_launch")


# launch layouts the counter passes measure: (key suffix, side-stream choice,
# open group width, x-face choice); each at every block order
COUNTER_LAYOUTS = {
    "f32": (("_in_bulk_walls", 3, 1, 0), ("_in_bulk", 1, 1, 0), ("_in_bulk_rotated", 1, 1, 3)),
    "f16": (("_in_bulk_walls", 3, 1, 0), ("_in_bulk", 1, 1, 0), ("_in_bulk_rotated", 1, 1, 3)),
    "bf16": (("_in_bulk_walls", 3, 1, 0), ("_in_bulk", 1, 1, 0), ("_in_bulk_rotated", 1, 1, 3)),
}
COUNTER_STEPS = 10


def counter_segments():
    """The measured blocks of the counters workload, in order:
    (storage, key, block order)."""
    return [(st, st + suffix, remap) for st, layouts in COUNTER_LAYOUTS.items()
            for suffix, _, _, _ in layouts for remap in range(4)]


def counters_workload():
    """A short sequence for rocprofv3 --pmc runs: the sphere workload in FP32,
    FP16 and BF16, each launch layout of COUNTER_LAYOUTS at every block order:
    3 unmeasured steps, then COUNTER_STEPS steps between two 4 KiB copies (the
    markers around a measured block); at the end three 1 GiB copies, the
    calibration kernel whose bytes are known."""
    from lbm3_amd import lib
    for st, layouts in COUNTER_LAYOUTS.items():
        case, dom, tris = workload()
        s = lib.Sim(case, st)
        s.build(dom, tris)
        s.init()
        s.run(2)
        for suffix, overlap, group, split in layouts:
            for remap in range(4):
                s.set_tune(0, remap, overlap, group, split, 0)
                s.run(3)
                s.sync()
                lib.copy_probe(4096, 1)
                s.run(COUNTER_STEPS)
                s.sync()
                lib.copy_probe(4096, 1)
                k = s.cell_kinds(0)
                print("counters workload: %s%s block order %d: bulk cells %d (open cells in it %d)"
                      % (st, suffix, remap, k["bulk"], k["open_in_bulk"]), flush=True)
        s.close()
    lib.copy_probe(1 << 30, 3)
    print("counters workload: 3 copies of 1 GiB", flush=True)


def _counter_dispatches(directory, counter):
    """[(dispatch id, kernel name, counter bytes)] in launch order from a
    rocprofv3 --pmc csv directory (the counter reads KiB, summed over its
    instances)."""
    import csv
    import glob
    files = glob.glob(os.path.join(directory, "*counter_collection.csv"))
    if len(files) != 1:
        raise SystemExit("expected one counter_collection.csv in %s, found %d" % (directory, len(files)))
    per, names = {}, {}
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] != counter:
                continue
            d = int(r["Dispatch_Id"])
            per[d] = per.get(d, 0.0) + float(r["Counter_Value"]) * 1024.0
            names[d] = r["Kernel_Name"]
    return [(d, names[d], per[d]) for d in sorted(per)]


def _segment_means(dispatches):
    """The runs of dispatches between runs of k_copy dispatches, in order, as
    the mean counter bytes per dispatch of each kernel of the run (what comes
    before the first copy is left out); and the last three k_copy dispatches
    (the 1 GiB calibration copies)."""
    segs, cur, copies = [], None, []
    for _, name, v in dispatches:
        if "k_copy" in name:
            copies.append(v)
            if cur:
                segs.append(cur)
            cur = {}
            continue
        if cur is not None:
            cur.setdefault(name, []).append(v)
    return [{k: sum(v) / len(v) for k, v in seg.items()} for seg in segs], copies[-3:]


def counters_summary(fetch_dir, write_dir, out_path):
    """HBM bytes per launch of each measured layout from the FETCH_SIZE and
    WRITE_SIZE passes of the counters workload, scaled by what each counter
    reads for the 1 GiB copies; written as JSON to out_path."""
    fe, fcal = _segment_means(_counter_dispatches(fetch_dir, "FETCH_SIZE"))
    wr, wcal = _segment_means(_counter_dispatches(write_dir, "WRITE_SIZE"))
    order = counter_segments()
    # measured blocks alternate with the unmeasured steps between them; the
    # calibration's own fill may add one run after the last block
    n = 2 * len(order) - 1
    if not (n <= len(fe) <= n + 1 and n <= len(wr) <= n + 1) or len(fcal) != 3 or len(wcal) != 3:
        raise SystemExit("counter passes do not match the workload: %d / %d runs between copies for %d blocks"
                         % (len(fe), len(wr), len(order)))
    fe, wr = fe[:n:2], wr[:n:2]
    gib = float(1 << 30)
    ff, fw = gib / (sum(fcal) / 3), gib / (sum(wcal) / 3)
    out = {"calibration": {"kernel": "k_copy (1 GiB read + 1 GiB written)", "fetch_factor": round(ff, 3),
                           "write_factor": round(fw, 3)},
           "workload": "bench.py --counters-workload: each layout at each block order, %d steps, "
                       "open cells completed inside the bulk kernel, LDS cap 0" % COUNTER_STEPS}
    for (st, key, remap), f, w in zip(order, fe, wr):
        # the step kernel of the block: the one moving the most bytes
        name = max(f, key=lambda k: f[k])
        fb, wb = f[name] * ff, w.get(name, 0.0) * fw
        e = out.setdefault(key, {"kernel": name.split("(")[0].replace("void ", ""), "by_block_order": {}})
        e["by_block_order"][str(remap)] = {"fetch_bytes": int(round(fb)), "write_bytes": int(round(wb)),
                                           "bytes_per_launch": int(round(fb + wb))}
    sb = {"f32": 4, "f16": 2, "bf16": 2}
    for st, key, _ in order:
        e = out[key]
        e["cells"] = BULK_CELLS_SPHERE
        walls = 5264 if key.endswith("_walls") else 0
        e["wall_cells"] = walls
        e["algorithmic_population_bytes"] = (BULK_CELLS_SPHERE + walls) * 2 * 27 * sb[st]
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    for key in dict.fromkeys(k for _, k, _ in order):
        e = out[key]
        print("%-22s %s" % (key, "  ".join("order %s: %.1f MB (%.3f)" % (r, v["bytes_per_launch"] / 1e6,
              v["bytes_per_launch"] / e["algorithmic_population_bytes"]) for r, v in sorted(e["by_block_order"].items()))))


def serving(n, steps_scale, storages, seed=0):
    """Serving workload (continuous batching vs one-at-a-time solves) on the
    same seeded Poisson arrival sequence; one entry per storage type."""
    from lbm3_amd import serving as sv
    out = {"requests": n, "seed": seed, "steps_scale": steps_scale}
    for st in storages:
        reqs = sv.make_requests(n=n, seed=seed, steps_scale=steps_scale)
        served = sv.run_served(reqs, storage=st)
        print("serving %s: served %s" % (st, served), file=sys.stderr, flush=True)
        seq = sv.run_sequential(reqs, storage=st)
        print("serving %s: sequential %s" % (st, seq), file=sys.stderr, flush=True)
        out[st] = {"served": {k: round(v, 4) for k, v in served.items()},
                   "sequential": {k: round(v, 4) for k, v in seq.items()},
                   "throughput_ratio": round(served["mlups"] / seq["mlups"], 3)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--variants", default="f16,bf16,f16c,k17")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=5000,
                    help="steps of the sphere runs compared with FP64 (velocity error, drag coefficient)")
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="CPU-baseline threads (default: every CPU this process may run on)")
    ap.add_argument("--serving", choices=["quick", "full", "none"], default="full",
                    help="serving workload: quick (32 requests, FP32), full (256 requests, "
                         "500-4000 steps, FP32 and FP16), none")
    ap.add_argument("--no-other-cases", action="store_true",
                    help="skip the multilevel-sphere and rotor-duct MLUPS lines")
    ap.add_argument("--counters-workload", action="store_true",
                    help="run only the short sequence that the HBM byte counters are collected on")
    ap.add_argument("--counters-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON"),
                    help="HBM bytes per launch from the FETCH_SIZE and WRITE_SIZE passes of --counters-workload")
    ap.add_argument("--case", choices=["multilevel", "rotor", "graph"],
                    help="measure only this configuration (FP32) and print its JSON object")
    args = ap.parse_args()
    if args.counters_workload:
        counters_workload()
        return
    if args.counters_summary:
        counters_summary(*args.counters_summary)
        return
    if args.case:
        run = {"multilevel": lambda: measure_multilevel("f32", steps=args.steps, warmup=args.warmup),
               "rotor": lambda: measure_rotor("f32", warmup=args.warmup),
               "graph": lambda: measure_graph("f32", steps=args.steps, warmup=args.warmup)}[args.case]
        print(json.dumps(run()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from lbm3_amd import lib
    name, mem, cus = lib.device_info()
    main_run, _ = measure("f32", args.steps, args.warmup)
    variants = {}
    for v in [x for x in args.variants.split(",") if x]:
        if v == "k17":   # FP32 storage, K17 with limiter 0.1 in the bulk (a different flow: no error vs FP64 here)
            r, _ = measure("f32", args.steps, args.warmup, k17=True)
            variants[v] = r
            continue
        r, _ = measure(v, args.steps, args.warmup)
        variants[v] = r
    accuracy = None if args.no_accuracy else accuracy_runs(args.accuracy_steps)
    copy_bytes = 1 << 30
    copy_ms = lib.copy_probe(copy_bytes, 20)
    # measured HBM bytes of the K15 bulk kernel (the K17 variant has no counter run of its own)
    for run in [main_run] + [v for k, v in variants.items() if k != "k17"]:
        run["roofline"]["traffic"] = traffic_from_profiles(run["storage"], run["bulk_cells"], run["tune"]["chosen"])
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_threads)
    other = None
    if not args.no_other_cases:
        other = {"multilevel_sphere": measure_multilevel("f32"), "rotor_duct": measure_rotor("f32"),
                 "graph_replay": measure_graph("f32")}
        print("other cases: %s" % other, file=sys.stderr, flush=True)
    serve = None
    if args.serving == "quick":
        serve = serving(32, 1.0, ["f32"])
    elif args.serving == "full":
        serve = serving(256, 1.0, ["f32", "f16"])
    line = {
        "metric": "MLUPS (sphere 256x128x128, FP32 population storage)",
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "flow around a sphere, uniform 256x128x128, D=32 at (64,64,64), u_in=0.05, Re=100, "
                               "K15 on level 0 as the reference case, IBB on a 57120-triangle STL sphere, FP32 storage",
                   "storage": "f32", "cells": main_run["cells"], "walls": main_run["walls"],
                   "fluid_updates_per_step": main_run["fluid_updates_per_step"]},
        "roofline": main_run["roofline"],
        "cpu_baseline": cpu,
        "tune": main_run["tune"],
        "kernel_us": round(main_run["kernel_us"], 2),
        "boundary_kernels_us": round(main_run["boundary_kernels_us"], 2),
        "kernel_mlups": round(main_run["kernel_mlups"], 1),
        "bulk_cells": main_run["bulk_cells"], "wall_adjacent": main_run["wall_adjacent"],
        "open_cells": main_run["open_cells"], "open_group_cells": main_run["open_group_cells"],
        "open_cells_in_bulk_kernel": main_run["open_cells_in_bulk_kernel"],
        "variants": {k: {"mlups": round(v["mlups"], 1), "ms_per_step": round(v["ms_per_step"], 4),
                         "tune": v["tune"]["chosen"], "bulk_cells": v["bulk_cells"],
                         "kernel_us": round(v["kernel_us"], 2), "boundary_kernels_us": round(v["boundary_kernels_us"], 2),
                         "roofline": v["roofline"],
                         "speedup_vs_f32": round(v["mlups"] / main_run["mlups"], 3)} for k, v in variants.items()},
        "accuracy": accuracy,
        "cases": other,
        "serving": serve,
        "copy_probe_gbs": round(2 * copy_bytes / (copy_ms / 1e3) / 1e9, 1),
        "device": {"name": name, "hbm_bytes": mem, "cus": cus},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
