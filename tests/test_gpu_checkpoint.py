"""Checkpoint and restart (lbm3_save_checkpoint / lbm3_load_checkpoint, no
counterpart in the reference) on the GPU: a simulation restored from a
checkpoint and stepped on reproduces the saved one bit for bit -- populations
through the macroscopic fields of every level, open-boundary statistics, wall
and rotor forces, iteration counters and esotwist parity -- and a checkpoint
that does not fit the simulation is refused without changing it."""
import os
import tempfile

import numpy as np
import pytest

from lbm3_amd import cases, geometry, lib
from tests.helpers import rel_max
from tests.oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def _pipe(levels, res=0.4):
    tri = geometry.read_binary_stl(os.path.join(lib_golden(), "pipeD20L200STL.STL"))
    return cases.pipe(tri, length=20.0, res=res, reynolds=100.0, u_mean=0.01, inlet_non_reflective=True,
                      profile_init=True, levels=levels)


def lib_golden():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _observe(s, levels, walls=True):
    out = {}
    for L in range(levels):
        out["macros%d" % L] = s.macros(L)
        I = s.info(L)
        out["info%d" % L] = (I.iterationsFinished, I.iterationsSinceTrackerReset, I.esotwistFlipper)
        if walls:
            out["walls%d" % L] = np.array(s.wall_forces(L))
    return out


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_restart_reproduces_the_pipe_run(tmp_path, storage):
    """Two-level pipe with a non-reflective inlet: 20 coarse steps, checkpoint
    with the open boundary condition still pending, 15 more steps; a second
    simulation built from the same case, restored and stepped 15 times ends
    identical."""
    case, dom, tris = _pipe(2)
    a = lib.Sim(case, storage)
    a.build(dom, tris)
    a.init()
    a.run(20)
    path = tmp_path / "pipe.ckpt"
    a.save_checkpoint(path)
    a.run(15)
    want = _observe(a, 2)
    want["open"] = [a.open_boundary_stats(b) for b in range(2)]
    a.close()
    b = lib.Sim(case, storage)
    b.build(dom, tris)
    b.init()
    b.run(3)                                   # state the checkpoint replaces
    b.load_checkpoint(path)
    b.run(15)
    got = _observe(b, 2)
    got["open"] = [b.open_boundary_stats(x) for x in range(2)]
    b.close()
    _same(want, got)
    assert want["info0"][0] == 35 and want["info1"][0] == 70


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_fresh_restart_with_a_fused_open_level(tmp_path, storage):
    """Sphere whose open faces touch no wall, so level 0 runs its open cells in
    the open-cell kernel (the boundary condition of a step completed when the
    next step loads the cell).  The checkpoint is taken with that condition
    pending and loaded into a freshly built simulation that has never run;
    the continued runs agree bit for bit."""
    case, dom, tris = cases.sphere(48, 24, 24, diameter=6.0, center=(14.0, 12.0, 12.0), segments=(16, 32))
    case.trackOpenBoundaries = 1
    a = lib.Sim(case, storage)
    a.build(dom, tris)
    a.init()
    a.run(17)
    assert a.cell_kinds(0)["open"] > 0                   # the open cells have their own kernel
    path = tmp_path / "sphere.ckpt"
    a.save_checkpoint(path)
    a.run(11)
    want = _observe(a, 1)
    want["open"] = [a.open_boundary_stats(x) for x in range(3)]
    want["ke"] = a.kinetic_energy()
    a.close()
    b = lib.Sim(case, storage)
    b.build(dom, tris)
    b.load_checkpoint(path)                              # no init, no run before
    assert b.cell_kinds(0)["open"] > 0
    b.run(11)
    got = _observe(b, 1)
    got["open"] = [b.open_boundary_stats(x) for x in range(3)]
    got["ke"] = b.kinetic_energy()
    b.close()
    _same(want, got)


def test_restored_fp64_run_equals_the_oracle():
    """FP64 sphere: 9 steps, checkpoint, a fresh simulation restored from it and
    stepped 13 more equals the oracle run of 22 steps (macroscopic fields and
    wall force within 1e-10)."""
    case, dom, tris = cases.sphere(40, 20, 20, diameter=6.0, center=(12.0, 10.0, 10.0), segments=(16, 32))
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    # the initial-condition hooks evaluate in float, whose last bit may differ
    # between libm and the device: start from the oracle's populations
    f0 = o.array("f", 0).reshape(27, -1).copy()
    o.run(22)
    a = lib.Sim(case, "f64")
    a.build(dom, tris)
    a.init()
    a.set_populations(0, f0)
    a.run(9)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "f64.ckpt")
        a.save_checkpoint(path)
        a.close()
        b = lib.Sim(case, "f64")
        b.build(dom, tris)
        b.load_checkpoint(path)
    b.run(13)
    mg, mo = b.macros(0), o.macros(0)
    err = max(rel_max(1 + mg[:, 0], 1 + mo[:, 0]), rel_max(mg[:, 1:], mo[:, 1:]))
    print("restored FP64 sphere vs oracle after 9 + 13 steps: %.2e" % err)
    assert err < 1e-10
    assert b.info(0).iterationsFinished == 22
    b.close()


def test_restart_reproduces_the_rotor_duct_f32(tmp_path):
    """Rotor in a duct (two levels, 4-blade rotor): the rotor angle and its
    force accumulators come back with the populations."""
    case, dom, statics, rotor, ri = cases.rotor_duct(48, 24, 24, seed=1, levels=2, n_theta=256, n_axial=8)

    def make():
        s = lib.Sim(case, "f32")
        s.build(dom, statics)
        s.build_rotors([rotor], 3)
        for L in range(2):
            s.set_rotor(L, 0, ri["omega"], ri["axis"], ri["origin"])
        s.init()
        return s

    a = make()
    a.run(12)
    path = tmp_path / "rotor.ckpt"
    a.save_checkpoint(path)
    a.run(9)
    want = _observe(a, 2, walls=False)
    want["rotor"] = np.concatenate([a.rotor_forces(L, 0) for L in range(2)] + [a.rotor_info(L, 0) for L in range(2)])
    a.close()
    b = make()
    b.load_checkpoint(path)
    b.run(9)
    got = _observe(b, 2, walls=False)
    got["rotor"] = np.concatenate([b.rotor_forces(L, 0) for L in range(2)] + [b.rotor_info(L, 0) for L in range(2)])
    b.close()
    _same(want, got)


def test_a_checkpoint_that_does_not_fit_is_refused(tmp_path):
    """Other storage, other grid, a short file, not a checkpoint: each refused
    with LBM3_E_ARG, the simulation unchanged."""
    case, dom, tris = _pipe(1)
    a = lib.Sim(case, "f32")
    a.build(dom, tris)
    a.init()
    a.run(5)
    path = tmp_path / "one.ckpt"
    a.save_checkpoint(path)
    before = a.macros(0)
    other = lib.Sim(case, "f16")
    other.build(dom, tris)
    other.init()
    with pytest.raises(lib.LBM3Error, match="storage"):
        other.load_checkpoint(path)
    other.close()
    case2, dom2, tris2 = _pipe(1, res=0.5)
    c = lib.Sim(case2, "f32")
    c.build(dom2, tris2)
    c.init()
    with pytest.raises(lib.LBM3Error, match="grid"):
        c.load_checkpoint(path)
    c.close()
    case3, dom3, tris3 = cases.pipe(geometry.read_binary_stl(os.path.join(lib_golden(), "pipeD20L200STL.STL")),
                                    length=20.0, res=0.4, reynolds=50.0, u_mean=0.01, inlet_non_reflective=True,
                                    profile_init=True, levels=1)
    v = lib.Sim(case3, "f32")                  # same grid, another viscosity
    v.build(dom3, tris3)
    v.init()
    with pytest.raises(lib.LBM3Error, match="grid"):
        v.load_checkpoint(path)
    v.close()
    data = path.read_bytes()
    short = tmp_path / "short.ckpt"
    short.write_bytes(data[:-8])
    a.run(2)
    mid = a.macros(0)
    with pytest.raises(lib.LBM3Error, match="size"):
        a.load_checkpoint(short)
    assert np.array_equal(a.macros(0), mid)
    junk = tmp_path / "junk.ckpt"
    junk.write_bytes(b"\0" * 64)
    with pytest.raises(lib.LBM3Error, match="not a checkpoint"):
        a.load_checkpoint(junk)
    a.load_checkpoint(path)
    assert np.array_equal(a.macros(0), before)
    a.close()
