"""lbm3_run_graph (the step sequence of two coarse steps captured into a HIP
graph and replayed) on the GPU: bit for bit the results of lbm3_run over
several calls in a row -- even and odd step counts, eager steps and host reads
between the replays -- on a two-level Taylor-Green grid (interface transfers,
periodic ghosts) and on the Couette cell (wall kernel on the second stream,
moving-wall bounceback, FP64 wall sums), the sphere with its open cells in
the open-cell kernel, a two-level pipe and a turning rotor (the values that
change between replays read from a device table)."""
import os

import numpy as np
import pytest

from lbm3_amd import cases, geometry, lib

pytestmark = pytest.mark.gpu


def _state(s, levels, walls):
    out = [s.macros(L) for L in range(levels)] + [np.array([s.kinetic_energy()])]
    if walls:
        out.append(s.wall_forces(0, bodies=2))
    out.append(np.array([[s.info(L).iterationsFinished, s.info(L).esotwistFlipper] for L in range(levels)]))
    return out


def _replays_equal_eager(make, levels, walls):
    a, b = make(), make()
    for graph_steps, eager_between in ((6, 1), (5, 2), (4, 0), (1, 0)):
        a.run_graph(graph_steps)
        b.run(graph_steps)
        for x, y in zip(_state(a, levels, walls), _state(b, levels, walls)):
            assert np.array_equal(x, y), graph_steps
        a.run(eager_between)
        b.run(eager_between)
    a.close()
    b.close()


def _with_overlap(s, overlap):
    t = s.tune()
    t["overlap"] = overlap
    s.set_tune(*(t[k] for k in lib.Sim.TUNE))


@pytest.mark.parametrize("storage,overlap", [("f32", 1), ("f16", 1), ("f32", 2), ("f32", 3)])
def test_graph_replay_two_level_taylor_green(storage, overlap):
    """overlap 2: the fine level sub-cycles on a stream of its own, captured
    into the same graph through the fork and join events."""
    case, dom, tris = cases.taylor_green(32, levels=2)

    def make():
        s = lib.Sim(case, storage)
        s.build(dom, tris)
        s.init()
        _with_overlap(s, overlap)
        return s

    _replays_equal_eager(make, 2, walls=False)


def test_graph_replay_couette_walls_f32():
    case, dom, bodies, info = cases.couette(r_inner=6.0, r_outer=10.0, length=4.0, n_theta=256)

    def make():
        s = lib.Sim(case, "f32")
        s.build(dom, bodies)
        s.init()
        return s

    _replays_equal_eager(make, 1, walls=True)


def _state_open(s, levels, n_open):
    out = _state(s, levels, walls=False)
    out.append(np.array([list(s.open_boundary_stats(b).values()) for b in range(n_open)]))
    out.append(s.array("openRefs", 0))
    return out


@pytest.mark.parametrize("group,overlap", [(8, 1), (1, 3)])
def test_graph_replay_sphere_open_cells(group, overlap):
    """The sphere with its open boundaries completed when the cells are next
    loaded (in the open-cell kernel, or in the bulk kernel with the wall cells
    in the same launch; a deferred condition pending across replays) and
    tracked flow: replays equal eager steps bit for bit, also around eager
    steps and reads."""
    case, dom, tris = cases.sphere(48, 24, 24, diameter=6.0, center=(14.0, 12.0, 12.0), segments=(16, 32))
    case.trackOpenBoundaries = 1

    def make():
        s = lib.Sim(case, "f32")
        s.build(dom, tris)
        s.init()
        s.set_tune(40960, 2, overlap, group, 0, 0)
        return s

    a, b = make(), make()
    for graph_steps, eager_between in ((6, 1), (5, 2), (4, 0), (1, 0)):
        a.run_graph(graph_steps)
        b.run(graph_steps)
        assert a.cell_kinds(0)["open"] > 0              # the open-cell kernel runs
        for x, y in zip(_state_open(a, 1, 3), _state_open(b, 1, 3)):
            assert np.array_equal(x, y), graph_steps
        a.run(eager_between)
        b.run(eager_between)
    a.close()
    b.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_graph_replay_two_level_pipe(storage):
    """Two-level pipe: the open boundary as its own pass (wall cells touch it)
    with the iteration-dependent reference update, on both levels."""
    tri = geometry.read_binary_stl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                                "pipeD20L200STL.STL"))
    case, dom, tris = cases.pipe(tri, length=20.0, res=0.5, reynolds=100.0, u_mean=0.01)

    def make():
        s = lib.Sim(case, storage)
        s.build(dom, tris)
        s.init()
        return s

    a, b = make(), make()
    levels = a.levels
    for graph_steps, eager_between in ((8, 1), (3, 0)):
        a.run_graph(graph_steps)
        b.run(graph_steps)
        for x, y in zip(_state(a, levels, walls=False), _state(b, levels, walls=False)):
            assert np.array_equal(x, y), graph_steps
        a.run(eager_between)
        b.run(eager_between)
    a.close()
    b.close()


@pytest.mark.parametrize("overlap", [1, 2])
def test_graph_replay_rotor_duct(overlap):
    """A turning rotor on two levels: each replay's rotor angles come from the
    device table; fields, FP64 rotor torque and angles equal eager steps."""
    case, dom, statics, rotor, info = cases.rotor_duct(48, 24, 24, seed=1, levels=2, n_theta=128, n_axial=8)

    def make():
        s = lib.Sim(case, "f32")
        s.build(dom, statics)
        s.build_rotors([rotor])
        for L in range(2):
            s.set_rotor(L, 0, info["omega"], info["axis"], info["origin"])
        s.init()
        _with_overlap(s, overlap)
        return s

    a, b = make(), make()
    for graph_steps, eager_between in ((6, 1), (5, 0)):
        a.run_graph(graph_steps)
        b.run(graph_steps)
        for x, y in zip(_state(a, 2, walls=False), _state(b, 2, walls=False)):
            assert np.array_equal(x, y), graph_steps
        for L in range(2):
            assert np.array_equal(a.rotor_forces(L, 0), b.rotor_forces(L, 0))
            assert np.array_equal(a.rotor_info(L, 0), b.rotor_info(L, 0))
        a.run(eager_between)
        b.run(eager_between)
    a.close()
    b.close()
