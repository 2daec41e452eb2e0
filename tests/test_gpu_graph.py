"""lbm3_run_graph (the step sequence of two coarse steps captured into a HIP
graph and replayed) on the GPU: bit for bit the results of lbm3_run over
several calls in a row -- even and odd step counts, eager steps and host reads
between the replays -- on a two-level Taylor-Green grid (interface transfers,
periodic ghosts) and on the Couette cell (wall kernel on the second stream,
moving-wall bounceback, FP64 wall sums); refused where open boundaries take
per-step arguments."""
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


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_graph_replay_two_level_taylor_green(storage):
    case, dom, tris = cases.taylor_green(32, levels=2)

    def make():
        s = lib.Sim(case, storage)
        s.build(dom, tris)
        s.init()
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


def test_graph_replay_refuses_open_boundaries():
    tri = geometry.read_binary_stl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                                "pipeD20L200STL.STL"))
    case, dom, tris = cases.pipe(tri, length=20.0, res=0.5, reynolds=100.0, u_mean=0.01)
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    s.init()
    with pytest.raises(lib.LBM3Error, match="open boundaries"):
        s.run_graph(4)
    s.run(2)
    s.close()
