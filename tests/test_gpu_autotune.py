"""lbm3_autotune: every launch choice timed, the fastest kept, the simulation's
state left as it was; the tuned run then equals a run that sets the same
choices by hand at the same step (each cell's arithmetic is the same, only the
kernels' launch parameters and placement differ; a new open-cell layout
completes the pending open-boundary condition at that step in both)."""
import numpy as np
import pytest

from lbm3_amd import cases, lib
from tests.helpers import rel_max

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_autotune_keeps_the_state(storage):
    case, dom, tris = cases.sphere(48, 24, 24, diameter=6.0, center=(14.0, 12.0, 12.0), segments=(16, 32))
    s = lib.Sim(case, storage)
    s.build(dom, tris)
    s.init()
    s.run(7)
    chosen, table = s.autotune(steps=3)
    assert s.info(0).iterationsFinished == 7
    assert len(table) in (48, 336)
    assert any(tuple(int(v) for v in r[:6]) == tuple(chosen[k] for k in lib.Sim.TUNE) for r in table)
    s.run(13)
    s.sync()
    got = s.macros(0)
    s.close()
    ref = lib.Sim(case, storage)
    ref.build(dom, tris)
    ref.init()
    ref.run(7)
    ref.set_tune(*(chosen[k] for k in lib.Sim.TUNE))
    ref.run(13)
    ref.sync()
    want = ref.macros(0)
    ref.close()
    err = max(rel_max(1 + got[:, 0], 1 + want[:, 0]), rel_max(got[:, 1:], want[:, 1:]))
    print("%s autotune chose %s (best %.4f ms/step of %d); tuned vs untuned run %.2e" %
          (storage, chosen, min(r[-1] for r in table), len(table), err))
    assert err < 1e-6


def test_set_tune_round_trip():
    case, dom, tris = cases.sphere(48, 24, 24, diameter=6.0, center=(14.0, 12.0, 12.0), segments=(16, 32))
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    s.init()
    s.run(2)
    s.set_tune(24576, 1, 0, 16, 1)
    assert s.tune() == {"bulk_lds": 24576, "xcd_remap": 1, "overlap": 0, "open_group": 16, "open_split": 1,
                        "rotor_in_bulk": 0}
    s.run(3)
    s.sync()
    assert np.isfinite(s.macros(0)).all()
    s.close()


def _fused_sphere():
    return cases.sphere(48, 24, 24, diameter=6.0, center=(14.0, 12.0, 12.0), segments=(16, 32))


@pytest.mark.parametrize("group", [8, 1])
def test_x_faces_in_a_separate_pass_equal_the_open_cell_kernel_f32(group):
    """FP32: the x-face open cells stepped by the bulk kernel and completed by a
    pass after streaming give the same bits as completing them when they are
    next loaded, in the open-cell kernel (groups of 8) or in the bulk kernel
    itself (group width 1); the completed values are stored exactly in FP32."""
    case, dom, tris = _fused_sphere()
    case.trackOpenBoundaries = 1
    runs = []
    for g, split in ((8, 0), (group, 1)):
        s = lib.Sim(case, "f32")
        s.build(dom, tris)
        s.init()
        s.run(1)
        s.set_tune(40960, 0, 1, g, split)
        s.run(25)
        s.sync()
        runs.append([s.macros(0), s.array("openRefs", 0), np.array([list(s.open_boundary_stats(b).values()) for b in range(3)])])
        assert s.tune()["open_split"] == split and s.tune()["open_group"] == g
        s.close()
    for x, y in zip(*runs):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_open_cells_in_the_bulk_kernel_equal_the_open_cell_kernel(storage):
    """Group width 1: the bulk kernel completes and steps the open cells
    itself (x-face choice 2: but those on the x faces, which stay in groups of
    the open-cell kernel; 3: with its lanes rotated so that a row's two x-face
    cells share a wave), with the same arithmetic as the open-cell kernel
    (the condition applied to the loaded populations, then the collision): the
    same bits in every storage type, with the side stream on or off or the
    wall cells in blocks of the bulk kernel's launch (side-stream choice 3);
    the skip and open masks count every fluid cell once."""
    case, dom, tris = _fused_sphere()
    case.trackOpenBoundaries = 1
    runs = []
    for g, overlap, split in ((0, 1, 0), (1, 1, 0), (1, 0, 0), (1, 1, 2), (1, 0, 2), (1, 1, 3), (1, 3, 0), (0, 3, 0)):
        s = lib.Sim(case, storage)
        s.build(dom, tris)
        s.init()
        if g:
            s.set_tune(40960, 0, overlap, g, split)
        s.run(30)
        s.sync()
        runs.append([s.macros(0), s.array("openRefs", 0), np.array([list(s.open_boundary_stats(b).values()) for b in range(3)]),
                     s.wall_forces(0)])
        k = s.cell_kinds(0)
        assert k["bulk"] + k["wall"] + k["open_group"] + k["solid"] + k["rotor"] + (k["open"] - k["open_in_bulk"]) == s.info(0).cellCount
        assert k["open_in_bulk"] == (0 if not g else (k["open"] if split in (0, 3) else k["open_in_bulk"])) and \
            (split != 2 or 0 < k["open_in_bulk"] < k["open"])
        s.close()
    for other in runs[1:]:
        for x, y in zip(runs[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("group,split", [(8, 1), (1, 0), (1, 2), (1, 3)])
def test_open_layouts_f64_equal_the_oracle(group, split):
    from tests.oracle_lib import OracleSim
    case, dom, tris = _fused_sphere()
    o = OracleSim(case, dom, tris, double=True)
    o.init()
    f0 = o.array("f", 0).reshape(27, -1).copy()
    o.run(30)
    g = lib.Sim(case, "f64")
    g.build(dom, tris)
    g.init()
    g.set_populations(0, f0)
    g.set_tune(40960, 0, 1, group, split)
    g.run(30)
    g.sync()
    mo, mg = o.macros(), g.macros()
    err = max(rel_max(1 + mg[:, 0], 1 + mo[:, 0]), rel_max(mg[:, 1:], mo[:, 1:]))
    g.close()
    print("open cells in groups of %d, x faces separate %d, FP64 sphere after 30 steps: GPU vs oracle %.2e"
          % (group, split, err))
    assert err < 1e-10


@pytest.mark.parametrize("save_group,save_split,load_group,load_split",
                         [(8, 1, 8, 0), (8, 0, 8, 1), (1, 0, 8, 0), (8, 0, 1, 0), (8, 1, 1, 0), (1, 2, 8, 1)])
def test_checkpoint_pending_across_split_choices(tmp_path, save_group, save_split, load_group, load_split):
    """A checkpoint taken with a condition pending under one grouping of the
    open cells, loaded into a simulation set to another: the loaded run
    continues bit for bit (a pending condition that covers the x faces or not
    is adopted with the file; groups of 8 and the bulk kernel's own open cells
    complete the same entries)."""
    case, dom, tris = _fused_sphere()
    a = lib.Sim(case, "f32")
    a.build(dom, tris)
    a.init()
    a.set_tune(40960, 0, 1, save_group, save_split)
    a.run(9)
    path = tmp_path / "split.ckpt"
    a.save_checkpoint(path)
    a.run(7)
    a.sync()
    want = a.macros(0)
    a.close()
    b = lib.Sim(case, "f32")
    b.build(dom, tris)
    b.init()
    b.set_tune(40960, 0, 1, load_group, load_split)
    b.load_checkpoint(path)
    b.run(7)
    b.sync()
    got = b.macros(0)
    b.close()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_finer_levels_on_their_own_streams_equal_in_order(storage):
    """Side-stream choice 2: each finer level sub-cycles on a stream of its own
    beside its coarser level's step; every level's fields and the FP64 wall
    force equal the in-order run (0) and the side-kernel-only run (1) bit for
    bit on the three-level sphere."""
    case, dom, tris = cases.multilevel_sphere(seed=0)
    runs = []
    for overlap in (0, 1, 2, 3):
        s = lib.Sim(case, storage)
        s.build(dom, tris)
        s.init()
        t = s.tune()
        t["overlap"] = overlap
        s.set_tune(*(t[k] for k in lib.Sim.TUNE))
        s.run(12)
        s.sync()
        runs.append([s.macros(L) for L in range(3)] + [s.wall_forces(2)])
        assert s.tune()["overlap"] == overlap
        s.close()
    for other in runs[1:]:
        for x, y in zip(runs[0], other):
            assert np.array_equal(x, y)
