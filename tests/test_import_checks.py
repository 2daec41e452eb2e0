"""The host-side checks of lbm3_import_level (lbm3_validate_level), on the CPU:
a level the oracle built passes; a row whose j+/k+/jk+ row starts inside the
level but whose last cell's neighbour (row neighbour + offset in the row,
cellFunctions.h:41-46) falls outside it is refused, as are rows that do not
hold consecutive cells and null open-boundary or periodic arrays with nonzero
counts.  No GPU is needed."""
import numpy as np
import pytest

from lbm3_amd import cases, lib
from tests.helpers import level_arrays
from tests.oracle_lib import OracleSim


@pytest.fixture(scope="module")
def sphere_level():
    case, dom, tris = cases.sphere(32, 16, 16, diameter=5.0, center=(10.0, 8.0, 8.0), segments=(12, 24))
    o = OracleSim(case, dom, tris)
    return level_arrays(o, 0), o.info(0)


@pytest.fixture(scope="module")
def tgv_level():
    case, dom, tris = cases.taylor_green(8)
    o = OracleSim(case, dom, tris)
    return level_arrays(o, 0), o.info(0)


def _row_lengths(shifter, rows):
    n = np.zeros(rows, np.int64)
    head = -1
    for c, sh in enumerate(shifter):
        head = c if sh >= 0 else head
        n[shifter[head]] += 1
    return n


def test_built_levels_pass(sphere_level, tgv_level):
    for arrays, info in (sphere_level, tgv_level):
        lib.validate_level(arrays, info)


@pytest.mark.parametrize("name", ["rowJPlus", "rowKPlus", "rowJKPlus"])
def test_row_neighbour_whose_tail_leaves_the_level_is_refused(sphere_level, name):
    arrays, info = sphere_level
    a = dict(arrays)
    n = info.cellCount
    lengths = _row_lengths(a["shifter"], len(a["rowI"]))
    r = int(np.argmax(lengths))
    assert lengths[r] > 2
    v = a[name].copy()
    v[r] = n - lengths[r] + 1             # first cell inside, last cell one past the end
    a[name] = v
    with pytest.raises(lib.LBM3Error, match="row neighbour out of range"):
        lib.validate_level(a, info)
    v[r] = n - lengths[r]                 # the whole row inside: accepted
    lib.validate_level(a, info)


def test_rows_must_hold_consecutive_cells(sphere_level):
    arrays, info = sphere_level
    a = dict(arrays)
    sh = a["shifter"].copy()
    c = int(np.nonzero(sh < -1)[0][0])    # a cell two or more past its row head
    sh[c] += 1                            # now points at a cell that is not a head
    a["shifter"] = sh
    with pytest.raises(lib.LBM3Error, match="row head"):
        lib.validate_level(a, info)


def test_null_open_and_periodic_arrays_with_counts_are_refused(sphere_level, tgv_level):
    arrays, info = sphere_level
    assert arrays["openCount"].sum() > 0
    a = dict(arrays)
    a["openIndex"] = np.zeros(0, np.int32)
    with pytest.raises(lib.LBM3Error, match="open-boundary cell lists"):
        lib.validate_level(a, info)
    b = dict(arrays)
    b["openCount"] = np.zeros(0, np.int32)
    with pytest.raises(lib.LBM3Error, match="open-boundary counts"):
        lib.validate_level(b, info, counts={"nOpenIDs": len(arrays["openCount"])})
    tarr, tinfo = tgv_level
    assert len(tarr["periodicIndex"]) > 0
    t = dict(tarr)
    t["periodicSource"] = np.zeros(0, np.int32)
    with pytest.raises(lib.LBM3Error, match="periodic arrays"):
        lib.validate_level(t, tinfo, counts={"nPeriodic": len(tarr["periodicIndex"])})
