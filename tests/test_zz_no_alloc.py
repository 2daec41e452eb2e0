"""lbm3_wall_forces_noalloc (no allocation, fixed device scratch) against
lbm3_wall_forces on two simulations in a row, the second with fewer
wall-adjacent cells, through the C ABI; plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from lbm3_amd import cases, lib

pytestmark = pytest.mark.gpu


def _forces(sim, fn, bodies=1, origin=(3.0, -2.0, 1.5)):
    o = (C.c_double * 3)(*origin)
    out = (C.c_double * (6 * bodies))()
    rc = getattr(lib.load(), fn)(sim.h, 0, o, bodies, out)
    return rc, np.array(out[:]).reshape(bodies, 6)


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_noalloc_wall_forces_equal_the_allocating_entry_point(storage):
    seen = []
    for n, d in ((96, 16.0), (48, 6.0)):               # second call: a smaller input
        case, dom, tris = cases.sphere(n, n // 2, n // 2, diameter=d, center=(n / 3.0, n / 4.0, n / 4.0),
                                       segments=(24, 48))
        s = lib.Sim(case, storage)
        s.build(dom, tris)
        s.init()
        s.run(30)
        W = len(s.array("wallIndex"))
        for bodies in (1, 3):
            rc0, want = _forces(s, "lbm3_wall_forces", bodies)
            rc1, got = _forces(s, "lbm3_wall_forces_noalloc", bodies)
            assert rc0 == 0 and rc1 == 0, lib.load().lbm3_last_error()
            scale = np.abs(want).max()
            err = np.abs(got - want).max() / scale
            print("%s sphere %d (W=%d), %d bodies: max |noalloc - alloc| / max %.2e" % (storage, n, W, bodies, err))
            assert scale > 0 and err < 1e-12
            assert np.all(got[1:] == 0.0)
        seen.append(W)
        s.close()
    assert seen[1] < seen[0]


def test_noalloc_refuses_more_than_its_scratch():
    case, dom, tris = cases.sphere(48, 24, 24, diameter=6.0, center=(16.0, 12.0, 12.0), segments=(16, 32))
    s = lib.Sim(case, "f32")
    s.build(dom, tris)
    rc, _ = _forces(s, "lbm3_wall_forces_noalloc", 32)     # beyond the body limit
    assert rc != 0
    s.close()
