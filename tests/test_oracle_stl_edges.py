"""The oracle's shared-edge counts (checkSTLEdges, STLFunctions.h:130-287) on
the CPU, pinned by brute force: for every edge of every kept triangle, the
number of triangles holding that edge in either orientation, itself included,
counted over the whole mesh with a hash table instead of the binned search.
A closed mesh gives 2 everywhere (the reference's pipe STL, a UV sphere); a
mesh with a triangle removed gives 1 on exactly the three edges around the
hole; a duplicated triangle gives 3."""
import collections
import os

import numpy as np

from lbm3_amd import geometry
from tests.oracle_lib import stl_edge_counts

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def brute_force(tri):
    t = np.asarray(tri, np.float32).reshape(-1, 3, 3)
    keep = [x for x in t if not (np.array_equal(x[0], x[1]) or np.array_equal(x[2], x[1]) or np.array_equal(x[2], x[0]))]
    directed = collections.Counter()
    for x in keep:
        for e in range(3):
            directed[(tuple(x[e]), tuple(x[(e + 1) % 3]))] += 1
    out = np.zeros((len(keep), 3), np.int32)
    for n, x in enumerate(keep):
        for e in range(3):
            a, b = tuple(x[e]), tuple(x[(e + 1) % 3])
            out[n, e] = directed[(a, b)] + directed[(b, a)]
    return out


def test_reference_pipe_stl_is_closed():
    tri = geometry.read_binary_stl(os.path.join(GOLDEN, "pipeD20L200STL.STL"))
    got = stl_edge_counts(tri, 0.4)
    assert got.shape == (376, 3)
    assert np.all(got == 2)
    assert np.array_equal(got, brute_force(tri))


def test_sphere_closed_then_opened():
    tri = geometry.uv_sphere(5.0, (10.0, 10.0, 10.0), 12, 24).reshape(-1, 9)
    assert np.all(stl_edge_counts(tri, 1.0) == 2)
    holed = np.delete(tri, 40, axis=0)
    got = stl_edge_counts(holed, 1.0)
    assert np.array_equal(got, brute_force(holed))
    assert int((got == 1).sum()) == 3 and int((got != 2).sum()) == 3


def test_duplicated_triangle_and_degenerate_triangle():
    tri = geometry.uv_sphere(5.0, (10.0, 10.0, 10.0), 12, 24).reshape(-1, 9)
    dup = np.concatenate([tri, tri[7:8]])
    got = stl_edge_counts(dup, 1.0)
    assert np.array_equal(got, brute_force(dup))
    assert np.all(got[7] == 3) and np.all(got[-1] == 3)
    degenerate = tri[3:4].copy()
    degenerate[0, 3:6] = degenerate[0, 0:3]            # vertex B repeats A: dropped as readSTL does
    with_degenerate = np.concatenate([tri, degenerate])
    assert stl_edge_counts(with_degenerate, 1.0).shape == (len(tri), 3)
