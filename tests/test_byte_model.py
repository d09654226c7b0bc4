"""The frozen byte model of the roofline (tests/golden/byte_model.json) and
the canonical BVH2 counter that made it (oracle/yo_bytemodel.c)."""
import json
import os

import numpy as np

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_counter_on_one_triangle():
    """A one-triangle tree is a single leaf: every ray visits it once and
    tests the triangle once if it enters its box, never otherwise."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    rays = np.array([[0.2, 0.2, 1, 0, 0, 0, -1, np.inf],     # hits
                     [5.0, 5.0, 1, 0, 0, 0, -1, np.inf]],    # misses the box
                    np.float32)
    n, t, b, nodes = po.canonical_counts(v, f, rays[:1], 0, threads=1)
    assert (n, t, nodes) == (1.0, 1.0, 1)
    n, t, b, _ = po.canonical_counts(v, f, rays[1:], 0, threads=1)
    assert (n, t) == (1.0, 0.0)


def test_counter_visits_shrink_with_closest_hit():
    """A stack of parallel quads seen head-on: the closest-hit traversal
    stops visiting boxes behind the first hit, the any-hit one stops at its
    first hit; both test fewer triangles than the tree holds."""
    quads = []
    for k in range(64):
        z = float(k)
        quads.append([[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]])
    v = np.array(quads, np.float32).reshape(-1, 3)
    f = np.array([[4 * k, 4 * k + 1, 4 * k + 2] for k in range(64)] +
                 [[4 * k, 4 * k + 2, 4 * k + 3] for k in range(64)], np.uint32)
    rays = np.array([[0.1, 0.2, -5, 0, 0, 0, 1, np.inf]], np.float32)
    nc, tc, _, _ = po.canonical_counts(v, f, rays, 0, threads=1)
    na, ta, _, _ = po.canonical_counts(v, f, rays, 1, threads=1)
    assert tc < 128 / 4 and ta <= tc and na <= nc


def test_fixture_is_consistent():
    bm = json.load(open(os.path.join(ROOT, "tests", "golden", "byte_model.json")))
    for kind, hit in (("extension", 16), ("shadow", 4)):
        k = bm[kind]
        assert abs(k["B_ray"] - (32 + hit + 32 * k["n_node"] + 36 * k["n_tri"])) < 1e-6
        assert k["n_box_tested"] > k["n_node"] > 1 and k["n_tri"] > 1
    assert bm["queries"]["sample_per_kind"] == 65536
