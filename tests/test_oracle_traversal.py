"""Pins of the oracle's ray/triangle tests and BVH traversal: hand-derived
Moller-Trumbore cases, and BVH results identical to brute force over all
triangles (closest hit: same triangle and t; any hit: same occlusion)."""
import json
import os

import numpy as np
import pytest

from firefly_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mt_cases.json")


def _single_tri_bvh(oracle, tri):
    nodes, ordered = oracle.bvh_build(np.asarray(tri, np.float32).reshape(1, 9))
    return nodes, oracle.ordered_tris(np.asarray(tri, np.float32).reshape(1, 9), ordered)


def test_mt_golden_cases(oracle):
    g = json.load(open(GOLDEN))
    nodes, tris = _single_tri_bvh(oracle, g["triangle"])
    for c in g["cases"]:
        ray = np.array([c["o"] + c["d"]], np.float32)
        for mode, two in (("cull", False), ("two", True)):
            t, u, v, i = oracle.trace_closest(nodes, tris, ray, two_sided=two)
            if c[mode] is None:
                assert i[0] == -1, (c["name"], mode)
            else:
                assert i[0] == 0, (c["name"], mode)
                assert [t[0], u[0], v[0]] == c[mode], (c["name"], mode, t[0], u[0], v[0])
            tb, ib = oracle.trace_closest_brute(tris, ray, two_sided=two)
            assert ib[0] == i[0] and (i[0] < 0 or tb[0] == t[0])
        assert bool(oracle.trace_any(nodes, tris, ray)[0]) == c["any"], c["name"]


def _soup(n, seed, edge=0.05):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    return (c + rng.normal(scale=edge, size=(n, 3, 3))).astype(np.float32)


@pytest.mark.parametrize("kind", ["spheres", "soup", "soup_large_tris"])
def test_bvh_traversal_equals_brute_force(oracle, kind):
    if kind == "spheres":
        pos = scenes.spheres_scene(segments=12, rings=6).positions
    elif kind == "soup":
        pos = _soup(3000, 1)
    else:
        pos = _soup(400, 2, edge=0.6)
    nodes, ordered = oracle.bvh_build(pos)
    tris = oracle.ordered_tris(pos, ordered)
    rays = scenes.random_rays(4000, 11)
    for two in (False, True):
        t, u, v, i = oracle.trace_closest(nodes, tris, rays, two_sided=two)
        tb, ib = oracle.trace_closest_brute(tris, rays, two_sided=two)
        assert np.array_equal(i, ib)
        assert np.array_equal(t.view(np.uint32), tb.view(np.uint32))
        assert (i >= 0).sum() > 100
    o = oracle.trace_any(nodes, tris, rays)
    assert np.array_equal(o, oracle.trace_any_brute(tris, rays))


@pytest.mark.parametrize("seed", [3, 4])
def test_pruned_two_child_walk_equals_reference_order(oracle, seed):
    """The GPU's visit scheme (both child boxes per interior node, pruning by
    the current hit) returns exactly the reference traversal's results."""
    sd = scenes.spheres_scene()
    nodes, ordered = oracle.bvh_build(sd.positions)
    tris = oracle.ordered_tris(sd.positions, ordered)
    rays = scenes.random_rays(20000, seed)
    t, u, v, i = oracle.trace_closest(nodes, tris, rays)
    tw, iw = oracle.trace_walk(nodes, tris, rays, prune=True)
    assert np.array_equal(i, iw) and np.array_equal(t.view(np.uint32), tw.view(np.uint32))
    occ = oracle.trace_any(nodes, tris, rays)
    _, ia = oracle.trace_walk(nodes, tris, rays, any_hit=True, prune=True)
    assert np.array_equal(occ.astype(bool), ia >= 0)


@pytest.mark.parametrize("scene", ["spheres", "icosphere5"])
def test_four_wide_walk_equals_reference_order(oracle, scene):
    """The 4-wide visit scheme (grandchild boxes per interior node, pruning by
    the current hit) returns exactly the reference traversal's results."""
    if scene == "spheres":
        pos = scenes.spheres_scene().positions
    else:
        pos = scenes.icosphere_scene(width=8, height=8, subdiv=5).positions
    nodes, ordered = oracle.bvh_build(pos)
    tris = oracle.ordered_tris(pos, ordered)
    rays = scenes.random_rays(20000, 7)
    t, u, v, i = oracle.trace_closest(nodes, tris, rays)
    tq, iq = oracle.trace_walk(nodes, tris, rays, prune=True, width=4)
    assert np.array_equal(i, iq) and np.array_equal(t.view(np.uint32), tq.view(np.uint32))
    assert (i >= 0).sum() > 1000
    occ = oracle.trace_any(nodes, tris, rays)
    _, ia = oracle.trace_walk(nodes, tris, rays, any_hit=True, prune=True, width=4)
    assert np.array_equal(occ.astype(bool), ia >= 0)


def test_four_wide_walk_reaches_the_same_leaves(oracle):
    """Without pruning a leaf is reached iff its box and every ancestor's box
    is hit (child boxes nest in their parent's, and the slab test is monotone
    in the plane coordinates), and both schemes reach leaves in the same
    order: the triangles tested, any-hit's early exit included, are the same
    for both record widths, while the 4-wide walk tests fewer boxes (the
    children's own boxes are skipped)."""
    pos = scenes.icosphere_scene(width=8, height=8, subdiv=5).positions
    nodes, ordered = oracle.bvh_build(pos)
    tris = oracle.ordered_tris(pos, ordered)
    rays = scenes.random_rays(5000, 8)
    for any_hit in (False, True):
        n2, t2 = oracle.count_traversal(nodes, tris, rays, any_hit=any_hit, prune=False, width=2)
        n4, t4 = oracle.count_traversal(nodes, tris, rays, any_hit=any_hit, prune=False, width=4)
        assert t4 == t2
        assert n4 < n2


def test_pruned_traversal_tests_fewer_boxes(oracle):
    sd = scenes.spheres_scene()
    nodes, ordered = oracle.bvh_build(sd.positions)
    tris = oracle.ordered_tris(sd.positions, ordered)
    rays = scenes.random_rays(5000, 3)
    n0, t0 = oracle.count_traversal(nodes, tris, rays, prune=False)
    n1, t1 = oracle.count_traversal(nodes, tris, rays, prune=True)
    assert n1 < n0 and t1 <= t0
    a0 = oracle.count_traversal(nodes, tris, rays, any_hit=True, prune=False)
    a1 = oracle.count_traversal(nodes, tris, rays, any_hit=True, prune=True)
    assert a1[0] <= a0[0]


def test_closest_hit_is_minimum_t_and_first_in_order(oracle):
    """Two coincident triangles: the first in BVH slot order wins the tie
    (strict '<' replacement)."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    pos = np.stack([tri, tri + np.float32([0, 0, 0])]).reshape(2, 9)
    nodes, ordered = oracle.bvh_build(pos)
    tris = oracle.ordered_tris(pos, ordered)
    t, u, v, i = oracle.trace_closest(nodes, tris, np.array([[0.2, 0.2, 1, 0, 0, -1]], np.float32))
    tb, ib = oracle.trace_closest_brute(tris, np.array([[0.2, 0.2, 1, 0, 0, -1]], np.float32))
    assert i[0] == ib[0] == 0 and t[0] == 1.0
