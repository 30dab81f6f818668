"""Pins of the SAH BVH build: structural invariants, hand-checkable split
decisions, and bit-identity of the product's host builder with the oracle."""
import numpy as np
import pytest

from firefly_amd import abi, scenes


def _check_invariants(nodes, ordered, pos, max_prims=8):
    pos = np.asarray(pos, np.float32).reshape(-1, 3, 3)
    n = len(pos)
    assert sorted(ordered.tolist()) == list(range(n)), "every primitive exactly once"
    seen = np.zeros(len(nodes), bool)
    stack = [0]
    while stack:
        i = stack.pop()
        seen[i] = True
        nd = nodes[i]
        if nd["n_prims"] > 0:
            sl = ordered[nd["offset"]: nd["offset"] + nd["n_prims"]]
            p = pos[sl].reshape(-1, 3)
            assert np.all(p.min(0) >= nd["bmin"]) and np.all(p.max(0) <= nd["bmax"])
            if nd["n_prims"] > max_prims:  # only for degenerate centroid extents
                tb = pos[sl]
                c = 0.5 * tb.min(1) + 0.5 * tb.max(1)
                assert np.any(np.ptp(c, axis=0) == 0)
        else:
            a, b = i + 1, nd["offset"]
            assert b > a, "pre-order layout: second child after the first subtree"
            for ch in (a, b):
                assert np.all(nodes[ch]["bmin"] >= nd["bmin"]) and np.all(nodes[ch]["bmax"] <= nd["bmax"])
            lo = np.minimum(nodes[a]["bmin"], nodes[b]["bmin"])
            hi = np.maximum(nodes[a]["bmax"], nodes[b]["bmax"])
            assert np.array_equal(lo, nd["bmin"]) and np.array_equal(hi, nd["bmax"])
            stack += [a, b]
    assert seen.all(), "every node reachable"


def _soup(n, seed, edge=0.05):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    return (c + rng.normal(scale=edge, size=(n, 3, 3))).astype(np.float32)


@pytest.mark.parametrize("seed,n", [(0, 1), (1, 2), (2, 3), (3, 17), (4, 1000), (5, 20000)])
def test_bvh_invariants(oracle, seed, n):
    pos = _soup(n, seed)
    nodes, ordered = oracle.bvh_build(pos)
    _check_invariants(nodes, ordered, pos)
    assert len(nodes) <= 2 * n - 1


def test_degenerate_centroids_make_one_leaf(oracle):
    tri = np.array([[-1, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    pos = np.stack([tri * s for s in (1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0)])
    pos[:, :, 1] -= pos[:, :, 1].max(1, keepdims=True) / 2  # same centroid for all
    nodes, ordered = oracle.bvh_build(pos)
    assert len(nodes) == 1 and nodes[0]["n_prims"] == 10


def test_two_primitives_sorted_by_centroid(oracle):
    """nth_element on two elements orders them by centroid[dim]."""
    a = np.array([[5, 0, 0], [6, 0, 0], [5, 1, 0]], np.float32)
    b = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nodes, ordered = oracle.bvh_build(np.stack([a, b]))
    assert len(nodes) == 3 and nodes[0]["axis"] == 0
    assert ordered.tolist() == [1, 0]


def test_sah_splits_two_clusters_at_the_gap(oracle):
    """Two clusters of 8 far apart: the root split separates them, and a
    cluster of 8 small triangles is split further (leaf cost 8 > SAH cost)."""
    rng = np.random.default_rng(0)
    left = rng.uniform(0, 1, (8, 3, 3)) * [1, 1, 1]
    right = left + [100, 0, 0]
    pos = np.concatenate([left, right]).astype(np.float32)
    nodes, ordered = oracle.bvh_build(pos)
    first_child = nodes[1]
    second = nodes[nodes[0]["offset"]]
    assert nodes[0]["axis"] == 0
    assert first_child["bmax"][0] < 50 and second["bmin"][0] > 50


def test_product_bvh_bit_identical_to_oracle(oracle):
    cases = [scenes.spheres_scene().positions, _soup(5000, 7), _soup(300, 8, edge=0.8),
             scenes.displaced_icosphere(4, (0, 1, 0), 1.0, 0.05, 0)[0]]
    for pos in cases:
        n_o, o_o = oracle.bvh_build(pos)
        n_p, o_p, depth = abi.bvh_build_host(pos)
        assert n_p.tobytes() == n_o.tobytes()
        assert np.array_equal(o_p, o_o)
        assert depth == oracle.bvh_max_depth(n_o)


def test_parallel_product_bvh_bit_identical_to_oracle(oracle):
    """Meshes large enough that the product builds subtrees on several
    threads (ranges of 16384 primitives and more) and splices them: the node
    array and slot order still equal the one-thread oracle build."""
    ico = scenes.displaced_icosphere(6, (0, 1, 0), 1.0, 0.05, 3)[0].reshape(-1, 9)
    cases = [ico, np.concatenate([ico, _soup(60000, 9).reshape(-1, 9)]), _soup(70000, 10, edge=0.05)]
    for pos in cases:
        n_o, o_o = oracle.bvh_build(pos)
        n_p, o_p, depth = abi.bvh_build_host(pos)
        assert len(n_o) > 60000
        assert n_p.tobytes() == n_o.tobytes()
        assert np.array_equal(o_p, o_o)


# ---- linear BVH (the GPU builder restated by the oracle) ----

def _leaf_sizes(nodes):
    return nodes["n_prims"][nodes["n_prims"] > 0]


@pytest.mark.parametrize("seed,n,max_prims", [(0, 1, 8), (1, 2, 8), (2, 5, 8), (3, 9, 8), (4, 1000, 8),
                                              (5, 20000, 4), (6, 3000, 1)])
def test_lbvh_invariants(oracle, seed, n, max_prims):
    """Every triangle once, boxes exact unions of their children, pre-order
    layout, leaves of at most max_prims triangles (Morton ranges of any
    size split down to that)."""
    pos = _soup(n, seed)
    nodes, ordered = oracle.lbvh_build(pos, max_prims)
    _check_invariants(nodes, ordered, pos, max_prims)
    assert len(nodes) <= 2 * n - 1 and _leaf_sizes(nodes).max() <= max_prims
    assert _leaf_sizes(nodes).sum() == n


def test_lbvh_identical_and_degenerate_triangles(oracle):
    """Equal Morton codes (identical triangles, one centroid for all) are
    split by the triangle index in the key: leaves still hold at most
    max_prims triangles, in index order."""
    tri = np.array([[-1, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    pos = np.repeat(tri[None], 37, axis=0)
    nodes, ordered = oracle.lbvh_build(pos, 4)
    _check_invariants(nodes, ordered, pos, 4)
    assert ordered.tolist() == list(range(37)) and _leaf_sizes(nodes).max() <= 4


def test_lbvh_morton_order_of_a_grid(oracle):
    """Triangles centred on a 4x4x4 grid: the slot order is the Morton (Z)
    order of the cells, x the most significant of each bit triple."""
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(g))
    tri = np.array([[-0.1, -0.1, 0.0], [0.1, -0.1, 0.0], [0.0, 0.1, 0.0]], np.float32)
    pos = (g[perm][:, None, :].astype(np.float32) + tri[None])
    pos[:, :, 2] = g[perm][:, None, 2]  # flat in z: the centre is the grid point
    nodes, ordered = oracle.lbvh_build(pos, 1)

    def morton(c):
        code = 0
        for b in range(2):
            for a, sh in ((0, 2), (1, 1), (2, 0)):
                code |= ((int(c[a]) >> b) & 1) << (3 * b + sh)
        return code
    want = sorted(range(len(g)), key=lambda i: (morton(g[perm][i]), i))
    assert ordered.tolist() == want


def test_lbvh_traversal_equals_brute_force(oracle):
    """Closest and any hits through the linear BVH are those of a test of
    every triangle (same t bits; the same triangle wherever t is unique)."""
    sd = scenes.spheres_scene(segments=16, rings=8)
    nodes, ordered = oracle.lbvh_build(sd.positions)
    tris = oracle.ordered_tris(sd.positions, ordered)
    rays = scenes.random_rays(4000, 21)
    t, u, v, i = oracle.trace_closest(nodes, tris, rays)
    tb, ib = oracle.trace_closest_brute(tris, rays)
    hit = ib >= 0
    assert np.array_equal(i >= 0, hit)
    assert np.array_equal(t[hit].view(np.uint32), tb[hit].view(np.uint32))
    assert np.mean(i[hit] == ib[hit]) > 0.999
    assert np.array_equal(oracle.trace_any(nodes, tris, rays), oracle.trace_any_brute(tris, rays))


# ---- PLOC (locally-ordered clustering, the FF_BVH_PLOC GPU builder) ----
def sah_cost(nodes):
    """SAH cost of a tree in the builders' units, relative to the root's
    area: 0.125 per interior node and 1 per leaf triangle, each weighted by
    its box's half area (CRT/BVH.cpp:237-262 prices splits the same way)."""
    lo, hi = nodes["bmin"].astype(np.float64), nodes["bmax"].astype(np.float64)
    d = hi - lo
    a = d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]
    leaf = nodes["n_prims"] > 0
    return float((np.where(leaf, nodes["n_prims"] * a, 0.125 * a)).sum() / a[0])


@pytest.mark.parametrize("seed,n,max_prims", [(0, 1, 8), (1, 2, 8), (2, 5, 8), (3, 9, 8), (4, 1000, 8),
                                              (5, 20000, 4), (6, 3000, 1), (7, 5000, 16)])
def test_ploc_invariants(oracle, seed, n, max_prims):
    """Every triangle once, boxes exact unions of their children, pre-order
    layout, leaves of at most max_prims triangles."""
    pos = _soup(n, seed)
    nodes, ordered = oracle.ploc_build(pos, max_prims)
    _check_invariants(nodes, ordered, pos, max_prims)
    assert len(nodes) <= 2 * n - 1 and _leaf_sizes(nodes).max() <= max_prims
    assert _leaf_sizes(nodes).sum() == n


def test_ploc_traversal_equals_brute_force(oracle):
    sd = scenes.spheres_scene(segments=16, rings=8)
    nodes, ordered = oracle.ploc_build(sd.positions)
    tris = oracle.ordered_tris(sd.positions, ordered)
    rays = scenes.random_rays(4000, 22)
    t, u, v, i = oracle.trace_closest(nodes, tris, rays)
    tb, ib = oracle.trace_closest_brute(tris, rays)
    hit = ib >= 0
    assert np.array_equal(i >= 0, hit)
    assert np.array_equal(t[hit].view(np.uint32), tb[hit].view(np.uint32))
    assert np.mean(i[hit] == ib[hit]) > 0.999
    assert np.array_equal(oracle.trace_any(nodes, tris, rays), oracle.trace_any_brute(tris, rays))


@pytest.mark.parametrize("name", ["soup20k", "icosphere6", "spheres"])
def test_ploc_tree_quality(oracle, name):
    """The clustered tree's SAH cost is below the linear BVH's and within
    10% of the reference's top-down SAH build on the same triangles."""
    if name == "soup20k":
        pos = _soup(20000, 9)
    elif name == "icosphere6":
        pos = scenes.icosphere_scene(width=8, height=8, subdiv=6).positions
    else:
        pos = scenes.spheres_scene().positions
    c_ploc = sah_cost(oracle.ploc_build(pos)[0])
    c_lbvh = sah_cost(oracle.lbvh_build(pos)[0])
    c_sah = sah_cost(oracle.bvh_build(pos)[0])
    print(f"{name}: SAH cost ploc {c_ploc:.2f}, lbvh {c_lbvh:.2f}, sah {c_sah:.2f}")
    assert c_ploc < c_lbvh
    assert c_ploc <= 1.10 * c_sah


def _strip(xs, size=0.01):
    """one small triangle centred at each x on the x axis"""
    pos = np.zeros((len(xs), 3, 3), np.float32)
    for k, x in enumerate(xs):
        pos[k] = [[x - size, 0, 0], [x + size, 0, 0], [x, size, 0]]
    return pos


def test_ploc_merges_nearest_pairs_first(oracle):
    """Triangles at x = 0, 1, 10, 11, 30, 31: mutual nearest neighbours pair
    up first ((0,1), (10,11), (30,31)), then the two closest pairs; the
    lower child on the split axis comes first, and with one triangle per
    leaf every leaf holds one."""
    xs = [30, 0, 11, 1, 31, 10]
    nodes, ordered = oracle.ploc_build(_strip(xs), 1)

    def leaves(i):
        nd = nodes[i]
        if nd["n_prims"] > 0:
            return [xs[ordered[nd["offset"]]]]
        return leaves(i + 1) + leaves(nd["offset"])

    def shape(i):
        nd = nodes[i]
        if nd["n_prims"] > 0:
            return xs[ordered[nd["offset"]]]
        return (shape(i + 1), shape(nd["offset"]))
    assert shape(0) == (((0, 1), (10, 11)), (30, 31))
    assert leaves(0) == sorted(xs)
    assert all(nodes[nodes["n_prims"] == 0]["axis"] == 0)  # separated along x


def test_ploc_collapses_clusters_into_leaves(oracle):
    """Eight tight clusters of four triangles far apart: each cluster is one
    leaf (its leaf cost 4A is below any split of it), the clusters are
    interior nodes above them."""
    xs = [c * 10 + k * 0.001 for c in range(8) for k in range(4)]
    nodes, ordered = oracle.ploc_build(_strip(xs, 0.05), 8)
    sizes = sorted(_leaf_sizes(nodes).tolist())
    assert sizes == [4] * 8
    for nd in nodes[nodes["n_prims"] > 0]:
        members = {round(xs[ordered[nd["offset"] + k]] // 10) for k in range(nd["n_prims"])}
        assert len(members) == 1


def test_ploc_identical_triangles(oracle):
    tri = np.array([[-1, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    pos = np.repeat(tri[None], 37, axis=0)
    nodes, ordered = oracle.ploc_build(pos, 4)
    _check_invariants(nodes, ordered, pos, 4)
    assert _leaf_sizes(nodes).max() <= 4 and sorted(ordered.tolist()) == list(range(37))
