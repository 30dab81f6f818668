"""Inputs at the edges of the domain, on the GPU against the oracle:
non-finite and zero-direction rays, axis-aligned rays, rays in box planes,
zero-area triangles, non-finite vertices, an all-zero environment, an empty
scene and invalid requests.
Every case must finish normally with the oracle's result or a clean error."""
import numpy as np
import pytest

from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _trace(sc, rays):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    h = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    o = torch.empty(len(rays), dtype=torch.uint8, device="cuda")
    sc.trace_closest(r, h)
    sc.trace_any(r, o)
    torch.cuda.synchronize()
    return h.cpu().numpy(), o.cpu().numpy()


def _check(h, occ, ps, oracle, rays, label):
    t, u, v, i = oracle.trace_closest(ps.nodes, ps.tris_ordered(), rays)
    gid = h[:, 3].view(np.int32)
    assert np.array_equal(gid, i), f"{label}: ids differ at {np.nonzero(gid != i)[0][:8]}"
    hit = i >= 0
    assert np.array_equal(h[hit, 0].view(np.uint32), t[hit].view(np.uint32)), label
    assert np.array_equal(occ, oracle.trace_any(ps.nodes, ps.tris_ordered(), rays)), label


def test_non_finite_zero_and_axis_aligned_rays(gpu, oracle):
    sd = scenes.spheres_scene(segments=16, rings=8)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    nodes, ordered = sc.bvh()
    ps = oracle.PreparedScene(sd, nodes, ordered)
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    rows = [
        [nan, 0.5, 0.0, 0, 0.0, -1.0, 0.0, 0],   # NaN origin
        [0.0, 0.5, 0.0, 0, nan, -1.0, 0.0, 0],   # NaN direction
        [0.0, 0.5, 0.0, 0, 0.0, 0.0, 0.0, 0],    # zero direction
        [0.0, 0.5, 0.0, 0, inf, 0.0, 0.0, 0],    # infinite direction
        [inf, 0.5, 0.0, 0, -1.0, 0.0, 0.0, 0],   # infinite origin
        [0.0, 2.0, 0.0, 0, 0.0, -1.0, 0.0, 0],   # straight down (two zero components)
        [-3.0, 0.3, 0.0, 0, 1.0, 0.0, 0.0, 0],   # along +x through the spheres' plane
        [0.0, 0.3, -3.0, 0, 0.0, 0.0, 1.0, 0],   # along +z
        [0.0, 0.3, 0.0, 0, -0.0, -1.0, -0.0, 0],  # negative zeros
    ]
    rng = np.random.default_rng(4)
    grid = np.zeros((400, 8), np.float32)  # axis-aligned rays on a grid of origins
    grid[:, 0] = rng.uniform(-1.5, 1.5, 400)
    grid[:, 1] = rng.uniform(0.0, 1.0, 400)
    grid[:, 2] = rng.uniform(-1.5, 1.5, 400)
    axis = rng.integers(0, 3, 400)
    grid[np.arange(400), 4 + axis] = rng.choice([-1.0, 1.0], 400)
    rays = np.concatenate([np.array(rows, np.float32), grid])
    h, occ = _trace(sc, rays)
    _check(h, occ, ps, oracle, rays, "edge rays")
    # NaN origin or direction, zero direction, infinite origin: nothing is hit
    assert np.all(h[[0, 1, 2, 4], 3].view(np.int32) == -1) and not occ[[0, 1, 2, 4]].any()
    assert (h[5:, 3].view(np.int32) >= 0).sum() > 100
    sc.close()


def test_rays_in_box_planes(gpu, oracle):
    """Rays whose origin lies exactly in a BVH box's plane while their
    direction has a zero component along that axis: the slab distance is
    0 * inf = NaN there.  Hits, occlusion and the per-ray box and triangle
    counts equal the oracle's (the reference's comparisons on NaN)."""
    import torch
    sd = scenes.spheres_scene(segments=16, rings=8)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    nodes, ordered = sc.bvh()
    ps = oracle.PreparedScene(sd, nodes, ordered)
    rng = np.random.default_rng(9)
    pick = rng.integers(0, len(nodes), 600)
    rays = np.zeros((600, 8), np.float32)
    for k, ni in enumerate(pick):
        lo, hi = nodes[ni]["bmin"], nodes[ni]["bmax"]
        o = rng.uniform(lo, hi).astype(np.float32)
        a = rng.integers(0, 3)
        o[a] = lo[a] if rng.random() < 0.5 else hi[a]   # on the box's plane
        d = rng.normal(size=3).astype(np.float32)
        d[a] = 0.0 if rng.random() < 0.5 else -0.0       # parallel to it
        if rng.random() < 0.3:                            # and axis-aligned
            d[(a + 1) % 3] = 0.0
        rays[k, :3], rays[k, 4:7] = o, d / np.linalg.norm(d)
    h, occ = _trace(sc, rays)
    _check(h, occ, ps, oracle, rays, "box-plane rays")
    abi.set_traversal_variant(gpu, 0)  # one step per iteration: counts equal the oracle's walk
    r = torch.from_numpy(rays).cuda()
    for any_hit in (False, True):
        cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
        if any_hit:
            sc.trace_any(r, torch.empty(len(rays), dtype=torch.uint8, device="cuda"), counts_dev=cnt)
        else:
            sc.trace_closest(r, torch.empty((len(rays), 4), dtype=torch.float32, device="cuda"), counts_dev=cnt)
        torch.cuda.synchronize()
        want = oracle.count_traversal(ps.nodes, ps.tris_ordered(), rays, any_hit=any_hit, prune=True)
        assert tuple(cnt.cpu().tolist()) == tuple(want), (any_hit, cnt, want)
    sc.close()


def test_non_finite_vertices_are_rejected(gpu):
    sd = scenes.spheres_scene(segments=8, rings=4)
    for bad in (np.nan, np.inf, -np.inf):
        pos = sd.positions.copy()
        pos.reshape(-1)[13] = bad  # triangle 1, second vertex, y
        with pytest.raises(abi.FFError) as e:
            abi.Scene(gpu, pos, sd.normals, sd.uvs, sd.material_ids, sd.materials)
        assert e.value.status == 1 and "non-finite" in str(e.value)


def test_texture_slots_are_validated(gpu):
    """A normal texture needs the scene's tangent frames; each texture slot
    takes its channel count (3 base colour / normal, 1 roughness / metalness)."""
    sd = scenes.spheres_scene(segments=8, rings=4)
    rgb, one = np.full((4, 4, 3), 0.5, np.float32), np.full((4, 4, 1), 0.5, np.float32)
    for slot, tex, needs in ((13, [rgb], "tangents"), (12, [rgb], "channels"), (13, [one], "channels"),
                             (11, [rgb], "channels")):
        mats = sd.materials.copy()
        mats[:, slot] = 0
        tan = scenes.tangent_frames(sd.normals) if needs == "channels" and slot == 13 else None
        with pytest.raises(abi.FFError) as e:
            abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, mats, tex, tangents=tan)
        assert e.value.status == 1 and needs in str(e.value), (slot, str(e.value))


def test_zero_area_triangles(gpu, oracle):
    """Collinear and collapsed triangles mixed into a scene: the BVH still
    equals the oracle's and they are never hit (det = 0)."""
    sd = scenes.spheres_scene(segments=16, rings=8)
    rng = np.random.default_rng(7)
    n = 300
    p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    q = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    deg = np.stack([p, q, p + np.float32(0.5) * (q - p)], 1)   # collinear
    deg[: n // 3, 1] = deg[: n // 3, 0]                       # two equal vertices
    deg[: n // 6, 2] = deg[: n // 6, 0]                       # a point
    pos = np.concatenate([sd.positions.reshape(-1, 3, 3), deg]).reshape(-1, 9).astype(np.float32)
    nrm = np.concatenate([sd.normals.reshape(-1, 9), np.tile(np.float32([0, 1, 0]), (n, 3))])
    uvs = np.concatenate([sd.uvs.reshape(-1, 6), np.zeros((n, 6), np.float32)])
    mid = np.concatenate([sd.material_ids, np.zeros(n, np.int32)])
    sd2 = scenes.SceneData(**{**sd.__dict__, "positions": pos.reshape(-1, 3, 3), "normals": nrm.reshape(-1, 3, 3),
                              "uvs": uvs.reshape(-1, 3, 2), "material_ids": mid})
    sc = abi.Scene(gpu, pos, nrm, uvs, mid, sd.materials)
    nodes, ordered = sc.bvh()
    n_o, o_o = oracle.bvh_build(pos)
    assert nodes.tobytes() == n_o.tobytes() and np.array_equal(ordered, o_o)
    ps = oracle.PreparedScene(sd2, nodes, ordered)
    rays = scenes.random_rays(20000, 8)
    h, occ = _trace(sc, rays)
    _check(h, occ, ps, oracle, rays, "zero-area")
    n_real = len(sd.positions)
    hit_slots = h[:, 3].view(np.int32)
    hit_prims = ordered[hit_slots[hit_slots >= 0]]
    assert np.all(hit_prims < n_real)
    sc.close()


def test_all_zero_environment_is_rejected(gpu):
    """An all-zero environment has no sampling distribution (the CDF
    denominator is 0 and every row would be 0/0): creation fails with a
    status code instead of producing NaN CDFs, and the context keeps
    working."""
    sd = scenes.spheres_scene(width=32, height=16, spp=2, max_depth=2, segments=8, rings=4)
    with pytest.raises(abi.FFError) as e:
        abi.EnvMap(gpu, np.zeros_like(sd.env))
    assert e.value.status == 1 and "all-zero" in str(e.value)
    env = abi.EnvMap(gpu, sd.env)
    env.close()


def test_empty_scene_and_invalid_requests_are_rejected(gpu):
    """Errors come back as status codes with a message; the context keeps
    working afterwards."""
    import torch
    sd = scenes.spheres_scene(width=32, height=16, spp=2, max_depth=2, segments=8, rings=4)
    with pytest.raises(abi.FFError) as e:
        abi.Scene(gpu, np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), np.zeros((0, 6), np.float32),
                  np.zeros(0, np.int32), sd.materials)
    assert e.value.status == 1
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    for w, h, spp, depth, off in ((0, 16, 2, 2, 0), (32, -1, 2, 2, 0), (32, 16, 0, 2, 0), (32, 16, 2, 0, 0),
                                  (32, 16, 2, 251, 0), (32, 16, 2, 2, -1)):
        with pytest.raises(abi.FFError) as e:
            abi.render(gpu, sc, env, cam, w, h, spp, depth, 0, film, sample_offset=off)
        assert e.value.status == 1, (w, h, spp, depth, off)
    abi.render(gpu, sc, env, cam, sd.width, sd.height, 2, 2, 0, film)
    torch.cuda.synchronize()
    g = film.cpu().numpy()
    assert np.isfinite(g).all() and np.all(g[..., 3] == 2)
    env.close()
    sc.close()
