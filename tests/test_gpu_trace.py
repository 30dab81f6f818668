"""GPU parity of the extend (closest hit) and shadow (any hit) stages and of
camera-ray generation against the CPU oracle: bit-exact on identical rays."""
import numpy as np
import pytest

from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu


KINDS = {"binary": (2, 0), "wide": (4, 0), "wide_ww": (4, 1), "binary_spec": (2, 2), "binary_batch": (2, 3),
         "binary_batch1": (2, 5), "binary_batch1_4": (2, 6), "binary_batch1_16": (2, 7)}


def _variant_of(kind):
    """First traversal variant of a kind: (BVH record width, loop shape)."""
    for v, _, _, _, ww, width in abi.traversal_variants():
        if (width, ww) == KINDS[kind]:
            return v
    raise AssertionError(f"no {kind} traversal variant")


@pytest.fixture(params=list(KINDS))
def trav(request, gpu):
    """Runs the test with the context's traversal variant set to each BVH
    record format and loop shape; yields (record width, loop shape)."""
    abi.set_traversal_variant(gpu, _variant_of(request.param))
    yield KINDS[request.param]
    abi.set_traversal_variant(gpu, 0)


@pytest.fixture(scope="module")
def cfg0(gpu, oracle):
    sd = scenes.spheres_scene()
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    nodes, ordered = sc.bvh()
    ps = oracle.PreparedScene(sd, nodes, ordered)
    yield sd, sc, ps
    sc.close()


def _closest(sc, rays, two_sided=False):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    h = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    sc.trace_closest(r, h, two_sided)
    torch.cuda.synchronize()
    return h.cpu().numpy()


def _any(sc, rays):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    o = torch.empty(len(rays), dtype=torch.uint8, device="cuda")
    sc.trace_any(r, o)
    torch.cuda.synchronize()
    return o.cpu().numpy()


def _assert_hits_equal(h, ref, label):
    t, u, v, i = ref
    gid = h[:, 3].view(np.int32)
    bad = np.nonzero(gid != i)[0]
    assert bad.size == 0, f"{label}: {bad.size} triangle id mismatches, first {bad[:5]}"
    hit = i >= 0
    for name, g, c in (("t", h[:, 0], t), ("u", h[:, 1], u), ("v", h[:, 2], v)):
        gb, cb = g[hit].view(np.uint32), c[hit].view(np.uint32)
        nb = np.count_nonzero(gb != cb)
        assert nb == 0, f"{label}: {nb} {name} values differ in bits"


def test_scene_bvh_equals_oracle_bvh(cfg0, oracle):
    sd, sc, _ = cfg0
    n_o, o_o = oracle.bvh_build(sd.positions)
    nodes, ordered = sc.bvh()
    assert nodes.tobytes() == n_o.tobytes()
    assert np.array_equal(ordered, o_o)


def test_camera_rays_bit_exact(cfg0, gpu, oracle):
    import torch
    sd, _, ps = cfg0
    W, H = sd.width, sd.height
    rays = torch.empty((W * H, 8), dtype=torch.float32, device="cuda")
    cam = abi.make_camera(sd.inv_vp)
    for sample in (0, 5):
        abi.generate_rays(gpu, cam, W, H, sd.seed, sample, rays)
        torch.cuda.synchronize()
        g = rays.cpu().numpy()
        rng = np.random.default_rng(sample)
        for pix in np.concatenate([[0, W - 1, W * H - 1], rng.integers(0, W * H, 400)]):
            ref = ps.gen_ray(int(pix % W), int(pix // W), sample)
            got = np.concatenate([g[pix, 0:3], g[pix, 4:7]])
            assert got.view(np.uint32).tolist() == ref.view(np.uint32).tolist(), f"pixel {pix}"


def test_thin_lens_camera_rays_bit_exact(gpu, oracle):
    """Camera rays through a lens of radius 0.1 focused at 3.0 equal the
    oracle's bit for bit (origin offset and refocused direction)."""
    import dataclasses
    import torch
    sd = dataclasses.replace(scenes.spheres_scene(width=96, height=64), lens_radius=0.1, focal=3.0)
    ps = oracle.PreparedScene(sd)
    W, H = sd.width, sd.height
    rays = torch.empty((W * H, 8), dtype=torch.float32, device="cuda")
    cam = abi.make_camera(sd.inv_vp, sd.lens_radius, sd.focal)
    for sample in (0, 3):
        abi.generate_rays(gpu, cam, W, H, sd.seed, sample, rays)
        torch.cuda.synchronize()
        g = rays.cpu().numpy()
        for pix in range(0, W * H, 7):
            ref = ps.gen_ray(int(pix % W), int(pix // W), sample)
            got = np.concatenate([g[pix, 0:3], g[pix, 4:7]])
            assert got.view(np.uint32).tolist() == ref.view(np.uint32).tolist(), f"pixel {pix}"


def test_primary_hits_bit_exact(cfg0, gpu, oracle):
    import torch
    sd, sc, ps = cfg0
    W, H = sd.width, sd.height
    rays = torch.empty((W * H, 8), dtype=torch.float32, device="cuda")
    abi.generate_rays(gpu, abi.make_camera(sd.inv_vp), W, H, sd.seed, 0, rays)
    torch.cuda.synchronize()
    r = rays.cpu().numpy()
    h = _closest(sc, r)
    ref = oracle.trace_closest(ps.nodes, ps.tris_ordered(), r)
    assert np.count_nonzero(ref[3] >= 0) > W * H // 4
    _assert_hits_equal(h, ref, "primary")


@pytest.mark.parametrize("seed", [1, 2])
def test_incoherent_hits_bit_exact(cfg0, oracle, seed, trav):
    sd, sc, ps = cfg0
    r = scenes.random_rays(65536, seed)
    h = _closest(sc, r)
    _assert_hits_equal(h, oracle.trace_closest(ps.nodes, ps.tris_ordered(), r), f"incoherent{seed}")


def test_two_sided_hits_bit_exact(cfg0, oracle, trav):
    sd, sc, ps = cfg0
    r = scenes.random_rays(32768, 3)
    h = _closest(sc, r, two_sided=True)
    _assert_hits_equal(h, oracle.trace_closest(ps.nodes, ps.tris_ordered(), r, two_sided=True), "two-sided")


def test_any_hit_bit_exact(cfg0, oracle, trav):
    sd, sc, ps = cfg0
    r = scenes.random_rays(65536, 4)
    o = _any(sc, r)
    ref = oracle.trace_any(ps.nodes, ps.tris_ordered(), r)
    assert 0 < ref.sum() < len(r)
    assert np.array_equal(o, ref)


def test_rays_from_surfaces_bit_exact(cfg0, oracle, trav):
    """Rays starting on surfaces (extension-ray origins) exercise the
    behind-origin pruning next to the origin triangle."""
    sd, sc, ps = cfg0
    r = scenes.random_rays(32768, 5)
    h0 = _closest(sc, r)
    hit = h0[:, 3].view(np.int32) >= 0
    o = r[hit, 0:3] + r[hit, 4:7] * h0[hit, 0:1]
    rng = np.random.default_rng(6)
    d = rng.normal(size=o.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r2 = np.zeros((o.shape[0], 8), np.float32)
    r2[:, 0:3] = o
    r2[:, 4:7] = d
    _assert_hits_equal(_closest(sc, r2), oracle.trace_closest(ps.nodes, ps.tris_ordered(), r2), "surface")
    assert np.array_equal(_any(sc, r2), oracle.trace_any(ps.nodes, ps.tris_ordered(), r2))


def test_empty_and_single_ray(cfg0, oracle):
    sd, sc, ps = cfg0
    sc.trace_closest(np.zeros((0, 8), np.float32), np.zeros((0, 4), np.float32))  # n = 0 is a no-op
    r = scenes.random_rays(1, 9)
    _assert_hits_equal(_closest(sc, r), oracle.trace_closest(ps.nodes, ps.tris_ordered(), r), "single")


def test_traversal_counters_match_oracle(cfg0, oracle, trav):
    """Boxes/triangles tested per ray (the roofline's N_node, N_tri) counted
    on the GPU equal the oracle's count of the same visit scheme, for each
    BVH record width.  The speculative and leaf-batching loops test leaves
    later than they reach them, so their pruning is weaker: their counts lie
    between the pruned and the unpruned walk's (any-hit: the same triangles,
    as leaves are still tested in order up to the first occluder)."""
    import torch
    width, ww = trav
    sd, sc, ps = cfg0
    r = scenes.random_rays(20000, 12)
    rd = torch.from_numpy(r).cuda()
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    h = torch.empty((len(r), 4), dtype=torch.float32, device="cuda")
    sc.trace_closest(rd, h, counts_dev=cnt)
    torch.cuda.synchronize()
    got = tuple(cnt.cpu().tolist())
    ref = oracle.count_traversal(ps.nodes, ps.tris_ordered(), r, prune=True, width=width)
    if ww >= 2:
        hi = oracle.count_traversal(ps.nodes, ps.tris_ordered(), r, prune=False, width=width)
        assert ref[0] <= got[0] <= hi[0] and ref[1] <= got[1] <= hi[1], (got, ref, hi)
    else:
        assert got == ref
    cnt.zero_()
    o = torch.empty(len(r), dtype=torch.uint8, device="cuda")
    sc.trace_any(rd, o, counts_dev=cnt)
    torch.cuda.synchronize()
    got = tuple(cnt.cpu().tolist())
    ref = oracle.count_traversal(ps.nodes, ps.tris_ordered(), r, any_hit=True, prune=True, width=width)
    if ww >= 2:
        assert got[0] >= ref[0] and got[1] == ref[1], (got, ref)
    else:
        assert got == ref


@pytest.fixture(scope="module")
def deep(gpu, oracle):
    """A deeper tree (icosphere, 81920 triangles): traversal stacks spill from
    LDS into the overflow area for the narrow LDS variants."""
    sd = scenes.icosphere_scene(width=8, height=8, subdiv=6)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    nodes, ordered = sc.bvh()
    ps = oracle.PreparedScene(sd, nodes, ordered)
    yield sd, sc, ps
    sc.close()


def test_deep_tree_hits_bit_exact_every_variant(deep, gpu, oracle):
    """Every traversal variant (all block sizes, LDS stack depths, loop
    shapes and record widths) returns the oracle's hits on a deep tree."""
    sd, sc, ps = deep
    r = scenes.random_rays(32768, 21)
    ref = oracle.trace_closest(ps.nodes, ps.tris_ordered(), r)
    occ = oracle.trace_any(ps.nodes, ps.tris_ordered(), r)
    assert np.count_nonzero(ref[3] >= 0) > 1000
    try:
        for v, block, lds, refill, ww, width in abi.traversal_variants():
            abi.set_traversal_variant(gpu, v)
            label = f"variant {v} (block {block}, lds {lds}, refill {refill}, ww {ww}, width {width})"
            print(label, flush=True)
            _assert_hits_equal(_closest(sc, r), ref, label)
            assert np.array_equal(_any(sc, r), occ), label
    finally:
        abi.set_traversal_variant(gpu, 0)
