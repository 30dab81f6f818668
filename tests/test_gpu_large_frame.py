"""Parity at the single-large-frame size: the 1.31 M-triangle displaced
icosphere (subdivision 8) with ground, 1920x1080, 2048x1024 sun environment.

The product's host BVH equals the oracle's on the full mesh, hits are
bit-exact on seeded camera and surface-bounce rays, and a full-resolution
one-sample frame matches the oracle's render of the same samples."""
import os

import numpy as np
import pytest

from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def big(gpu, oracle):
    sd = scenes.icosphere_scene(1920, 1080, 1, 8, subdiv=8)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    nodes, ordered = sc.bvh()
    ps = oracle.PreparedScene(sd, nodes, ordered)
    yield sd, sc, ps
    sc.close()


def _closest(sc, rays):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    h = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    sc.trace_closest(r, h)
    torch.cuda.synchronize()
    return h.cpu().numpy()


def _assert_hits_equal(h, ref, label):
    t, u, v, i = ref
    gid = h[:, 3].view(np.int32)
    assert np.array_equal(gid, i), f"{label}: {np.count_nonzero(gid != i)} triangle id mismatches"
    hit = i >= 0
    for name, g, c in (("t", h[:, 0], t), ("u", h[:, 1], u), ("v", h[:, 2], v)):
        assert np.array_equal(g[hit].view(np.uint32), c[hit].view(np.uint32)), f"{label}: {name} bits differ"


def test_full_mesh_bvh_equals_oracle(big, oracle):
    sd, sc, ps = big
    n_o, o_o = oracle.bvh_build(sd.positions)
    nodes, ordered = sc.bvh()
    assert len(nodes) > 2_000_000
    assert nodes.tobytes() == n_o.tobytes()
    assert np.array_equal(ordered, o_o)


def test_full_mesh_ploc_tree_equals_oracle(big, gpu, oracle):
    """The clustering build of the 1.31 M-triangle mesh on the GPU equals the
    oracle's node for node."""
    sd, _, _ = big
    nodes, ordered, depth, ms = abi.bvh_build_gpu(gpu, sd.positions, 8, builder=abi.FF_BVH_PLOC)
    n_o, o_o = oracle.ploc_build(sd.positions)
    assert len(nodes) == len(n_o) and nodes.tobytes() == n_o.tobytes()
    assert np.array_equal(ordered, o_o)
    print(f"PLOC build of {sd.n_tris} triangles: {ms:.2f} ms of kernels, depth {depth}")


def test_camera_and_bounce_hits_bit_exact(big, gpu, oracle):
    import torch
    sd, sc, ps = big
    W, H = sd.width, sd.height
    rays = torch.empty((W * H, 8), dtype=torch.float32, device="cuda")
    abi.generate_rays(gpu, abi.make_camera(sd.inv_vp), W, H, sd.seed, 0, rays)
    torch.cuda.synchronize()
    pick = np.random.default_rng(5).choice(W * H, 65536, replace=False)
    r = rays.cpu().numpy()[pick]
    h = _closest(sc, r)
    ref = oracle.trace_closest(ps.nodes, ps.tris_ordered(), r)
    assert np.count_nonzero(ref[3] >= 0) > 30000
    _assert_hits_equal(h, ref, "camera")
    # bounce rays leaving the hit points (the grazing, surface-bound rays that
    # dominate the traversal cost), upper hemisphere
    ok = ref[3] >= 0
    o = r[ok, 0:3] + r[ok, 4:7] * ref[0][ok, None]
    d = np.random.default_rng(6).normal(size=o.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 1] = np.abs(d[:, 1])
    b = np.zeros((len(o), 8), np.float32)
    b[:, 0:3] = o + d * 1e-3
    b[:, 4:7] = d
    _assert_hits_equal(_closest(sc, b), oracle.trace_closest(ps.nodes, ps.tris_ordered(), b), "bounce")
    occ = torch.empty(len(b), dtype=torch.uint8, device="cuda")
    sc.trace_any(torch.from_numpy(b).cuda(), occ)
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy(), oracle.trace_any(ps.nodes, ps.tris_ordered(), b))


def _assert_exact_with_tables(ps, tables, g, st, **kw):
    """The oracle's render (of kw's rows) with the GPU's environment tables
    (FP64 sums in tree order; the oracle's own are the reference's serial
    sums, so a light-sample uniform within a few ulps of a CDF entry can pick
    the neighbouring texel) equals the GPU film in every bit."""
    import sys
    from oracle import oracle as O
    with O.env_tables(*tables):
        c, cst = ps.render(**kw)
    r0, r1 = kw.get("rows", (0, g.shape[0]))
    same = float(np.mean(np.all(g[r0:r1] == c[r0:r1], axis=-1)))
    print(f"rows {r0}-{r1} with the GPU's CDF tables: {same:.6f} of pixels bit-identical", file=sys.stderr,
          flush=True)
    assert np.array_equal(g[r0:r1], c[r0:r1])
    if st is not None:
        for k in ("rays_primary", "rays_extension", "rays_shadow", "rays_shadow_gated", "paths"):
            assert getattr(st, k) == getattr(cst, k), k


def test_full_frame_one_sample_matches_oracle(big, gpu):
    """All 2.07 M paths of one sample per pixel at depth 8, every material
    queue and the importance-sampled sun: relative RMSE <= 1% of the oracle,
    and nearly every pixel bit-identical (same random numbers, same hits)."""
    import torch
    sd, sc, ps = big
    env = abi.EnvMap(gpu, sd.env)
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    abi.render(gpu, sc, env, abi.make_camera(sd.inv_vp), sd.width, sd.height, 1, 8, sd.seed, film)
    g = film.cpu().numpy()
    st = abi.last_render_stats(gpu)
    tables = env.read()[:4]
    env.close()
    c, _ = ps.render(spp=1, max_depth=8, n_threads=THREADS)
    _assert_exact_with_tables(ps, tables, g, st, spp=1, max_depth=8, n_threads=THREADS)
    assert np.isfinite(g).all() and np.all(g[..., 3] == 1)
    err = float(np.sqrt(np.mean((g[..., :3] - c[..., :3]) ** 2)) / np.sqrt(np.mean(c[..., :3] ** 2)))
    exact = float(np.mean(np.all(g[..., :3] == c[..., :3], axis=-1)))
    print(f"1080p icosphere8 1 spp depth 8: rel RMSE {err:.3e}; bit-identical pixels {exact:.5f}")
    assert err < 0.01
    assert exact > 0.999


def test_benchmark_frame_matches_oracle(big, gpu):
    """The benchmarked frame at its own launch shape: 64 spp with all 64
    samples of every pixel in flight (133 M paths), traversal variant 23
    (leaf batching) for bounce iterations 1-3 and variant 0 from iteration 4
    on, against the oracle's 64-spp render with the same layer grouping:
    relative RMSE <= 1% and nearly every pixel bit-identical."""
    import sys
    import torch
    sd, sc, ps = big
    env = abi.EnvMap(gpu, sd.env)
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    print("1080p frame, 64 spp, 64 in flight, variant 23 + tail 0 from iteration 4", file=sys.stderr, flush=True)
    abi.set_traversal_variant(gpu, 23)
    abi.set_traversal_tail(gpu, 0, 4)
    try:
        abi.render(gpu, sc, env, abi.make_camera(sd.inv_vp), sd.width, sd.height, 64, 8, sd.seed, film,
                   spp_batch=64)
        torch.cuda.synchronize()
    finally:
        abi.set_traversal_tail(gpu, -1, 1)
        abi.set_traversal_variant(gpu, 0)
    g = film.cpu().numpy()
    del film
    tables = env.read()[:4]
    env.close()
    c, _ = ps.render(spp=64, max_depth=8, spp_batch=64, n_threads=THREADS)
    _assert_exact_with_tables(ps, tables, g, None, spp=64, max_depth=8, spp_batch=64, n_threads=THREADS,
                              rows=(500, 540))
    assert np.isfinite(g).all() and np.all(g[..., 3] == 64)
    err = float(np.sqrt(np.mean((g[..., :3] - c[..., :3]) ** 2)) / np.sqrt(np.mean(c[..., :3] ** 2)))
    exact = float(np.mean(np.all(g[..., :3] == c[..., :3], axis=-1)))
    print(f"1080p icosphere8 64 spp depth 8 (benchmark shape): rel RMSE {err:.3e}; bit-identical pixels "
          f"{exact:.5f}", file=sys.stderr, flush=True)
    assert err < 0.01
    assert exact > 0.999
