"""Batched requests: several scenes/cameras of uneven resolution, spp and depth
rendered in one wavefront (optionally replayed from a hipGraph) give each
request exactly the film a plain per-request render gives it."""
import numpy as np
import pytest

from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _requests():
    specs = [  # (seed, width, height, spp, depth, material)
        (0, 64, 48, 3, 2, abi.FF_MAT_MIX),
        (1, 40, 40, 5, 5, abi.FF_MAT_DIFFUSE),
        (2, 96, 32, 1, 3, abi.FF_MAT_CONDUCTOR),
        (3, 24, 36, 4, 7, abi.FF_MAT_DIELECTRIC),
    ]
    out = []
    for seed, w, h, spp, depth, mt in specs:
        sd = scenes.spheres_scene(width=w, height=h, spp=spp, max_depth=depth, seed=seed, segments=16, rings=8)
        mats = sd.materials.copy()
        mats[1:, 0] = mt
        if mt == abi.FF_MAT_CONDUCTOR:
            mats[1:, 5] = 1.0
            mats[1:, 4] = 0.3
        sd.materials = mats
        out.append(sd)
    return out


@pytest.mark.parametrize("flags,spp_batch", [(0, 1), (abi.FF_RENDER_USE_GRAPH, 1), (0, [2, 4, 1, 3]),
                                             (abi.FF_RENDER_USE_GRAPH, 3)])
def test_batch_equals_individual_renders(gpu, flags, spp_batch):
    """Each request's film is bit-identical to its plain render with the same
    samples in flight (one value for all items, or one per item)."""
    import torch
    kb = spp_batch if isinstance(spp_batch, list) else [spp_batch] * 4
    reqs = _requests()
    objs = []
    for sd in reqs:
        sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
        env = abi.EnvMap(gpu, sd.env)
        objs.append((sd, sc, env, abi.make_camera(sd.inv_vp)))
    singles = []
    for (sd, sc, env, cam), k in zip(objs, kb):
        f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        abi.render(gpu, sc, env, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f, spp_batch=k)
        singles.append(f)
    torch.cuda.synchronize()
    films = [torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda") for sd, *_ in objs]
    abi.render_batch(gpu, [(sc, env, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f)
                           for (sd, sc, env, cam), f in zip(objs, films)], flags=flags, spp_batch=spp_batch)
    torch.cuda.synchronize()
    for (sd, *_), f, s in zip(objs, films, singles):
        a, b = f.cpu().numpy(), s.cpu().numpy()
        assert np.all(a[..., 3] == sd.spp)
        assert np.array_equal(a, b), f"request seed {sd.seed}: batched film differs"
    st = abi.last_render_stats(gpu)
    assert st.paths == sum(sd.width * sd.height * sd.spp for sd, *_ in objs)
    for _, sc, env, _ in objs:
        sc.close()
        env.close()


def test_graph_render_equals_plain_render(gpu):
    """FF_RENDER_USE_GRAPH on a single request replays the captured batch."""
    import torch
    sd = scenes.spheres_scene(width=64, height=40, spp=6, max_depth=4)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    a = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    abi.render(gpu, sc, env, cam, sd.width, sd.height, 6, 4, 0, a, spp_batch=2)
    abi.render(gpu, sc, env, cam, sd.width, sd.height, 6, 4, 0, b, spp_batch=2, flags=abi.FF_RENDER_USE_GRAPH)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_overlapped_and_serial_schedules_agree(gpu):
    """The shadow traversal of each iteration runs on a second stream beside
    the extension traversal; with per-launch timing the stages run one after
    another.  Eager, graph-replayed and serial renders give the same film and
    the same ray and traversal counts, over three batches of samples."""
    import torch
    sd = scenes.spheres_scene(width=96, height=64, spp=12, max_depth=5)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    keys = ("rays_primary", "rays_extension", "rays_shadow", "rays_shadow_gated", "paths", "ext_nodes",
            "ext_tris", "shadow_nodes", "shadow_tris")
    out = []
    for flags in (0, abi.FF_RENDER_USE_GRAPH, abi.FF_RENDER_TIMING):
        f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        abi.render(gpu, sc, env, cam, sd.width, sd.height, 12, 5, 3, f, spp_batch=4, flags=flags)
        st = abi.last_render_stats(gpu)
        out.append((f.cpu(), {k: getattr(st, k) for k in keys}))
    sc.close()
    env.close()
    for f, st in out[1:]:
        assert torch.equal(f, out[0][0])
        assert st == out[0][1]
    assert out[0][1]["rays_shadow"] > 0 and out[0][1]["rays_shadow_gated"] > 0


def test_traversal_variants_agree(gpu):
    """Every traversal variant (workgroup, LDS stack, refill, loop shape, BVH
    record width) gives the same image."""
    import torch
    sd = scenes.spheres_scene(width=48, height=32, spp=2, max_depth=3)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    ref = None
    for v, block, lds, refill, ww, width in abi.traversal_variants():
        abi.set_traversal_variant(gpu, v)
        f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        abi.render(gpu, sc, env, cam, sd.width, sd.height, 2, 3, 0, f)
        torch.cuda.synchronize()
        if ref is None:
            ref = f
        assert torch.equal(f, ref), (f"variant {v} (block {block}, lds {lds}, refill {refill}, while-while {ww}, "
                                    f"width {width})")
    abi.set_traversal_variant(gpu, 0)


def test_tail_variant_gives_the_same_image(gpu):
    """A second traversal variant for the late bounces (ff_set_traversal_tail)
    changes launch shapes only: the film is the one-variant film."""
    import torch
    sd = scenes.spheres_scene(width=48, height=32, spp=4, max_depth=6)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    films = []
    for tail, start in ((-1, 1), (1, 3), (0, 2), (25, 5)):
        abi.set_traversal_variant(gpu, 23)
        abi.set_traversal_tail(gpu, tail, start)
        f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        abi.render(gpu, sc, env, cam, sd.width, sd.height, 4, 6, 0, f, spp_batch=4, flags=abi.FF_RENDER_USE_GRAPH)
        torch.cuda.synchronize()
        films.append(f)
    for f in films[1:]:
        assert torch.equal(f, films[0])
    with pytest.raises(abi.FFError):
        abi.set_traversal_tail(gpu, 999, 3)
    with pytest.raises(abi.FFError):
        abi.set_traversal_tail(gpu, 1, 0)
    abi.set_traversal_tail(gpu, -1, 1)
    abi.set_traversal_variant(gpu, 0)
    sc.close()
    env.close()


def test_concurrent_scene_creation_deep_trees(gpu):
    """Scenes built from several host threads on one context, with BVHs deeper
    than the default LDS traversal stack (so the overflow area is used), render
    the same films batched (graph) as one at a time."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    env = scenes.sun_env(256, 128, 0)
    specs = [{"index": i, "width": 48 + 8 * i, "height": 40, "spp": 1 + i % 3, "max_depth": 3 + i % 4,
              "tris": t, "seed": 70 + i} for i, t in enumerate((12000, 60000, 250000, 30000))]
    sds = [scenes.batch_request_scene(s, env) for s in specs]

    def make(sd):
        return abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)

    with ThreadPoolExecutor(4) as ex:
        scs = list(ex.map(make, sds))
    depths = [abi.bvh_build_host(sd.positions)[2] for sd in sds]
    assert max(depths) > 16  # deeper than the default 16-entry LDS stack
    e = abi.EnvMap(gpu, env)
    cams = [abi.make_camera(sd.inv_vp) for sd in sds]
    singles, films = [], []
    for sd, sc, cam in zip(sds, scs, cams):
        f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        abi.render(gpu, sc, e, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f)
        singles.append(f)
        films.append(torch.zeros_like(f))
    abi.render_batch(gpu, [(sc, e, cam, sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f)
                           for sd, sc, cam, f in zip(sds, scs, cams, films)], flags=abi.FF_RENDER_USE_GRAPH)
    torch.cuda.synchronize()
    for sd, a, b in zip(sds, films, singles):
        assert torch.equal(a, b), sd.name
    for sc in scs:
        sc.close()
    e.close()


def test_uneven_request_renders_finite(gpu):
    """Batched request 19 (727^2, 2 spp, depth 4) once rendered a NaN pixel:
    a GGX sample normalised to |y| = 1 + 1 ulp reached asin in the spherical
    map.  The lookup clamps the argument now; the film is finite."""
    import torch
    spec = scenes.batch_request_specs(64, seed=0)[19]
    sd = scenes.batch_request_scene(spec, scenes.sun_env(2048, 1024, 3))
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    abi.render(gpu, sc, env, abi.make_camera(sd.inv_vp), sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(f).all())
    sc.close()
    env.close()
