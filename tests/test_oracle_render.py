"""Pins of the oracle's full path logic by closed forms and invariants:
escaped camera rays return the environment radiance exactly, results are
independent of thread count, and film layering only regroups the sums."""
import numpy as np

from firefly_amd import scenes


def _far_scene(w=24, h=16):
    sd = scenes.spheres_scene(width=w, height=h)
    # one small triangle behind the camera: every camera ray escapes
    sd.positions = np.array([[[0, 0, 50], [1, 0, 50], [0, 1, 50]]], np.float32)
    sd.normals = np.tile(np.array([0, 0, 1], np.float32), (1, 3, 1))
    sd.uvs = np.zeros((1, 3, 2), np.float32)
    sd.material_ids = np.zeros(1, np.int32)
    sd.materials = sd.materials[:1]
    sd.env = np.full((8, 16, 3), 0.625, np.float32)
    return sd


def test_escaped_primary_rays_return_environment(oracle):
    sd = _far_scene()
    film, st = oracle.PreparedScene(sd).render(spp=3, max_depth=4, n_threads=2)
    assert np.all(film[..., 3] == 3)
    np.testing.assert_allclose(film[..., :3] / 3, 0.625, rtol=1e-6)
    assert st.rays_extension == 0 and st.rays_shadow == 0 and st.rays_primary == 24 * 16 * 3


def test_render_independent_of_threads(oracle):
    sd = scenes.spheres_scene(width=40, height=24)
    ps = oracle.PreparedScene(sd)
    a, _ = ps.render(spp=4, n_threads=1)
    b, _ = ps.render(spp=4, n_threads=7)
    assert np.array_equal(a, b)


def test_film_layers_only_regroup_sums(oracle):
    sd = scenes.spheres_scene(width=32, height=20)
    ps = oracle.PreparedScene(sd)
    a, _ = ps.render(spp=8, spp_batch=1, n_threads=4)
    b, _ = ps.render(spp=8, spp_batch=4, n_threads=4)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


def test_depth_one_is_direct_lighting_only(oracle):
    """max_depth 1: the path ends at its second logic step, so a pixel holds
    the escaped-ray radiance or the first vertex's NEE (never deeper light)."""
    sd = scenes.spheres_scene(width=32, height=20)
    ps = oracle.PreparedScene(sd)
    f1, s1 = ps.render(spp=4, max_depth=1, n_threads=4)
    f3, s3 = ps.render(spp=4, max_depth=3, n_threads=4)
    assert s1.rays_extension <= s1.rays_primary
    assert s3.rays_extension > s1.rays_extension
    assert np.isfinite(f1).all() and np.isfinite(f3).all()


def test_constant_textures_equal_their_factors(oracle):
    """A texture of one value in place of a material factor renders the
    image of that factor (bilinear lookups of a constant give the constant
    up to rounding): base colour, roughness and metalness each feed the
    field they stand for."""
    sd = scenes.spheres_scene(width=40, height=24)
    alb, rough, metal = (0.3, 0.6, 0.45), 0.35, 0.7
    sd.textures = [np.tile(np.array(alb, np.float32), (5, 7, 1)),
                   np.full((4, 3, 1), rough, np.float32),
                   np.full((6, 2, 1), metal, np.float32)]
    fac = sd.materials.copy()
    fac[1:, 1:4] = alb
    fac[1:, 4] = rough
    fac[1:, 5] = metal
    tex = sd.materials.copy()
    tex[1:, 1:4] = 0.9
    tex[1:, 4] = 1.0
    tex[1:, 5] = 0.0
    tex[1:, 10:13] = (0, 1, 2)
    films = []
    for m in (fac, tex):
        sd.materials = m
        films.append(oracle.PreparedScene(sd).render(spp=4, max_depth=3, n_threads=4)[0])
    np.testing.assert_allclose(films[1], films[0], rtol=1e-4, atol=1e-6)
    # without the metalness texture the image differs: the lookup is not inert
    tex[1:, 12] = -1
    sd.materials = tex
    other = oracle.PreparedScene(sd).render(spp=4, max_depth=3, n_threads=4)[0]
    assert np.abs(other - films[0]).max() > 1e-2


def test_normal_map_equals_the_mapped_shading_normal(oracle):
    """On faceted triangles (one normal and one tangent frame per face) a
    constant normal texture m gives the image of shading normals
    normalize(T m.x + B m.y + N m.z) without a map, up to the light shadow
    ray's offset, which keeps the interpolated normal; with T and B swapped
    the image differs, and a texel of (0.5, 0.5, 1) leaves the normal as is."""
    sd = scenes.spheres_scene(width=40, height=24)
    e1 = sd.positions[:, 1] - sd.positions[:, 0]
    e2 = sd.positions[:, 2] - sd.positions[:, 0]
    fn = np.cross(e1, e2).astype(np.float64)
    fn /= np.linalg.norm(fn, axis=-1, keepdims=True)
    sd.normals = np.repeat(fn[:, None], 3, axis=1).astype(np.float32)
    frames = scenes.tangent_frames(sd.normals)
    sd.tangents = frames
    m = np.array([0.35, -0.25, 0.9])
    m /= np.linalg.norm(m)
    tex = ((m + 1) / 2).astype(np.float32)
    sd.textures = [np.tile(tex, (4, 4, 1)), np.tile(np.array([0.5, 0.5, 1.0], np.float32), (3, 5, 1))]
    mats = sd.materials.copy()
    mats[:, 13] = 0
    sd.materials = mats

    def render(s):
        return oracle.PreparedScene(s).render(spp=4, max_depth=3, n_threads=4)[0][..., :3]
    mapped = render(sd)

    def baked(t, b):
        nb = t * m[0] + b * m[1] + fn[:, None] * m[2]
        nb /= np.linalg.norm(nb, axis=-1, keepdims=True)
        s = scenes.SceneData(**{**sd.__dict__})
        s.normals = nb.astype(np.float32)
        s.materials = sd.materials.copy()
        s.materials[:, 13] = -1
        s.tangents = None
        return render(s)
    t, b = frames[..., :3].astype(np.float64), frames[..., 3:].astype(np.float64)
    want = baked(t, b)
    rmse = lambda a, c: float(np.sqrt(np.mean((a - c) ** 2)) / np.sqrt(np.mean(c ** 2)))
    assert rmse(mapped, want) < 3e-3
    assert rmse(baked(b, t), want) > 3e-2
    sd.materials[:, 13] = 1
    flat = render(sd)
    sd.materials[:, 13] = -1
    np.testing.assert_allclose(flat, render(sd), rtol=1e-4, atol=1e-6)


def test_gated_brdf_visibility_rays_change_no_pixel(oracle):
    """The BRDF-sample visibility rays are traced only where the light ray is
    unoccluded (its visibility gates the whole NEE sum).  Tracing them
    everywhere, as the reference's BRDF_L does (test hook 8), gives the same
    film bit for bit with more shadow rays."""
    sd = scenes.spheres_scene(width=48, height=32)
    ps = oracle.PreparedScene(sd)
    a, sa = ps.render(spp=4, max_depth=4, n_threads=4)
    oracle.lib().or_set_mutation(8)
    try:
        b, sb = ps.render(spp=4, max_depth=4, n_threads=4)
    finally:
        oracle.lib().or_set_mutation(0)
    assert np.array_equal(a, b)
    assert sb.rays_shadow > sa.rays_shadow
    assert sb.rays_shadow == sa.rays_shadow + sa.rays_shadow_gated and sb.rays_shadow_gated == 0
    assert (sa.rays_primary, sa.rays_extension, sa.paths) == (sb.rays_primary, sb.rays_extension, sb.paths)


def test_environment_tables_hook(oracle):
    """Renders inside oracle.env_tables use the given CDF tables: the oracle's
    own tables give its film bit for bit, conditional CDFs scaled by 1 + 1e-3
    move light samples to other texels, and leaving the block restores the
    oracle's own build."""
    sd = scenes.spheres_scene(width=40, height=30, spp=2, max_depth=3)
    ps = oracle.PreparedScene(sd)
    a, sa = ps.render(n_threads=4)
    my, mp, conds, pdf, _ = oracle.env_build(sd.env)
    with oracle.env_tables(my, mp, conds, pdf):
        b, sb = ps.render(n_threads=4)
    assert np.array_equal(a, b) and sa.rays_shadow == sb.rays_shadow
    with oracle.env_tables(my, mp, conds * (1 + 1e-3), pdf):
        c, _ = ps.render(n_threads=4)
    assert np.mean(np.all(a == c, axis=-1)) < 0.95
    with oracle.env_tables(my[:-1], mp[:-1], conds[:-1], pdf[:-1]):  # another map size: not used
        d, _ = ps.render(n_threads=4)
    assert np.array_equal(a, d)
    e, _ = ps.render(n_threads=4)
    assert np.array_equal(a, e)
